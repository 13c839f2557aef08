// Internal data model of the MI355X FFT-SVD BEM library.
//
// Host side: panels, octrees and region operators are built on the CPU and
// flattened into structure-of-arrays blocks that are uploaded once.  Device
// side: every per-panel integral, every compressed block, the matvec and the
// Krylov vectors live in HBM; the host only drives launches.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace fsvd {

// YUKAWA_DKAPPA: d/dkappa of the Yukawa panel integrals (gradients only;
// no operator is built with it)
enum Kernel { LAPLACE = 0, YUKAWA = 1, YUKAWA_DKAPPA = 2 };
enum Precision { PREC_F64 = 0, PREC_F32 = 1 };

#define FSVD_CHECK(x)                                                                                  \
   do {                                                                                                \
      hipError_t e_ = (x);                                                                             \
      if (e_ != hipSuccess) {                                                                          \
         fprintf(stderr, "fftsvd: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
         throw std::runtime_error(hipGetErrorString(e_));                                              \
      }                                                                                                \
   } while (0)

// Flat panel geometry in the two frames the integrators use.  Layout is
// fixed (80 doubles) so the device reads one contiguous record per panel.
struct PanelGeom {
   double c[3];      // centroid
   double ax[9];     // analytic frame, rows x,y,z (z = (v2-v0)x(v1-v0) normalised)
   double pv[9];     // vertices in the analytic frame
   double cC[3];     // edge direction cosines in the analytic frame
   double cS[3];     // edge direction sines
   double el[3];     // edge lengths
   double mom[16];   // moments, 1-based: mom[1..15]
   double maxd;
   double axn[9];    // numerical frame, z = (v1-v0)x(v2-v0) normalised
   double pvn[9];    // vertices in the numerical frame
   double en[9];     // in-plane edge normals (numerical frame)
   double erhs[3];
   double eov[3];
   double area;
   double pad_[1];
};
static_assert(sizeof(PanelGeom) == 82 * 8, "PanelGeom layout");

void panel_geometry(const double* v0, const double* v1, const double* v2, PanelGeom* p);

// One cube of a flattened octree.  Panels and points of every cube are
// contiguous ranges of the tree order (the depth-first concatenation of the
// leaves), so every gather in the matvec reads contiguous segments.
struct Cube {
   int level;
   int idx[3];
   double b0[3], b1[3], ctr[3];
   int parent;
   int child[8];
   int leaf;
   int p0, np;    // panel range in tree order
   int q0, nq;    // point range in tree order
   std::vector<int> loc, itr;
   // filled later
   int rowrank = 0, colrank = 0;
};

struct Tree {
   int kernel = LAPLACE;
   double kappa = 0.0;
   int npan = 0, npt = 0;
   int depth = 0;
   int gridpoints = 4;
   int maxpanels = 32;
   double svdeps = 1e-5;
   std::vector<int> gpl;          // grid points per level
   std::vector<Cube> cubes;       // cubes[0] = root, parents before children
   std::vector<int> pan_perm;     // tree order -> caller panel index
   std::vector<int> pt_perm;      // tree order -> caller point index
   // per cube, its panels / points in caller-index ascending order
   std::vector<int> cube_pan_sorted_off, cube_pan_sorted;
   std::vector<int> cube_pt_sorted_off, cube_pt_sorted;
};

// Builds the octree exactly as the reference decides it (cubic root box,
// centroid split, depth by the 90%-of-leaves rule) and the lists, from the
// panel centroids (3 per panel) and the points.
void build_tree(Tree& t, const double* pan_cent, int npan, const double* pts, int npt, int maxpanels,
                int gridpoints);

// Device-resident compressed operator of one tree in one storage precision.
struct DevTreeOp;
// Device-resident region operator (both trees, precond, solver scratch).
struct Problem;

}  // namespace fsvd
