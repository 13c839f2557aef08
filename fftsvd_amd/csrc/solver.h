// Nested region operator over 2N unknowns (phi and dphi/dn per panel), its
// block preconditioner and the GMRES driver, all device resident.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "op.h"

namespace fsvd {

struct SurfaceSpec {
   // Surfaces in the order the caller numbers its panels: each dielectric
   // surface, then each dielectric cavity.  Cavity parent = index of the
   // dielectric that encloses it.
   std::vector<std::vector<double>> dielectric_verts;   // per dielectric, npan*9
   std::vector<std::vector<double>> cavity_verts;       // per cavity, npan*9
   std::vector<int> cavity_parent;
   // Stern-layer formulation when salt_verts is not empty: the salt
   // (ion-exclusion) surfaces, the salt surface enclosing each dielectric,
   // and the salt cavities with the dielectric cavity enclosing each
   std::vector<std::vector<double>> salt_verts;
   std::vector<int> dielectric_parent;
   std::vector<std::vector<double>> salt_cavity_verts;
   std::vector<int> salt_cavity_parent;
   std::vector<double> qpos, q;                         // all atoms
   double eps_in = 4.0, eps_out = 80.0, kappa = 0.0;
   int maxpanels = 32, gridpoints = 4;
   double svdeps = 1e-5;
   bool want32 = true;
   // accelerateM3: reaction potentials through a panels -> charges tree per
   // region with charges (FFTSVDpbeAPI.c:1503-1513) instead of direct sums
   bool m3 = false;
};

// Qualocation operator (QualocationOperator.c): one Laplace tree over all
// dielectric (and cavity) panels with double-layer-only bases, one unknown
// per panel, Laplace only.
struct QualSpec {
   std::vector<double> verts;   // npan*9, caller order
   std::vector<double> qpos, q;
   double eps_in = 4.0, eps_out = 80.0;
   int maxpanels = 32, gridpoints = 4;
   double svdeps = 1e-5;
   bool want32 = true;
   bool m3 = false;             // accelerateM3: the M1M3 panels -> charges tree
};

struct Region {
   int kernel = LAPLACE;
   double eps = 1.0, kappa = 0.0;
   int mynpan = 0;
   int startphi = 0, startdphi = 0;
   int parent = -1;
   std::vector<int> child;
   std::vector<int> panel_ids;       // global panel ids of the tree panels (own then children's)
   std::unique_ptr<TreeOp> op;
   DBuf<PanelGeom> d_panels;         // tree-local panel geometry
   DBuf<int> phi_idx, dphi_idx, out_idx;
   DBuf<double> s_phi, s_dphi, corr;
   DBuf<double> xs, xd, yt;
   DBuf<double> xsB, xdB, ytB;       // the same for the columns of a batched matvec
   bool has_charges = false;
   int shares_tree_with = -1;        // reuse another region's operator (same panels and kernel)
   // M3 tree (region panels -> charges) and the region join in its panel order
   std::unique_ptr<TreeOp> m3;
   DBuf<PanelGeom> m3_panels;
   DBuf<int> m3_phi_idx, m3_dphi_idx;
   DBuf<double> m3_s_phi, m3_s_dphi, m3_xs, m3_xd, m3_y;
   int lane = 0;                     // matvec stream: 0 = main, k > 0 = aux[k-1]
};

struct SolveInfo {
   int iterations = 0;
   int refinements = 0;
   double final_residual = 0.0;
   double ms = 0.0;
   int matvecs64 = 0, matvecs32 = 0;
};

// One column of a lockstep batched solve: its GMRES work vectors, Givens
// state and Krylov basis (slabs of Problem::BASIS_SLAB vectors, with device
// copies of their pointers), the refinement vectors of the mixed mode, and
// the column's host-side state.
struct GmCol {
   DBuf<double> w, r, partial, partial_m, htmp, hbuf, pnorm, H, v;
   DBuf<double> res, e, tmp;
   std::vector<std::unique_ptr<DBuf<double>>> basis64;
   std::vector<std::unique_ptr<DBuf<float>>> basis32;
   DBuf<const double*> ptr64;
   DBuf<const float*> ptr32;
   int nptr64 = 0, nptr32 = 0;
   double normr = 0.0, bn = 0.0, rn = 0.0, tol = 0.0;
   int done = -1;
   bool run = false, refine = false;
   SolveInfo info;
};

struct Problem {
   hipStream_t stream = nullptr;
   // the distinct tree operators run concurrently: operator k of the matvec
   // on stream 0 (k = 0) or aux[k-1]; forked from / joined into `stream`
   static constexpr int MAX_AUX = 3;
   hipStream_t aux[MAX_AUX] = {nullptr, nullptr, nullptr};
   hipEvent_t ev_fork = nullptr, ev_join[MAX_AUX] = {nullptr, nullptr, nullptr};
   int naux = 0;
   int ntotal = 0;                       // total panels
   long long nvec = 0;                   // unknowns: 2N (surface operator) or N (qualocation)
   // qualocation operator (regions[0] holds its tree): per tree position the
   // panel, its area and its D_double diagonal; the diagonal preconditioner
   // per panel; GMRES tests ||P^-1 b|| < tol before iterating unless
   // gmres_zero_test is off (GMRES_solv_ecf_qual has no such test)
   bool qual = false;
   double eps_out = 80.0;
   DBuf<int> qual_perm;
   DBuf<double> qual_area_t, qual_diag_t, qual_pdiag;
   bool gmres_zero_test = true;
   // while set, problem_matvec applies A^T and problem_precond P^-T (the
   // adjoint solves of problem_solve_adjoint)
   bool adjoint = false;
   std::vector<double> cent;             // panel centroids (3 per panel, global order)
   std::vector<PanelGeom> panels;        // host panel geometry (qualocation problems only)
   DBuf<PanelGeom> d_allpanels;          // device panel geometry, global order (direct potentials, gradients)
   std::vector<Region> regions;          // regions[0] = exterior (root)
   std::vector<double> qpos, q;
   double eps_in = 4.0;
   // preconditioner
   int nblocks = 0;
   DBuf<int> pc_a, pc_b;
   DBuf<double> pc_m;
   // device charges
   DBuf<double> d_qpos, d_q;
   // scratch
   DBuf<double> w, r, tmp, partial, hbuf, htmp, partial_m;
   // Krylov bases: FP64 for FP64 solves, FP32 for the FP32-operator solves
   // of the mixed mode, in slabs of BASIS_SLAB vectors (the first slab of
   // each is reserved when the problem is built, so a solve of up to
   // BASIS_SLAB iterations allocates nothing); device copies of their
   // pointers
   static constexpr int BASIS_SLAB = 128;
   std::vector<std::unique_ptr<DBuf<double>>> basis;
   std::vector<std::unique_ptr<DBuf<float>>> basis32;
   DBuf<const double*> d_basis_ptr;
   DBuf<const float*> d_basis_ptr32;
   int basis_ptr_n = 0, basis_ptr_n32 = 0;
   DBuf<double> gm_H, gm_v;             // GMRES Givens state (H; rotations, g, history)
   DBuf<double> pnorm;                  // block partials of |w|^2 from the last multi-axpy
   double* h_pinned = nullptr;
   // lockstep batched solves: per-column state and a pinned staging area of
   // PIN_COL doubles per column
   static constexpr int PIN_COL = 2048;
   std::vector<std::unique_ptr<GmCol>> gm_cols;
   double* h_pinned_b = nullptr;
   int h_pinned_bcols = 0;
   // small uploads of the batched paths (pointer arrays, column
   // descriptors): RING slots of RING_BYTES, pinned and on the device, each
   // with the event of its last copy
   static constexpr int RING = 8;
   static constexpr size_t RING_BYTES = 65536;
   unsigned char* ring_h = nullptr;
   DBuf<unsigned char> ring_d;
   hipEvent_t ring_ev[RING] = {};
   int ring_next = 0;
   double setup_ms = 0.0;
   ~Problem();
};

std::unique_ptr<Problem> build_problem(const SurfaceSpec& spec);
std::unique_ptr<Problem> build_qual_problem(const QualSpec& spec);

// y = A x over the 2N unknowns (reference SurfaceOperator_topmultiply,
// SurfaceOperator.c:90-97); device pointers.
void problem_matvec(Problem& pr, int prec, const double* x, double* y);
void problem_matvec(Problem& pr, int prec, const float* x, double* y);
void problem_precond(Problem& pr, const double* b, double* x);
void problem_rhs(Problem& pr, double* rhs_dev);
// GMRES (reference GMRES_SurfaceOperator, GMRES.c:446-539); device vectors.
SolveInfo problem_gmres(Problem& pr, int prec, const double* rhs, double* sol, double tol, int maxit,
                        std::vector<double>* hist);
// Mixed precision: FP32-operator GMRES with FP64 residual refinement.
// ncol right-hand sides (device, column c at rhs + c * nvec) solved in
// lockstep: every column runs problem_solve's GMRES (and, in mixed mode, its
// FP64 refinement) on its own, and the columns still iterating share one
// column-batched matvec per step, so each column's solution, iteration
// count and residual equal problem_solve's on that column bit for bit.
// Surface operators only.  info[c]: that column's counts; ms is the batch's.
void problem_solve_batch(Problem& pr, int mode, int ncol, const double* rhs, double* sol, double tol, int maxit,
                         std::vector<SolveInfo>& info);
// y_c = A x_c for ncol vectors (device, column c at c * nvec) by the
// column-batched matvec.
void problem_matvec_batch(Problem& pr, int prec, int ncol, const double* x, double* y);
SolveInfo problem_solve(Problem& pr, int mode, const double* rhs, double* sol, double tol, int maxit,
                        std::vector<double>* hist);
// Reaction potentials at the charges: direct sums, or through the M3 trees
// where built (use_m3).  The batched potentials and the charge gradient use
// the direct sums.
void problem_potentials(Problem& pr, const double* sol_host, double* phi_host, bool use_m3 = true);
// Many charge vectors / solutions at once (column-major; q and rhs on the
// device, sol and phi on the host).
void problem_rhs_batch(Problem& pr, int ncol, const double* q, double* rhs);
// Adjoint: A^T solve (device vectors) and the gradient of
// E = 1/2 sum_j q_j phi_j with respect to the charges,
// dE/dq = 1/2 (phi + B^T lambda), A^T lambda = C^T q (B: charges -> RHS,
// C: solution -> reaction potentials); sol and grad on the host.
SolveInfo problem_solve_adjoint(Problem& pr, int mode, const double* rhs, double* sol, double tol, int maxit);
SolveInfo problem_grad_charges(Problem& pr, int mode, const double* sol_host, double tol, int maxit,
                               double* grad_host);
// dE/dkappa (energy units of grad_charges) at the forward solution sol by
// the adjoint: -1/2 lambda^T (dA/dkappa) x with lambda = A^-T C^T q.
SolveInfo problem_grad_kappa(Problem& pr, int mode, const double* sol_host, double tol, int maxit, double* dedk);
void problem_potentials_batch(Problem& pr, int ncol, const double* sol_host, double* phi_host);

}  // namespace fsvd
