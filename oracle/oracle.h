/*
 * CPU oracle for the FFTSVD Poisson-Boltzmann BEM solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, single-threaded
 * restatement of the reference algorithm (jbardhan/fftsvd), used as the
 * checker for the HIP implementation in fftsvd_amd/.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library never links, loads or calls anything in here.
 *
 * Storage precision is FP64 everywhere (the reference's SREAL_IS_DOUBLE
 * variant, FFTSVD.h:23-26); FFTs are direct separable DFTs in FP64.
 *
 * Each function cites the reference file:line it restates.  Functions
 * whose parity is only pinned by mathematics (not by reference golden
 * vectors, of which the reference ships none) say so in the test that
 * pins them.
 */
#ifndef FFTSVD_ORACLE_H
#define FFTSVD_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_LAPLACE = 0, ORC_YUKAWA = 1 };

/* Flat triangular panel: FlatPanel.h:3-27 (fields restated, own layout). */
typedef struct {
   double v[3][3];       /* vertices                                   */
   double c[3];          /* centroid                                   */
   double el[3];         /* edge lengths |v1-v0|, |v2-v1|, |v0-v2|     */
   double ax[3][3];      /* analytic frame, ax[2] ~ (v2-v0)x(v1-v0)    */
   double pv[3][3];      /* vertices in analytic frame                 */
   double cC[3], cS[3];  /* edge direction cos/sin in analytic frame   */
   double axn[3][3];     /* numerical frame, axn[2] ~ (v1-v0)x(v2-v0)  */
   double pvn[3][3];     /* vertices in numerical frame                */
   double en[3][3];      /* in-plane edge normals (numerical frame)    */
   double erhs[3];       /* edge line offsets                          */
   double eov[3];        /* opposite-vertex side indicator             */
   double area;
   double mom[16];       /* FASTCAP moments, 1-based (mom[1..15])      */
   double maxd, mind;
} orc_panel;

void orc_panel_init(orc_panel* p, const double* v0, const double* v1, const double* v2);
void orc_laplace(const double* x, const orc_panel* p, double* slp, double* dlp);
void orc_yukawa(const double* x, const orc_panel* p, double kappa, double* slp, double* dlp);
void orc_yukawa_rule(const double* x, const orc_panel* p, double kappa, int nq, const double* gx, const double* gw,
                     double* slp, double* dlp);
/* calcp_LogIntDoLine's radial line integral (FlatIntegration.c:473-635) alone */
void orc_yukawa_line(double z, double x, double y1, double y2, double kappa, int nq, const double* gx,
                     const double* gw, double* sl, double* dl);
void orc_yukawa_rule_v9(const double* x, const double* v9, double kappa, int nq, const double* gx, const double* gw,
                        double* slp, double* dlp);
void orc_integrate(int kernel, double kappa, const double* x, const orc_panel* p, double* slp, double* dlp);
/* kappa-derivatives of the Yukawa single and double layer of orc_yukawa */
void orc_yukawa_dk(const double* x, const orc_panel* p, double kappa, double* dslp, double* ddlp);
void orc_yukawa_dk_batch(double kappa, const double* v9, int npan, const double* pts, int npt, double* ds,
                         double* dd);   /* npt x npan */
double orc_green(int kernel, double kappa, const double* dest, const double* src);
void orc_gl4(double* x, double* w);

/* ------------------------------------------------------------------ */
/* Octree of one compressed operator ("Tree", Tree.h:3-38).            */
typedef struct orc_cube {
   int level;
   int idx[3];
   double b0[3], b1[3], ctr[3];
   int parent;
   int child[8];           /* cx*4 + cy*2 + cz, -1 if absent */
   int leaf;
   int npan, npt;
   int* pan;               /* panel indices */
   int* pt;                /* point indices */
   int nloc, nint;
   int* loc;               /* local cube ids */
   int* itr;               /* interacting cube ids */
   int drows, dcols;
   double* Ds;             /* drows x dcols row-major */
   double* Dd;
   int rowrank, colrank;
   double* VT;             /* rowrank x npan */
   double* U;              /* npt x colrank */
   double* PVs;            /* gp3 x rowrank */
   double* PVd;
   double* UTI;            /* colrank x gp3 */
   double* spec_src;       /* pad complex (re,im interleaved) */
   double* spec_acc;
} orc_cube;

typedef struct {
   int kernel;
   double kappa;
   int npan, npt;
   const orc_panel* panels;   /* borrowed */
   double* pts;               /* npt x 3, owned copy */
   int maxpanels;
   int gridpoints;
   double svdeps;
   int depth;
   int* gpl;                  /* grid points per level */
   int ncubes, cap;
   orc_cube* cubes;           /* cubes[0] = root */
   double** pinv;             /* per level, gp3 x 49 row-major */
   double** T;                /* per level, 343 offsets x pad complex (NULL for inner offsets) */
   int double_only;           /* DOUBLE_LAYER_INT tree: bases sampled from the double layer only */
   double* dS;                /* dense mode: all-pairs integrals (npt x npan each), or NULL */
   double* dD;
} orc_tree;

orc_tree* orc_tree_build(const orc_panel* panels, int npan, const double* pts, int npt,
                         int kernel, double kappa, int maxpanels, int gridpoints, double svdeps);
void orc_tree_fill(orc_tree* t);
/* Sample the bases from the double layer only (a DOUBLE_LAYER_INT tree,
 * Cube.c:1171-1412); call before orc_tree_fill. */
void orc_tree_set_double_only(orc_tree* t);
void orc_tree_free(orc_tree* t);
/* y = S*xs + D*xd over all (point, panel) pairs, compressed (Tree.c:597-683) */
void orc_tree_apply_both(orc_tree* t, const double* xs, const double* xd, double* y);
/* y = A_layer^T x (Tree.c:769-832); layer 0 = single, 1 = double */
void orc_tree_apply_transpose(orc_tree* t, int layer, const double* x, double* y);

/* ------------------------------------------------------------------ */
/* Nested region operator ("SurfaceOperator", SurfaceOperator.h:5-23). */
typedef struct orc_region {
   int kernel;
   double eps, kappa;
   int mynpan;
   int startphi, startdphi;
   orc_tree* tree;            /* over own panels followed by children's panels */
   orc_panel* pan_store;      /* the tree's panels when gathered from several surfaces, or NULL */
   int nchild;
   struct orc_region** child;
   struct orc_region* parent;
   int ncharge;
   double* qpos;              /* ncharge x 3 */
   double* q;
   double* resint;
   double* resext;
   orc_tree* m3;              /* panels -> charges tree (accelerateM3), or NULL */
} orc_region;

typedef struct {
   orc_region* root;
   int ntotal;                /* total panels N; unknowns = 2N */
   orc_panel* panels;         /* all dielectric panels */
   double* precond;           /* 2x2 inverse blocks per global panel: 4 per panel */
   int* precond_phi;          /* startphi + i for each block */
   int* precond_dphi;
   int nblocks;
} orc_problem;

/* One dielectric surface, no cavities (FFTSVDpbeAPI.c:1381-1554 salt,
 * 1557-1730 no-salt) with all charges inside. */
orc_problem* orc_problem_single(const double* verts9, int npan, const double* qpos, const double* q,
                                int ncharge, double eps_in, double eps_out, double kappa,
                                int maxpanels, int gridpoints, double svdeps);
orc_problem* orc_problem_stern(const double* salt9, int nsalt, const double* diel9, int ndiel, const double* qpos,
                               const double* q, int ncharge, double eps_in, double eps_out, double kappa,
                               int maxpanels, int gridpoints, double svdeps);
/* Every surface layout generateSurfaceOperator builds (FFTSVDpbeAPI.c:1064-1730):
 * surfaces concatenated salt, dielectric, dielectric cavity, salt cavity in
 * verts9 with counts[nsalt + ndiel + ncav + nscav]; diel_parent indexes the
 * salt surfaces (Stern layout only), cav_parent the dielectrics,
 * scav_parent the dielectric cavities; a parent of 999 drops the cavity
 * (readSRF, FFTSVDpbeAPI.c:1006-1029).  stern != 0 selects the Stern branch
 * (needs nsalt > 0 and kappa > 0); otherwise salt surfaces and salt
 * cavities are ignored and kappa > 0 selects the salt branch. */
orc_problem* orc_problem_layout(const double* verts9, const int* counts, int nsalt, int ndiel, int ncav, int nscav,
                                const int* diel_parent, const int* cav_parent, const int* scav_parent,
                                const double* qpos, const double* q, int ncharge, double eps_in, double eps_out,
                                double kappa, int stern, int maxpanels, int gridpoints, double svdeps);
void orc_problem_free(orc_problem* pr);
/* Dense mode: every region's tree applies its all-pairs integrals
 * (orc_tree_dense) instead of the compressed operator -- the uncompressed
 * BEM operator, a checker for derivatives of the energy. */
void orc_problem_make_dense(orc_problem* pr);
void orc_so_matvec(orc_problem* pr, const double* x, double* y);
void orc_rhs(orc_problem* pr, double* rhs);
void orc_precond_build(orc_problem* pr);
void orc_precond_solve(orc_problem* pr, const double* b, double* x);
int orc_gmres(orc_problem* pr, const double* rhs, double* sol, double tol, int maxit, double* resid_hist);
void orc_reaction_potentials(orc_problem* pr, const double* sol, double* phi);

/* ------------------------------------------------------------------ */
/* ctypes-friendly accessors */
int orc_tree_depth(orc_tree* t);
int orc_tree_ncubes(orc_tree* t);
void orc_tree_cube_info(orc_tree* t, int c, int* info /* level,i,j,k,leaf,npan,npt,nloc,nint,rowrank,colrank,parent */);
void orc_tree_cube_lists(orc_tree* t, int c, int* pan, int* pt, int* loc, int* itr);
void orc_tree_cube_mats(orc_tree* t, int c, double* Ds, double* Dd, double* VT, double* U,
                        double* PVs, double* PVd, double* UTI);
void orc_tree_gpl(orc_tree* t, int* gpl);
void orc_tree_pinv(orc_tree* t, int level, double* out);
void orc_tree_T(orc_tree* t, int level, int off, double* out);
/* Replace one cube's compressed blocks (for applying an externally built operator). */
void orc_tree_set_cube_mats(orc_tree* t, int c, int rowrank, int colrank, const double* Ds, const double* Dd,
                            const double* VT, const double* U, const double* PVs, const double* PVd,
                            const double* UTI);
orc_tree* orc_problem_tree(orc_problem* pr, int which /* region number in pre-order, 0 = root */);
int orc_problem_ntotal(orc_problem* pr);   /* panels in the operator; unknowns = twice that */
/* M3 trees (accelerateM3, FFTSVDpbeAPI.c:1503-1513): per region with charges,
 * a Laplace tree from the region's panels to its charges; the reaction
 * potentials through it (SurfaceOperator.c:347-363). */
void orc_m3_build(orc_problem* pr);
void orc_m3_potentials(orc_problem* pr, const double* sol, double* phi);
orc_tree* orc_problem_m3_tree(orc_problem* pr, int which);
void orc_tree_dense(orc_tree* t, double* S, double* D);   /* npt x npan each, direct integrals */

/* Unit-level helpers */
void orc_panels_from_verts(const double* verts9, int n, orc_panel* out);
size_t orc_panel_size(void);
void orc_integrate_batch(int kernel, double kappa, const double* verts9, int npan, const double* pts,
                         int npt, double* slp, double* dlp);   /* npt x npan */
void orc_panel_geom(const double* verts9, double* out /* c[3], area, maxd, mind, mom[15], ax[9], axn[9] */);
/* 3D r2c DFT of an N^3 real grid into N x N x (N/2+1) complex, and the inverse (unnormalised) */
void orc_dft_r2c(int N, const double* in, double* out);
void orc_dft_c2r(int N, const double* in, double* out);
void orc_pinv(const double* X, int m, int n, double* XI);
int orc_colbasis_mgs(const double* A, int rows, int cols, double eps, double* Q /* rows x cols */);
/* Reading MSMS-style .vert/.face (VertFace.c:26-202): returns number of panels kept. */
int orc_read_vertface(const char* vertfile, const char* facefile, int flip, int accessible, double** verts9);
void orc_free(void* p);

/* ------------------------------------------------------------------ */
/* Qualocation (second-kind, apparent-surface-charge) operator on one
 * Laplace tree of double-layer integrals (QualocationOperator.c;
 * Solv_ecf_multiply_qual, GMRES.c:159-181; GMRES_QualocationOperator,
 * GMRES.c:541-640; Preconditioner_fill_diagonal_solv_ecf_qual,
 * Preconditioner.c:49-60; Charge_makerhs_ecf_qual, Charge.c:129-142). */
typedef struct {
   int npan, nq;
   orc_panel* panels;
   double* cent;
   orc_tree* tree;
   double* diag;              /* D_double self entries (Tree_extractdiagonal) */
   double* qpos;
   double* q;
   double idiel, odiel;
   orc_tree* m1m3;            /* panels -> charges tree (accelerateM3), or NULL */
} orc_qual;

orc_qual* orc_qual_create(const double* verts9, int npan, const double* qpos, const double* q, int nq,
                          double idiel, double odiel, int maxpanels, int gridpoints, double svdeps);
void orc_qual_free(orc_qual* o);
void orc_qual_matvec(orc_qual* o, const double* x, double* b);
void orc_qual_precond_diag(orc_qual* o, double* p);
void orc_qual_rhs(orc_qual* o, double* rhs);
/* early_exit: return zeros when ||P^-1 b|| < tol (GMRES_QualocationOperator);
 * 0 for GMRES_solv_ecf_qual, which has no such test.  Returns the iterations. */
int orc_qual_gmres(orc_qual* o, const double* rhs, double* sol, double tol, int maxit, int early_exit, double* hist);
void orc_qual_potentials(orc_qual* o, const double* sol, double* phi);
orc_tree* orc_qual_tree(orc_qual* o);
/* The M1M3 tree (generateQualocationOperatorA1A3, FFTSVDpbeAPI.c:1841-1851):
 * once built, orc_qual_rhs and orc_qual_potentials go through it
 * (QualocationOperator.c:33-70). */
void orc_qual_m1m3_build(orc_qual* o);
orc_tree* orc_qual_m1m3_tree(orc_qual* o);

#ifdef __cplusplus
}
#endif
#endif
