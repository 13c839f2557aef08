/*
 * Extended C-ABI of libfftsvd_hip.so: array-based entry points with status
 * codes, precision selection, device-resident benchmarking and the views
 * the parity tests need.  None of these exist in the reference; each says
 * which reference routine it mirrors.  Return value: 0 on success,
 * negative on error (message in fftsvd_last_error()).
 *
 * Vectors of unknowns have length 2N in the reference's global order
 * (per region phi block then dphi/dn block, regions in pre-order,
 * SurfaceOperator.c:39-47).  Precision: 0 = FP64 storage and arithmetic,
 * 1 = FP32 storage/arithmetic of the compressed operator with FP64 Krylov
 * vectors ("mixed").
 */
#ifndef FFTSVD_HIP_H
#define FFTSVD_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fftsvd_problem fftsvd_problem;

const char* fftsvd_last_error(void);
int fftsvd_device_info(char* name, int len, int* num_cu);
/* Device-to-device copy of `bytes` bytes repeated nrep times on a private
 * stream; gbs = (read + written bytes) / time.  A bandwidth probe and the
 * reference kernel for calibrating memory counters. */
int fftsvd_copy_bandwidth(long long bytes, int nrep, double* gbs);
/* LDS read rate of the whole device in GB/s: every lane of eight
 * 256-thread blocks per CU reads width-byte words (8 or 16; a wave one
 * contiguous 64 * width bytes, no bank conflicts) from a 32 KiB LDS array,
 * repeated nrep times.  The measured ceiling for the LDS-bound
 * translation, whose lanes read 8 bytes at a time. */
int fftsvd_lds_bandwidth(int width, int nrep, double* gbs);

/* One dielectric surface (panels as v0,v1,v2 triples, the vertex order the
 * reference builds from a flipped MSMS face), all charges inside; salt if
 * kappa > 0 (mirrors generateSurfaceOperator + generateSurfaceOperatorPreconditioner,
 * FFTSVDpbeAPI.c:1381-1730,1853-1861). */
int fftsvd_problem_create(const double* verts9, int npan, const double* qpos, const double* q, int ncharge,
                          double eps_in, double eps_out, double kappa, int maxpanels, int gridpoints, double svdeps,
                          int want_f32, fftsvd_problem** out);
/* General surfaces: ndiel dielectric surfaces and ncav cavities, panel
 * counts in npan_diel / npan_cav, vertices concatenated in that order,
 * cavity parents index the dielectrics. */
int fftsvd_problem_create_multi(int ndiel, const int* npan_diel, int ncav, const int* npan_cav,
                                const int* cav_parent, const double* verts9, const double* qpos, const double* q,
                                int ncharge, double eps_in, double eps_out, double kappa, int maxpanels,
                                int gridpoints, double svdeps, int want_f32, fftsvd_problem** out);
/* Stern-layer formulation (mirrors generateSurfaceOperator's Stern branch,
 * FFTSVDpbeAPI.c:1064-1377): nsalt salt (ion-exclusion) surfaces, ndiel
 * dielectric surfaces each inside salt surface diel_parent[i], ncav
 * dielectric cavities (parent: a dielectric), nscav salt cavities (parent:
 * a dielectric cavity); vertices concatenated salt, dielectric, cavity,
 * salt cavity.  Needs kappa > 0.  Unknowns follow the reference's pre-order
 * start indices (salt surfaces' phi and dphi first). */
int fftsvd_problem_create_stern(int nsalt, const int* npan_salt, int ndiel, const int* npan_diel,
                                const int* diel_parent, int ncav, const int* npan_cav, const int* cav_parent,
                                int nscav, const int* npan_scav, const int* scav_parent, const double* verts9,
                                const double* qpos, const double* q, int ncharge, double eps_in, double eps_out,
                                double kappa, int maxpanels, int gridpoints, double svdeps, int want_f32,
                                fftsvd_problem** out);
/* Qualocation operator (mirrors generateQualocationOperator(Minimal) and
 * generateQualocationOperatorPreconditioner, FFTSVDpbeAPI.c:1741-1839,1863-1869;
 * QualocationOperator.c): Laplace only, one tree over the npan panels whose
 * bases are sampled from the double layer alone (DOUBLE_LAYER_INT), one
 * unknown per panel, diagonal preconditioner.  The problem_* calls below
 * then take N-vectors: matvec is Solv_ecf_multiply_qual (GMRES.c:159-181),
 * rhs Charge_makerhs_ecf_qual (Charge.c:129-142), solve
 * GMRES_QualocationOperator (GMRES.c:541-640) and potentials
 * QualocationOperator_calculateReactionPotentials (QualocationOperator.c:55-70). */
int fftsvd_qual_create(const double* verts9, int npan, const double* qpos, const double* q, int ncharge,
                       double eps_in, double eps_out, int maxpanels, int gridpoints, double svdeps, int want_f32,
                       fftsvd_problem** out);
int fftsvd_problem_destroy(fftsvd_problem* p);
int fftsvd_problem_size(fftsvd_problem* p, int* npanels, int* ncharges);
/* Length of the problem's vectors: 2N for a surface operator, N for qualocation. */
int fftsvd_problem_unknowns(fftsvd_problem* p, long long* n);

/* y = A x (SurfaceOperator_topmultiply, SurfaceOperator.c:90-97) */
int fftsvd_problem_matvec(fftsvd_problem* p, int precision, const double* x, double* y);
/* x = P^{-1} b (Preconditioner_solve, Preconditioner.c:272-294) */
int fftsvd_problem_precond(fftsvd_problem* p, const double* b, double* x);
/* rhs (SurfaceOperator_makeRHS, SurfaceOperator.c:291-321) */
int fftsvd_problem_rhs(fftsvd_problem* p, double* rhs);
/* GMRES to tol (GMRES_SurfaceOperator, GMRES.c:446-539).  precision 1
 * adds FP64 residual refinement; iters (optional) receives the iteration
 * count, resid_hist (optional, maxit entries) the residual history, info
 * (optional, 6 doubles): iterations, refinements, final residual, ms,
 * FP64 matvecs, FP32 matvecs. */
int fftsvd_problem_solve(fftsvd_problem* p, int precision, const double* rhs, double* sol, double tol, int maxit,
                         int* iters, double* resid_hist, double* info);
/* Several problems (an uneven batch of molecules, each with its own
 * operator) solved at once: problem k solves rhs[k] -> sol[k] (host arrays
 * of its nvec) exactly as fftsvd_problem_solve would, iters[k] and info[6k
 * ..] likewise; the solves run concurrently, each on its own problem's
 * streams, largest first, so small molecules share the GPU instead of
 * taking it in turn.  The reference solves one molecule per program run
 * (FFTSVDpbe.c:101-268); there is no counterpart. */
int fftsvd_problems_solve(int nprob, fftsvd_problem* const* probs, int precision, const double* const* rhs,
                          double* const* sol, double tol, int maxit, int* iters, double* info);
/* reaction potentials at the charges (SurfaceOperator.c:323-401 direct branch) */
int fftsvd_problem_potentials(fftsvd_problem* p, const double* sol, double* phi);
/* Many charge vectors at once (the multi-right-hand-side pattern of the
 * reference's hessian(), FFTSVDpbe.c:52-99, which loops generateRHS ->
 * solveSurfaceOperator -> collectPotentials per charge vector).  Arrays are
 * column-major: q ncharge x ncol, rhs and sol nvec x ncol, phi ncharge x
 * ncol, info 6 per column as fftsvd_problem_solve.  Column c equals the
 * single-vector call with the problem's charges replaced by q_c; the
 * integrals of the right-hand sides and potentials are formed once for all
 * columns.  fftsvd_problem_solve_batch runs one GMRES per column (and, in
 * mixed mode, its FP64 refinement) in lockstep, the columns still
 * iterating sharing one column-batched matvec per step, so every column's
 * solution and counts equal fftsvd_problem_solve's bit for bit; info[3]
 * (ms) is the whole batch's time.  Up to 64 columns share a step; more
 * are taken 64 at a time. */
int fftsvd_problem_rhs_batch(fftsvd_problem* p, int ncol, const double* q, double* rhs);
int fftsvd_problem_solve_batch(fftsvd_problem* p, int precision, int ncol, const double* rhs, double* sol,
                               double tol, int maxit, int* iters, double* info);
/* y_c = A x_c for ncol column-major vectors (nvec x ncol), the operator
 * applied to all columns in one launch per stage (each column's result is
 * the single-vector fftsvd_problem_matvec's, bit for bit).  Surface
 * operators.  No counterpart in the reference, whose hessian() applies the
 * operator once per charge vector. */
int fftsvd_problem_matvec_batch(fftsvd_problem* p, int precision, int ncol, const double* x, double* y);
int fftsvd_problem_potentials_batch(fftsvd_problem* p, int ncol, const double* sol, double* phi);
/* Adjoint (surface operators): x = A^T y, the transposes of the region
 * join / tree apply / split (Tree_multiply_transpose, Tree.c:769-832, per
 * layer); A^T x = b by the same GMRES with the transposed 2x2 blocks as
 * preconditioner; and the gradient of E = 1/2 sum_j q_j phi_j with respect
 * to the problem's charges at the forward solution sol (2N, host):
 * dE/dq = 1/2 (phi + B^T lambda) with A^T lambda = C^T q, B the charges ->
 * right-hand-side map and C the solution -> reaction-potential map.  The
 * reference has no adjoint path; FFTSVDpbe.c's hessian() forms the same
 * derivatives column by column from forward solves. */
int fftsvd_problem_matvec_transpose(fftsvd_problem* p, int precision, const double* y, double* x);
int fftsvd_problem_solve_adjoint(fftsvd_problem* p, int precision, const double* rhs, double* sol, double tol,
                                 int maxit, int* iters, double* info);
int fftsvd_problem_grad_charges(fftsvd_problem* p, int precision, const double* sol, double tol, int maxit,
                                double* grad, int* iters);
/* dE/dkappa of the same energy at the forward solution sol (2N, host), in
 * the units of fftsvd_problem_grad_charges: -1/2 lambda^T (dA/dkappa) x
 * with A^T lambda = C^T q, evaluated in tangent form as -1/2 q^T C w with
 * A w = (dA/dkappa) x (one forward solve; iters reports it).  dA/dkappa is the derivative of each Yukawa
 * region's compressed operator as it is applied, at fixed tree, ranks and
 * bases: near-field blocks of the kappa-derivative of the reference's
 * Yukawa scheme (FlatIntegration.c:473-949 differentiated term by term),
 * and the derivatives of the equivalent densities (through the
 * pseudoinverse), of the translation spectra and of the target-side
 * projections.  Built once per operator on the first call.  No counterpart
 * in the reference (FFTSVDpbe has no kappa derivative). */
int fftsvd_problem_grad_kappa(fftsvd_problem* p, int precision, const double* sol, double tol, int maxit,
                              double* dedk, int* iters);
/* Operator statistics, up to 32 doubles: 0 panels, 1 regions, 2 setup ms,
 * 3 near entries, 4 far entries, 5 spectrum bins, 6 translation MACs,
 * 7 depth, 8 leaves, 9 tree applications per matvec, 10/11 algorithmic
 * bytes per matvec (FP64 / FP32 storage), 12/13 near-field bytes, 14 near
 * column indices, 15 U entries, 16 VT entries, 17 PV entries, 18 UTI
 * entries, 19 T spectrum bins, 20 leaf far-field blocks, 21 projection
 * chunks, 22 points applied, 23 source spectrum bins, 24 target spectrum
 * bins (all counts summed over tree applications), 25..31 host wall ms of
 * the setup phases summed over built trees: tree, near field, pinv + T
 * spectra, sampling integrals, pivoted MGS, bases/PV/UTI, lists + plan. */
int fftsvd_problem_stats(fftsvd_problem* p, double* out, int nout);
/* Times `nrep` surface matvecs on device-resident vectors with HIP events
 * on the library stream: total ms, and in ms_parts[0..3] the per-stage ms
 * (leaf: near field + far-field output; projection + source transforms;
 * translation; target transforms) from separately timed passes. */
int fftsvd_problem_bench_matvec(fftsvd_problem* p, int precision, int nrep, double* ms_total, double* ms_parts);
/* The same for nrep column-batched matvecs of ncol columns each
 * (fftsvd_problem_matvec_batch on device-resident vectors). */
int fftsvd_problem_bench_matvec_batch(fftsvd_problem* p, int precision, int ncol, int nrep, double* ms_total,
                                      double* ms_parts);

/* Per-tree views for parity tests.  which = region index (0 = exterior). */
int fftsvd_tree_info(fftsvd_problem* p, int which, int* depth, int* ncubes, int* npanels);
/* info[13] per cube: level, i, j, k, leaf, npan, npt, nloc, nint, rowrank, colrank, parent, dcols */
int fftsvd_tree_cube_info(fftsvd_problem* p, int which, int cube, int* info);
/* caller (tree-local) panel indices of the cube in tree order, and of the
 * leaf's near-field columns */
int fftsvd_tree_cube_panels(fftsvd_problem* p, int which, int cube, int* pan, int* dcols);
/* Caller indices of a cube's points (its nq entries).  For every
 * fftsvd_tree_* call, which >= 0 selects the tree of region `which` and
 * which < 0 the M3 tree (panels -> charges, built when accelerateM3 is set
 * at creation) of region -which-1. */
int fftsvd_tree_cube_points(fftsvd_problem* p, int which, int cube, int* pt);
/* FP64 blocks of one cube: VT (r x np), U (np x r'), PVs/PVd (gp3 x r),
 * UTI (r' x gp3), Ds/Dd (p x m); rows/columns in tree order. */
int fftsvd_tree_cube_mats(fftsvd_problem* p, int which, int cube, double* VT, double* U, double* PVs, double* PVd,
                          double* UTI, double* Ds, double* Dd);
/* y = S xs + D xd of one tree (Tree_multiplyboth, Tree.c:597-683), caller order */
int fftsvd_tree_apply(fftsvd_problem* p, int which, int precision, const double* xs, const double* xd, double* y);
/* y = A_layer^T x of one tree (Tree_multiply_transpose, Tree.c:769-832) */
int fftsvd_tree_apply_transpose(fftsvd_problem* p, int which, int precision, int layer, const double* x, double* y);

/* Panel integrals on the GPU: slp, dlp (npt x npan) for kernel 0 = Laplace,
 * 1 = Yukawa (Integration, Integration.c:32-185), 2 = the kappa-derivatives
 * of the Yukawa pair (the gradient kernel of fftsvd_problem_grad_kappa). */
int fftsvd_integrate(int kernel, double kappa, const double* verts9, int npan, const double* pts, int npt,
                     double* slp, double* dlp);

/* Allocation-free variant of fftsvd_integrate.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  verts9 (npan x 9), pts (npt x 3), slp and dlp (npt x npan,
 * row-major) are DEVICE arrays of the caller; the call enqueues two kernels
 * on `stream` (a hipStream_t, NULL for the default stream) and returns
 * without waiting.  The per-panel geometry records live in a fixed device
 * array of FFTSVD_NOALLOC_MAX_PANELS records (10.75 MB per GPU, held until
 * the process ends): larger npan is rejected with an error.  Calls on one
 * GPU share that array, so calls that overlap in time on one GPU race;
 * npan * npt must stay below 2^31.  The results equal fftsvd_integrate's. */
#define FFTSVD_NOALLOC_MAX_PANELS 16384
int fftsvd_integrate_noalloc(int kernel, double kappa, const double* verts9, int npan, const double* pts, int npt,
                             double* slp, double* dlp, void* stream);

/* Solvent-excluded surface of atoms (x, y, z, r per atom) with a probe of
 * radius `probe`, contoured on a grid of spacing h; the library's own
 * replacement for the MSMS run behind the reference's generateFlatSRF
 * (FFTSVDpbeAPI.c:1034-1040, meshmaker.c:119-270).  Faces are 0-based,
 * counter-clockwise seen from outside; free both arrays with fftsvd_free. */
int fftsvd_mesh_ses(const double* xyzr, int natoms, double probe, double h, double** verts, int* nverts,
                    int** faces, int* nfaces);
/* Surface of the union of the atom spheres grown by `grow` (x, y, z, r per
 * atom), contoured like fftsvd_mesh_ses: the ion-exclusion (Stern) surface
 * the reference's meshmaker asks MSMS for with radii + ionexclusionradius
 * and a vanishing probe (meshmaker.c:105-133). */
int fftsvd_mesh_union(const double* xyzr, int natoms, double grow, double h, double** verts, int* nverts,
                      int** faces, int* nfaces);
/* Every connected component of the SES (union_mode 0, probe) or of the
 * grown-sphere union (union_mode 1, grow), largest first: the faces and the
 * vertices of a component are contiguous and comp_faces[k] counts the
 * faces of component k (ncomp of them).  Faces keep the MSMS convention
 * (counter-clockwise seen from the solvent side), so an outer surface
 * encloses a positive volume and a cavity a negative one, the split the
 * reference's meshmaker makes from MSMS's -all_components output
 * (meshmaker.c:119-262).  Free the three arrays with fftsvd_free. */
int fftsvd_mesh_components(const double* xyzr, int natoms, double probe, double h, int union_mode, double** verts,
                           int* nverts, int** faces, int* nfaces, int** comp_faces, int* ncomp);
void fftsvd_free(void* p);

/* Autotuning of the FP32 translation (staging mode, tile source budget,
 * targets per tile): with on != 0 (or FFTSVD_AUTOTUNE=1 in the
 * environment) the next FP32 operator built times each candidate on its own
 * translation and the fastest is kept for the rest of the process; off
 * restores the hand-tuned defaults.  fftsvd_translation_params reports the
 * parameters in use, whether they were tuned and the chosen candidate's
 * time per launch.  No counterpart in the reference. */
int fftsvd_set_autotune(int on);
int fftsvd_translation_params(int* mode, int* budget, int* tile, int* tuned, double* tuned_ms);
/* Precision used by solveSurfaceOperator (0 FP64, 1 mixed); default 0. */
int fftsvd_set_precision(int precision);

#ifdef __cplusplus
}
#endif
#endif
