// Device-resident compressed operator of one tree, and its region-operator
// and solver wrappers.
#pragma once

#include <memory>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "device.h"
#include "internal.h"

namespace fsvd {

// A device buffer that frees itself.
template <typename T>
struct DBuf {
   T* p = nullptr;
   size_t n = 0;
   DBuf() = default;
   DBuf(const DBuf&) = delete;
   DBuf& operator=(const DBuf&) = delete;
   DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) {
      o.p = nullptr;
      o.n = 0;
   }
   DBuf& operator=(DBuf&& o) noexcept {
      if (this != &o) {
         reset();
         p = o.p;
         n = o.n;
         o.p = nullptr;
         o.n = 0;
      }
      return *this;
   }
   ~DBuf() { reset(); }
   void reset() {
      if (p) (void)hipFree(p);
      p = nullptr;
      n = 0;
   }
   void alloc(size_t count) {
      reset();
      n = count;
      if (count) FSVD_CHECK(hipMalloc(&p, sizeof(T) * count));
   }
   void upload(const T* h, size_t count, hipStream_t s) {
      alloc(count);
      if (count) FSVD_CHECK(hipMemcpyAsync(p, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
   }
   void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
};

struct FarLevel {
   int level = 0, n = 0, pad = 0, gp3 = 0;
   int nsrc = 0, ntgt = 0, maxr_src = 0, maxr_tgt = 0;
   std::vector<SrcDesc> h_src;
   std::vector<TgtDesc> h_tgt;
   std::vector<int> h_isrc, h_itoff;     // target-side interaction lists
   // the transposed apply's descriptors and lists (transposed: panel-side
   // receivers, one per source cube), made on its first use
   std::vector<TgtDesc> h_tgtT;
   std::vector<int> h_isrcT, h_itoffT;
   DBuf<SrcDesc> src;
   DBuf<TgtDesc> tgt, tgtT;
   DBuf<int> isrcT, itoffT;
   DBuf<double> T64;     // 343 * pad complex
   DBuf<float> T32;
   DBuf<double> specT, accT;     // ntgt*pad, nsrc*pad complex: transposed apply only, allocated on first use
   // compressed blocks of this level (offsets in SrcDesc / TgtDesc)
   DBuf<double> VT64, PV64, UT64, UTI64;
   DBuf<float> VT32, PV32, UT32, UTI32;
};

// Parameters of the pipelined FP32 translation (the autotuner's choices):
// staging mode (0 through registers, 1 direct global->LDS), the source
// union at which tile growth stops, targets per tile at most.
struct PipeParams {
   int mode = 1, budget = TP_SRC_BUDGET, tile = TP_TILE;
   float tuned_ms = 0.f;   // the chosen candidate's time per launch, when tuned
};
void set_autotune(bool on);
PipeParams current_pipe_params(bool* tuned);

struct TreeOpStats {
   long long near_entries = 0;     // stored near-field entries (both layers)
   long long far_entries = 0;      // VT + PV + U^T + UTI entries
   long long spectra_bins = 0;     // sum over levels of (#src + #tgt) * pad
   long long translate_macs = 0;   // sum over targets of #interactions * pad
   int depth = 0, nleaf = 0, ncubes = 0;
   double setup_ms_near = 0, setup_ms_far = 0;
   // host wall time of the setup phases: 0 tree, 1 near-field fill,
   // 2 pinv + T spectra, 3 sampling integrals, 4 pivoted MGS, 5 bases,
   // PV and UTI, 6 lists, FP32 copies and apply plan
   double phase_ms[8] = {0};
   PipeParams pipe;   // the translation parameters used
};

// Pipelined FP32 translation plan of one tree, host form: tiles whose
// pointers are not set yet (t_off holds each tile's T offset in complex
// elements), octets, octet outputs, source offsets, interaction words, and
// the tasks (tile, first bin, first octet, octets) in list order with
// their modelled cost.
struct PipeHost {
   std::vector<TpTile> tiles;
   std::vector<long long> t_off;
   std::vector<TpOct> octs;
   std::vector<long long> oacc;
   std::vector<int> gsrc, gint;
   std::vector<int4> tasks;
   std::vector<double> cost;
};
// The device form: nblk persistent blocks, block b takes the tasks
// range[b]; one launch of k_translate_pipe.  It may cover several trees.
struct PipeDev {
   int nblk = 0, mode = 1;   // mode: 0 staging through registers, 1 direct global->LDS transfers
   DBuf<int2> range;
   DBuf<int4> tasks;
   DBuf<TpTile> tiles;
   DBuf<TpOct> octs;
   DBuf<long long> oacc;
   DBuf<int> gsrc, gint;
};

// Per-level data the translation plans need.
struct TrLevel {
   int t0, nt;         // the level's targets: tgt[t0 .. t0 + nt)
   int pad, pst, nsrc;
   long long t_off;    // its T spectra (complex index; code c at t_off + c * pst)
   long long sbase;    // its forward spectra: source slot k at sbase + k * pst
};

// Forward-apply plan of a tree: every level's blocks concatenated, with
// absolute offsets, so each stage is one launch.
struct ApplyPlan {
   int nchunk = 0, nsrc = 0, ntgt = 0, nublk = 0;
   int nsrc_big = 0, ntgt_big = 0;   // cubes with n > 4, first in src / tgt
   int proj_maxcols = 0, maxr_src = 0, max_n = 4;
   DBuf<ProjChunk> chunks;
   DBuf<SrcPlan> src;
   DBuf<TgtPlan> tgt;
   int ntask_d = 0, tr_max_src = 0;   // FP64 translation (k_translate<double>)
   DBuf<int4> task_d;
   DBuf<TrTile> tiles;
   DBuf<GTgt> gtgt;
   DBuf<int> gsrc, gint;
   PipeHost pipe_h;   // FP32 translation (k_translate_pipe) of this tree, host form
   PipeDev pipe;      // ... on the device
   PipeDev pipe2;     // the two-column form (4-bin tasks), made at the first column-batched apply
   // the translation lists the pipelined plan is built from (kept for the
   // autotuner's re-planning)
   std::vector<TrLevel> tr_lv;
   std::vector<TgtPlan> tr_tgt;
   std::vector<int> tr_isrc, tr_itoff;
   DBuf<int> map_leaf;   // XCD block map of the leaf launch
   int nmap_leaf = 0;
   DBuf<LeafU> ublk;
   DBuf<double> tpart, spec, acc, w;   // work vectors, sized for FP64
   // the same for ncolB columns of a column-batched apply, column c at c
   // times the single-column size
   int ncolB = 0;
   DBuf<double> tpartB, specB, accB, wB;
   DBuf<double> VTc64, PV64, Tsp64, UTI64, UT64;
   DBuf<float> VTc32, PV32, Tsp32, UTI32, UT32;
};

// d/dkappa of a Yukawa operator at fixed tree, ranks and bases, in the
// layouts of the FP64 apply plan: near-field blocks of the kappa-derivative
// integrals, the derivatives of PV, of the T spectra and of UTI (VT and U
// unchanged).  Built on demand for the adjoint kappa gradient.
struct DkData {
   double kappa = 0.0;
   DBuf<double> D64, PV64, Tsp64, UTI64;
};

struct TreeOp {
   Tree tree;
   int kernel = LAPLACE;
   double kappa = 0.0;
   bool have32 = false;
   bool haveT = false;   // the transposed apply's lists are on the device
   // geometry
   DBuf<double> pts_tree;     // tree-order point coordinates
   DBuf<int> pan_perm;        // tree order -> caller panel index
   // near field
   int nleaf = 0, maxm = 0, maxp = 0;
   std::vector<LeafDesc> h_leaves;
   std::vector<int> h_dcol;   // tree-order panel positions
   DBuf<LeafDesc> leaves;
   DBuf<int> dcol;
   DBuf<double> D64;
   DBuf<float> D32;
   // far field
   std::vector<FarLevel> levels;
   std::vector<std::vector<double>> h_pinv;   // per level gp3 x 49 (levels >= 2)
   DBuf<double> yfar;
   TreeOpStats stats;
   std::vector<int> cube_src_slot, cube_tgt_slot, cube_leaf;   // per cube: slot in its level / leaf index, -1
   ApplyPlan plan;
   std::unique_ptr<DkData> dk;
};

// The device plan of one launch covering the FP32 translations of the given
// trees (each with its own spectra), tasks dealt to the persistent blocks.
// pair: the tasks of the two-column kernel (each 8-bin task split into its
// two 4-bin halves, every level's first halves before its second halves).
void make_pipe(PipeDev& d, const std::vector<const struct TreeOp*>& ops, hipStream_t s, int mode = 1,
               bool pair = false);
void launch_translate_pipe2(int nblk, const int2* range, const int4* tasks, const TpTile* tiles, const TpOct* octs,
                            const long long* oacc, const int* gsrc, const int* gint, hipStream_t s, int ncol,
                            long long sdel, long long adel, long long scs, long long acs);
// ncol > 1: every column's spectra and accumulators at byte offsets sdel +
// c * scs and adel + c * acs from the buffers the tiles point at.
void launch_translate_pipe(int nblk, int mode, const int2* range, const int4* tasks, const TpTile* tiles,
                           const TpOct* octs, const long long* oacc, const int* gsrc, const int* gint, hipStream_t s,
                           int ncol = 1, long long sdel = 0, long long adel = 0, long long scs = 0,
                           long long acs = 0);
inline void launch_pipe(const PipeDev& d, hipStream_t s) {
   launch_translate_pipe(d.nblk, d.mode, d.range.p, d.tasks.p, d.tiles.p, d.octs.p, d.oacc.p, d.gsrc.p, d.gint.p, s);
}

// Builds tree, near field and far field for `npan` panels (caller-order
// geometry already on the device in `d_panels`) and `npt` points.
// `pre`, if given, is the octree of exactly these panels and points (it
// depends on neither the kernel nor kappa) and is copied instead of rebuilt.
// `layers` = 2 samples the bases from both layers (SINGLE_AND_DOUBLE_LAYER_INT),
// 1 from the double layer only (DOUBLE_LAYER_INT, Cube.c:1171-1412).
// pan_cent: the npan panel centroids (3 each, host), d_panels their device
// geometry records, in the same order.
std::unique_ptr<TreeOp> build_tree_op(const double* pan_cent, int npan, const PanelGeom* d_panels,
                                      const double* pts, int npt, int kernel, double kappa, int maxpanels,
                                      int gridpoints, double svdeps, bool want32, hipStream_t s,
                                      const Tree* pre = nullptr, int layers = 2);

// y_t = S xs_t + D xd_t; all vectors in tree order, device pointers.
void tree_apply(TreeOp& op, int prec, const double* xs, const double* xd, double* y, hipStream_t s);
// One stage of the apply only (0 leaf: near field + U w, 1 projection and
// source transforms, 2 translation, 3 target transforms), for per-kernel
// timing.
void tree_apply_part(TreeOp& op, int prec, int part, const double* xs, const double* xd, double* y, hipStream_t s);
// The same part for ncol columns in one launch per stage: column c of xs and
// xd at c * ldx, of y at c * ldy.  Every column is computed as the
// single-column apply computes it, so its result is the same bit for bit;
// the columns' copies of a block read the operator together (one XCD's L2).
void tree_apply_part_batch(TreeOp& op, int prec, int part, int ncol, const double* xs, const double* xd,
                           long long ldx, double* y, long long ldy, hipStream_t s);
// The leaf stage (part 0) with the region split fused in: instead of y,
// writes yout[out_idx[k]] = y_k + corr[k] xd[k] (k_split's result, bit for
// bit).
void tree_apply_leaf_split(TreeOp& op, int prec, const double* xs, const double* xd, const int* out_idx,
                           const double* corr, double* yout, hipStream_t s);
// Sizes the operator's column work vectors for ncol columns (before a
// batched apply is enqueued on several streams).
void tree_batch_reserve(TreeOp& op, int ncol, hipStream_t s);
// kappa-derivative of a Yukawa tree's operator, at fixed ranks and bases:
// build_dkappa fills op.dk (once per kappa); tree_apply_dkappa applies it
// (FP64), y_t = S' xs + D' xd in tree order.
void build_dkappa(TreeOp& op, const PanelGeom* d_panels, hipStream_t s);
void tree_apply_dkappa(TreeOp& op, const double* xs, const double* xd, double* y, hipStream_t s);
// y_t = A_layer^T x_t (panels, tree order).
void tree_apply_transpose(TreeOp& op, int prec, int layer, const double* x, double* y, hipStream_t s);
// Both layers at once, ys = S^T x and yd = D^T x: the points-side
// transforms and the translation are shared by the two layers.
void tree_apply_transpose_both(TreeOp& op, int prec, const double* x, double* ys, double* yd, hipStream_t s);

// Block map of a launch over n logical blocks with the given work: block b
// of the returned grid runs logical block map[b] (-1: idle); XCD x (blocks
// b = x mod 8) gets a contiguous range of about 1/8 of the total work.
std::vector<int> xcd_block_map(const std::vector<double>& work);

// launchers (apply.hip / fill.hip)
template <typename T>
void launch_far_translate(int nblk, const TgtDesc* tgt, const int* isrc, const int* itoff, const void* Tspec,
                          const void* spec, void* acc, int pad, hipStream_t s);
// One stage of the forward apply (0 project, 1 source, 2 translate,
// 3 target, 4 leaf) over every level of a tree.
template <typename T>
void launch_apply_stage(int stage, const ApplyArgs& a, hipStream_t s);
template <typename T>
void launch_transpose_batched(int n, const TransDesc* d, hipStream_t s);
template <typename T>
void launch_farT_source(int n, int nblk, const TgtDesc* tgt, const T* UT, const T* UTI, const double* x, void* spec,
                        int maxr, hipStream_t s);
template <typename T>
void launch_farT_target(int n, int nblk, const SrcDesc* src, const T* VT, const T* PV, const void* acc, int layer,
                        double* yfar, double* yfar2, int maxr, hipStream_t s);
template <typename T>
void launch_nearT(int nleaf, const LeafDesc* leaves, const int* dcol, const T* D, const double* x, int layer,
                  double* y, double* y2, hipStream_t s);
void launch_zero(double* a, long long n, hipStream_t s);
void launch_integral_blocks(int nblk, const IntBlock* blocks, const double* bpts, const int* bpan,
                            const PanelGeom* panels, int kernel, double kappa, double* out, hipStream_t s);
// Device panel geometry records: out[k] = record of panel ids[k] (panel k if
// ids is null) of the 9-double vertex records verts9, equal to the host's.
void launch_panel_records(const double* verts9, const int* ids, int n, PanelGeom* out, hipStream_t s);
void launch_green_blocks(int nblk, const GreenBlock* blocks, const double* dest, const double* src, int kernel,
                         double kappa, double* out, hipStream_t s);
void launch_gemm_batched(int nbatch, int maxtiles, const GemmDesc* descs, hipStream_t s);
// Pivoted MGS of every problem in `descs` (fills seg/wave0/ss_off);
// ranks[rank_slot] receives each problem's rank.  Synchronous.
void run_mgs(std::vector<MgsDesc>& descs, double* A, double* Q, std::vector<int>& ranks, double eps,
             hipStream_t s);
void launch_f64_to_f32(const double* a, float* b, long long n, hipStream_t s);

}  // namespace fsvd
