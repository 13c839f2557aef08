// Kernel descriptors shared by host orchestration and device code.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.h"

namespace fsvd {

// A rectangular block of point x panel integrals: one workgroup evaluates
// every (point, panel) pair and writes slp at out[off + i*ld_pt + j*ld_pan]
// and dlp at the same place + layer_stride.
struct IntBlock {
   int npt, npan;
   int pt_off;        // into the block-point coordinate array (xyz triples)
   int pan_off;       // into the block-panel index array (caller panel indices)
   long long out_off;
   int ld_pt, ld_pan;
   long long layer_stride;   // double layer at +layer_stride; < 0: write the double layer only, at out_off
};

// Point-kernel block: out[off + i*ld_pt + j*ld_src] = G(dest_i, src_j)
struct GreenBlock {
   int ndest, nsrc;
   int dest_off, src_off;   // into coordinate arrays
   long long out_off;
   int ld_dest, ld_src;
};

// Dense small GEMM C = A * B with arbitrary element strides (one per batch).
struct GemmDesc {
   int M, N, K;
   const double* A;
   const double* B;
   double* C;
   int a_rs, a_cs, b_rs, b_cs, c_rs, c_cs;
};

// Pivoted-MGS basis problem: A (rows x cols, column-major, column j at
// a_off + j*rows, modified in place) -> rank and the first `rank`
// orthonormal pivot columns at q_off (column k at q_off + k*rows).
// seg lanes of a wave share one column; the problem's columns occupy the
// update grid's waves wave0.. and its squared column norms sit at
// ss_off.. (filled by the host scheduler).
struct MgsDesc {
   int rows, cols;
   long long a_off;
   long long q_off;
   int rank_slot;
   int seg, wave0, ss_off, pid;   // pid: index of the problem's MgsState
};

// Per-problem progress of the stepwise MGS.
struct MgsState {
   double fro;
   int rank, done;
};

// Near-field leaf: rows q0..q0+p-1 (tree-order points), m columns listed at
// dcol[col_off..], interleaved column-major block of 2*p*m values at d_off
// (column c: p single-layer then p double-layer entries).
// u0/nu: the leaf's far-field output blocks (LeafU) in the apply plan.
struct LeafDesc {
   int q0, p, m, col_off;
   long long d_off;
   int u0, nu;
};

// Far-field source cube (panels side) at one level.
struct SrcDesc {
   int p0, np, r;
   int spec_slot;
   long long vt_off;   // VT: r x np row-major
   long long pv_off;   // PVs^T then PVd^T: r x gp3 each, row-major
};

// Far-field target cube (points side) at one level.
struct TgtDesc {
   int q0, nq, r;
   int spec_slot;
   int int_off, nint;  // into the interaction arrays (source spectrum slot, T offset index)
   long long u_off;    // U^T: r x nq row-major
   long long uti_off;  // UTI: r x gp3 row-major
};

// ---- apply plan: one launch per stage over every level of a tree ----

// Columns [c0, c0+ncol) of a source cube's row basis, stored panel-major
// (column i holds the r coefficients of panel p0+i) at vt_off; the block
// writes 2r partial projections (layer 0 then layer 1) at out_off.
struct ProjChunk {
   long long vt_off;
   int p0, ncol, r, out_off;
};

// Source cube: n grid points per side, rank r, its chunk partials at
// part_off + j*2r (j < nchunk), PVs then PVd (r x gp3 each, k-major) at
// pv_off, forward spectrum (complex index) at spec_off.
struct SrcPlan {
   int n, r, part_off, nchunk;
   long long pv_off;
   long long spec_off;
};

// Target cube: interactions int_off..+nint (absolute source spectrum
// offsets, T offset codes), its level's T spectra
// at t_off (offset code c at t_off + c*pad), accumulated spectrum at
// acc_off, UTI (r x gp3) at uti_off, r outputs w at w_off.
struct TgtPlan {
   int n, r, pad, nint, int_off, w_off;
   long long t_off, acc_off, uti_off;
};

// FP64 translation tiles (k_translate<double>).  A block handles TR_BINS_F64
// frequency bins, one 64-byte row per spectrum, of up to TR_TILE tree-order
// neighbouring targets, four lanes per target; tiles grow while their
// source union stays within TR_SRC_BUDGET, so the block's LDS ((343 +
// sources) rows) stays under 85 KiB.  TR_MAX_SRC bounds a tile's sources
// (local source row offsets must fit 16 bits).
constexpr int TR_BINS_F64 = 4;
constexpr int TR_LPT_F64 = 4;
constexpr int TR_ROW = 64;
constexpr int TR_TILE = 128;
constexpr int TR_SRC_BUDGET = 800;
constexpr int TR_MAX_SRC = 1023;

// Translation tile: nt <= TR_TILE consecutive (tree-order) target cubes of one
// level, GTgt entries at tgt0, their union of nsrc source spectra
// (absolute offsets at src_off), the level's pad and padded stride pst, its
// T spectra at t_off.  Each target's interactions are packed as byte
// offsets of LDS rows, (T code * TR_ROW) | (local source * TR_ROW) << 16, in
// interaction-list order from a 16-byte aligned int_off; every list is
// followed by at least four readable words (the next list or padding).
struct TrTile {
   int nt, tgt0, nsrc, src_off, pad, pst;
   long long t_off;
};
struct GTgt {
   long long acc_off;
   int int_off, nint;
};

// Spectrum rows in HBM start every TR_PST_ALIGN bins: a level's spectra
// (forward, accumulated, T) are stored at stride pst = pad rounded up to
// that multiple, so every 8-bin (64-byte FP32) row is 64-byte aligned.
constexpr int TR_PST_ALIGN = 8;
__host__ __device__ constexpr int tr_pst(int pad) { return (pad + TR_PST_ALIGN - 1) / TR_PST_ALIGN * TR_PST_ALIGN; }

// Pipelined FP32 translation (k_translate_pipe).  Persistent blocks (one
// per CU) of TP_CW consumer waves and TP_PW producer waves work through
// their list of tasks (tile x 8-bin chunk).  While the consumers sum task k
// from one LDS buffer, the producers stage task k+1 into the other (the T
// rows only when the level or chunk changes); the consumers claim the
// task's octets one at a time from an LDS counter.  A buffer holds TP_TROWS T
// rows (the 343 offsets, rows the staging may fill with copies, and a zero
// row for padding words) and up to TP_MAX_SRC source rows, 64 bytes each.
// A consumer wave sums eight targets of similar list length at once (an
// octet): lane (t, b) = (lane >> 3, lane & 7) takes target t of the octet
// and bin b of the chunk, so every lane owns one output bin.
constexpr int TP_CW = 13, TP_PW = 3, TP_NT = 64 * (TP_CW + TP_PW);
constexpr int TP_TROWS = 356;        // 352 staged rows (343 used), the zero row 352, 3 spare
constexpr int TP_ZERO_ROW = 352;
constexpr int TP_MAX_SRC = 912;
constexpr int TP_BUF = (TP_TROWS + TP_MAX_SRC) * 64;
constexpr int TP_SRC_BUDGET = 880;   // tile growth stops above this union size
constexpr int TP_TILE = 512;         // targets per tile at most
// a tile of more than TP_QUANTUM targets is cut to a multiple of it (whole
// octets)
constexpr int TP_QUANTUM = 8;
constexpr unsigned TP_NULL_WORD = TP_ZERO_ROW * 64u;   // zero T row x source row 0
static_assert(2 * TP_BUF <= 163840, "two pipeline buffers must fit the LDS of a CU");
static_assert(TP_MAX_SRC % 16 == 0 && (TP_MAX_SRC - 1) * 64 < 65536, "source rows and 16-bit offsets");
static_assert(TP_TROWS % 4 == 0, "source rows start on an LDS bank-quarter boundary");

// Tile of the pipelined translation: its level's FP32 T spectra (code c at
// T + c * pst), its tree's forward spectra (source k of the tile at
// spec + gsrc[src_off + k]) and accumulated spectra (octet target
// outputs at acc + oacc[...]), the union size nsrc, the level's pad and
// stride.  Tiles carry their pointers, so one launch can cover the
// translations of several trees.
struct TpTile {
   const float2* T;
   const float2* spec;
   float2* acc;
   int nsrc, src_off, pad, pst;
};
// An octet of the pipelined translation: nk steps (the longest list,
// rounded up to 8); its words from int_off (16-byte aligned) in groups of
// four steps, group i holding target t's words of steps 4i..4i+3 at
// int_off + 4 * (8 i + t) (null words past the end of a list); target t's
// accumulated spectrum at oacc[acc_idx + t] (-1: no target).
struct TpOct {
   int int_off, nk, acc_idx, pad_;
};

// Far-field output block of a leaf: rows of one target cube's U^T
// (r x ld, k-major) starting at the leaf's first point, weights at w_off.
struct LeafU {
   long long u_off;
   int ld, r, w_off, pad_;
};

// Element-wise transpose job: dst (cols x rows, row-major) = src^T.
struct TransDesc {
   const void* src;
   void* dst;
   int rows, cols;
};

// Everything one stage launch of the forward apply needs (device pointers;
// matrix pointers are of the storage type of the launch).
struct ApplyArgs {
   int nchunk, nsrc, ntgt, nleaf;
   int nsrc_big, ntgt_big;  // src / tgt entries with n > 4 come first
   int proj_maxcols, maxr_src, max_n, maxm, maxp;
   const ProjChunk* chunks;
   const SrcPlan* src;
   const TgtPlan* tgt;
   int ntask_d, tr_max_src;          // FP64 translation tasks (tile, 1, first bin)
   const int4* task_d;
   const TrTile* tiles;
   const GTgt* gtgt;
   const int *gsrc, *gint;
   // pipelined FP32 translation: nblk_p persistent blocks, block b takes the
   // tasks (tile, first bin, first TpOct, octets) range_p[b]
   int nblk_p, mode_p;   // mode: 0 staging through registers, 1 direct global->LDS transfers
   const int2* range_p;
   const int4* task_p;
   const TpTile* tiles_p;
   const TpOct* oct_p;
   const long long* oacc_p;
   const int *gsrc_p, *gint_p;
   // its two-column form (column-batched applies)
   int nblk_p2 = 0;
   const int2* range_p2 = nullptr;
   const int4* task_p2 = nullptr;
   const TpTile* tiles_p2 = nullptr;
   const TpOct* oct_p2 = nullptr;
   const long long* oacc_p2 = nullptr;
   const int *gsrc_p2 = nullptr, *gint_p2 = nullptr;
   int nmap_leaf;                    // XCD block map of the leaf launch
   const int* map_leaf;
   const LeafDesc* leaves;
   const int* dcol;
   const LeafU* ublk;
   const void *VTc, *PV, *Tsp, *UTI, *UT, *D;
   void *tpart, *spec, *acc, *w;
   const double *xs, *xd;
   double* y;
   // ncol columns at once: column c of every vector at c times its column
   // stride (in elements of the vector's own type); the pipelined
   // translation's tiles point at the single-column spectra, so its
   // column-c buffers are at byte offsets spec_delta / acc_delta + c times
   // the stride from them
   int ncol = 1;
   long long cs_x = 0, cs_y = 0, cs_tpart = 0, cs_spec = 0, cs_acc = 0, cs_w = 0;
   long long spec_delta = 0, acc_delta = 0;
   // single column: the region split fused into the leaf stage (output
   // sp_y[sp_out[k]] = y_k + sp_corr[k] xd[k] instead of y)
   const int* sp_out = nullptr;
   const double* sp_corr = nullptr;
   double* sp_y = nullptr;
};

// One column of a lockstep GMRES step (device copy, one per active
// column): its basis pointer array (const VT* const*), the next basis
// vector, the work vectors and Givens state, and its initial residual norm.
struct GmDev {
   const void* V;
   void* vnext;
   double *w, *partial_m, *htmp, *hbuf, *pnorm, *H, *cs, *sn, *g, *hist;
   double normr;
};

// Column-batched launches: the ncol copies of virtual block b (one per
// column) run on b's XCD one after another (workgroups are dealt to the
// eight XCDs round-robin by index), so the copies share that XCD's L2 for
// the operator rows they all read.  A grid of col_grid(nb, ncol) blocks;
// copies of virtual blocks b >= nb have no work.
struct ColBlk {
   int b, c;
};
__device__ __forceinline__ ColBlk col_block(int ncol) {
   const int lin = (int)blockIdx.x;
   if (ncol == 1) return {lin, 0};
   const int k = lin >> 3;
   return {(k / ncol) * 8 + (lin & 7), k % ncol};
}
inline int col_grid(int nb, int ncol) { return ncol == 1 ? nb : (nb + 7) / 8 * 8 * ncol; }
// The same grid taken column by column (all blocks of column 0, then of
// column 1, ...), block b still on XCD b % 8: for the persistent
// translation, whose copies share little but the T rows.
__device__ __forceinline__ ColBlk col_block_major(int nb, int ncol) {
   const int lin = (int)blockIdx.x;
   if (ncol == 1) return {lin, 0};
   const int nbp = (nb + 7) / 8 * 8;
   return {lin % nbp, lin / nbp};
}

}  // namespace fsvd
