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

// Translation tiles.  A block handles TR_BINS_F32 (FP32) or TR_BINS_F64
// (FP64) frequency bins, one 64-byte row per spectrum, of up to TR_TILE
// tree-order neighbouring targets; tiles grow while their source union
// stays within TR_SRC_BUDGET, so the block's LDS ((343 + sources) rows)
// stays under 72 KiB and two blocks (32 FP32 waves) share a CU.  Larger
// tiles reuse each staged source row for more targets: on the 100k-panel
// globule (both trees), 32-target tiles ran the FP32 translation in
// 0.30 ms, 64 in 0.23 ms and 128 in 0.20 ms; one tile per block instead of
// three (more blocks, shorter tail) then 0.16 ms.  TR_MAX_SRC bounds a
// single target's sources (local source row offsets must fit 16 bits).
constexpr int TR_BINS_F32 = 8, TR_BINS_F64 = 4;
// lanes per target: FP32 one bin (8 bytes) per lane, FP64 one bin (16 bytes)
constexpr int TR_LPT_F32 = 8, TR_LPT_F64 = 4;
constexpr int TR_ROW = 64;
constexpr int TR_TILE = 128;
constexpr int TR_SRC_BUDGET = 800;
// Tiles per translation block (they share the block's staged T spectra).
constexpr int TR_RUN = 1;
constexpr int TR_MAX_SRC = 1023;

// Translation tile: nt <= TR_TILE consecutive (tree-order) target cubes of one
// level, GTgt entries at tgt0, their union of nsrc source spectra
// (absolute offsets at src_off).  Each target's interactions are packed as
// byte offsets of LDS rows, (T code * TR_ROW) | (local source * TR_ROW) << 16,
// in a fixed order chosen per half-wave of four targets (their rows of one
// step spread over the LDS bank quarters), from a 16-byte aligned int_off; every list is
// followed by at least four readable words (the next list or padding).
struct TrTile {
   int nt, tgt0, nsrc, src_off, pad, pad_;
   long long t_off;
};
struct GTgt {
   long long acc_off;
   int int_off, nint;
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
   int ntask_f, ntask_d, tr_max_src;  // translation tasks (first tile, tiles, first bin), FP32 / FP64
   const int4 *task_f, *task_d;
   const TrTile* tiles;
   const GTgt* gtgt;
   const int *gsrc, *gint;
   int nmap_leaf;                    // XCD block map of the leaf launch
   const int* map_leaf;
   const LeafDesc* leaves;
   const int* dcol;
   const LeafU* ublk;
   const void *VTc, *PV, *Tsp, *UTI, *UT, *D;
   void *tpart, *spec, *acc, *w;
   const double *xs, *xd;
   double* y;
};

}  // namespace fsvd
