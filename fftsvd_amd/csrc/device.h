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
   long long layer_stride;
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

// Pivoted-MGS basis problem: A (rows x cols, row-major, modified in place)
// -> rank and the first `rank` orthonormal pivot columns at q_off
// (column-major, column k at q_off + k*rows).
struct MgsDesc {
   int rows, cols;
   long long a_off;
   long long q_off;
   int rank_slot;
};

// Near-field leaf: rows q0..q0+p-1 (tree-order points), m columns listed at
// dcol[col_off..], interleaved column-major block of 2*p*m values at d_off
// (column c: p single-layer then p double-layer entries).
struct LeafDesc {
   int q0, p, m, col_off;
   long long d_off;
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

}  // namespace fsvd
