// Matrix-vector product of one compressed FFT-SVD operator on the GPU.
//
//   y = D_s xs + D_d xd  (near field, dense per-leaf blocks)
//     + U UTI IFFT( sum_c T[d-c] . FFT(PV_s VT xs_c + PV_d VT xd_c) )  (far field)
//
// Storage precision T is float (mixed mode) or double (FP64 mode); vectors
// enter and leave in FP64.  The forward apply is five launches per tree
// (k_project, k_source, k_translate, k_target, k_leaf; each covers every
// level).  The transposed operator (k_farT_*, k_far_translate, k_nearT)
// runs level by level with the roles of the two sides swapped.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <type_traits>

#include "device.h"

namespace fsvd {

template <typename T>
struct Cplx;
template <>
struct Cplx<float> {
   using type = float2;
};
template <>
struct Cplx<double> {
   using type = double2;
};

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
   return v;
}

// Marks a loaded value as used here, so the compiler cannot sink its load
// into a later conditional block (where it would wait for it alone).
template <typename T>
__device__ __forceinline__ void keep(const T& v) {
   asm volatile("" ::"v"(v));
}

// Workgroups are dealt round-robin to the 8 XCDs.  Launches whose blocks
// share data across tree-order neighbours use a host-built block map
// (xcd_block_map): block b runs logical block map[b], and each XCD gets a
// contiguous tree-order range of about equal work, so neighbouring work
// shares one L2.  -1 marks padding blocks.
// The threads that share one cube's work: a whole block (__syncthreads
// between phases) or one wavefront.  A wavefront's LDS accesses are
// processed in issue order, so between its phases only the compiler's
// reordering has to be fenced.
struct BlockGroup {
   int tid, nt;
   __device__ void sync() const { __syncthreads(); }
};
struct WaveGroup {
   int tid;
   static constexpr int nt = 64;
   __device__ void sync() const {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
   }
};
__device__ __forceinline__ BlockGroup block_group() { return BlockGroup{(int)threadIdx.x, (int)blockDim.x}; }

// e^{-2 pi i m / N} table
template <typename T, int N, typename G = BlockGroup>
__device__ __forceinline__ void twiddle_table(T* cs, T* sn, const G& gr = block_group()) {
   for (int m = gr.tid; m < N; m += gr.nt) {
      double s, c;
      sincospi(2.0 * m / N, &s, &c);
      cs[m] = (T)c;
      sn[m] = (T)s;
   }
}

// Forward transform of an n^3 grid zero-padded at offset n/2 into an
// N^3 box (N = 2n-1), real-to-complex with the half last dimension
// (reference FFT_forwardGridTransform, FFT.c:66-113).  g: n^3 in LDS,
// out: N*N*H complex in global memory.  Work arrays A (n*n*H) and
// B (n*N*H) in LDS.
template <int n, typename T, typename G = BlockGroup>
__device__ void grid_forward(const T* g, typename Cplx<T>::type* A, typename Cplx<T>::type* B,
                             typename Cplx<T>::type* out, const T* cs, const T* sn, const G& gr = block_group()) {
   constexpr int N = 2 * n - 1, H = N / 2 + 1, OFF = n / 2;
   const int tid = gr.tid, nt = gr.nt;
   // along k
   for (int e = tid; e < n * n * H; e += nt) {
      int c = e % H, ij = e / H;
      T re = 0, im = 0;
#pragma unroll
      for (int k = 0; k < n; k++) {
         int m = (c * (k + OFF)) % N;
         T v = g[ij * n + k];
         re += v * cs[m];
         im -= v * sn[m];
      }
      A[e].x = re;
      A[e].y = im;
   }
   gr.sync();
   // along j
   for (int e = tid; e < n * N * H; e += nt) {
      int c = e % H, b = (e / H) % N, i = e / (H * N);
      T re = 0, im = 0;
#pragma unroll
      for (int j = 0; j < n; j++) {
         int m = (b * (j + OFF)) % N;
         auto a = A[(i * n + j) * H + c];
         re += a.x * cs[m] + a.y * sn[m];
         im += a.y * cs[m] - a.x * sn[m];
      }
      B[e].x = re;
      B[e].y = im;
   }
   gr.sync();
   // along i
   for (int e = tid; e < N * N * H; e += nt) {
      int c = e % H, b = (e / H) % N, a = e / (H * N);
      T re = 0, im = 0;
#pragma unroll
      for (int i = 0; i < n; i++) {
         int m = (a * (i + OFF)) % N;
         auto v = B[(i * N + b) * H + c];
         re += v.x * cs[m] + v.y * sn[m];
         im += v.y * cs[m] - v.x * sn[m];
      }
      out[e].x = re;
      out[e].y = im;
   }
}

// Inverse transform of an N*N*H half spectrum, cropped back to the n^3
// grid at offset n/2 and scaled by 1/N^3 (reference
// FFT_backwardGridTransform, FFT.c:117-150).  S in LDS (N*N*H),
// C1 (n*N*H), C2 (n*n*H) in LDS, h (n^3) in LDS.
template <int n, typename T, typename G = BlockGroup>
__device__ void grid_backward(const typename Cplx<T>::type* S, typename Cplx<T>::type* C1,
                              typename Cplx<T>::type* C2, T* h, const T* cs, const T* sn,
                              const G& gr = block_group()) {
   constexpr int N = 2 * n - 1, H = N / 2 + 1, OFF = n / 2;
   const int tid = gr.tid, nt = gr.nt;
   for (int e = tid; e < n * N * H; e += nt) {
      int c = e % H, b = (e / H) % N, i = e / (H * N);
      T re = 0, im = 0;
#pragma unroll
      for (int a = 0; a < N; a++) {
         int m = (a * (i + OFF)) % N;
         auto v = S[(a * N + b) * H + c];
         re += v.x * cs[m] - v.y * sn[m];
         im += v.y * cs[m] + v.x * sn[m];
      }
      C1[e].x = re;
      C1[e].y = im;
   }
   gr.sync();
   for (int e = tid; e < n * n * H; e += nt) {
      int c = e % H, j = (e / H) % n, i = e / (H * n);
      T re = 0, im = 0;
#pragma unroll
      for (int b = 0; b < N; b++) {
         int m = (b * (j + OFF)) % N;
         auto v = C1[(i * N + b) * H + c];
         re += v.x * cs[m] - v.y * sn[m];
         im += v.y * cs[m] + v.x * sn[m];
      }
      C2[e].x = re;
      C2[e].y = im;
   }
   gr.sync();
   const T scale = (T)(1.0 / ((double)N * N * N));
   for (int e = tid; e < n * n * n; e += nt) {
      int k = e % n, ij = e / n;
      const auto* row = C2 + ij * H;
      T v = row[0].x;
#pragma unroll
      for (int c = 1; c < H; c++) {
         int m = (c * (k + OFF)) % N;
         v += (T)2 * (row[c].x * cs[m] - row[c].y * sn[m]);
      }
      h[e] = v * scale;
   }
}

template <int n>
struct FarSizes {
   static constexpr int N = 2 * n - 1, H = N / 2 + 1, PAD = N * N * H, GP3 = n * n * n;
};

// Translation: S_d[b] = sum_e T[toff_e][b] * G[src_e][b].  Blocks
// (target cube, slice of gridDim.y over the bins), threads over bins; each
// bin's sum is the same whatever the slicing.
template <typename T>
__global__ void __launch_bounds__(256) k_far_translate(const TgtDesc* __restrict__ tgt,
                                                       const int* __restrict__ int_src,
                                                       const int* __restrict__ int_toff,
                                                       const typename Cplx<T>::type* __restrict__ Tspec,
                                                       const typename Cplx<T>::type* __restrict__ spec,
                                                       typename Cplx<T>::type* __restrict__ acc, int pad) {
   const TgtDesc d = tgt[blockIdx.x];
   for (int b = threadIdx.x + (int)(blockIdx.y * blockDim.x); b < pad; b += (int)(blockDim.x * gridDim.y)) {
      T re = 0, im = 0;
      for (int e = 0; e < d.nint; e++) {
         auto t = Tspec[(long long)int_toff[d.int_off + e] * pad + b];
         auto s = spec[(long long)int_src[d.int_off + e] * pad + b];
         re += t.x * s.x - t.y * s.y;
         im += t.x * s.y + t.y * s.x;
      }
      acc[(long long)d.spec_slot * pad + b] = {re, im};
   }
}

// ---------------------------------------------------------------------------
// Forward apply, one launch per stage over all levels of a tree:
//   k_project    chunk of a source cube's panels: partial VT.xs, VT.xd
//   k_source     source cube: sum of chunk partials, PV interpolation onto
//                the n^3 grid, zero-padded r2c DFT -> spectrum
//   k_translate  tile of <= TR_TILE target cubes x a bin chunk: sum over each
//                interaction list of T . spectrum, from LDS
//   k_target     target cube: cropped c2r DFT, w = UTI h
//   k_leaf       leaf: near-field block apply plus U w of every target cube
//                containing the leaf, writes y
// Every sum runs in a fixed order: the result does not depend on scheduling.

// Operator entries read once per launch are loaded nontemporal (NT), so
// they do not evict reusable lines from L2.  The column-batched launches
// read each block once per column copy and load it normally, so that the
// copies after the first hit L2.
template <bool NT, typename T>
__device__ __forceinline__ T ld_op(const T* p) {
   if constexpr (NT)
      return __builtin_nontemporal_load(p);
   else
      return *p;
}

// Rows e of a column-major E x m block (column c at blk + c*E), streamed
// with unit stride by G = 256/Ec column groups per row chunk; returns the
// thread's partial sum for (row e, layer inputs x0/x1 selected by e >= split).
template <typename T, bool NT>
__device__ __forceinline__ T stream_rows(const T* __restrict__ colp, int E, int m, int g, int G, const T* xl) {
   // eight columns in flight per thread
   const long long step = (long long)G * E;
   T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
   int c = g;
   for (; c + 7 * G < m; c += 8 * G) {
      const T* pc = colp + (long long)c * E;
      T d[8];
#pragma unroll
      for (int u = 0; u < 8; u++) d[u] = ld_op<NT>(pc + u * step);
      a0 += d[0] * xl[c];
      a1 += d[1] * xl[c + G];
      a2 += d[2] * xl[c + 2 * G];
      a3 += d[3] * xl[c + 3 * G];
      a0 += d[4] * xl[c + 4 * G];
      a1 += d[5] * xl[c + 5 * G];
      a2 += d[6] * xl[c + 6 * G];
      a3 += d[7] * xl[c + 7 * G];
   }
   if (c < m) {
      // the last partial batch, all of its loads in flight together
      const T* pc = colp + (long long)c * E;
      T d[7];
#pragma unroll
      for (int u = 0; u < 7; u++) d[u] = c + u * G < m ? ld_op<NT>(pc + u * step) : (T)0;
#pragma unroll
      for (int u = 0; u < 7; u++)
         if (c + u * G < m) a0 += d[u] * xl[c + u * G];
   }
   return (a0 + a1) + (a2 + a3);
}

// Projection: one chunk of a source cube's panels per wavefront (four per
// block), tpart[out_off + e] = sum_i VT[e, i] xs[i] and
// tpart[out_off + r + e] = sum_i VT[e, i] xd[i] over the chunk's columns.
// Lane (g, e) owns row e and the columns g, g + G, ... (G = 64 / rows of the
// row chunk); the G partial sums of a row are added in group order.
template <typename T, bool NT>
__global__ void __launch_bounds__(256) k_project(const ProjChunk* __restrict__ chunks, int nchunk,
                                                 const T* __restrict__ VTc, const double* __restrict__ xs,
                                                 const double* __restrict__ xd, T* __restrict__ tpart, int wave_lds,
                                                 int ncol, long long cs_x, long long cs_tpart) {
   extern __shared__ __align__(16) unsigned char smem[];
   const ColBlk cb = col_block(ncol);
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, ci = cb.b * 4 + wave;
   if (ci >= nchunk) return;
   xs += cb.c * cs_x;
   xd += cb.c * cs_x;
   tpart += cb.c * cs_tpart;
   const WaveGroup gr{lane};
   const ProjChunk c = chunks[ci];
   const int m = c.ncol, E = c.r;
   T* part = (T*)(smem + (size_t)wave * wave_lds);   // 2*64
   T* xl = part + 128;                                // 2*m
   // x staged four columns per lane at a time, loads before stores
   for (int i0 = lane; i0 < m; i0 += 256) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
         const int i = i0 + 64 * u;
         a[u] = i < m ? xs[c.p0 + i] : 0.0;
         b[u] = i < m ? xd[c.p0 + i] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
         const int i = i0 + 64 * u;
         if (i < m) {
            xl[i] = (T)a[u];
            xl[m + i] = (T)b[u];
         }
      }
   }
   gr.sync();
   const T* blk = VTc + c.vt_off;
   for (int e0 = 0; e0 < E; e0 += 64) {
      const int Ec = min(64, E - e0), G = 64 / Ec;
      const int g = lane / Ec, e = e0 + lane - g * Ec;
      T as = 0, ad = 0;
      if (g < G) {
         const T* colp = blk + e;
         const long long step = (long long)G * E;
         T s0 = 0, s1 = 0, d0 = 0, d1 = 0;
         int i = g;
         // eight columns in flight per lane
         for (; i + 7 * G < m; i += 8 * G) {
            T v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = ld_op<NT>(colp + (long long)i * E + u * step);
#pragma unroll
            for (int u = 0; u < 8; u += 2) {
               s0 += v[u] * xl[i + u * G];
               d0 += v[u] * xl[m + i + u * G];
               s1 += v[u + 1] * xl[i + (u + 1) * G];
               d1 += v[u + 1] * xl[m + i + (u + 1) * G];
            }
         }
         if (i < m) {
            // the last partial batch, all of its loads in flight together
            T v[7];
#pragma unroll
            for (int u = 0; u < 7; u++)
               v[u] = i + u * G < m ? ld_op<NT>(colp + (long long)i * E + u * step) : (T)0;
#pragma unroll
            for (int u = 0; u < 7; u++)
               if (i + u * G < m) {
                  s0 += v[u] * xl[i + u * G];
                  d0 += v[u] * xl[m + i + u * G];
               }
         }
         as = s0 + s1;
         ad = d0 + d1;
      }
      part[lane] = as;
      part[64 + lane] = ad;
      gr.sync();
      if (lane < Ec) {
         T ss = 0, sd = 0;
         for (int gg = 0; gg < G; gg++) {
            ss += part[gg * Ec + lane];
            sd += part[64 + gg * Ec + lane];
         }
         tpart[c.out_off + e0 + lane] = ss;
         tpart[c.out_off + E + e0 + lane] = sd;
      }
      gr.sync();
   }
}

// k_project for NC columns per wave (column-batched applies): each VT entry
// is loaded once and multiplied into the NC columns' accumulators, each in
// k_project's order (inputs interleaved, input i of column j at
// xl[i * NC + j]), so every column's partials are k_project's bit for bit.
// A partial last group repeats its last column and stores only its own.
// columns per wave: four FP32, two FP64 (whose four would take the whole
// LDS of a CU for the widest chunk)
template <typename T>
constexpr int proj_nc() { return sizeof(T) == 4 ? 4 : 2; }
template <typename T, int NC>
__global__ void __launch_bounds__(256) k_project_mc(const ProjChunk* __restrict__ chunks, int nchunk,
                                                    const T* __restrict__ VTc, const double* __restrict__ xs,
                                                    const double* __restrict__ xd, T* __restrict__ tpart,
                                                    int wave_lds, int ncol, long long cs_x, long long cs_tpart) {
   extern __shared__ __align__(16) unsigned char smem[];
   const int ngrp = (ncol + NC - 1) / NC;
   const ColBlk cb = col_block(ngrp);
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, ci = cb.b * 4 + wave;
   if (ci >= nchunk) return;
   const int c0 = cb.c * NC, nc = min(NC, ncol - c0);
   const WaveGroup gr{lane};
   const ProjChunk c = chunks[ci];
   const int m = c.ncol, E = c.r;
   T* part = (T*)(smem + (size_t)wave * wave_lds);   // 2 * 64 * NC
   T* xl = part + 128 * NC;                           // 2 * m * NC
   for (int i = lane; i < m; i += 64) {
#pragma unroll
      for (int j = 0; j < NC; j++) {
         const long long o = (long long)(c0 + min(j, nc - 1)) * cs_x + c.p0 + i;
         xl[(size_t)i * NC + j] = (T)xs[o];
         xl[(size_t)(m + i) * NC + j] = (T)xd[o];
      }
   }
   gr.sync();
   const T* blk = VTc + c.vt_off;
   for (int e0 = 0; e0 < E; e0 += 64) {
      const int Ec = min(64, E - e0), G = 64 / Ec;
      const int g = lane / Ec, e = e0 + lane - g * Ec;
      T as[NC], ad[NC];
#pragma unroll
      for (int j = 0; j < NC; j++) as[j] = ad[j] = 0;
      if (g < G) {
         const T* colp = blk + e;
         const long long step = (long long)G * E;
         T s0[NC], s1[NC], d0[NC], d1[NC];
#pragma unroll
         for (int j = 0; j < NC; j++) s0[j] = s1[j] = d0[j] = d1[j] = 0;
         int i = g;
         for (; i + 7 * G < m; i += 8 * G) {
            T v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = colp[(long long)i * E + u * step];
#pragma unroll
            for (int u = 0; u < 8; u += 2) {
               const T* xa = xl + (size_t)(i + u * G) * NC;
               const T* xb = xl + (size_t)(i + (u + 1) * G) * NC;
               const T* ya = xl + (size_t)(m + i + u * G) * NC;
               const T* yb = xl + (size_t)(m + i + (u + 1) * G) * NC;
#pragma unroll
               for (int j = 0; j < NC; j++) {
                  s0[j] += v[u] * xa[j];
                  d0[j] += v[u] * ya[j];
                  s1[j] += v[u + 1] * xb[j];
                  d1[j] += v[u + 1] * yb[j];
               }
            }
         }
         if (i < m) {
            T v[7];
#pragma unroll
            for (int u = 0; u < 7; u++) v[u] = i + u * G < m ? colp[(long long)i * E + u * step] : (T)0;
#pragma unroll
            for (int u = 0; u < 7; u++)
               if (i + u * G < m) {
                  const T* xa = xl + (size_t)(i + u * G) * NC;
                  const T* ya = xl + (size_t)(m + i + u * G) * NC;
#pragma unroll
                  for (int j = 0; j < NC; j++) {
                     s0[j] += v[u] * xa[j];
                     d0[j] += v[u] * ya[j];
                  }
               }
         }
#pragma unroll
         for (int j = 0; j < NC; j++) {
            as[j] = s0[j] + s1[j];
            ad[j] = d0[j] + d1[j];
         }
      }
#pragma unroll
      for (int j = 0; j < NC; j++) {
         part[(2 * j) * 64 + lane] = as[j];
         part[(2 * j + 1) * 64 + lane] = ad[j];
      }
      gr.sync();
      if (lane < Ec) {
         for (int j = 0; j < nc; j++) {
            T ss = 0, sd = 0;
            for (int gg = 0; gg < G; gg++) {
               ss += part[(2 * j) * 64 + gg * Ec + lane];
               sd += part[(2 * j + 1) * 64 + gg * Ec + lane];
            }
            T* tp = tpart + (long long)(c0 + j) * cs_tpart;
            tp[c.out_off + e0 + lane] = ss;
            tp[c.out_off + E + e0 + lane] = sd;
         }
      }
      gr.sync();
   }
}

template <int n, typename T, bool NT, typename G>
__device__ void source_body(const SrcPlan& d, const T* __restrict__ tpart, const T* __restrict__ PV,
                            typename Cplx<T>::type* __restrict__ spec, unsigned char* smem, const G& gr) {
   using C2 = typename Cplx<T>::type;
   constexpr int N = FarSizes<n>::N, H = FarSizes<n>::H, GP3 = FarSizes<n>::GP3;
   C2* A = (C2*)smem;            // n*n*H
   C2* B = A + n * n * H;        // n*N*H
   T* cs = (T*)(B + n * N * H);  // N
   T* sn = cs + N;               // N
   T* g = sn + N;                // GP3
   T* part = g + GP3;            // partial grid values (block groups)
   const int tid = gr.tid, nt = gr.nt, r = d.r;
   T* t = part + (nt > GP3 ? nt : 0);   // 2r: ts then td
   twiddle_table<T, N>(cs, sn, gr);
   // sum of the chunk partials in chunk order, eight loads in flight
   for (int k = tid; k < 2 * r; k += nt) {
      const T* pp = tpart + d.part_off + k;
      T a = 0;
      int j = 0;
      for (; j + 7 < d.nchunk; j += 8) {
         T v[8];
#pragma unroll
         for (int u = 0; u < 8; u++) v[u] = pp[(long long)(j + u) * 2 * r];
#pragma unroll
         for (int u = 0; u < 8; u++) a += v[u];
      }
      if (j < d.nchunk) {
         T v[7];
#pragma unroll
         for (int u = 0; u < 7; u++) v[u] = j + u < d.nchunk ? pp[(long long)(j + u) * 2 * r] : (T)0;
#pragma unroll
         for (int u = 0; u < 7; u++)
            if (j + u < d.nchunk) a += v[u];
      }
      t[k] = a;
   }
   gr.sync();
   // g = PV t over the 2r rows (PVs then PVd, k-major): thread (kg, q) sums
   // the rows k = kg, kg + KG, ... so every thread streams PV
   const T* pv = PV + d.pv_off;
   const int R2 = 2 * r;
   const int KG = nt >= GP3 ? nt / GP3 : 1, QW = nt / KG;
   for (int q0 = 0; q0 < GP3; q0 += QW) {
      const int kg = tid / QW, q = q0 + tid - kg * QW;
      T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
      if (q < GP3) {
         // eight rows in flight per thread
         int k = kg;
         for (; k + 7 * KG < R2; k += 8 * KG) {
            T v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = ld_op<NT>(pv + (long long)(k + u * KG) * GP3 + q);
            a0 += v[0] * t[k];
            a1 += v[1] * t[k + KG];
            a2 += v[2] * t[k + 2 * KG];
            a3 += v[3] * t[k + 3 * KG];
            a0 += v[4] * t[k + 4 * KG];
            a1 += v[5] * t[k + 5 * KG];
            a2 += v[6] * t[k + 6 * KG];
            a3 += v[7] * t[k + 7 * KG];
         }
         if (k < R2) {
            // the last partial batch, all of its loads in flight together
            T v[7];
#pragma unroll
            for (int u = 0; u < 7; u++)
               v[u] = k + u * KG < R2 ? ld_op<NT>(pv + (long long)(k + u * KG) * GP3 + q) : (T)0;
#pragma unroll
            for (int u = 0; u < 7; u++)
               if (k + u * KG < R2) a0 += v[u] * t[k + u * KG];
         }
      }
      const T sum = (a0 + a1) + (a2 + a3);
      if (KG == 1) {
         if (q < GP3) g[q] = sum;
      } else {
         part[kg * QW + (q - q0)] = sum;
         gr.sync();
         if (tid < QW && q0 + tid < GP3) {
            T acc = 0;
            for (int gg = 0; gg < KG; gg++) acc += part[gg * QW + tid];
            g[q0 + tid] = acc;
         }
      }
      gr.sync();
   }
   grid_forward<n, T>(g, A, B, spec + d.spec_off, cs, sn, gr);
}

// LDS of one source cube's work (block groups of nt threads; a wave group
// of 64 lanes uses no partial-sum rows when GP3 = 64)
template <typename T>
__host__ __device__ constexpr size_t source_cube_lds(int n, int maxr, int nt) {
   return sizeof(typename Cplx<T>::type) * (size_t)(n * n * ((2 * n - 1) / 2 + 1) + n * (2 * n - 1) * ((2 * n - 1) / 2 + 1)) +
          sizeof(T) * (size_t)(2 * (2 * n - 1) + n * n * n + (nt > n * n * n ? nt : 0) + 2 * maxr) + 64;
}

// Source cubes, one launch: the nbig cubes with other than four grid points
// per side first, one block of CUBE_NT threads each (these cubes are few,
// so each gets many loads in flight), then the cubes of four grid points
// per side (n^3 = 64, one per lane), one per wavefront, CUBE_NT/64 per
// block, each wave in its own LDS slice of wave_lds bytes.  The small cubes
// run while the large ones finish.
constexpr int CUBE_NT = 512;
template <typename T, bool NT>
__global__ void __launch_bounds__(CUBE_NT) k_source(const SrcPlan* __restrict__ src, int nbig, int nsrc,
                                                    const T* __restrict__ tpart, const T* __restrict__ PV,
                                                    typename Cplx<T>::type* __restrict__ spec, int wave_lds, int nb,
                                                    int ncol, long long cs_tpart, long long cs_spec) {
   extern __shared__ __align__(16) unsigned char smem[];
   const ColBlk cb = col_block(ncol);
   if (cb.b >= nb) return;
   tpart += cb.c * cs_tpart;
   spec += cb.c * cs_spec;
   if (cb.b < nbig) {
      const SrcPlan d = src[cb.b];
      const BlockGroup gr = block_group();
      switch (d.n) {
         case 3: source_body<3, T, NT>(d, tpart, PV, spec, smem, gr); break;
         case 5: source_body<5, T, NT>(d, tpart, PV, spec, smem, gr); break;
         case 6: source_body<6, T, NT>(d, tpart, PV, spec, smem, gr); break;
         case 7: source_body<7, T, NT>(d, tpart, PV, spec, smem, gr); break;
         case 8: source_body<8, T, NT>(d, tpart, PV, spec, smem, gr); break;
         case 9: source_body<9, T, NT>(d, tpart, PV, spec, smem, gr); break;
         default: break;
      }
      return;
   }
   const int wave = threadIdx.x >> 6, i = nbig + (cb.b - nbig) * (CUBE_NT / 64) + wave;
   if (i >= nsrc) return;
   const SrcPlan d = src[i];
   source_body<4, T, NT>(d, tpart, PV, spec, smem + (size_t)wave * wave_lds, WaveGroup{(int)(threadIdx.x & 63)});
}

// FP64 translation by tiles: a block owns BC frequency bins (one TR_ROW-byte
// row per spectrum) of one tile (up to TR_TILE neighbouring target cubes of
// one level).  It stages the 343 T rows of those bins and the rows of every
// source spectrum the tile's targets interact with in LDS; LPT lanes share a
// target, each owning TR_ROW / LPT bytes of the row, and sum its interaction
// list from LDS in list order; the packed interaction word holds the two
// row byte offsets, so an address is one AND or shift away.
template <typename T, int LPT>
struct TrVec;
template <>
struct TrVec<double, 4> {
   using type = double2;  // bin l
   static constexpr int NB = 1;
};

template <typename T, int LPT>
__device__ __forceinline__ void tr_mac(const typename TrVec<T, LPT>::type& t, const typename TrVec<T, LPT>::type& v,
                                       T* re, T* im) {
   // (t * v) accumulated per bin as four FMAs
   re[0] = fma(t.x, v.x, re[0]);
   re[0] = fma(-t.y, v.y, re[0]);
   im[0] = fma(t.x, v.y, im[0]);
   im[0] = fma(t.y, v.x, im[0]);
   if constexpr (TrVec<T, LPT>::NB == 2) {
      re[1] = fma(t.z, v.z, re[1]);
      re[1] = fma(-t.w, v.w, re[1]);
      im[1] = fma(t.z, v.w, im[1]);
      im[1] = fma(t.w, v.z, im[1]);
   }
}

template <typename T, int BC, int LPT>
__global__ void __launch_bounds__(TR_TILE * LPT) k_translate(const int4* __restrict__ tasks,
                                                           const TrTile* __restrict__ tiles,
                                                           const GTgt* __restrict__ gtgt,
                                                           const int* __restrict__ gsrc, const int* __restrict__ gint,
                                                           const typename Cplx<T>::type* __restrict__ Tspec,
                                                           const typename Cplx<T>::type* __restrict__ spec,
                                                           typename Cplx<T>::type* __restrict__ acc, int nb, int ncol,
                                                           long long cs_spec, long long cs_acc) {
   using C2 = typename Cplx<T>::type;
   using V = typename TrVec<T, LPT>::type;
   constexpr int NB = TrVec<T, LPT>::NB;
   static_assert(BC * sizeof(C2) == TR_ROW && NB * sizeof(C2) * LPT == TR_ROW, "translation row layout");
   extern __shared__ __align__(16) unsigned char smem[];
   const ColBlk cb = col_block(ncol);
   if (cb.b >= nb) return;   // uniform over the block: no barrier is reached
   spec += cb.c * cs_spec;
   acc += cb.c * cs_acc;
   const int4 tk = tasks[cb.b];   // first tile, tile count, first bin
   const int b0 = tk.z, tid = threadIdx.x, nthr = TR_TILE * LPT;
   const int pad = tiles[tk.x].pad, pst = tiles[tk.x].pst;
   const long long t_off = tiles[tk.x].t_off;
   C2* Ts = (C2*)smem;        // 343 rows
   C2* Ss = Ts + 343 * BC;    // source rows
   // staging: eight loads in flight per thread before their LDS stores
   constexpr int U = 8;
   for (int i0 = tid; i0 < 343 * BC; i0 += U * nthr) {
      C2 v[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
         const int i = i0 + u * nthr, c = i / BC, b = i - c * BC;
         v[u].x = 0;
         v[u].y = 0;
         if (i < 343 * BC && b0 + b < pad) v[u] = Tspec[t_off + (long long)c * pst + b0 + b];
      }
#pragma unroll
      for (int u = 0; u < U; u++)
         if (i0 + u * nthr < 343 * BC) Ts[i0 + u * nthr] = v[u];
   }
   const int j = tid / LPT, l = tid % LPT;
   const int bb = l * NB;   // lane's first bin in the row
   // lane's bytes of T row 0 and source row 0
   const unsigned char* Tl = smem + l * (TR_ROW / LPT);
   const unsigned char* Sl = Tl + 343 * TR_ROW;
   for (int ti = tk.x; ti < tk.x + tk.y; ti++) {
      const TrTile tl = tiles[ti];
      __syncthreads();   // the previous tile's sums are done with Ss
      const int ns = tl.nsrc * BC;
      for (int i0 = tid; i0 < ns; i0 += U * nthr) {
         int so[U];
#pragma unroll
         for (int u = 0; u < U; u++) {
            const int i = i0 + u * nthr;
            so[u] = i < ns ? gsrc[tl.src_off + i / BC] : 0;
         }
         C2 v[U];
#pragma unroll
         for (int u = 0; u < U; u++) {
            const int i = i0 + u * nthr, b = i - (i / BC) * BC;
            v[u].x = 0;
            v[u].y = 0;
            if (i < ns && b0 + b < pad) v[u] = spec[(long long)so[u] + b0 + b];
         }
#pragma unroll
         for (int u = 0; u < U; u++)
            if (i0 + u * nthr < ns) Ss[i0 + u * nthr] = v[u];
      }
      __syncthreads();
      if (j < tl.nt && b0 + bb < pad) {
         const GTgt tt = gtgt[tl.tgt0 + j];
         const int* it = gint + tt.int_off;
         T re0[NB] = {}, im0[NB] = {}, re1[NB] = {}, im1[NB] = {};
         int e = 0;
         // the next four interaction words load while the current four are
         // summed (a list is followed by at least four readable words)
         int4 q = *(const int4*)it;
         for (; e + 3 < tt.nint; e += 4) {
            int4 qn = *(const int4*)(it + e + 4);
            const unsigned u0 = (unsigned)q.x, u1 = (unsigned)q.y, u2 = (unsigned)q.z, u3 = (unsigned)q.w;
            const V t0 = *(const V*)(Tl + (u0 & 0xffffu)), s0 = *(const V*)(Sl + (u0 >> 16));
            const V t1 = *(const V*)(Tl + (u1 & 0xffffu)), s1 = *(const V*)(Sl + (u1 >> 16));
            const V t2 = *(const V*)(Tl + (u2 & 0xffffu)), s2 = *(const V*)(Sl + (u2 >> 16));
            const V t3 = *(const V*)(Tl + (u3 & 0xffffu)), s3 = *(const V*)(Sl + (u3 >> 16));
            tr_mac<T, LPT>(t0, s0, re0, im0);
            tr_mac<T, LPT>(t1, s1, re1, im1);
            tr_mac<T, LPT>(t2, s2, re0, im0);
            tr_mac<T, LPT>(t3, s3, re1, im1);
            // consumed here, after the sums: keeps the read-ahead load from
            // being moved to the top of the next iteration
            asm volatile("" : "+v"(qn.x), "+v"(qn.y), "+v"(qn.z), "+v"(qn.w));
            q = qn;
         }
         for (; e < tt.nint; e++) {
            const unsigned u0 = (unsigned)it[e];
            const V t0 = *(const V*)(Tl + (u0 & 0xffffu)), s0 = *(const V*)(Sl + (u0 >> 16));
            tr_mac<T, LPT>(t0, s0, re0, im0);
         }
#pragma unroll
         for (int k = 0; k < NB; k++) {
            if (b0 + bb + k < pad) {
               C2 out;
               out.x = re0[k] + re1[k];
               out.y = im0[k] + im1[k];
               acc[tt.acc_off + b0 + bb + k] = out;
            }
         }
      }
   }
}

// ---------------------------------------------------------------------------
// Pipelined FP32 translation.  See TpTile / TpOct in device.h for the layout.
// Producer waves stage task tk (tile tl, nsrc >= 1) into the LDS buffer
// buf: the tile's source rows, and the T rows of the task's level and chunk
// when stage_t.  The rows are 64 bytes (8 bins), moved in 16-byte pieces:
// producer lane pl takes pieces pl, pl + 64 TP_PW, ...  Loads are issued
// in batches before their stores (addresses clamped to valid pieces,
// stores only for real ones); the source row offsets are loaded first,
// the T rows moved last.
constexpr int TP_PL = 64 * TP_PW;                            // producer lanes
constexpr int TP_SPIECES = (TP_MAX_SRC * 4 + TP_PL - 1) / TP_PL;   // source pieces per lane
constexpr int TP_TPIECES = (343 * 4 + TP_PL - 1) / TP_PL;          // T pieces per lane
constexpr int TP_BATCH = 10;
__device__ __forceinline__ void tp_stage(const int4& tk, const TpTile& tl, bool stage_t, unsigned char* buf, int pl,
                                         const int* __restrict__ gsrc) {
   const int b0 = tk.y;
   const int npiece = tl.nsrc * 4;
   // byte offsets of the pieces in the spectrum array (below 4 GiB, checked
   // when the plan is built), so each load is one 32-bit register offset
   unsigned so[TP_SPIECES];
#pragma unroll
   for (int u = 0; u < TP_SPIECES; u++) {
      const int p = pl + TP_PL * u;
      so[u] = 8u * (unsigned)(gsrc[tl.src_off + min(p >> 2, tl.nsrc - 1)] + b0 + 2 * (p & 3));
   }
   unsigned char* sbuf = buf + TP_TROWS * 64;
#pragma unroll
   for (int u0 = 0; u0 < TP_SPIECES; u0 += TP_BATCH) {
      float4 v[TP_BATCH];
#pragma unroll
      for (int u = 0; u < TP_BATCH; u++)
         if (u0 + u < TP_SPIECES) v[u] = *(const float4*)((const char*)tl.spec + so[u0 + u]);
#pragma unroll
      for (int u = 0; u < TP_BATCH; u++) {
         const int p = pl + TP_PL * (u0 + u);
         if (u0 + u < TP_SPIECES && p < npiece) *(float4*)(sbuf + 16 * p) = v[u];
      }
      __builtin_amdgcn_sched_barrier(0);   // one batch of registers at a time
   }
   if (stage_t) {
      float4 v[TP_TPIECES];
#pragma unroll
      for (int u = 0; u < TP_TPIECES; u++) {
         const int p = min(pl + TP_PL * u, 343 * 4 - 1);
         v[u] = *(const float4*)(tl.T + (long long)(p >> 2) * tl.pst + b0 + 2 * (p & 3));
      }
#pragma unroll
      for (int u = 0; u < TP_TPIECES; u++)
         if (pl + TP_PL * u < 343 * 4) *(float4*)(buf + 16 * (pl + TP_PL * u)) = v[u];
   }
}

// The same staging by direct global->LDS transfers (16 bytes per lane, lane
// i of a transfer writes base + 16 i): no registers hold the rows.  The
// transfers are issued unconditionally, in straight-line code, so the
// compiler can keep them all in flight: source transfers cover all
// TP_MAX_SRC rows (rows past the tile's sources repeat its last one and are
// never read), T transfers 352 rows (rows past 342 repeat row 342), which
// leaves the zero row TP_ZERO_ROW untouched.
typedef __attribute__((address_space(3))) void* lds_ptr_t;
__device__ __forceinline__ void tp_dma16(const void* g, unsigned char* base) {
   __builtin_amdgcn_global_load_lds(g, (lds_ptr_t)base, 16, 0, 0);
}
__device__ __forceinline__ void tp_stage_dma(const int4& tk, const TpTile& tl, bool stage_t, unsigned char* buf,
                                             int pw, int lane, const int* __restrict__ gsrc) {
   constexpr int NS = TP_MAX_SRC * 4 / 64, NT = 352 * 4 / 64;   // transfers per buffer
   constexpr int MS = (NS + TP_PW - 1) / TP_PW, MT = (NT + TP_PW - 1) / TP_PW;
   const int b0 = tk.y, q2 = 2 * (lane & 3);
   int so[MS];
#pragma unroll
   for (int u = 0; u < MS; u++) {
      const int i = min(pw + u * TP_PW, NS - 1);
      so[u] = gsrc[tl.src_off + min(i * 16 + (lane >> 2), tl.nsrc - 1)];
   }
   unsigned char* sbuf = buf + TP_TROWS * 64;
#pragma unroll
   for (int u = 0; u < MS; u++) {
      const int i = min(pw + u * TP_PW, NS - 1);   // a repeated transfer rewrites the same rows
      tp_dma16(tl.spec + so[u] + b0 + q2, sbuf + i * 1024);
   }
   if (stage_t) {
#pragma unroll
      for (int u = 0; u < MT; u++) {
         const int i = min(pw + u * TP_PW, NT - 1);
         const int row = min(i * 16 + (lane >> 2), 342);
         tp_dma16(tl.T + (long long)row * tl.pst + b0 + q2, buf + i * 1024);
      }
   }
   __builtin_amdgcn_s_waitcnt(0);
}


// Word group i (steps 4i..4i+3) of target t of octet o; reads past the
// octet's last group are clamped to it.
__device__ __forceinline__ int4 tp_group(const int* __restrict__ gint, const TpOct& o, int i, int t) {
   const int last = max(o.nk / 4 - 1, 0);
   return *((const int4*)(gint + o.int_off) + 8 * min(i, last) + t);
}

// Two word groups (eight steps) of a lane: all sixteen LDS reads are issued
// before the first sum (a scheduling barrier keeps the compiler from
// interleaving them), then the sums, steps alternating over four
// accumulator pairs.
__device__ __forceinline__ void tp_groups2(const int4& x, const int4& y, const unsigned char* Tl,
                                           const unsigned char* Sl, float* re, float* im) {
   const unsigned u[8] = {(unsigned)x.x, (unsigned)x.y, (unsigned)x.z, (unsigned)x.w,
                          (unsigned)y.x, (unsigned)y.y, (unsigned)y.z, (unsigned)y.w};
   float2 t[8], v[8];
#pragma unroll
   for (int q = 0; q < 8; q++) {
      t[q] = *(const float2*)(Tl + (u[q] & 0xffffu));
      v[q] = *(const float2*)(Sl + (u[q] >> 16));
   }
   __builtin_amdgcn_sched_barrier(0);
#pragma unroll
   for (int q = 0; q < 8; q++) {
      float& r = re[q & 3];
      float& m = im[q & 3];
      r = fmaf(t[q].x, v[q].x, r);
      r = fmaf(-t[q].y, v[q].y, r);
      m = fmaf(t[q].x, v[q].y, m);
      m = fmaf(t[q].y, v[q].x, m);
   }
}

// Word-group pair p (steps 8p..8p+7) of target t of octet o (clamped to the
// last pair).
struct TpPair {
   int4 x, y;
};
__device__ __forceinline__ TpPair tp_pair(const int* __restrict__ gint, const TpOct& o, int p, int t) {
   return {tp_group(gint, o, 2 * p, t), tp_group(gint, o, 2 * p + 1, t)};
}

// The lane's steps of octet o (nk a multiple of 8, uniform over the wave),
// one word-group pair (eight steps, sixteen LDS reads) at a time; the
// words load two pairs ahead of their use, in three rotating register sets
// (no copies, so no wait on a load in flight).  A: pair 0, loaded; B: pair
// 1, issued by the caller.
__device__ __forceinline__ void tp_sum(const TpOct& o, int t, TpPair A, TpPair B, const int* __restrict__ gint,
                                       const unsigned char* Tl, const unsigned char* Sl, float* re, float* im) {
   const int np = o.nk / 8;
   for (int p = 0; p < np; p += 3) {
      const TpPair C = tp_pair(gint, o, p + 2, t);
      tp_groups2(A.x, A.y, Tl, Sl, re, im);
      if (p + 1 >= np) break;
      A = tp_pair(gint, o, p + 3, t);
      tp_groups2(B.x, B.y, Tl, Sl, re, im);
      if (p + 2 >= np) break;
      B = tp_pair(gint, o, p + 4, t);
      tp_groups2(C.x, C.y, Tl, Sl, re, im);
   }
}

// Producer waves: stage task 0, then task k+1 while the consumers sum task
// k; one barrier per task, as many as the consumers pass.
template <int MODE>
__device__ __forceinline__ void tp_producer(const int4* __restrict__ tasks, int n, const TpTile* __restrict__ tiles,
                                            const int* __restrict__ gsrc, unsigned char* smem, int* ctr, int pw,
                                            int lane, long long sdel) {
   long long key0, key1 = -1;   // the T rows (level and chunk) each buffer holds
   {
      const int4 t0 = tasks[0];
      TpTile tl = tiles[t0.x];
      tl.spec = (const float2*)((const char*)tl.spec + sdel);
      key0 = (long long)(tl.T + t0.y);
      if constexpr (MODE == 1)
         tp_stage_dma(t0, tl, true, smem, pw, lane, gsrc);
      else
         tp_stage(t0, tl, true, smem, 64 * pw + lane, gsrc);
   }
   __syncthreads();
   for (int k = 0; k < n; k++) {
      // task k+2's counter was task k-1's, done at the last barrier; it is
      // claimed from no earlier than the barrier that ends this iteration
      if (pw == 0 && lane == 0) ctr[(k + 2) % 3] = 0;
      if (k + 1 < n) {
         const int4 tn = tasks[k + 1];
         TpTile tl = tiles[tn.x];
         tl.spec = (const float2*)((const char*)tl.spec + sdel);
         const long long key = (long long)(tl.T + tn.y);
         const int nb = (k + 1) & 1;
         const bool st = (nb ? key1 : key0) != key;
         if (nb)
            key1 = key;
         else
            key0 = key;
         if constexpr (MODE == 1)
            tp_stage_dma(tn, tl, st, smem + nb * TP_BUF, pw, lane, gsrc);
         else
            tp_stage(tn, tl, st, smem + nb * TP_BUF, 64 * pw + lane, gsrc);
      }
      __syncthreads();
   }
}

// One claim on task k's octet counter (three counters in rotation, in LDS
// after the two buffers): the next unclaimed octet index of the task.  One
// lane of the wave adds, the wave reads its result.
__device__ __forceinline__ int tp_claim(int* ctr, int k, int lane) {
   int v = 0;
   if (lane == 0) v = atomicAdd(ctr + k % 3, 1);
   return __builtin_amdgcn_readfirstlane(v);
}

// Consumer waves claim the octets of task k one at a time, longest first
// (the plan sorts them), so the waves of a block finish a task together.
// While an octet is summed, the next is claimed and its first words load:
// the next octet of task k, or, when task k has none left, one of task k+1
// (its counter was cleared before the previous barrier).  One barrier per
// task.
__device__ __forceinline__ void tp_consumer(const int4* __restrict__ tasks, int n, const TpTile* __restrict__ tiles,
                                            const TpOct* __restrict__ octs, const long long* __restrict__ oacc,
                                            const int* __restrict__ gint, const unsigned char* smem, int* ctr,
                                            int lane, long long adel) {
   const int t = lane >> 3, b = lane & 7;
   int kc = -1, ic = 0;   // the claimed octet (task kc, index ic), none when kc < k
   TpOct o = {0, 0, 0, 0};
   TpPair A = {make_int4(0, 0, 0, 0), make_int4(0, 0, 0, 0)};
   long long ao = -1;
   __syncthreads();
   for (int k = 0; k < n; k++) {
      const int4 tk = tasks[k];
      const int pad = tiles[tk.x].pad;
      float2* __restrict__ acc = (float2*)((char*)tiles[tk.x].acc + adel);
      const unsigned char* Tl = smem + (k & 1) * TP_BUF + b * 8;
      const unsigned char* Sl = Tl + TP_TROWS * 64;
      if (kc < k) {
         ic = tp_claim(ctr, k, lane);
         if (ic < tk.w) {
            kc = k;
            o = octs[tk.z + ic];
            A = tp_pair(gint, o, 0, t);
            ao = oacc[o.acc_idx + t];
         }
      }
      while (kc == k) {
         int kn = k, jn = tp_claim(ctr, k, lane);
         if (jn >= tk.w) {
            kn = -1;
            if (k + 1 < n) {
               jn = tp_claim(ctr, k + 1, lane);
               if (jn < tasks[k + 1].w) kn = k + 1;
            }
         }
         // pair 1 of this octet, then the next octet's pair 0 (issued in
         // this order, so waiting for the former never waits for the latter)
         const TpPair B = tp_pair(gint, o, 1, t);
         TpOct on = {0, 0, 0, 0};
         TpPair An = {make_int4(0, 0, 0, 0), make_int4(0, 0, 0, 0)};
         long long aon = -1;
         if (kn >= 0) {
            on = octs[tasks[kn].z + jn];
            An = tp_pair(gint, on, 0, t);
            aon = oacc[on.acc_idx + t];
         }
         {
            float re[4] = {0.f, 0.f, 0.f, 0.f}, im[4] = {0.f, 0.f, 0.f, 0.f};
            tp_sum(o, t, A, B, gint, Tl, Sl, re, im);
            const float rs = (re[0] + re[1]) + (re[2] + re[3]), is = (im[0] + im[1]) + (im[2] + im[3]);
            if (ao >= 0 && tk.y + b < pad) acc[ao + tk.y + b] = make_float2(rs, is);
         }
         kc = kn;
         ic = jn;
         o = on;
         A = An;
         ao = aon;
      }
      __syncthreads();
   }
}

template <int MODE>
__global__ void __launch_bounds__(TP_NT) k_translate_pipe(const int2* __restrict__ ranges,
                                                          const int4* __restrict__ tasks_all,
                                                          const TpTile* __restrict__ tiles,
                                                          const TpOct* __restrict__ octs,
                                                          const long long* __restrict__ oacc,
                                                          const int* __restrict__ gsrc, const int* __restrict__ gint,
                                                          int nb, int ncol, long long sdel, long long adel,
                                                          long long scs, long long acs) {
   extern __shared__ __align__(16) unsigned char smem[];
   const ColBlk cb = col_block(ncol);
   if (cb.b >= nb) return;   // uniform over the block: no barrier is reached
   const int2 rg = ranges[cb.b];
   const int tid = threadIdx.x, lane = tid & 63;
   const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
   // the zero T row (padding words) of both buffers, which no transfer
   // writes, and the three octet counters
   int* ctr = (int*)(smem + 2 * TP_BUF);
   if (tid < 16) ((float2*)(smem + (tid >> 3) * TP_BUF + TP_ZERO_ROW * 64))[tid & 7] = make_float2(0.f, 0.f);
   if (tid < 3) ctr[tid] = 0;
   if (rg.y == 0) return;   // uniform over the block: no barrier is reached
   // both roles pass rg.y + 1 barriers
   if (wave >= TP_CW)
      tp_producer<MODE>(tasks_all + rg.x, rg.y, tiles, gsrc, smem, ctr, wave - TP_CW, lane, sdel + cb.c * scs);
   else
      tp_consumer(tasks_all + rg.x, rg.y, tiles, octs, oacc, gint, smem, ctr, lane, adel + cb.c * acs);
}

// ---------------------------------------------------------------------------
// The pipelined translation for two columns at once (column-batched
// applies): the same tiles, octets, interaction words and per-target
// summation order as k_translate_pipe, with 4-bin chunks.  A source row
// holds both columns' bins interleaved (bin j: column c0 then c1, 16
// bytes), a T row 4 bins (32 bytes); lane (target, bin j) reads its T bin
// once and both columns' source bins with one 16-byte read, so every T read
// serves two complex multiply-adds.  Each column's accumulators see the
// single-column sequence of operations, so both outputs equal
// k_translate_pipe's bit for bit.  A wave sums two octets (16 targets, 4
// lanes each); staging goes through registers (the interleave rules out
// direct transfers).
constexpr int TP2_BUF = TP_TROWS * 32 + TP_MAX_SRC * 64;
static_assert(2 * TP2_BUF + 16 <= 163840, "two pair buffers must fit the LDS of a CU");
constexpr int TP2_SPIECES = (TP_MAX_SRC * 2 + TP_PL - 1) / TP_PL;   // source half-rows per producer lane
constexpr int TP2_TPIECES = (343 * 2 + TP_PL - 1) / TP_PL;          // T half-rows per producer lane

__device__ __forceinline__ void tp_stage2(const int4& tk, const TpTile& tl, const float2* spec1, bool stage_t,
                                          unsigned char* buf, int pl, const int* __restrict__ gsrc) {
   const int b0 = tk.y;
   const int npiece = tl.nsrc * 2;
   unsigned so[TP2_SPIECES];
#pragma unroll
   for (int u = 0; u < TP2_SPIECES; u++) {
      const int p = pl + TP_PL * u;
      so[u] = 8u * (unsigned)(gsrc[tl.src_off + min(p >> 1, tl.nsrc - 1)] + b0 + 2 * (p & 1));
   }
   unsigned char* sbuf = buf + TP_TROWS * 32;
   constexpr int B = 6;
#pragma unroll
   for (int u0 = 0; u0 < TP2_SPIECES; u0 += B) {
      float4 a[B], b[B];
#pragma unroll
      for (int u = 0; u < B; u++)
         if (u0 + u < TP2_SPIECES) {
            a[u] = *(const float4*)((const char*)tl.spec + so[u0 + u]);
            b[u] = *(const float4*)((const char*)spec1 + so[u0 + u]);
         }
#pragma unroll
      for (int u = 0; u < B; u++) {
         const int p = pl + TP_PL * (u0 + u);
         if (u0 + u < TP2_SPIECES && p < npiece) {
            unsigned char* row = sbuf + 64 * (p >> 1) + 32 * (p & 1);
            *(float4*)row = make_float4(a[u].x, a[u].y, b[u].x, b[u].y);
            *(float4*)(row + 16) = make_float4(a[u].z, a[u].w, b[u].z, b[u].w);
         }
      }
      __builtin_amdgcn_sched_barrier(0);
   }
   if (stage_t) {
      float4 v[TP2_TPIECES];
#pragma unroll
      for (int u = 0; u < TP2_TPIECES; u++) {
         const int p = min(pl + TP_PL * u, 343 * 2 - 1);
         v[u] = *(const float4*)(tl.T + (long long)(p >> 1) * tl.pst + b0 + 2 * (p & 1));
      }
#pragma unroll
      for (int u = 0; u < TP2_TPIECES; u++)
         if (pl + TP_PL * u < 343 * 2) *(float4*)(buf + 16 * (pl + TP_PL * u)) = v[u];
   }
}

__device__ __forceinline__ void tp_producer2(const int4* __restrict__ tasks, int n, const TpTile* __restrict__ tiles,
                                             const int* __restrict__ gsrc, unsigned char* smem, int* ctr, int pl,
                                             long long sdel0, long long sdel1) {
   long long key0, key1 = -1;
   {
      const int4 t0 = tasks[0];
      TpTile tl = tiles[t0.x];
      const float2* s1 = (const float2*)((const char*)tl.spec + sdel1);
      tl.spec = (const float2*)((const char*)tl.spec + sdel0);
      key0 = (long long)(tl.T + t0.y);
      tp_stage2(t0, tl, s1, true, smem, pl, gsrc);
   }
   __syncthreads();
   for (int k = 0; k < n; k++) {
      if (pl == 0) ctr[(k + 2) % 3] = 0;
      if (k + 1 < n) {
         const int4 tn = tasks[k + 1];
         TpTile tl = tiles[tn.x];
         const float2* s1 = (const float2*)((const char*)tl.spec + sdel1);
         tl.spec = (const float2*)((const char*)tl.spec + sdel0);
         const long long key = (long long)(tl.T + tn.y);
         const int nb = (k + 1) & 1;
         const bool st = (nb ? key1 : key0) != key;
         if (nb)
            key1 = key;
         else
            key0 = key;
         tp_stage2(tn, tl, s1, st, smem + nb * TP2_BUF, pl, gsrc);
      }
      __syncthreads();
   }
}

// Words of group i (steps 4i..4i+3) of target t of octet o; null words past
// the octet's steps (the other octet of the wave may run longer).
__device__ __forceinline__ int4 tp_group2(const int* __restrict__ gint, const TpOct& o, int i, int t) {
   if (4 * i >= o.nk) return make_int4((int)TP_NULL_WORD, (int)TP_NULL_WORD, (int)TP_NULL_WORD, (int)TP_NULL_WORD);
   return *((const int4*)(gint + o.int_off) + 8 * i + t);
}

// Two word groups (eight steps) for both columns: all sixteen LDS reads
// issued before the first multiply-add, then the steps in list order, step
// q on accumulator q & 3, as the single-column sum.
__device__ __forceinline__ void tp_steps2(const int4& x, const int4& y, const unsigned char* Tl,
                                          const unsigned char* Sl, float* re0, float* im0, float* re1, float* im1) {
   const unsigned u[8] = {(unsigned)x.x, (unsigned)x.y, (unsigned)x.z, (unsigned)x.w,
                          (unsigned)y.x, (unsigned)y.y, (unsigned)y.z, (unsigned)y.w};
   float2 t[8];
   float4 v[8];
#pragma unroll
   for (int q = 0; q < 8; q++) {
      t[q] = *(const float2*)(Tl + ((u[q] & 0xffffu) >> 1));
      v[q] = *(const float4*)(Sl + (u[q] >> 16));
   }
   __builtin_amdgcn_sched_barrier(0);
#pragma unroll
   for (int q = 0; q < 8; q++) {
      float& r0 = re0[q & 3];
      float& m0 = im0[q & 3];
      float& r1 = re1[q & 3];
      float& m1 = im1[q & 3];
      r0 = fmaf(t[q].x, v[q].x, r0);
      r0 = fmaf(-t[q].y, v[q].y, r0);
      m0 = fmaf(t[q].x, v[q].y, m0);
      m0 = fmaf(t[q].y, v[q].x, m0);
      r1 = fmaf(t[q].x, v[q].z, r1);
      r1 = fmaf(-t[q].y, v[q].w, r1);
      m1 = fmaf(t[q].x, v[q].w, m1);
      m1 = fmaf(t[q].y, v[q].z, m1);
   }
}

__device__ __forceinline__ void tp_consumer2(const int4* __restrict__ tasks, int n, const TpTile* __restrict__ tiles,
                                             const TpOct* __restrict__ octs, const long long* __restrict__ oacc,
                                             const int* __restrict__ gint, const unsigned char* smem, int* ctr,
                                             int lane, long long adel0, long long adel1) {
   const int h = lane >> 5, t = (lane >> 2) & 7, j = lane & 3;
   __syncthreads();
   for (int k = 0; k < n; k++) {
      const int4 tk = tasks[k];
      const int pad = tiles[tk.x].pad;
      float2* acc0 = (float2*)((char*)tiles[tk.x].acc + adel0);
      float2* acc1 = (float2*)((char*)tiles[tk.x].acc + adel1);
      const unsigned char* Tl = smem + (k & 1) * TP2_BUF + j * 8;
      const unsigned char* Sl = smem + (k & 1) * TP2_BUF + TP_TROWS * 32 + j * 16;
      for (;;) {
         int ic = 0;
         if (lane == 0) ic = atomicAdd(ctr + k % 3, 2);
         ic = __builtin_amdgcn_readfirstlane(ic);
         if (ic >= tk.w) break;
         // octets ic (lanes 0-31) and ic + 1 (lanes 32-63, if any); the wave
         // runs the longer one's steps, the other half reads null words
         const TpOct oa = octs[tk.z + ic];
         const bool hb = ic + 1 < tk.w;
         const TpOct ob = hb ? octs[tk.z + ic + 1] : TpOct{0, 0, 0, 0};
         const TpOct o = h ? ob : oa;
         const int nkw = max(oa.nk, ob.nk);
         const long long ao = (h == 0 || hb) ? oacc[o.acc_idx + t] : -1;
         float re0[4] = {0.f, 0.f, 0.f, 0.f}, im0[4] = {0.f, 0.f, 0.f, 0.f};
         float re1[4] = {0.f, 0.f, 0.f, 0.f}, im1[4] = {0.f, 0.f, 0.f, 0.f};
         // word-group pairs, loaded two pairs ahead of their use in three
         // rotating register sets
         const int np = nkw / 8;
         TpPair A = {tp_group2(gint, o, 0, t), tp_group2(gint, o, 1, t)};
         TpPair B = {tp_group2(gint, o, 2, t), tp_group2(gint, o, 3, t)};
         for (int q = 0; q < np; q += 3) {
            const TpPair C = {tp_group2(gint, o, 2 * q + 4, t), tp_group2(gint, o, 2 * q + 5, t)};
            tp_steps2(A.x, A.y, Tl, Sl, re0, im0, re1, im1);
            if (q + 1 >= np) break;
            A = {tp_group2(gint, o, 2 * q + 6, t), tp_group2(gint, o, 2 * q + 7, t)};
            tp_steps2(B.x, B.y, Tl, Sl, re0, im0, re1, im1);
            if (q + 2 >= np) break;
            B = {tp_group2(gint, o, 2 * q + 8, t), tp_group2(gint, o, 2 * q + 9, t)};
            tp_steps2(C.x, C.y, Tl, Sl, re0, im0, re1, im1);
         }
         if (ao >= 0 && tk.y + j < pad) {
            acc0[ao + tk.y + j] = make_float2((re0[0] + re0[1]) + (re0[2] + re0[3]), (im0[0] + im0[1]) + (im0[2] + im0[3]));
            acc1[ao + tk.y + j] = make_float2((re1[0] + re1[1]) + (re1[2] + re1[3]), (im1[0] + im1[1]) + (im1[2] + im1[3]));
         }
      }
      __syncthreads();
   }
}

__global__ void __launch_bounds__(TP_NT) k_translate_pipe2(const int2* __restrict__ ranges,
                                                           const int4* __restrict__ tasks_all,
                                                           const TpTile* __restrict__ tiles,
                                                           const TpOct* __restrict__ octs,
                                                           const long long* __restrict__ oacc,
                                                           const int* __restrict__ gsrc, const int* __restrict__ gint,
                                                           int nb, int ncol, long long sdel, long long adel,
                                                           long long scs, long long acs) {
   extern __shared__ __align__(16) unsigned char smem[];
   const ColBlk cb = col_block_major(nb, (ncol + 1) / 2);
   if (cb.b >= nb) return;   // uniform over the block: no barrier is reached
   const int c0 = 2 * cb.c, c1 = min(c0 + 1, ncol - 1);
   const int2 rg = ranges[cb.b];
   const int tid = threadIdx.x, lane = tid & 63;
   const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
   // the zero T row of both buffers (4 bins) and the three octet counters
   int* ctr = (int*)(smem + 2 * TP2_BUF);
   if (tid < 8) ((float2*)(smem + (tid >> 2) * TP2_BUF + TP_ZERO_ROW * 32))[tid & 3] = make_float2(0.f, 0.f);
   if (tid < 3) ctr[tid] = 0;
   if (rg.y == 0) return;
   if (wave >= TP_CW)
      tp_producer2(tasks_all + rg.x, rg.y, tiles, gsrc, smem, ctr, 64 * (wave - TP_CW) + lane, sdel + c0 * scs,
                   sdel + c1 * scs);
   else
      tp_consumer2(tasks_all + rg.x, rg.y, tiles, octs, oacc, gint, smem, ctr, lane, adel + c0 * acs,
                   adel + c1 * acs);
}

void launch_translate_pipe2(int nblk, const int2* range, const int4* tasks, const TpTile* tiles, const TpOct* octs,
                            const long long* oacc, const int* gsrc, const int* gint, hipStream_t s, int ncol,
                            long long sdel, long long adel, long long scs, long long acs) {
   constexpr size_t lds = 2 * (size_t)TP2_BUF + 16;
   static std::once_flag attr;
   std::call_once(attr, [] {
      FSVD_CHECK(hipFuncSetAttribute((const void*)k_translate_pipe2, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds));
   });
   if (nblk <= 0) return;
   hipLaunchKernelGGL(k_translate_pipe2, dim3(col_grid(nblk, (ncol + 1) / 2)), dim3(TP_NT), lds, s, range, tasks,
                      tiles, octs, oacc, gsrc, gint, nblk, ncol, sdel, adel, scs, acs);
}

template <int n, typename T, typename G>
__device__ void target_body(const TgtPlan& d, const T* __restrict__ UTI,
                            const typename Cplx<T>::type* __restrict__ acc, T* __restrict__ w, unsigned char* smem,
                            const G& gr) {
   using C2 = typename Cplx<T>::type;
   constexpr int N = FarSizes<n>::N, H = FarSizes<n>::H, PAD = FarSizes<n>::PAD, GP3 = FarSizes<n>::GP3;
   C2* S = (C2*)smem;             // PAD
   C2* C1 = S + PAD;              // n*N*H
   C2* Cc = C1 + n * N * H;       // n*n*H
   T* cs = (T*)(Cc + n * n * H);  // N
   T* sn = cs + N;
   T* h = sn + N;                 // GP3
   const int tid = gr.tid, nt = gr.nt;
   twiddle_table<T, N>(cs, sn, gr);
   // four bins per thread at a time, loads before stores
   for (int b0 = tid; b0 < PAD; b0 += 4 * nt) {
      C2 v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
         const int b = b0 + u * nt;
         v[u] = acc[d.acc_off + (b < PAD ? b : 0)];   // valid address, no branch
         keep(v[u].x);
         keep(v[u].y);
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
         const int b = b0 + u * nt;
         if (b < PAD) S[b] = v[u];
      }
   }
   gr.sync();
   grid_backward<n, T>(S, C1, Cc, h, cs, sn, gr);
   gr.sync();
   const T* uti = UTI + d.uti_off;
   // w = UTI h: 16 lanes per row, nt/16 rows per pass, four rows per lane in
   // flight (each row summed over q in order, then a fixed butterfly)
   const int kr = tid >> 4, ql = tid & 15, RP = nt >> 4;
   for (int k = kr; k < d.r; k += 4 * RP) {
      const T* row[4];
      bool ok[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
         ok[u] = k + u * RP < d.r;
         row[u] = uti + (long long)(ok[u] ? k + u * RP : k) * GP3;   // a valid row either way
      }
      T a[4] = {0, 0, 0, 0};
#pragma unroll
      for (int q0 = 0; q0 < GP3; q0 += 16) {
         // clamped column and selects: no branch around the loads
         const int q = q0 + ql < GP3 ? q0 + ql : 0;
         const bool okq = q0 + ql < GP3;
         T x[4];
#pragma unroll
         for (int u = 0; u < 4; u++) x[u] = row[u][q];
         const T hq = h[q];
#pragma unroll
         for (int u = 0; u < 4; u++) a[u] += okq && ok[u] ? x[u] * hq : (T)0;
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
         for (int o = 8; o > 0; o >>= 1) a[u] += __shfl_xor(a[u], o, 64);
      if (ql == 0)
#pragma unroll
         for (int u = 0; u < 4; u++)
            if (ok[u]) w[d.w_off + k + u * RP] = a[u];
   }
}

// target_body for NC columns of a group (c0 .. c0 + nc - 1): each column's
// spectrum is transformed as target_body does (the LDS transform buffers
// reused column after column, grid values kept per column), then every UTI
// row is loaded once for all columns, each column's sums in target_body's
// order, so w equals the single-column result bit for bit.
template <int n, typename T, int NC, typename G>
__device__ void target_body_mc(const TgtPlan& d, const T* __restrict__ UTI,
                               const typename Cplx<T>::type* __restrict__ acc, long long cs_acc,
                               T* __restrict__ w, long long cs_w, int c0, int nc, unsigned char* smem, const G& gr) {
   using C2 = typename Cplx<T>::type;
   constexpr int N = FarSizes<n>::N, H = FarSizes<n>::H, PAD = FarSizes<n>::PAD, GP3 = FarSizes<n>::GP3;
   C2* S = (C2*)smem;             // PAD
   C2* C1 = S + PAD;              // n*N*H
   C2* Cc = C1 + n * N * H;       // n*n*H
   T* cs = (T*)(Cc + n * n * H);  // N
   T* sn = cs + N;
   T* h = sn + N;                 // NC x GP3
   const int tid = gr.tid, nt = gr.nt;
   twiddle_table<T, N>(cs, sn, gr);
   for (int j = 0; j < nc; j++) {
      const C2* accj = acc + (long long)(c0 + j) * cs_acc;
      for (int b0 = tid; b0 < PAD; b0 += 4 * nt) {
         C2 v[4];
#pragma unroll
         for (int u = 0; u < 4; u++) {
            const int b = b0 + u * nt;
            v[u] = accj[d.acc_off + (b < PAD ? b : 0)];
            keep(v[u].x);
            keep(v[u].y);
         }
#pragma unroll
         for (int u = 0; u < 4; u++) {
            const int b = b0 + u * nt;
            if (b < PAD) S[b] = v[u];
         }
      }
      gr.sync();
      grid_backward<n, T>(S, C1, Cc, h + j * GP3, cs, sn, gr);
      gr.sync();
   }
   const T* uti = UTI + d.uti_off;
   const int kr = tid >> 4, ql = tid & 15, RP = nt >> 4;
   for (int k = kr; k < d.r; k += 2 * RP) {
      const T* row0 = uti + (long long)k * GP3;
      const bool two = k + RP < d.r;
      const T* row1 = two ? row0 + (long long)RP * GP3 : row0;
      T a[NC], b[NC];
#pragma unroll
      for (int j = 0; j < NC; j++) a[j] = b[j] = 0;
#pragma unroll
      for (int q0 = 0; q0 < GP3; q0 += 16) {
         const int q = q0 + ql < GP3 ? q0 + ql : 0;
         const bool okq = q0 + ql < GP3;
         const T x0 = row0[q], x1 = row1[q];
#pragma unroll
         for (int j = 0; j < NC; j++) {
            const T hq = h[(j < nc ? j : 0) * GP3 + q];
            a[j] += okq ? x0 * hq : (T)0;
            b[j] += okq && two ? x1 * hq : (T)0;
         }
      }
#pragma unroll
      for (int j = 0; j < NC; j++)
#pragma unroll
         for (int o = 8; o > 0; o >>= 1) {
            a[j] += __shfl_xor(a[j], o, 64);
            b[j] += __shfl_xor(b[j], o, 64);
         }
      if (ql == 0)
         for (int j = 0; j < nc; j++) {
            T* wj = w + (long long)(c0 + j) * cs_w;
            wj[d.w_off + k] = a[j];
            if (two) wj[d.w_off + k + RP] = b[j];
         }
   }
}

// LDS of one target cube's work
template <typename T>
__host__ __device__ constexpr size_t target_cube_lds(int n) {
   return sizeof(typename Cplx<T>::type) *
              (size_t)((2 * n - 1) * (2 * n - 1) * ((2 * n - 1) / 2 + 1) + n * (2 * n - 1) * ((2 * n - 1) / 2 + 1) +
                       n * n * ((2 * n - 1) / 2 + 1)) +
          sizeof(T) * (size_t)(2 * (2 * n - 1) + n * n * n) + 64;
}

// Target cubes, one launch: the nbig cubes with other than four grid points
// per side first, one block of CUBE_NT threads each, then the cubes of four
// grid points per side, one per wavefront in its own LDS slice.
template <typename T>
__global__ void __launch_bounds__(CUBE_NT) k_target(const TgtPlan* __restrict__ tgt, int nbig, int ntgt,
                                                    const T* __restrict__ UTI,
                                                    const typename Cplx<T>::type* __restrict__ acc,
                                                    T* __restrict__ w, int wave_lds, int nb, int ncol,
                                                    long long cs_acc, long long cs_w) {
   extern __shared__ __align__(16) unsigned char smem[];
   const ColBlk cb = col_block(ncol);
   if (cb.b >= nb) return;
   acc += cb.c * cs_acc;
   w += cb.c * cs_w;
   if (cb.b < nbig) {
      const TgtPlan d = tgt[cb.b];
      const BlockGroup gr = block_group();
      switch (d.n) {
         case 3: target_body<3, T>(d, UTI, acc, w, smem, gr); break;
         case 5: target_body<5, T>(d, UTI, acc, w, smem, gr); break;
         case 6: target_body<6, T>(d, UTI, acc, w, smem, gr); break;
         case 7: target_body<7, T>(d, UTI, acc, w, smem, gr); break;
         case 8: target_body<8, T>(d, UTI, acc, w, smem, gr); break;
         case 9: target_body<9, T>(d, UTI, acc, w, smem, gr); break;
         default: break;
      }
      return;
   }
   const int wave = threadIdx.x >> 6, i = nbig + (cb.b - nbig) * (CUBE_NT / 64) + wave;
   if (i >= ntgt) return;
   const TgtPlan d = tgt[i];
   target_body<4, T>(d, UTI, acc, w, smem + (size_t)wave * wave_lds, WaveGroup{(int)(threadIdx.x & 63)});
}

// k_target for TGT_NC columns per block or wavefront (column-batched
// applies), with the extra grid values of each column in the LDS slice.
constexpr int TGT_NC = 4;
template <typename T>
__global__ void __launch_bounds__(CUBE_NT) k_target_mc(const TgtPlan* __restrict__ tgt, int nbig, int ntgt,
                                                       const T* __restrict__ UTI,
                                                       const typename Cplx<T>::type* __restrict__ acc,
                                                       T* __restrict__ w, int wave_lds, int nb, int ncol,
                                                       long long cs_acc, long long cs_w) {
   extern __shared__ __align__(16) unsigned char smem[];
   const int ngrp = (ncol + TGT_NC - 1) / TGT_NC;
   const ColBlk cb = col_block(ngrp);
   if (cb.b >= nb) return;
   const int c0 = cb.c * TGT_NC, nc = min(TGT_NC, ncol - c0);
   if (cb.b < nbig) {
      const TgtPlan d = tgt[cb.b];
      const BlockGroup gr = block_group();
      switch (d.n) {
         case 3: target_body_mc<3, T, TGT_NC>(d, UTI, acc, cs_acc, w, cs_w, c0, nc, smem, gr); break;
         case 5: target_body_mc<5, T, TGT_NC>(d, UTI, acc, cs_acc, w, cs_w, c0, nc, smem, gr); break;
         case 6: target_body_mc<6, T, TGT_NC>(d, UTI, acc, cs_acc, w, cs_w, c0, nc, smem, gr); break;
         case 7: target_body_mc<7, T, TGT_NC>(d, UTI, acc, cs_acc, w, cs_w, c0, nc, smem, gr); break;
         case 8: target_body_mc<8, T, TGT_NC>(d, UTI, acc, cs_acc, w, cs_w, c0, nc, smem, gr); break;
         case 9: target_body_mc<9, T, TGT_NC>(d, UTI, acc, cs_acc, w, cs_w, c0, nc, smem, gr); break;
         default: break;
      }
      return;
   }
   const int wave = threadIdx.x >> 6, i = nbig + (cb.b - nbig) * (CUBE_NT / 64) + wave;
   if (i >= ntgt) return;
   const TgtPlan d = tgt[i];
   target_body_mc<4, T, TGT_NC>(d, UTI, acc, cs_acc, w, cs_w, c0, nc, smem + (size_t)wave * wave_lds,
                                WaveGroup{(int)(threadIdx.x & 63)});
}

// Leaf output: y = D_s xs + D_d xd (interleaved 2p x m block) + sum over the
// leaf's far-field blocks of U w.  Thread (g, e) owns block row e and the
// columns c = g, g+G, ...; the G partial sums of each row are added in group
// order, then the two layers and the far field.
template <typename T, bool NT>
__global__ void __launch_bounds__(256) k_leaf(const LeafDesc* __restrict__ leaves, const int* __restrict__ dcol,
                                              const T* __restrict__ D, const double* __restrict__ xs,
                                              const double* __restrict__ xd, const LeafU* __restrict__ ublk,
                                              const T* __restrict__ UT, const T* __restrict__ w,
                                              double* __restrict__ y, const int* __restrict__ bmap, int nb, int ncol,
                                              long long cs_x, long long cs_w, long long cs_y,
                                              const int* __restrict__ sp_out, const double* __restrict__ sp_corr,
                                              double* __restrict__ sp_y) {
   extern __shared__ __align__(16) unsigned char smem[];
   const ColBlk cb = col_block(ncol);
   if (cb.b >= nb) return;
   const int li = bmap[cb.b];
   if (li < 0) return;
   xs += cb.c * cs_x;
   xd += cb.c * cs_x;
   w += cb.c * cs_w;
   y += cb.c * cs_y;
   const LeafDesc L = leaves[li];
   const int E = 2 * L.p, m = L.m, p = L.p;
   T* xg = (T*)smem;       // 2*m
   T* part = xg + 2 * m;   // 256
   T* rows = part + 256;   // 2p near rows, then p far rows
   const int tid = threadIdx.x;
   for (int c = tid; c < m; c += 256) {
      int col = dcol[L.col_off + c];
      xg[c] = (T)xs[col];
      xg[m + c] = (T)xd[col];
   }
   __syncthreads();
   const T* blk = D + L.d_off;
   for (int e0 = 0; e0 < E; e0 += 256) {
      const int Ec = min(256, E - e0), G = 256 / Ec;
      const int g = tid / Ec, e = e0 + tid - g * Ec;
      T acc = 0;
      if (g < G) acc = stream_rows<T, NT>(blk + e, E, m, g, G, xg + (e >= p ? m : 0));
      part[tid] = acc;
      __syncthreads();
      if (tid < Ec) {
         T sum = 0;
         for (int gg = 0; gg < G; gg++) sum += part[gg * Ec + tid];
         rows[e0 + tid] = sum;
      }
      __syncthreads();
   }
   // far field: rows e < p of every block, the k index split over G groups
   for (int e0 = 0; e0 < p; e0 += 256) {
      const int Ec = min(256, p - e0), G = 256 / Ec;
      const int g = tid / Ec, e = e0 + tid - g * Ec;
      T acc = 0;
      if (g < G)
         for (int u = 0; u < L.nu; u++) {
            const LeafU b = ublk[L.u0 + u];
            const T* col = UT + b.u_off + e;
            const T* wb = w + b.w_off;
            // up to eight k in flight (a masked last batch)
            for (int k = g; k < b.r; k += 8 * G) {
               T cv[8], wv[8];
#pragma unroll
               for (int v = 0; v < 8; v++) {
                  const int kk = k + v * G;
                  cv[v] = kk < b.r ? col[(long long)kk * b.ld] : (T)0;
                  wv[v] = kk < b.r ? wb[kk] : (T)0;
               }
#pragma unroll
               for (int v = 0; v < 8; v++) acc += cv[v] * wv[v];
            }
         }
      part[tid] = acc;
      __syncthreads();
      if (tid < Ec) {
         T sum = 0;
         for (int gg = 0; gg < G; gg++) sum += part[gg * Ec + tid];
         rows[E + e0 + tid] = sum;
      }
      __syncthreads();
   }
   // with sp_y: the region split fused in (k_split's y[out] = yt + corr xd)
   for (int r = tid; r < p; r += 256) {
      const double v = (double)((rows[r] + rows[p + r]) + rows[E + r]);
      if (sp_y)
         sp_y[sp_out[L.q0 + r]] = v + sp_corr[L.q0 + r] * xd[L.q0 + r];
      else
         y[L.q0 + r] = v;
   }
}

// columns per block of the multi-column leaf kernel
constexpr int LEAF_NC = 4;

// stream_rows for NC columns at once: every entry of the block is loaded
// once and multiplied into each column's accumulators; xl holds the
// columns' inputs interleaved (input i of column j at xl[i * NC + j]).  Each
// column's accumulators see the single-column sequence of operations, so
// out[j] is stream_rows's result for column j bit for bit.
template <typename T, int NC>
__device__ __forceinline__ void stream_rows_mc(const T* __restrict__ colp, int E, int m, int g, int G,
                                               const T* xl, T* out) {
   const long long step = (long long)G * E;
   T a[4][NC];
#pragma unroll
   for (int q = 0; q < 4; q++)
#pragma unroll
      for (int j = 0; j < NC; j++) a[q][j] = 0;
   int c = g;
   for (; c + 7 * G < m; c += 8 * G) {
      const T* pc = colp + (long long)c * E;
      T d[8];
#pragma unroll
      for (int u = 0; u < 8; u++) d[u] = pc[u * step];
#pragma unroll
      for (int u = 0; u < 8; u++) {
         const T* xv = xl + (size_t)(c + u * G) * NC;
#pragma unroll
         for (int j = 0; j < NC; j++) a[u & 3][j] += d[u] * xv[j];
      }
   }
   if (c < m) {
      const T* pc = colp + (long long)c * E;
      T d[7];
#pragma unroll
      for (int u = 0; u < 7; u++) d[u] = c + u * G < m ? pc[u * step] : (T)0;
#pragma unroll
      for (int u = 0; u < 7; u++)
         if (c + u * G < m) {
            const T* xv = xl + (size_t)(c + u * G) * NC;
#pragma unroll
            for (int j = 0; j < NC; j++) a[0][j] += d[u] * xv[j];
         }
   }
#pragma unroll
   for (int j = 0; j < NC; j++) out[j] = (a[0][j] + a[1][j]) + (a[2][j] + a[3][j]);
}

// k_leaf for NC columns per block: the near-field block and the U rows are
// read once for the NC columns of the block's group (columns c0 .. c0+NC-1
// of ncol; a partial last group repeats its last column and stores only its
// own).  Per column the arithmetic is k_leaf's, so y equals the
// single-column result bit for bit.  Grid: col_grid(nb, number of groups).
template <typename T, int NC>
__global__ void __launch_bounds__(256) k_leaf_mc(const LeafDesc* __restrict__ leaves, const int* __restrict__ dcol,
                                                 const T* __restrict__ D, const double* __restrict__ xs,
                                                 const double* __restrict__ xd, const LeafU* __restrict__ ublk,
                                                 const T* __restrict__ UT, const T* __restrict__ w,
                                                 double* __restrict__ y, const int* __restrict__ bmap, int nb,
                                                 int ncol, long long cs_x, long long cs_w, long long cs_y) {
   extern __shared__ __align__(16) unsigned char smem[];
   const int ngrp = (ncol + NC - 1) / NC;
   const ColBlk cb = col_block(ngrp);
   if (cb.b >= nb) return;
   const int li = bmap[cb.b];
   if (li < 0) return;
   const int c0 = cb.c * NC, nc = min(NC, ncol - c0);
   const LeafDesc L = leaves[li];
   const int E = 2 * L.p, m = L.m, p = L.p;
   T* xg = (T*)smem;                 // 2m x NC
   T* part = xg + 2 * m * NC;        // NC x 256
   T* rows = part + 256 * NC;        // NC x (2p near rows, then p far rows)
   const int tid = threadIdx.x, R3 = 3 * p;
   for (int c = tid; c < m; c += 256) {
      const int col = dcol[L.col_off + c];
#pragma unroll
      for (int j = 0; j < NC; j++) {
         const long long o = (long long)(c0 + min(j, nc - 1)) * cs_x + col;
         xg[(size_t)c * NC + j] = (T)xs[o];
         xg[(size_t)(m + c) * NC + j] = (T)xd[o];
      }
   }
   __syncthreads();
   const T* blk = D + L.d_off;
   for (int e0 = 0; e0 < E; e0 += 256) {
      const int Ec = min(256, E - e0), G = 256 / Ec;
      const int g = tid / Ec, e = e0 + tid - g * Ec;
      T acc[NC];
#pragma unroll
      for (int j = 0; j < NC; j++) acc[j] = 0;
      if (g < G) stream_rows_mc<T, NC>(blk + e, E, m, g, G, xg + (e >= p ? (size_t)m * NC : 0), acc);
#pragma unroll
      for (int j = 0; j < NC; j++) part[j * 256 + tid] = acc[j];
      __syncthreads();
      if (tid < Ec) {
#pragma unroll
         for (int j = 0; j < NC; j++) {
            T sum = 0;
            for (int gg = 0; gg < G; gg++) sum += part[j * 256 + gg * Ec + tid];
            rows[j * R3 + e0 + tid] = sum;
         }
      }
      __syncthreads();
   }
   for (int e0 = 0; e0 < p; e0 += 256) {
      const int Ec = min(256, p - e0), G = 256 / Ec;
      const int g = tid / Ec, e = e0 + tid - g * Ec;
      T acc[NC];
#pragma unroll
      for (int j = 0; j < NC; j++) acc[j] = 0;
      if (g < G)
         for (int u = 0; u < L.nu; u++) {
            const LeafU b = ublk[L.u0 + u];
            const T* col = UT + b.u_off + e;
            for (int k = g; k < b.r; k += 8 * G) {
               T cv[8];
#pragma unroll
               for (int v = 0; v < 8; v++) {
                  const int kk = k + v * G;
                  cv[v] = kk < b.r ? col[(long long)kk * b.ld] : (T)0;
               }
#pragma unroll
               for (int j = 0; j < NC; j++) {
                  const T* wb = w + (long long)(c0 + min(j, nc - 1)) * cs_w + b.w_off;
                  T wv[8];
#pragma unroll
                  for (int v = 0; v < 8; v++) {
                     const int kk = k + v * G;
                     wv[v] = kk < b.r ? wb[kk] : (T)0;
                  }
#pragma unroll
                  for (int v = 0; v < 8; v++) acc[j] += cv[v] * wv[v];
               }
            }
         }
#pragma unroll
      for (int j = 0; j < NC; j++) part[j * 256 + tid] = acc[j];
      __syncthreads();
      if (tid < Ec) {
#pragma unroll
         for (int j = 0; j < NC; j++) {
            T sum = 0;
            for (int gg = 0; gg < G; gg++) sum += part[j * 256 + gg * Ec + tid];
            rows[j * R3 + E + e0 + tid] = sum;
         }
      }
      __syncthreads();
   }
   for (int j = 0; j < nc; j++) {
      const T* rj = rows + j * R3;
      double* yj = y + (long long)(c0 + j) * cs_y;
      for (int r = tid; r < p; r += 256) yj[L.q0 + r] = (double)((rj[r] + rj[p + r]) + rj[E + r]);
   }
}

__global__ void k_zero(double* __restrict__ a, long long n) {
   for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
      a[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Transposed operator of one layer (reference Tree_multiply_transpose,
// Tree.c:769-832).  Points side forward: u = U^T x, h = UTI^T u, DFT.  Any
// block size gives the same results (one wave per row of U^T, one thread
// per grid value); the coarse levels, with few large cubes, run 16 waves
// per cube.
template <int n, typename T>
__global__ void __launch_bounds__(1024) k_farT_source(const TgtDesc* __restrict__ tgt, const T* __restrict__ UT,
                                                     const T* __restrict__ UTI, const double* __restrict__ x,
                                                     typename Cplx<T>::type* __restrict__ spec, int maxr) {
   using C2 = typename Cplx<T>::type;
   constexpr int N = FarSizes<n>::N, H = FarSizes<n>::H, PAD = FarSizes<n>::PAD, GP3 = FarSizes<n>::GP3;
   extern __shared__ __align__(16) unsigned char smem[];
   C2* A = (C2*)smem;
   C2* B = A + n * n * H;
   T* cs = (T*)(B + n * N * H);
   T* sn = cs + N;
   T* g = sn + N;
   T* u = g + GP3;
   const TgtDesc d = tgt[blockIdx.x];
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
   twiddle_table<T, N>(cs, sn);
   const T* ut = UT + d.u_off;
   for (int k = wave; k < d.r; k += nw) {
      const T* row = ut + (long long)k * d.nq;
      T a = 0;
      for (int i = lane; i < d.nq; i += 64) a += row[i] * (T)x[d.q0 + i];
      a = wave_sum(a);
      if (lane == 0) u[k] = a;
   }
   __syncthreads();
   const T* uti = UTI + d.uti_off;
   for (int q = tid; q < GP3; q += blockDim.x) {
      T a = 0;
      for (int k = 0; k < d.r; k++) a += uti[(long long)k * GP3 + q] * u[k];
      g[q] = a;
   }
   __syncthreads();
   grid_forward<n, T>(g, A, B, spec + (long long)d.spec_slot * PAD, cs, sn);
}

// Panels side backward: h = IDFT, w = PV_layer^T h, y[panels] += VT^T w.
// The cube's panels are cut into gridDim.y contiguous chunks, one block
// each (every block forms the same h and w; each output is written by one
// block with the arithmetic of the unsplit kernel), so that the coarse
// levels' few large cubes occupy more than one CU each.
template <int n, typename T>
__global__ void __launch_bounds__(256) k_farT_target(const SrcDesc* __restrict__ src, const T* __restrict__ VT,
                                                     const T* __restrict__ PV,
                                                     const typename Cplx<T>::type* __restrict__ acc, int layer,
                                                     double* __restrict__ yfar, double* __restrict__ yfar2,
                                                     int maxr) {
   using C2 = typename Cplx<T>::type;
   constexpr int N = FarSizes<n>::N, H = FarSizes<n>::H, PAD = FarSizes<n>::PAD, GP3 = FarSizes<n>::GP3;
   extern __shared__ __align__(16) unsigned char smem[];
   C2* S = (C2*)smem;
   C2* C1 = S + PAD;
   C2* Cc = C1 + n * N * H;
   T* cs = (T*)(Cc + n * n * H);
   T* sn = cs + N;
   T* h = sn + N;
   T* w = h + GP3;
   const SrcDesc d = src[blockIdx.x];
   const int chunk = (d.np + (int)gridDim.y - 1) / (int)gridDim.y;
   const int i0 = (int)blockIdx.y * chunk, i1 = i0 + chunk < d.np ? i0 + chunk : d.np;
   if (i0 >= i1) return;   // uniform over the block
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
   twiddle_table<T, N>(cs, sn);
   for (int b = tid; b < PAD; b += blockDim.x) S[b] = acc[(long long)d.spec_slot * PAD + b];
   __syncthreads();
   grid_backward<n, T>(S, C1, Cc, h, cs, sn);
   __syncthreads();
   // yfar2 set: both layers from the one inverse transform (single layer
   // into yfar, double layer into yfar2), each summed as alone
   const int nl = yfar2 ? 2 : 1;
   for (int li = 0; li < nl; li++) {
      const T* pv = PV + d.pv_off + ((yfar2 ? li : layer) ? (long long)d.r * GP3 : 0);
      for (int k = wave; k < d.r; k += nw) {
         const T* row = pv + (long long)k * GP3;
         T a = 0;
         for (int q = lane; q < GP3; q += 64) a += row[q] * h[q];
         a = wave_sum(a);
         if (lane == 0) w[li * d.r + k] = a;
      }
   }
   __syncthreads();
   const T* vt = VT + d.vt_off;
   for (int li = 0; li < nl; li++) {
      double* out = li ? yfar2 : yfar;
      for (int i = i0 + tid; i < i1; i += blockDim.x) {
         T a = 0;
         for (int k = 0; k < d.r; k++) a += vt[(long long)k * d.np + i] * w[li * d.r + k];
         out[d.p0 + i] += (double)a;
      }
   }
}

// Near field transposed: y[cols] += D_layer^T x[rows] (global atomics; columns
// are shared between neighbouring leaves).
// With y2: both layers (single into y, double into y2) from one pass over
// the block.
template <typename T>
__global__ void __launch_bounds__(256) k_nearT(const LeafDesc* __restrict__ leaves, const int* __restrict__ dcol,
                                               const T* __restrict__ D, const double* __restrict__ x, int layer,
                                               double* __restrict__ y, double* __restrict__ y2) {
   const LeafDesc L = leaves[blockIdx.x];
   const T* blk = D + L.d_off + ((layer && !y2) ? L.p : 0);
   for (int c = threadIdx.x; c < L.m; c += blockDim.x) {
      const T* colp = blk + (long long)c * 2 * L.p;
      double a = 0.0;
      for (int r = 0; r < L.p; r++) a += (double)colp[r] * x[L.q0 + r];
      atomicAdd(&y[dcol[L.col_off + c]], a);
      if (y2) {
         double b = 0.0;
         for (int r = 0; r < L.p; r++) b += (double)colp[L.p + r] * x[L.q0 + r];
         atomicAdd(&y2[dcol[L.col_off + c]], b);
      }
   }
}

// ---------------------------------------------------------------------------
// Explicit instantiations and launchers
template <typename T>
size_t far_source_lds(int n, int maxr) {
   int N = 2 * n - 1, H = N / 2 + 1;
   return sizeof(typename Cplx<T>::type) * (size_t)(n * n * H + n * N * H) +
          sizeof(T) * (size_t)(2 * N + n * n * n + 256 + 2 * maxr) + 64;
}
template <typename T>
size_t far_target_lds(int n, int maxr) {
   int N = 2 * n - 1, H = N / 2 + 1;
   return sizeof(typename Cplx<T>::type) * (size_t)(N * N * H + n * N * H + n * n * H) + sizeof(T) * (size_t)(2 * N + n * n * n + maxr) + 64;
}

#define FSVD_DISPATCH_N(n, CALL) \
   switch (n) {                  \
      case 3: CALL(3); break;    \
      case 4: CALL(4); break;    \
      case 5: CALL(5); break;    \
      case 6: CALL(6); break;    \
      case 7: CALL(7); break;    \
      case 8: CALL(8); break;    \
      case 9: CALL(9); break;    \
      default: throw std::runtime_error("unsupported grid size"); \
   }

// gfx950: 160 KiB of LDS per workgroup
static void check_lds(size_t lds) {
   if (lds > 163840) throw std::runtime_error("LDS request above 160 KiB per workgroup");
}

template <typename T>
void launch_far_translate(int nblk, const TgtDesc* tgt, const int* isrc, const int* itoff, const void* Tspec,
                          const void* spec, void* acc, int pad, hipStream_t s) {
   // one block per 256 bins of each cube (the coarse levels have few cubes)
   const int slices = std::max(1, std::min(8, (pad + 255) / 256));
   hipLaunchKernelGGL((k_far_translate<T>), dim3(nblk, slices), dim3(256), 0, s, tgt, isrc, itoff,
                      (const typename Cplx<T>::type*)Tspec, (const typename Cplx<T>::type*)spec,
                      (typename Cplx<T>::type*)acc, pad);
}

void launch_translate_pipe(int nblk, int mode, const int2* range, const int4* tasks, const TpTile* tiles,
                           const TpOct* octs, const long long* oacc, const int* gsrc, const int* gint, hipStream_t s,
                           int ncol, long long sdel, long long adel, long long scs, long long acs) {
   // two LDS buffers per persistent block, three counters
   constexpr size_t lds = 2 * (size_t)TP_BUF + 16;
   check_lds(lds);
   static std::once_flag attr;
   std::call_once(attr, [] {
      FSVD_CHECK(hipFuncSetAttribute((const void*)k_translate_pipe<0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds));
      FSVD_CHECK(hipFuncSetAttribute((const void*)k_translate_pipe<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds));
   });
   if (nblk <= 0) return;
   const int g = col_grid(nblk, ncol);
   if (mode == 1)
      hipLaunchKernelGGL(k_translate_pipe<1>, dim3(g), dim3(TP_NT), lds, s, range, tasks, tiles, octs, oacc, gsrc,
                         gint, nblk, ncol, sdel, adel, scs, acs);
   else
      hipLaunchKernelGGL(k_translate_pipe<0>, dim3(g), dim3(TP_NT), lds, s, range, tasks, tiles, octs, oacc, gsrc,
                         gint, nblk, ncol, sdel, adel, scs, acs);
}

template <typename T>
void launch_apply_stage(int stage, const ApplyArgs& a, hipStream_t s) {
   using C2 = typename Cplx<T>::type;
   if (a.max_n > 9) throw std::runtime_error("grid points per side above 9");
   switch (stage) {
      case 0: {
         const int wl = (int)((sizeof(T) * (size_t)(2 * a.proj_maxcols + 128) + 15) / 16 * 16);
         check_lds(4 * (size_t)wl);
         if (a.nchunk && a.ncol == 1) {
            hipLaunchKernelGGL((k_project<T, true>), dim3((a.nchunk + 3) / 4), dim3(256), 4 * (size_t)wl, s,
                               a.chunks, a.nchunk, (const T*)a.VTc, a.xs, a.xd, (T*)a.tpart, wl, 1, a.cs_x,
                               a.cs_tpart);
         } else if (a.nchunk) {
            constexpr int NC = proj_nc<T>();
            const int wlm = (int)((sizeof(T) * (size_t)NC * (2 * a.proj_maxcols + 128) + 15) / 16 * 16);
            check_lds(4 * (size_t)wlm);
            static std::once_flag attr;
            std::call_once(attr, [] {
               FSVD_CHECK(hipFuncSetAttribute((const void*)k_project_mc<T, NC>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, 163840));
            });
            const int ngrp = (a.ncol + NC - 1) / NC;
            hipLaunchKernelGGL((k_project_mc<T, NC>), dim3(col_grid((a.nchunk + 3) / 4, ngrp)), dim3(256),
                               4 * (size_t)wlm, s, a.chunks, a.nchunk, (const T*)a.VTc, a.xs, a.xd, (T*)a.tpart,
                               wlm, a.ncol, a.cs_x, a.cs_tpart);
         }
         break;
      }
      case 1: {
         const int wl = (int)((source_cube_lds<T>(4, a.maxr_src, 64) + 15) / 16 * 16);
         size_t lds = (size_t)(CUBE_NT / 64) * wl;
         for (int n = 3; n <= a.max_n; n++)
            if (n != 4) lds = std::max(lds, source_cube_lds<T>(n, a.maxr_src, CUBE_NT));
         check_lds(lds);
         static std::once_flag attr;   // large n needs more than the default dynamic LDS
         std::call_once(attr, [] {
            FSVD_CHECK(hipFuncSetAttribute((const void*)k_source<T, true>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 163840));
            FSVD_CHECK(hipFuncSetAttribute((const void*)k_source<T, false>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 163840));
         });
         const int nb = a.nsrc_big + (a.nsrc - a.nsrc_big + CUBE_NT / 64 - 1) / (CUBE_NT / 64);
         if (nb) {
            auto kern = a.ncol == 1 ? k_source<T, true> : k_source<T, false>;
            hipLaunchKernelGGL(kern, dim3(col_grid(nb, a.ncol)), dim3(CUBE_NT), lds, s, a.src, a.nsrc_big, a.nsrc,
                               (const T*)a.tpart, (const T*)a.PV, (C2*)a.spec, wl, nb, a.ncol, a.cs_tpart,
                               a.cs_spec);
         }
         break;
      }
      case 2: {
         if constexpr (std::is_same<T, float>::value) {
            if (a.ncol >= 2 && a.nblk_p2 > 0)
               launch_translate_pipe2(a.nblk_p2, a.range_p2, a.task_p2, a.tiles_p2, a.oct_p2, a.oacc_p2, a.gsrc_p2,
                                      a.gint_p2, s, a.ncol, a.spec_delta, a.acc_delta,
                                      a.cs_spec * (long long)sizeof(C2), a.cs_acc * (long long)sizeof(C2));
            else
               launch_translate_pipe(a.nblk_p, a.mode_p, a.range_p, a.task_p, a.tiles_p, a.oct_p, a.oacc_p,
                                     a.gsrc_p, a.gint_p, s, a.ncol, a.spec_delta, a.acc_delta,
                                     a.cs_spec * (long long)sizeof(C2), a.cs_acc * (long long)sizeof(C2));
         } else {
            // LDS: T spectra and at most TR_MAX_SRC source spectra of BC bins
            constexpr int BC = TR_BINS_F64, LPT = TR_LPT_F64;
            const size_t lds_max = sizeof(C2) * (size_t)(343 + TR_MAX_SRC) * BC;
            const size_t lds = sizeof(C2) * (size_t)(343 + a.tr_max_src) * BC;
            check_lds(lds_max);
            if (a.tr_max_src > TR_MAX_SRC) throw std::runtime_error("translation tile above its LDS bound");
            static std::once_flag attr;
            std::call_once(attr, [lds_max] {
               FSVD_CHECK(hipFuncSetAttribute((const void*)k_translate<T, BC, LPT>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_max));
            });
            if (a.ntask_d)
               hipLaunchKernelGGL((k_translate<T, BC, LPT>), dim3(col_grid(a.ntask_d, a.ncol)), dim3(TR_TILE * LPT),
                                  lds, s, a.task_d, a.tiles, a.gtgt, a.gsrc, a.gint, (const C2*)a.Tsp,
                                  (const C2*)a.spec, (C2*)a.acc, a.ntask_d, a.ncol, a.cs_spec, a.cs_acc);
         }
         break;
      }
      case 3: {
         const int wl = (int)((target_cube_lds<T>(4) + 15) / 16 * 16);
         size_t lds = (size_t)(CUBE_NT / 64) * wl;
         for (int n = 3; n <= a.max_n; n++)
            if (n != 4) lds = std::max(lds, target_cube_lds<T>(n));
         check_lds(lds);
         static std::once_flag attr;   // large n needs more than the default dynamic LDS
         std::call_once(attr, [] {
            FSVD_CHECK(hipFuncSetAttribute((const void*)k_target<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           163840));
         });
         const int nb = a.ntgt_big + (a.ntgt - a.ntgt_big + CUBE_NT / 64 - 1) / (CUBE_NT / 64);
         if (nb && (a.ncol == 1 || sizeof(T) == 4)) {
            // FP32 batches run one block copy per column: the FP32 UTI rows
            // are cheap to re-read from L2, and the transforms run in
            // parallel (0.033 against 0.042 ms per column for four columns
            // per block on the 100k-panel globule)
            hipLaunchKernelGGL((k_target<T>), dim3(col_grid(nb, a.ncol)), dim3(CUBE_NT), lds, s, a.tgt, a.ntgt_big,
                               a.ntgt, (const T*)a.UTI, (const C2*)a.acc, (T*)a.w, wl, nb, a.ncol, a.cs_acc, a.cs_w);
         } else if (nb) {
            // FP64: TGT_NC columns per block, their grid values beside the
            // transform buffers (0.052 against 0.063 ms per column)
            const size_t extra4 = sizeof(T) * (size_t)(TGT_NC - 1) * 64;
            const int wlm = (int)((target_cube_lds<T>(4) + extra4 + 15) / 16 * 16);
            size_t lmc = (size_t)(CUBE_NT / 64) * wlm;
            for (int n = 3; n <= a.max_n; n++)
               if (n != 4)
                  lmc = std::max(lmc, target_cube_lds<T>(n) + sizeof(T) * (size_t)(TGT_NC - 1) * n * n * n);
            check_lds(lmc);
            static std::once_flag attr_mc;
            std::call_once(attr_mc, [] {
               FSVD_CHECK(hipFuncSetAttribute((const void*)k_target_mc<T>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, 163840));
            });
            const int ngrp = (a.ncol + TGT_NC - 1) / TGT_NC;
            hipLaunchKernelGGL((k_target_mc<T>), dim3(col_grid(nb, ngrp)), dim3(CUBE_NT), lmc, s, a.tgt, a.ntgt_big,
                               a.ntgt, (const T*)a.UTI, (const C2*)a.acc, (T*)a.w, wlm, nb, a.ncol, a.cs_acc,
                               a.cs_w);
         }
         break;
      }
      case 4: {
         size_t lds = sizeof(T) * (size_t)(2 * a.maxm + 256 + 3 * a.maxp) + 64;
         check_lds(lds);
         if (a.nleaf && a.ncol == 1) {
            hipLaunchKernelGGL((k_leaf<T, true>), dim3(a.nmap_leaf), dim3(256), lds, s, a.leaves, a.dcol,
                               (const T*)a.D, a.xs, a.xd, a.ublk, (const T*)a.UT, (const T*)a.w, a.y, a.map_leaf,
                               a.nmap_leaf, 1, a.cs_x, a.cs_w, a.cs_y, a.sp_out, a.sp_corr, a.sp_y);
         } else if (a.nleaf && sizeof(T) * (size_t)LEAF_NC * (2 * a.maxm + 256 + 3 * a.maxp) + 64 > 163840) {
            // leaves too wide for LEAF_NC columns in LDS: one block copy per column
            hipLaunchKernelGGL((k_leaf<T, false>), dim3(col_grid(a.nmap_leaf, a.ncol)), dim3(256), lds, s, a.leaves,
                               a.dcol, (const T*)a.D, a.xs, a.xd, a.ublk, (const T*)a.UT, (const T*)a.w, a.y,
                               a.map_leaf, a.nmap_leaf, a.ncol, a.cs_x, a.cs_w, a.cs_y, nullptr, nullptr, nullptr);
         } else if (a.nleaf) {
            // LEAF_NC columns per block
            const size_t lmc = sizeof(T) * (size_t)LEAF_NC * (2 * a.maxm + 256 + 3 * a.maxp) + 64;
            check_lds(lmc);
            static std::once_flag attr;
            std::call_once(attr, [] {
               FSVD_CHECK(hipFuncSetAttribute((const void*)k_leaf_mc<T, LEAF_NC>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, 163840));
            });
            const int ngrp = (a.ncol + LEAF_NC - 1) / LEAF_NC;
            hipLaunchKernelGGL((k_leaf_mc<T, LEAF_NC>), dim3(col_grid(a.nmap_leaf, ngrp)), dim3(256), lmc, s,
                               a.leaves, a.dcol, (const T*)a.D, a.xs, a.xd, a.ublk, (const T*)a.UT, (const T*)a.w,
                               a.y, a.map_leaf, a.nmap_leaf, a.ncol, a.cs_x, a.cs_w, a.cs_y);
         }
         break;
      }
      default: break;
   }
}

template <typename T>
__global__ void k_transpose_batched(const TransDesc* __restrict__ d) {
   const TransDesc t = d[blockIdx.x];
   const T* src = (const T*)t.src;
   T* dst = (T*)t.dst;
   const int total = t.rows * t.cols;
   for (int e = threadIdx.x; e < total; e += blockDim.x) {
      int j = e / t.rows, i = e - j * t.rows;
      dst[e] = src[(long long)i * t.cols + j];
   }
}

template <typename T>
void launch_transpose_batched(int n, const TransDesc* d, hipStream_t s) {
   for (int b0 = 0; b0 < n; b0 += 65535)
      hipLaunchKernelGGL((k_transpose_batched<T>), dim3(std::min(65535, n - b0)), dim3(256), 0, s, d + b0);
}

template <typename T>
void launch_farT_source(int n, int nblk, const TgtDesc* tgt, const T* UT, const T* UTI, const double* x, void* spec,
                        int maxr, hipStream_t s) {
   size_t lds = far_source_lds<T>(n, maxr);
   check_lds(lds);
   const int threads = n >= 5 ? 1024 : 256;
#define CALL(NN)                                                                                                    \
   hipLaunchKernelGGL((k_farT_source<NN, T>), dim3(nblk), dim3(threads), lds, s, tgt, UT, UTI, x,                  \
                      (typename Cplx<T>::type*)spec, maxr)
   FSVD_DISPATCH_N(n, CALL)
#undef CALL
}

template <typename T>
void launch_farT_target(int n, int nblk, const SrcDesc* src, const T* VT, const T* PV, const void* acc, int layer,
                        double* yfar, double* yfar2, int maxr, hipStream_t s) {
   size_t lds = far_target_lds<T>(n, yfar2 ? 2 * maxr : maxr);
   check_lds(lds);
   // panel chunks per cube: the coarse levels (more grid points) have few,
   // large cubes
   const int split = n >= 6 ? 8 : (n == 5 ? 4 : 1);
#define CALL(NN)                                                                                                    \
   hipLaunchKernelGGL((k_farT_target<NN, T>), dim3(nblk, split), dim3(256), lds, s, src, VT, PV,                   \
                      (const typename Cplx<T>::type*)acc, layer, yfar, yfar2, maxr)
   FSVD_DISPATCH_N(n, CALL)
#undef CALL
}

template <typename T>
void launch_nearT(int nleaf, const LeafDesc* leaves, const int* dcol, const T* D, const double* x, int layer,
                  double* y, double* y2, hipStream_t s) {
   hipLaunchKernelGGL((k_nearT<T>), dim3(nleaf), dim3(256), 0, s, leaves, dcol, D, x, layer, y, y2);
}

void launch_zero(double* a, long long n, hipStream_t s) {
   int blocks = (int)std::min<long long>((n + 255) / 256, 4096);
   if (blocks < 1) blocks = 1;
   hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(256), 0, s, a, n);
}

#define FSVD_INST(T)                                                                                           \
   template void launch_far_translate<T>(int, const TgtDesc*, const int*, const int*, const void*, const void*,      \
                                         void*, int, hipStream_t);                                                   \
   template void launch_apply_stage<T>(int, const ApplyArgs&, hipStream_t);                                          \
   template void launch_transpose_batched<T>(int, const TransDesc*, hipStream_t);                                    \
   template void launch_farT_source<T>(int, int, const TgtDesc*, const T*, const T*, const double*, void*, int,       \
                                       hipStream_t);                                                                  \
   template void launch_farT_target<T>(int, int, const SrcDesc*, const T*, const T*, const void*, int, double*,     \
                                       double*, int,                                                              \
                                       hipStream_t);                                                                  \
   template void launch_nearT<T>(int, const LeafDesc*, const int*, const T*, const double*, int, double*, double*,   \
                                 hipStream_t);
FSVD_INST(float)
FSVD_INST(double)

}  // namespace fsvd
