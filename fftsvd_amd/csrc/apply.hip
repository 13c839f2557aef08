// Matrix-vector product of one compressed FFT-SVD operator on the GPU.
//
//   y = D_s xs + D_d xd  (near field, dense per-leaf blocks)
//     + U UTI IFFT( sum_c T[d-c] . FFT(PV_s VT xs_c + PV_d VT xd_c) )  (far field)
//
// Storage precision T is float (mixed mode) or double (FP64 mode); vectors
// enter and leave in FP64.  Kernels:
//   k_far_source<n,T>  per source cube: VT projection, PV interpolation onto
//                      the n^3 grid, zero-padded r2c DFT -> spectrum
//   k_far_translate<T> per target cube: sum over the interaction list of
//                      T[offset] * spectrum (complex multiply-accumulate)
//   k_far_target<n,T>  per target cube: cropped c2r DFT, UTI, U, += y_far
//   k_near<T>          per leaf: dense block apply, adds y_far, writes y
// The transposed operator reuses the same pieces with the roles swapped.
#include <hip/hip_runtime.h>

#include <stdexcept>

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

// e^{-2 pi i m / N} table
template <typename T, int N>
__device__ __forceinline__ void twiddle_table(T* cs, T* sn) {
   for (int m = threadIdx.x; m < N; m += blockDim.x) {
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
template <int n, typename T>
__device__ void grid_forward(const T* g, typename Cplx<T>::type* A, typename Cplx<T>::type* B,
                             typename Cplx<T>::type* out, const T* cs, const T* sn) {
   constexpr int N = 2 * n - 1, H = N / 2 + 1, OFF = n / 2;
   const int tid = threadIdx.x, nt = blockDim.x;
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
   __syncthreads();
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
   __syncthreads();
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
template <int n, typename T>
__device__ void grid_backward(const typename Cplx<T>::type* S, typename Cplx<T>::type* C1,
                              typename Cplx<T>::type* C2, T* h, const T* cs, const T* sn) {
   constexpr int N = 2 * n - 1, H = N / 2 + 1, OFF = n / 2;
   const int tid = threadIdx.x, nt = blockDim.x;
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
   __syncthreads();
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
   __syncthreads();
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

// Source side.  Block per source cube; LDS: twiddles, t_s/t_d (r), grid g,
// DFT work arrays.
template <int n, typename T>
__global__ void __launch_bounds__(256) k_far_source(const SrcDesc* __restrict__ src, const T* __restrict__ VT,
                                                    const T* __restrict__ PV, const double* __restrict__ xs,
                                                    const double* __restrict__ xd,
                                                    typename Cplx<T>::type* __restrict__ spec, int maxr) {
   using C2 = typename Cplx<T>::type;
   constexpr int N = FarSizes<n>::N, H = FarSizes<n>::H, PAD = FarSizes<n>::PAD, GP3 = FarSizes<n>::GP3;
   extern __shared__ __align__(16) unsigned char smem[];
   C2* A = (C2*)smem;                     // n*n*H
   C2* B = A + n * n * H;                 // n*N*H
   T* cs = (T*)(B + n * N * H);           // N
   T* sn = cs + N;                        // N
   T* g = sn + N;                         // GP3
   T* ts = g + GP3;                       // maxr
   T* td = ts + maxr;                     // maxr
   const SrcDesc d = src[blockIdx.x];
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
   twiddle_table<T, N>(cs, sn);
   const T* vt = VT + d.vt_off;
   for (int k = wave; k < d.r; k += nw) {
      const T* row = vt + (long long)k * d.np;
      T as = 0, ad = 0;
      for (int i = lane; i < d.np; i += 64) {
         T v = row[i];
         as += v * (T)xs[d.p0 + i];
         ad += v * (T)xd[d.p0 + i];
      }
      as = wave_sum(as);
      ad = wave_sum(ad);
      if (lane == 0) {
         ts[k] = as;
         td[k] = ad;
      }
   }
   __syncthreads();
   const T* pvs = PV + d.pv_off;
   const T* pvd = pvs + (long long)d.r * GP3;
   for (int q = tid; q < GP3; q += blockDim.x) {
      T acc = 0;
      for (int k = 0; k < d.r; k++) acc += pvs[(long long)k * GP3 + q] * ts[k];
      for (int k = 0; k < d.r; k++) acc += pvd[(long long)k * GP3 + q] * td[k];
      g[q] = acc;
   }
   __syncthreads();
   grid_forward<n, T>(g, A, B, spec + (long long)d.spec_slot * PAD, cs, sn);
}

// Translation: S_d[b] = sum_e T[toff_e][b] * G[src_e][b].  One block per
// target cube, threads over frequency bins.
template <typename T>
__global__ void __launch_bounds__(256) k_far_translate(const TgtDesc* __restrict__ tgt,
                                                       const int* __restrict__ int_src,
                                                       const int* __restrict__ int_toff,
                                                       const typename Cplx<T>::type* __restrict__ Tspec,
                                                       const typename Cplx<T>::type* __restrict__ spec,
                                                       typename Cplx<T>::type* __restrict__ acc, int pad) {
   const TgtDesc d = tgt[blockIdx.x];
   for (int b = threadIdx.x; b < pad; b += blockDim.x) {
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

// Target side.  Block per target cube: inverse DFT, UTI, U, accumulate.
template <int n, typename T>
__global__ void __launch_bounds__(256) k_far_target(const TgtDesc* __restrict__ tgt, const T* __restrict__ UT,
                                                    const T* __restrict__ UTI,
                                                    const typename Cplx<T>::type* __restrict__ acc,
                                                    double* __restrict__ yfar, int maxr) {
   using C2 = typename Cplx<T>::type;
   constexpr int N = FarSizes<n>::N, H = FarSizes<n>::H, PAD = FarSizes<n>::PAD, GP3 = FarSizes<n>::GP3;
   extern __shared__ __align__(16) unsigned char smem[];
   C2* S = (C2*)smem;             // PAD
   C2* C1 = S + PAD;              // n*N*H
   C2* Cc = C1 + n * N * H;       // n*n*H
   T* cs = (T*)(Cc + n * n * H);  // N
   T* sn = cs + N;
   T* h = sn + N;                 // GP3
   T* w = h + GP3;                // maxr
   const TgtDesc d = tgt[blockIdx.x];
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
   twiddle_table<T, N>(cs, sn);
   for (int b = tid; b < PAD; b += blockDim.x) S[b] = acc[(long long)d.spec_slot * PAD + b];
   __syncthreads();
   grid_backward<n, T>(S, C1, Cc, h, cs, sn);
   __syncthreads();
   const T* uti = UTI + d.uti_off;
   for (int k = wave; k < d.r; k += nw) {
      const T* row = uti + (long long)k * GP3;
      T a = 0;
      for (int q = lane; q < GP3; q += 64) a += row[q] * h[q];
      a = wave_sum(a);
      if (lane == 0) w[k] = a;
   }
   __syncthreads();
   const T* ut = UT + d.u_off;
   for (int i = tid; i < d.nq; i += blockDim.x) {
      T a = 0;
      for (int k = 0; k < d.r; k++) a += ut[(long long)k * d.nq + i] * w[k];
      yfar[d.q0 + i] += (double)a;
   }
}

// Near field: block per leaf.  x gathered to LDS in storage precision; the
// 2*p*m block is streamed with unit stride and reduced into LDS rows.
template <typename T>
__global__ void __launch_bounds__(256) k_near(const LeafDesc* __restrict__ leaves, const int* __restrict__ dcol,
                                              const T* __restrict__ D, const double* __restrict__ xs,
                                              const double* __restrict__ xd, const double* __restrict__ yfar,
                                              double* __restrict__ y) {
   extern __shared__ __align__(16) unsigned char smem[];
   const LeafDesc L = leaves[blockIdx.x];
   T* xg = (T*)smem;        // 2*m
   T* yr = xg + 2 * L.m;    // p
   const int tid = threadIdx.x, nt = blockDim.x;
   for (int c = tid; c < L.m; c += nt) {
      int col = dcol[L.col_off + c];
      xg[c] = (T)xs[col];
      xg[L.m + c] = (T)xd[col];
   }
   for (int r = tid; r < L.p; r += nt) yr[r] = 0;
   __syncthreads();
   const int E = 2 * L.p;
   const long long total = (long long)E * L.m;
   const T* blk = D + L.d_off;
   int c = tid / E, r2 = tid - (tid / E) * E;
   const int dc = nt / E, dr = nt - dc * E;
   for (long long e = tid; e < total; e += nt) {
      int layer = r2 >= L.p;
      int row = r2 - layer * L.p;
      atomicAdd(&yr[row], blk[e] * xg[layer * L.m + c]);
      c += dc;
      r2 += dr;
      if (r2 >= E) {
         r2 -= E;
         c++;
      }
   }
   __syncthreads();
   for (int r = tid; r < L.p; r += nt) y[L.q0 + r] = (double)yr[r] + yfar[L.q0 + r];
}

__global__ void k_zero(double* __restrict__ a, long long n) {
   for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
      a[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Transposed operator of one layer (reference Tree_multiply_transpose,
// Tree.c:769-832).  Points side forward: u = U^T x, h = UTI^T u, DFT.
template <int n, typename T>
__global__ void __launch_bounds__(256) k_farT_source(const TgtDesc* __restrict__ tgt, const T* __restrict__ UT,
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
template <int n, typename T>
__global__ void __launch_bounds__(256) k_farT_target(const SrcDesc* __restrict__ src, const T* __restrict__ VT,
                                                     const T* __restrict__ PV,
                                                     const typename Cplx<T>::type* __restrict__ acc, int layer,
                                                     double* __restrict__ yfar, int maxr) {
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
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
   twiddle_table<T, N>(cs, sn);
   for (int b = tid; b < PAD; b += blockDim.x) S[b] = acc[(long long)d.spec_slot * PAD + b];
   __syncthreads();
   grid_backward<n, T>(S, C1, Cc, h, cs, sn);
   __syncthreads();
   const T* pv = PV + d.pv_off + (layer ? (long long)d.r * GP3 : 0);
   for (int k = wave; k < d.r; k += nw) {
      const T* row = pv + (long long)k * GP3;
      T a = 0;
      for (int q = lane; q < GP3; q += 64) a += row[q] * h[q];
      a = wave_sum(a);
      if (lane == 0) w[k] = a;
   }
   __syncthreads();
   const T* vt = VT + d.vt_off;
   for (int i = tid; i < d.np; i += blockDim.x) {
      T a = 0;
      for (int k = 0; k < d.r; k++) a += vt[(long long)k * d.np + i] * w[k];
      yfar[d.p0 + i] += (double)a;
   }
}

// Near field transposed: y[cols] += D_layer^T x[rows] (global atomics; columns
// are shared between neighbouring leaves).
template <typename T>
__global__ void __launch_bounds__(256) k_nearT(const LeafDesc* __restrict__ leaves, const int* __restrict__ dcol,
                                               const T* __restrict__ D, const double* __restrict__ x, int layer,
                                               double* __restrict__ y) {
   const LeafDesc L = leaves[blockIdx.x];
   const T* blk = D + L.d_off + (layer ? L.p : 0);
   for (int c = threadIdx.x; c < L.m; c += blockDim.x) {
      const T* colp = blk + (long long)c * 2 * L.p;
      double a = 0.0;
      for (int r = 0; r < L.p; r++) a += (double)colp[r] * x[L.q0 + r];
      atomicAdd(&y[dcol[L.col_off + c]], a);
   }
}

// ---------------------------------------------------------------------------
// Explicit instantiations and launchers
template <typename T>
size_t far_source_lds(int n, int maxr) {
   int N = 2 * n - 1, H = N / 2 + 1;
   return sizeof(typename Cplx<T>::type) * (size_t)(n * n * H + n * N * H) + sizeof(T) * (size_t)(2 * N + n * n * n + 2 * maxr) + 64;
}
template <typename T>
size_t far_target_lds(int n, int maxr) {
   int N = 2 * n - 1, H = N / 2 + 1;
   return sizeof(typename Cplx<T>::type) * (size_t)(N * N * H + n * N * H + n * n * H) + sizeof(T) * (size_t)(2 * N + n * n * n + maxr) + 64;
}

#define FSVD_DISPATCH_N(n, CALL) \
   switch (n) {                  \
      case 4: CALL(4); break;    \
      case 5: CALL(5); break;    \
      case 6: CALL(6); break;    \
      case 7: CALL(7); break;    \
      case 8: CALL(8); break;    \
      case 9: CALL(9); break;    \
      default: throw std::runtime_error("unsupported grid size"); \
   }

static void check_lds(size_t lds) {
   if (lds > 65536) throw std::runtime_error("LDS request above 64 KiB per workgroup");
}

template <typename T>
void launch_far_source(int n, int nblk, const SrcDesc* src, const T* VT, const T* PV, const double* xs,
                       const double* xd, void* spec, int maxr, hipStream_t s) {
   size_t lds = far_source_lds<T>(n, maxr);
   check_lds(lds);
#define CALL(NN)                                                                                                    \
   hipLaunchKernelGGL((k_far_source<NN, T>), dim3(nblk), dim3(256), lds, s, src, VT, PV, xs, xd,                  \
                      (typename Cplx<T>::type*)spec, maxr)
   FSVD_DISPATCH_N(n, CALL)
#undef CALL
}

template <typename T>
void launch_far_translate(int nblk, const TgtDesc* tgt, const int* isrc, const int* itoff, const void* Tspec,
                          const void* spec, void* acc, int pad, hipStream_t s) {
   hipLaunchKernelGGL((k_far_translate<T>), dim3(nblk), dim3(256), 0, s, tgt, isrc, itoff,
                      (const typename Cplx<T>::type*)Tspec, (const typename Cplx<T>::type*)spec,
                      (typename Cplx<T>::type*)acc, pad);
}

template <typename T>
void launch_far_target(int n, int nblk, const TgtDesc* tgt, const T* UT, const T* UTI, const void* acc,
                       double* yfar, int maxr, hipStream_t s) {
   size_t lds = far_target_lds<T>(n, maxr);
   check_lds(lds);
#define CALL(NN)                                                                                                    \
   hipLaunchKernelGGL((k_far_target<NN, T>), dim3(nblk), dim3(256), lds, s, tgt, UT, UTI,                          \
                      (const typename Cplx<T>::type*)acc, yfar, maxr)
   FSVD_DISPATCH_N(n, CALL)
#undef CALL
}

template <typename T>
void launch_near(int nleaf, const LeafDesc* leaves, const int* dcol, const T* D, const double* xs, const double* xd,
                 const double* yfar, double* y, int maxm, int maxp, hipStream_t s) {
   size_t lds = sizeof(T) * (size_t)(2 * maxm + maxp) + 64;
   check_lds(lds);
   hipLaunchKernelGGL((k_near<T>), dim3(nleaf), dim3(256), lds, s, leaves, dcol, D, xs, xd, yfar, y);
}

template <typename T>
void launch_farT_source(int n, int nblk, const TgtDesc* tgt, const T* UT, const T* UTI, const double* x, void* spec,
                        int maxr, hipStream_t s) {
   size_t lds = far_source_lds<T>(n, maxr);
   check_lds(lds);
#define CALL(NN)                                                                                                    \
   hipLaunchKernelGGL((k_farT_source<NN, T>), dim3(nblk), dim3(256), lds, s, tgt, UT, UTI, x,                      \
                      (typename Cplx<T>::type*)spec, maxr)
   FSVD_DISPATCH_N(n, CALL)
#undef CALL
}

template <typename T>
void launch_farT_target(int n, int nblk, const SrcDesc* src, const T* VT, const T* PV, const void* acc, int layer,
                        double* yfar, int maxr, hipStream_t s) {
   size_t lds = far_target_lds<T>(n, maxr);
   check_lds(lds);
#define CALL(NN)                                                                                                    \
   hipLaunchKernelGGL((k_farT_target<NN, T>), dim3(nblk), dim3(256), lds, s, src, VT, PV,                          \
                      (const typename Cplx<T>::type*)acc, layer, yfar, maxr)
   FSVD_DISPATCH_N(n, CALL)
#undef CALL
}

template <typename T>
void launch_nearT(int nleaf, const LeafDesc* leaves, const int* dcol, const T* D, const double* x, int layer,
                  double* y, hipStream_t s) {
   hipLaunchKernelGGL((k_nearT<T>), dim3(nleaf), dim3(256), 0, s, leaves, dcol, D, x, layer, y);
}

void launch_zero(double* a, long long n, hipStream_t s) {
   int blocks = (int)std::min<long long>((n + 255) / 256, 4096);
   if (blocks < 1) blocks = 1;
   hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(256), 0, s, a, n);
}

#define FSVD_INST(T)                                                                                                   \
   template void launch_far_source<T>(int, int, const SrcDesc*, const T*, const T*, const double*, const double*, void*, \
                                      int, hipStream_t);                                                              \
   template void launch_far_translate<T>(int, const TgtDesc*, const int*, const int*, const void*, const void*, void*, \
                                         int, hipStream_t);                                                           \
   template void launch_far_target<T>(int, int, const TgtDesc*, const T*, const T*, const void*, double*, int,        \
                                      hipStream_t);                                                                   \
   template void launch_near<T>(int, const LeafDesc*, const int*, const T*, const double*, const double*,              \
                                const double*, double*, int, int, hipStream_t);                                       \
   template void launch_farT_source<T>(int, int, const TgtDesc*, const T*, const T*, const double*, void*, int,       \
                                       hipStream_t);                                                                  \
   template void launch_farT_target<T>(int, int, const SrcDesc*, const T*, const T*, const void*, int, double*, int,  \
                                       hipStream_t);                                                                  \
   template void launch_nearT<T>(int, const LeafDesc*, const int*, const T*, const double*, int, double*, hipStream_t);
FSVD_INST(float)
FSVD_INST(double)

}  // namespace fsvd
