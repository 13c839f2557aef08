// Setup kernels: collocation integrals, point kernels, small batched GEMMs
// and the pivoted modified Gram-Schmidt that builds the sampled bases.
// All of it runs in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "device.h"
#include "geometry.h"
#include "integrals.h"
#include "op.h"

namespace fsvd {

// One workgroup per IntBlock; threads stride over the (point, panel) pairs.
__global__ void __launch_bounds__(256) k_integral_blocks(const IntBlock* __restrict__ blocks,
                                                         const double* __restrict__ bpts,
                                                         const int* __restrict__ bpan,
                                                         const PanelGeom* __restrict__ panels, int kernel,
                                                         double kappa, double* __restrict__ out) {
   const IntBlock b = blocks[blockIdx.x];
   const int total = b.npt * b.npan;
   for (int e = threadIdx.x; e < total; e += blockDim.x) {
      int i = e / b.npan;
      int j = e - i * b.npan;
      const double* x = bpts + 3 * (size_t)(b.pt_off + i);
      const PanelGeom& P = panels[bpan[b.pan_off + j]];
      double s, d;
      panel_integrals(kernel, kappa, x, P, s, d);
      long long o = b.out_off + (long long)i * b.ld_pt + (long long)j * b.ld_pan;
      if (b.layer_stride >= 0) {
         out[o] = s;
         out[o + b.layer_stride] = d;
      } else {
         out[o] = d;   // double layer only
      }
   }
}

__global__ void __launch_bounds__(256) k_green_blocks(const GreenBlock* __restrict__ blocks,
                                                      const double* __restrict__ dest,
                                                      const double* __restrict__ src, int kernel, double kappa,
                                                      double* __restrict__ out) {
   const GreenBlock b = blocks[blockIdx.x];
   const int total = b.ndest * b.nsrc;
   for (int e = threadIdx.x; e < total; e += blockDim.x) {
      int i = e / b.nsrc;
      int j = e - i * b.nsrc;
      double g = green(kernel, kappa, dest + 3 * (size_t)(b.dest_off + i), src + 3 * (size_t)(b.src_off + j));
      out[b.out_off + (long long)i * b.ld_dest + (long long)j * b.ld_src] = g;
   }
}

// C = A*B with arbitrary element strides, one descriptor per grid.y.
// 32x32 output tiles (grid.x strides over a problem's tiles), K staged in
// LDS chunks of 32; each thread owns a 2x2 block of the tile.  Each output
// is summed over k in increasing order.  Tile loads read along whichever
// operand dimension has unit stride.
__global__ void __launch_bounds__(256) k_gemm_batched(const GemmDesc* __restrict__ descs) {
   const GemmDesc d = descs[blockIdx.y];
   __shared__ double As[32][33], Bs[32][33];   // As[i][k], Bs[k][j]
   const int tm = (d.M + 31) / 32, tn = (d.N + 31) / 32;
   const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
   for (int tile = blockIdx.x; tile < tm * tn; tile += gridDim.x) {
      const int i0 = (tile / tn) * 32, j0 = (tile % tn) * 32;
      double c00 = 0.0, c01 = 0.0, c10 = 0.0, c11 = 0.0;
      for (int k0 = 0; k0 < d.K; k0 += 32) {
         for (int e = tid; e < 1024; e += 256) {
            int ii, kk;
            if (d.a_cs == 1) {
               ii = e >> 5;
               kk = e & 31;
            } else {
               kk = e >> 5;
               ii = e & 31;
            }
            const int gi = i0 + ii, gk = k0 + kk;
            As[ii][kk] = (gi < d.M && gk < d.K) ? d.A[(long long)gi * d.a_rs + (long long)gk * d.a_cs] : 0.0;
            int jj;
            if (d.b_rs == 1) {
               jj = e >> 5;
               kk = e & 31;
            } else {
               kk = e >> 5;
               jj = e & 31;
            }
            const int gj = j0 + jj, gk2 = k0 + kk;
            Bs[kk][jj] = (gj < d.N && gk2 < d.K) ? d.B[(long long)gk2 * d.b_rs + (long long)gj * d.b_cs] : 0.0;
         }
         __syncthreads();
#pragma unroll 8
         for (int kk = 0; kk < 32; kk++) {
            const double a0 = As[2 * ty][kk], a1 = As[2 * ty + 1][kk];
            const double b0 = Bs[kk][2 * tx], b1 = Bs[kk][2 * tx + 1];
            c00 += a0 * b0;
            c01 += a0 * b1;
            c10 += a1 * b0;
            c11 += a1 * b1;
         }
         __syncthreads();
      }
      const int i = i0 + 2 * ty, j = j0 + 2 * tx;
      if (i < d.M && j < d.N) d.C[(long long)i * d.c_rs + (long long)j * d.c_cs] = c00;
      if (i < d.M && j + 1 < d.N) d.C[(long long)i * d.c_rs + (long long)(j + 1) * d.c_cs] = c01;
      if (i + 1 < d.M && j < d.N) d.C[(long long)(i + 1) * d.c_rs + (long long)j * d.c_cs] = c10;
      if (i + 1 < d.M && j + 1 < d.N) d.C[(long long)(i + 1) * d.c_rs + (long long)(j + 1) * d.c_cs] = c11;
   }
}

// Column basis by pivoted MGS with two re-orthogonalisations and the rank
// cut ||col|| < eps*||A||_F (reference Matrix_columnbasis_pivotedmgs,
// Matrix.c:215-307), run as one pair of launches per pivot step over every
// problem of a batch:
// (problems small enough for LDS run whole in k_mgs_lds instead)
//   k_mgs_pivot   workgroup per problem: argmax of the column norms (ties to
//                 the lowest index), rank cut, column swap, q_k = a_k / |a_k|
//   k_mgs_update  seg lanes per remaining column: two passes of
//                 a -= (q_k . a) q_k, then the fresh squared norm of a
// Columns are contiguous (column-major A), so each segment reads and writes
// its column with unit stride; the dot products are segment reductions.

__device__ __forceinline__ double seg_sum(double v, int S) {
   for (int o = S >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
   return v;
}

template <int NR>
__global__ void __launch_bounds__(256) k_mgs_update(const MgsDesc* __restrict__ descs, int nprob, int total_waves,
                                                    double* __restrict__ Abuf, const double* __restrict__ Qbuf,
                                                    double* __restrict__ ss, const MgsState* __restrict__ st,
                                                    int k) {
   const int gw = (int)((blockIdx.x * (unsigned)blockDim.x + threadIdx.x) >> 6);
   if (gw >= total_waves) return;
   int lo = 0, hi = nprob - 1;
   while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (descs[mid].wave0 <= gw) lo = mid;
      else hi = mid - 1;
   }
   const MgsDesc d = descs[lo];
   if (k >= 0 && st[d.pid].done) return;  // uniform over the wave
   const int lane = threadIdx.x & 63, S = d.seg, sub = lane / S, sl = lane - sub * S;
   const int j = (gw - d.wave0) * (64 / S) + sub;
   const bool active = j < d.cols && j > k;
   const int rows = d.rows;
   double* a = Abuf + d.a_off + (long long)(active ? j : 0) * rows;
   const double* q = Qbuf + d.q_off + (long long)(k > 0 ? k : 0) * rows;
   const int nr = (rows + S - 1) / S;
   double s2 = 0.0;
   if (nr <= NR) {
      double v[NR];
#pragma unroll
      for (int t = 0; t < NR; t++) {
         int i = sl + t * S;
         v[t] = (active && t < nr && i < rows) ? a[i] : 0.0;
      }
      if (k >= 0) {
         for (int pass = 0; pass < 2; pass++) {
            double r = 0.0;
#pragma unroll
            for (int t = 0; t < NR; t++) {
               int i = sl + t * S;
               if (t < nr && i < rows) r = fma(q[i], v[t], r);
            }
            r = seg_sum(r, S);
#pragma unroll
            for (int t = 0; t < NR; t++) {
               int i = sl + t * S;
               if (t < nr && i < rows) v[t] = fma(-q[i], r, v[t]);
            }
         }
         if (active) {
#pragma unroll
            for (int t = 0; t < NR; t++) {
               int i = sl + t * S;
               if (t < nr && i < rows) a[i] = v[t];
            }
         }
      }
#pragma unroll
      for (int t = 0; t < NR; t++) s2 = fma(v[t], v[t], s2);
   } else {
      // long columns (S == 64): stream the column three times
      if (k >= 0) {
         double r1 = 0.0;
         for (int i = sl; i < rows; i += S) r1 = fma(q[i], active ? a[i] : 0.0, r1);
         r1 = seg_sum(r1, S);
         double r2 = 0.0;
         for (int i = sl; i < rows; i += S) r2 = fma(q[i], fma(-q[i], r1, active ? a[i] : 0.0), r2);
         r2 = seg_sum(r2, S);
         for (int i = sl; i < rows; i += S) {
            double v = fma(-q[i], r2, fma(-q[i], r1, active ? a[i] : 0.0));
            if (active) a[i] = v;
            s2 = fma(v, v, s2);
         }
      } else {
         for (int i = sl; i < rows; i += S) {
            double v = active ? a[i] : 0.0;
            s2 = fma(v, v, s2);
         }
      }
   }
   s2 = seg_sum(s2, S);
   if (active && sl == 0) ss[d.ss_off + j] = s2;
}

__global__ void __launch_bounds__(256) k_mgs_pivot(const MgsDesc* __restrict__ descs, double* __restrict__ Abuf,
                                                   double* __restrict__ Qbuf, const double* __restrict__ ss,
                                                   MgsState* __restrict__ st, int k, double eps) {
   const MgsDesc d = descs[blockIdx.x];
   const int p = d.pid;
   if (st[p].done) return;
   const int tid = threadIdx.x, nt = blockDim.x;
   __shared__ double rv[256];
   __shared__ int ri[256];
   double fro;
   if (k == 0) {
      double part = 0.0;
      for (int j = tid; j < d.cols; j += nt) part += ss[d.ss_off + j];
      rv[tid] = part;
      __syncthreads();
      for (int w = nt / 2; w > 0; w >>= 1) {
         if (tid < w) rv[tid] += rv[tid + w];
         __syncthreads();
      }
      fro = sqrt(rv[0]);
      __syncthreads();
      if (tid == 0) st[p].fro = fro;
   } else {
      fro = st[p].fro;
   }
   if (k >= d.cols || k >= d.rows) {
      // every column taken, or `rows` orthonormal columns found (the rest is
      // rounding noise far below any positive cut)
      if (tid == 0) {
         st[p].rank = k;
         st[p].done = 1;
      }
      return;
   }
   double bv = 0.0;
   int bi = 0x7fffffff;
   for (int j = k + tid; j < d.cols; j += nt) {
      double v = sqrt(ss[d.ss_off + j]);
      if (v > bv || (v == bv && j < bi)) {
         bv = v;
         bi = j;
      }
   }
   rv[tid] = bv;
   ri[tid] = bi;
   __syncthreads();
   for (int w = nt / 2; w > 0; w >>= 1) {
      if (tid < w) {
         double ov = rv[tid + w];
         int oi = ri[tid + w];
         if (ov > rv[tid] || (ov == rv[tid] && oi < ri[tid])) {
            rv[tid] = ov;
            ri[tid] = oi;
         }
      }
      __syncthreads();
   }
   const double mx = rv[0];
   const int best = mx > 0.0 ? ri[0] : k;
   if (k >= 1 && mx < eps * fro) {
      if (tid == 0) {
         st[p].rank = k;
         st[p].done = 1;
      }
      return;
   }
   double* A = Abuf + d.a_off;
   double* ak = A + (long long)k * d.rows;
   double* ab = A + (long long)best * d.rows;
   double* q = Qbuf + d.q_off + (long long)k * d.rows;
   for (int i = tid; i < d.rows; i += nt) {
      double vb = ab[i], vk = ak[i];
      ab[i] = vk;
      ak[i] = vb;
      q[i] = vb / mx;
   }
   if (tid == 0) st[p].rank = k + 1;
}

// Whole pivoted MGS of one small problem in LDS (one workgroup per
// problem): A is loaded once, with an odd column stride against bank
// conflicts; each thread owns whole columns, so every dot product and norm
// is a serial sum over the rows, as in the reference.  Writes the pivot
// columns q_k to Q and the rank to st.
__global__ void __launch_bounds__(256) k_mgs_lds(const MgsDesc* __restrict__ descs, const int* __restrict__ list,
                                                 const double* __restrict__ Abuf, double* __restrict__ Qbuf,
                                                 MgsState* __restrict__ st, double eps) {
   extern __shared__ double sm[];
   __shared__ double rv[256];
   __shared__ int ri[256];
   const int p = list[blockIdx.x];
   const MgsDesc d = descs[p];
   const int rows = d.rows, cols = d.cols, lda = rows | 1, tid = threadIdx.x, nt = blockDim.x;
   double* As = sm;                         // lda * cols
   double* cn = As + (size_t)lda * cols;    // cols: squared column norms
   const double* A = Abuf + d.a_off;
   for (int e = tid; e < rows * cols; e += nt) {
      const int j = e / rows, i = e - j * rows;
      As[j * lda + i] = A[e];
   }
   __syncthreads();
   double part = 0.0;
   for (int j = tid; j < cols; j += nt) {
      const double* a = As + j * lda;
      double sq = 0.0;
      for (int i = 0; i < rows; i++) sq += a[i] * a[i];
      cn[j] = sq;
      part += sq;
   }
   rv[tid] = part;
   __syncthreads();
   for (int w = nt / 2; w > 0; w >>= 1) {
      if (tid < w) rv[tid] += rv[tid + w];
      __syncthreads();
   }
   const double fro = sqrt(rv[0]);
   __syncthreads();
   int rank = 0;
   for (int k = 0;; k++) {
      if (k >= cols || k >= rows) {
         rank = k;
         break;
      }
      double bv = 0.0;
      int bi = 0x7fffffff;
      for (int j = k + tid; j < cols; j += nt) {
         const double v = sqrt(cn[j]);
         if (v > bv || (v == bv && j < bi)) {
            bv = v;
            bi = j;
         }
      }
      rv[tid] = bv;
      ri[tid] = bi;
      __syncthreads();
      for (int w = nt / 2; w > 0; w >>= 1) {
         if (tid < w) {
            const double ov = rv[tid + w];
            const int oi = ri[tid + w];
            if (ov > rv[tid] || (ov == rv[tid] && oi < ri[tid])) {
               rv[tid] = ov;
               ri[tid] = oi;
            }
         }
         __syncthreads();
      }
      const double mx = rv[0];
      const int best = mx > 0.0 ? ri[0] : k;
      __syncthreads();
      if (k >= 1 && mx < eps * fro) {
         rank = k;
         break;
      }
      double* ak = As + k * lda;
      double* ab = As + best * lda;
      double* q = Qbuf + d.q_off + (long long)k * rows;
      for (int i = tid; i < rows; i += nt) {
         const double vb = ab[i], vk = ak[i];
         ab[i] = vk;
         ak[i] = vb / mx;   // column k now holds q_k
         q[i] = vb / mx;
      }
      __syncthreads();
      for (int j = k + 1 + tid; j < cols; j += nt) {
         double* a = As + j * lda;
         for (int pass = 0; pass < 2; pass++) {
            double r = 0.0;
            for (int i = 0; i < rows; i++) r += ak[i] * a[i];
            for (int i = 0; i < rows; i++) a[i] -= ak[i] * r;
         }
         double sq = 0.0;
         for (int i = 0; i < rows; i++) sq += a[i] * a[i];
         cn[j] = sq;
      }
      __syncthreads();
   }
   if (tid == 0) {
      st[p].fro = fro;
      st[p].rank = rank;
      st[p].done = 1;
   }
}

__global__ void k_mgs_init(MgsState* __restrict__ st, int n) {
   int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) {
      st[i].fro = 0.0;
      st[i].rank = 0;
      st[i].done = 0;
   }
}

__global__ void k_f64_to_f32(const double* __restrict__ a, float* __restrict__ b, long long n) {
   for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
      b[i] = (float)a[i];
}

}  // namespace fsvd

namespace fsvd {

void launch_integral_blocks(int nblk, const IntBlock* blocks, const double* bpts, const int* bpan,
                            const PanelGeom* panels, int kernel, double kappa, double* out, hipStream_t s) {
   if (nblk > 0)
      hipLaunchKernelGGL(k_integral_blocks, dim3(nblk), dim3(256), 0, s, blocks, bpts, bpan, panels, kernel, kappa, out);
}

void launch_green_blocks(int nblk, const GreenBlock* blocks, const double* dest, const double* src, int kernel,
                         double kappa, double* out, hipStream_t s) {
   if (nblk > 0)
      hipLaunchKernelGGL(k_green_blocks, dim3(nblk), dim3(256), 0, s, blocks, dest, src, kernel, kappa, out);
}

void launch_gemm_batched(int nbatch, int maxtiles, const GemmDesc* descs, hipStream_t s) {
   if (nbatch > 0)
      hipLaunchKernelGGL(k_gemm_batched, dim3(maxtiles, nbatch), dim3(256), 0, s, descs);
}

void run_mgs(std::vector<MgsDesc>& descs, double* A, double* Q, std::vector<int>& ranks, double eps,
             hipStream_t s) {
   // Problems whose matrix fits in 96 KiB of LDS run whole in k_mgs_lds.
   // The rest go step by step over two update classes: columns up to 16*seg
   // rows (seg <= 64) held in 16 registers per lane, longer ones in 32
   // (streamed beyond 2048 rows).
   constexpr int NRS = 16, NRL = 32;
   constexpr size_t LDS_MAX = 96 * 1024;
   const int nprob = (int)descs.size();
   ranks.assign(nprob, 0);
   if (nprob == 0) return;
   std::vector<MgsDesc> cls[2], step;
   std::vector<int> small;
   size_t lds_small = 0;
   int waves[2] = {0, 0}, ncol = 0, kmax = 0;
   for (int p = 0; p < nprob; p++) {
      MgsDesc& d = descs[p];
      d.pid = p;
      d.ss_off = ncol;
      ncol += d.cols;
      const size_t lds = sizeof(double) * ((size_t)(d.rows | 1) + 1) * d.cols;
      if (lds <= LDS_MAX) {
         small.push_back(p);
         lds_small = std::max(lds_small, lds);
         continue;
      }
      const int c = d.rows > 64 * NRS ? 1 : 0;
      const int NR = c ? NRL : NRS;
      int S = 4;
      while (S < 64 && (d.rows + S - 1) / S > NR) S <<= 1;
      d.seg = S;
      d.wave0 = waves[c];
      waves[c] += (d.cols + 64 / S - 1) / (64 / S);
      kmax = std::max(kmax, std::min(d.rows, d.cols));
      cls[c].push_back(d);
      step.push_back(d);
   }
   DBuf<MgsDesc> dall, dd, dc[2];
   dall.upload(descs, s);
   dd.upload(step, s);
   for (int c = 0; c < 2; c++) dc[c].upload(cls[c], s);
   DBuf<double> ss;
   ss.alloc(std::max(ncol, 1));
   DBuf<MgsState> st;
   st.alloc(nprob);
   hipLaunchKernelGGL(k_mgs_init, dim3((nprob + 255) / 256), dim3(256), 0, s, st.p, nprob);
   DBuf<int> dsmall;
   if (!small.empty()) {
      static std::once_flag attr;
      std::call_once(attr, [] {
         FSVD_CHECK(hipFuncSetAttribute((const void*)k_mgs_lds, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)LDS_MAX));
      });
      dsmall.upload(small, s);
      for (size_t b0 = 0; b0 < small.size(); b0 += 65535)
         hipLaunchKernelGGL(k_mgs_lds, dim3((unsigned)std::min<size_t>(65535, small.size() - b0)), dim3(256),
                            lds_small, s, dall.p, dsmall.p + b0, A, Q, st.p, eps);
   }
   auto update = [&](int k) {
      if (!cls[0].empty())
         hipLaunchKernelGGL((k_mgs_update<NRS>), dim3((waves[0] + 3) / 4), dim3(256), 0, s, dc[0].p,
                            (int)cls[0].size(), waves[0], A, Q, ss.p, st.p, k);
      if (!cls[1].empty())
         hipLaunchKernelGGL((k_mgs_update<NRL>), dim3((waves[1] + 3) / 4), dim3(256), 0, s, dc[1].p,
                            (int)cls[1].size(), waves[1], A, Q, ss.p, st.p, k);
   };
   MgsState* h = nullptr;
   FSVD_CHECK(hipHostMalloc((void**)&h, sizeof(MgsState) * nprob, hipHostMallocDefault));
   int steps_done = 0;
   if (!step.empty()) {
      update(-1);
      for (int k = 0; k <= kmax; k++) {
         steps_done = k + 1;
         hipLaunchKernelGGL(k_mgs_pivot, dim3((unsigned)step.size()), dim3(256), 0, s, dd.p, A, Q, ss.p, st.p, k,
                            eps);
         if (k < kmax) update(k);
         if ((k & 7) == 7 || k == kmax) {
            FSVD_CHECK(hipMemcpyAsync(h, st.p, sizeof(MgsState) * nprob, hipMemcpyDeviceToHost, s));
            FSVD_CHECK(hipStreamSynchronize(s));
            bool all = true;
            for (const MgsDesc& d : step) all = all && h[d.pid].done != 0;
            if (all) break;
         }
      }
   }
   FSVD_CHECK(hipMemcpyAsync(h, st.p, sizeof(MgsState) * nprob, hipMemcpyDeviceToHost, s));
   FSVD_CHECK(hipStreamSynchronize(s));
   FSVD_CHECK(hipGetLastError());
   for (int i = 0; i < nprob; i++) ranks[descs[i].rank_slot] = h[i].rank;
   if (const char* e = std::getenv("FFTSVD_TIMING"); e && e[0] == '1') {
      int mr = 0, mc = 0, rk = 0;
      double bytes = 0;
      for (const MgsDesc& d : step) {
         mr = std::max(mr, d.rows);
         mc = std::max(mc, d.cols);
         rk = std::max(rk, h[d.pid].rank);
         bytes += 8.0 * d.rows * d.cols;
      }
      std::fprintf(stderr, "[fftsvd mgs] %d problems: %zu in LDS, %zu stepwise (max %d x %d, %.1f MB, max rank %d, steps %d)\n",
                   nprob, small.size(), step.size(), mr, mc, bytes / 1e6, rk, steps_done);
   }
   FSVD_CHECK(hipHostFree(h));
}

void launch_f64_to_f32(const double* a, float* b, long long n, hipStream_t s) {
   if (n <= 0) return;
   int blocks = (int)std::min<long long>((n + 255) / 256, 8192);
   hipLaunchKernelGGL(k_f64_to_f32, dim3(blocks), dim3(256), 0, s, a, b, n);
}

}  // namespace fsvd

namespace fsvd {
struct CopyDesc {
   const double* src;
   double* dst;
   long long n;
};
__global__ void k_batched_copy(const CopyDesc* __restrict__ d) {
   const CopyDesc c = d[blockIdx.x];
   for (long long i = threadIdx.x; i < c.n; i += blockDim.x) c.dst[i] = c.src[i];
}
void launch_batched_copy(int n, const void* descs, hipStream_t s) {
   if (n > 0) hipLaunchKernelGGL(k_batched_copy, dim3(n), dim3(256), 0, s, (const CopyDesc*)descs);
}
}  // namespace fsvd

namespace fsvd {
// Panel geometry records on the device (FlatPanel_allocate +
// FlatPanel_moments, FlatPanel.c:5-179,210-298) with the host's arithmetic,
// so every field equals the host record's bit for bit: record k is that of
// panel ids[k] (panel k without ids) of the 9-double vertex records verts9.
__global__ void __launch_bounds__(64) k_panel_records(const double* __restrict__ verts9, const int* __restrict__ ids,
                                                      int n, PanelGeom* __restrict__ out) {
   const int k = blockIdx.x * blockDim.x + threadIdx.x;
   if (k >= n) return;
   const double* v = verts9 + 9 * (long long)(ids ? ids[k] : k);
   double v0[3], v1[3], v2[3];
   for (int d = 0; d < 3; d++) {
      v0[d] = v[d];
      v1[d] = v[3 + d];
      v2[d] = v[6 + d];
   }
   PanelGeom g;
   geom::panel_geometry_hd(v0, v1, v2, &g);
   const double* src = reinterpret_cast<const double*>(&g);
   double* dst = reinterpret_cast<double*>(out + k);
   for (int q = 0; q < (int)(sizeof(PanelGeom) / sizeof(double)); q++) dst[q] = src[q];
}
void launch_panel_records(const double* verts9, const int* ids, int n, PanelGeom* out, hipStream_t s) {
   if (n > 0) hipLaunchKernelGGL(k_panel_records, dim3((n + 63) / 64), dim3(64), 0, s, verts9, ids, n, out);
}
}  // namespace fsvd
