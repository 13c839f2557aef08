// Setup kernels: collocation integrals, point kernels, small batched GEMMs
// and the pivoted modified Gram-Schmidt that builds the sampled bases.
// All of it runs in FP64.
#include <hip/hip_runtime.h>

#include "device.h"
#include "integrals.h"

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
      out[o] = s;
      out[o + b.layer_stride] = d;
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

// C = A*B, one thread per output element, sum over k in increasing order.
// grid.x = output tiles of 256, grid.y = batch.
__global__ void __launch_bounds__(256) k_gemm_batched(const GemmDesc* __restrict__ descs) {
   const GemmDesc d = descs[blockIdx.y];
   const int total = d.M * d.N;
   for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
      int i = e / d.N;
      int j = e - i * d.N;
      const double* a = d.A + (long long)i * d.a_rs;
      const double* b = d.B + (long long)j * d.b_cs;
      double s = 0.0;
      for (int k = 0; k < d.K; k++) s += a[(long long)k * d.a_cs] * b[(long long)k * d.b_rs];
      d.C[(long long)i * d.c_rs + (long long)j * d.c_cs] = s;
   }
}

// Column basis by pivoted MGS with two re-orthogonalisations and the rank
// cut ||col|| < eps*||A||_F (reference Matrix_columnbasis_pivotedmgs,
// Matrix.c:215-307).  One workgroup per problem; every thread owns whole
// columns, so each column's dot products and norms are sequential sums over
// the rows (the reference's summation order).
__global__ void __launch_bounds__(256) k_pivoted_mgs(const MgsDesc* __restrict__ descs, double* __restrict__ Abuf,
                                                     double* __restrict__ Qbuf, int* __restrict__ ranks,
                                                     double eps) {
   const MgsDesc d = descs[blockIdx.x];
   double* A = Abuf + d.a_off;
   double* Q = Qbuf + d.q_off;
   const int rows = d.rows, cols = d.cols;
   const int tid = threadIdx.x, nt = blockDim.x;
   extern __shared__ double sh[];
   double* red_v = sh;                     // nt
   int* red_i = (int*)(red_v + nt);        // nt
   __shared__ double s_fro;
   __shared__ int s_best;
   __shared__ double s_max;
   __shared__ int s_done;

   // Frobenius norm: per-column sums, then the column sums in order.
   double part = 0.0;
   for (int j = tid; j < cols; j += nt) {
      double s = 0.0;
      for (int i = 0; i < rows; i++) {
         double v = A[(long long)i * cols + j];
         s += v * v;
      }
      part += s;
   }
   red_v[tid] = part;
   __syncthreads();
   if (tid == 0) {
      double f = 0.0;
      for (int t = 0; t < nt; t++) f += red_v[t];
      s_fro = sqrt(f);
      s_done = 0;
   }
   __syncthreads();
   const double thresh = eps * s_fro;
   int rank = 0;
   for (int k = 0; k < cols; k++) {
      // at most `rows` orthonormal columns exist: the remaining ones are
      // rounding noise, far below any positive cut
      if (k >= rows) {
         rank = k;
         break;
      }
      // column norms for l >= k, argmax with ties to the lowest index
      double bv = 0.0;
      int bi = 0x7fffffff;
      for (int l = k + tid; l < cols; l += nt) {
         double s = 0.0;
         for (int i = 0; i < rows; i++) {
            double v = A[(long long)i * cols + l];
            s += v * v;
         }
         s = sqrt(s);
         if (s > bv || (s == bv && l < bi)) {
            bv = s;
            bi = l;
         }
      }
      red_v[tid] = bv;
      red_i[tid] = bi;
      __syncthreads();
      for (int w = nt / 2; w > 0; w >>= 1) {
         if (tid < w) {
            double ov = red_v[tid + w];
            int oi = red_i[tid + w];
            if (ov > red_v[tid] || (ov == red_v[tid] && oi < red_i[tid])) {
               red_v[tid] = ov;
               red_i[tid] = oi;
            }
         }
         __syncthreads();
      }
      if (tid == 0) {
         double mx = red_v[0];
         int best = red_i[0];
         if (!(mx > 0.0)) best = k;  // all remaining columns zero: reference keeps index 0 -> k
         s_max = mx;
         s_best = best;
         if (k >= 1 && mx < thresh) s_done = 1;
      }
      __syncthreads();
      if (s_done) {
         rank = k;
         break;
      }
      const int best = s_best;
      const double mx = s_max;
      if (best != k)
         for (int i = tid; i < rows; i += nt) {
            double t = A[(long long)i * cols + best];
            A[(long long)i * cols + best] = A[(long long)i * cols + k];
            A[(long long)i * cols + k] = t;
         }
      __syncthreads();
      for (int i = tid; i < rows; i += nt) Q[(long long)k * rows + i] = A[(long long)i * cols + k] / mx;
      __syncthreads();
      const double* q = Q + (long long)k * rows;
      if (k < cols - 1) {
         for (int j = k + 1 + tid; j < cols; j += nt) {
            for (int pass = 0; pass < 2; pass++) {
               double r = 0.0;
               for (int i = 0; i < rows; i++) r += q[i] * A[(long long)i * cols + j];
               for (int i = 0; i < rows; i++) A[(long long)i * cols + j] -= q[i] * r;
            }
         }
      }
      __syncthreads();
      if (k == cols - 1) rank = cols;
   }
   if (tid == 0) ranks[d.rank_slot] = rank;
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

void launch_mgs(int nprob, int maxrows, const MgsDesc* descs, double* A, double* Q, int* ranks, double eps,
                hipStream_t s) {
   (void)maxrows;
   size_t lds = sizeof(double) * 256 + sizeof(int) * 256 + 64;
   if (nprob > 0) hipLaunchKernelGGL(k_pivoted_mgs, dim3(nprob), dim3(256), lds, s, descs, A, Q, ranks, eps);
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
