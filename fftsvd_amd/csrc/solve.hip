// Region-operator glue, block preconditioner, Krylov vector kernels, the
// charge right-hand side and the reaction potentials at the charges.  All
// FP64.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "device.h"
#include "integrals.h"

namespace fsvd {

// Joined potentials of one region in its tree order (reference
// SurfaceOperator_joinpotentials, SurfaceOperator.c:51-65):
// xd = s_phi * x[phi_idx] (double-layer input), xs = s_dphi * x[dphi_idx].
template <typename XT>
__global__ void k_join(int n, const int* __restrict__ phi_idx, const int* __restrict__ dphi_idx,
                       const double* __restrict__ s_phi, const double* __restrict__ s_dphi, const XT* __restrict__ x,
                       double* __restrict__ xs, double* __restrict__ xd) {
   int k = blockIdx.x * blockDim.x + threadIdx.x;
   if (k < n) {
      xd[k] = s_phi[k] * (double)x[phi_idx[k]];
      xs[k] = s_dphi[k] * (double)x[dphi_idx[k]];
   }
}
// k_join for the columns blockIdx.y: input x[c], outputs at c * ld
template <typename XT>
__global__ void k_join_b(int n, const int* __restrict__ phi_idx, const int* __restrict__ dphi_idx,
                         const double* __restrict__ s_phi, const double* __restrict__ s_dphi,
                         const XT* const* __restrict__ xc, double* __restrict__ xs, double* __restrict__ xd,
                         long long ld) {
   const int c = blockIdx.y;
   const XT* x = xc[c];
   int k = blockIdx.x * blockDim.x + threadIdx.x;
   if (k < n) {
      xd[c * ld + k] = s_phi[k] * (double)x[phi_idx[k]];
      xs[c * ld + k] = s_dphi[k] * (double)x[dphi_idx[k]];
   }
}

// Split back to the global result with the -4 pi jump on child panels
// (SurfaceOperator.c:141-147 and splitpotentials/joinglobal, :78-107).
__global__ void k_split(int n, const int* __restrict__ out_idx, const double* __restrict__ corr,
                        const double* __restrict__ yt, const double* __restrict__ xd, double* __restrict__ y) {
   int k = blockIdx.x * blockDim.x + threadIdx.x;
   if (k < n) y[out_idx[k]] = yt[k] + corr[k] * xd[k];
}
// k_split for the columns blockIdx.y: inputs at c * ld, output yc[c]
__global__ void k_split_b(int n, const int* __restrict__ out_idx, const double* __restrict__ corr,
                          const double* __restrict__ yt, const double* __restrict__ xd, long long ld,
                          double* const* __restrict__ yc) {
   const int c = blockIdx.y;
   int k = blockIdx.x * blockDim.x + threadIdx.x;
   if (k < n) yc[c][out_idx[k]] = yt[c * ld + k] + corr[k] * xd[c * ld + k];
}

// Transposed region operator, per region on one stream: gather
// yh[k] = y[out_idx[k]], then (after the transposed tree applies into
// vs = S^T yh and vd = D^T yh) scatter-add x[dphi_idx[k]] += s_dphi vs and
// x[phi_idx[k]] += s_phi (vd + corr yh): the transpose of k_join / tree
// apply / k_split.  Within a region the targets are distinct.
template <typename XT>
__global__ void k_gather_out(int n, const int* __restrict__ out_idx, const XT* __restrict__ y,
                             double* __restrict__ yh) {
   for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
      yh[k] = (double)y[out_idx[k]];
}

__global__ void k_scatter_adj(int n, const int* __restrict__ phi_idx, const int* __restrict__ dphi_idx,
                              const double* __restrict__ s_phi, const double* __restrict__ s_dphi,
                              const double* __restrict__ corr, const double* __restrict__ yh,
                              const double* __restrict__ vs, const double* __restrict__ vd, double* __restrict__ x) {
   for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
      x[dphi_idx[k]] += s_dphi[k] * vs[k];
      x[phi_idx[k]] += s_phi[k] * (vd[k] + corr[k] * yh[k]);
   }
}

// x = P^{-T} b for the transposed 2x2 blocks; in place safe.
__global__ void k_precond_t(int nb, const int* __restrict__ ia, const int* __restrict__ ib,
                            const double* __restrict__ M, const double* __restrict__ b, double* __restrict__ x) {
   int k = blockIdx.x * blockDim.x + threadIdx.x;
   if (k < nb) {
      int i = ia[k], j = ib[k];
      double bi = b[i], bj = b[j];
      const double* m = M + 4 * (size_t)k;
      x[i] = m[0] * bi + m[2] * bj;
      x[j] = m[1] * bi + m[3] * bj;
   }
}

// x = P^{-1} b for disjoint 2x2 blocks (factored on the host); in place safe.
__device__ __forceinline__ void precond_body(int nb, const int* __restrict__ ia, const int* __restrict__ ib,
                                             const double* __restrict__ M, const double* b, double* x) {
   int k = blockIdx.x * blockDim.x + threadIdx.x;
   if (k < nb) {
      int i = ia[k], j = ib[k];
      double bi = b[i], bj = b[j];
      const double* m = M + 4 * (size_t)k;
      x[i] = m[0] * bi + m[1] * bj;
      x[j] = m[2] * bi + m[3] * bj;
   }
}
__global__ void k_precond(int nb, const int* __restrict__ ia, const int* __restrict__ ib,
                          const double* __restrict__ M, const double* __restrict__ b, double* __restrict__ x) {
   precond_body(nb, ia, ib, M, b, x);
}
// in place on the w of column blockIdx.y
__global__ void k_precond_b(int nb, const int* __restrict__ ia, const int* __restrict__ ib,
                            const double* __restrict__ M, const GmDev* __restrict__ d) {
   double* w = d[blockIdx.y].w;
   precond_body(nb, ia, ib, M, w, w);
}

__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
   int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   if (lane == 0) red[wave] = v;
   __syncthreads();
   double s = 0.0;
   if (threadIdx.x == 0) {
      for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[w];
   }
   __syncthreads();
   return s;
}

// partial[b] = sum over this block's chunk of a.*b
__global__ void __launch_bounds__(256) k_dot_partial(long long n, const double* __restrict__ a,
                                                     const double* __restrict__ b, double* __restrict__ partial) {
   __shared__ double red[8];
   double s = 0.0;
   for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
      s += a[i] * b[i];
   s = block_sum(s, red);
   if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__device__ __forceinline__ double sum_partials(const double* partial, int np) {
   // every thread sums the same partials in the same order: identical result
   double h = 0.0;
   for (int k = 0; k < np; k++) h += partial[k];
   return h;
}

// The same sum by one wavefront: lane l adds partials l, l+64, ... in order,
// then a fixed butterfly; every lane of every block gets the same value.
__device__ __forceinline__ double wave_sum_partials(const double* partial, int np) {
   const int lane = threadIdx.x & 63;
   double h = 0.0;
   for (int k = lane; k < np; k += 64) h += partial[k];
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o, 64);
   return h;
}

// v = w / sqrt(sum(pin)); hout[0] = the norm.
template <typename VT>
__global__ void __launch_bounds__(256) k_normalize(long long n, const double* __restrict__ pin, int np,
                                                   double* __restrict__ hout, const double* __restrict__ w,
                                                   VT* __restrict__ v) {
   __shared__ double hs;
   if (threadIdx.x < 64) {
      const double t = sqrt(wave_sum_partials(pin, np));
      if (threadIdx.x == 0) hs = t;
   }
   __syncthreads();
   const double h = hs;
   if (blockIdx.x == 0 && threadIdx.x == 0) hout[0] = h;
   const double inv = 1.0 / h;
   for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
      v[i] = (VT)(w[i] * inv);
}

// Classical Gram-Schmidt pass over m basis vectors (two passes make CGS2).
// Block (b, jg) owns the contiguous chunk b of the vectors and the columns
// MD_COLS*jg ..: partial[j*nch + b] = chunk sum of V_j . w.  (Finishing
// the sums in the last block to arrive would need a device-scope release
// fence in every block, which writes back the L2 of each XCD: far slower
// than the separate finish launch.)
constexpr int MD_COLS = 8;
template <typename VT>
__device__ __forceinline__ void multidot_body(long long n, int m, int nch, const VT* const* __restrict__ V,
                                              const double* __restrict__ w, double* __restrict__ partial) {
   __shared__ double red[4][MD_COLS];
   const long long chunk = (n + nch - 1) / nch;
   const long long i0 = blockIdx.x * chunk, i1 = i0 + chunk < n ? i0 + chunk : n;
   const int j0 = blockIdx.y * MD_COLS;
   const int nj = m - j0 < MD_COLS ? m - j0 : MD_COLS;
   const VT* v[MD_COLS];
#pragma unroll
   for (int u = 0; u < MD_COLS; u++) v[u] = V[j0 + (u < nj ? u : 0)];
   double acc[MD_COLS];
#pragma unroll
   for (int u = 0; u < MD_COLS; u++) acc[u] = 0.0;
   for (long long i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
      const double wi = w[i];
#pragma unroll
      for (int u = 0; u < MD_COLS; u++) acc[u] += (double)v[u][i] * wi;
   }
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
   for (int u = 0; u < MD_COLS; u++) {
      double t = acc[u];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
      if (lane == 0) red[wave][u] = t;
   }
   __syncthreads();
   if (threadIdx.x < nj) {
      const int u = threadIdx.x;
      partial[(long long)(j0 + u) * nch + blockIdx.x] = ((red[0][u] + red[1][u]) + red[2][u]) + red[3][u];
   }
}
template <typename VT>
__global__ void __launch_bounds__(256) k_multidot(long long n, int m, int nch, const VT* const* __restrict__ V,
                                                  const double* __restrict__ w, double* __restrict__ partial) {
   multidot_body<VT>(n, m, nch, V, w, partial);
}
// the column blockIdx.z
template <typename VT>
__global__ void __launch_bounds__(256) k_multidot_b(long long n, int m, int nch, const GmDev* __restrict__ d) {
   const GmDev& a = d[blockIdx.z];
   multidot_body<VT>(n, m, nch, (const VT* const*)a.V, a.w, a.partial_m);
}

// h[j] = sum of column j's nch partials (one wave per column, lanes over
// the partials, then a fixed-order wave reduction); hacc[j] = h[j] (first
// pass) or hacc[j] + h[j] (second pass).
__device__ __forceinline__ void multidot_finish_body(int m, int nch, const double* __restrict__ partial,
                                                     double* __restrict__ h, double* __restrict__ hacc, int add) {
   const int j = blockIdx.x;
   if (j >= m) return;
   double sum = 0.0;
   for (int k = threadIdx.x; k < nch; k += 64) sum += partial[(long long)j * nch + k];
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
   if (threadIdx.x == 0) {
      h[j] = sum;
      hacc[j] = add ? hacc[j] + sum : sum;
   }
}
__global__ void __launch_bounds__(64) k_multidot_finish(int m, int nch, const double* __restrict__ partial,
                                                        double* __restrict__ h, double* __restrict__ hacc, int add) {
   multidot_finish_body(m, nch, partial, h, hacc, add);
}
__global__ void __launch_bounds__(64) k_multidot_finish_b(int m, int nch, const GmDev* __restrict__ d, int add) {
   const GmDev& a = d[blockIdx.y];
   multidot_finish_body(m, nch, a.partial_m, a.htmp, a.hbuf, add);
}

// One Arnoldi column's Givens step (reference GMRES.c:3-22, 495-510): the
// previous rotations applied to H[:, i] = hcol, the new rotation from
// (H[i,i], H[i+1,i]), g updated, |g[i+1]|/normr recorded.  Run by the first
// wave of a block: the column and the rotations are staged in LDS in
// parallel, lane 0 applies the rotations in order, the column is written
// back in parallel.  H is column-major with leading dimension W.
__device__ __forceinline__ void wave_lds_sync() {
   __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
   __builtin_amdgcn_wave_barrier();
   __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ void arnoldi_givens(int i, int W, const double* __restrict__ hcol, double hlast, double* __restrict__ H,
                               double* __restrict__ cs, double* __restrict__ sn, double* __restrict__ g,
                               double* __restrict__ hist, double normr, double* gsm) {
   double* hs = gsm;            // i + 2
   double* cl = hs + i + 2;     // i
   double* sl = cl + i;         // i
   const int lane = threadIdx.x & 63;
   for (int j = lane; j <= i; j += 64) hs[j] = hcol[j];
   if (lane == 0) hs[i + 1] = hlast;
   for (int k = lane; k < i; k += 64) {
      cl[k] = cs[k];
      sl[k] = sn[k];
   }
   wave_lds_sync();
   if (lane == 0) {
      double a = hs[0];
      for (int k = 0; k < i; k++) {
         const double b = hs[k + 1], c = cl[k], s = sl[k];
         hs[k] = a * c - b * s;
         a = c * b + s * a;
      }
      double b = hs[i + 1];
      double c = 1.0, s = 0.0;
      if (b != 0.0) {
         const double r1 = fabs(a), r2 = fabs(b);
         const double L = sqrt(r1 * r1 + r2 * r2);
         c = a / L;
         s = -b / L;
      }
      cs[i] = c;
      sn[i] = s;
      hs[i] = a * c - b * s;
      hs[i + 1] = c * b + s * a;
      const double gi = g[i];
      g[i] = gi * c;   // g[i+1] was zero
      g[i + 1] = s * gi;
      hist[i] = fabs(g[i + 1]) / normr;
   }
   wave_lds_sync();
   double* Hi = H + (long long)i * W;
   for (int j = lane; j <= i + 1; j += 64) Hi[j] = hs[j];
}

// End of Arnoldi step i: nrm = sqrt(sum of the np partials of |w|^2),
// v = w / nrm, H[i+1, i] = nrm (hcol[i+1]), and block 0's first wave does
// the Givens step of column i.
template <typename VT>
__device__ __forceinline__ void normalize_givens_body(long long n, const double* __restrict__ pin, int np,
                                                      const double* __restrict__ w, VT* __restrict__ v, int i, int W,
                                                      double* __restrict__ hcol, double* __restrict__ H,
                                                      double* __restrict__ cs, double* __restrict__ sn,
                                                      double* __restrict__ g, double* __restrict__ hist, double normr) {
   extern __shared__ double gsm[];
   __shared__ double hs;
   if (threadIdx.x < 64) {
      const double t = sqrt(wave_sum_partials(pin, np));
      if (threadIdx.x == 0) hs = t;
   }
   __syncthreads();
   const double nrm = hs;
   const double inv = 1.0 / nrm;
   for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
      v[k] = (VT)(w[k] * inv);
   if (blockIdx.x == 0 && threadIdx.x < 64) {
      if (threadIdx.x == 0) hcol[i + 1] = nrm;
      arnoldi_givens(i, W, hcol, nrm, H, cs, sn, g, hist, normr, gsm);
   }
}
template <typename VT>
__global__ void __launch_bounds__(256) k_normalize_givens(long long n, const double* __restrict__ pin, int np,
                                                          const double* __restrict__ w, VT* __restrict__ v, int i,
                                                          int W, double* __restrict__ hcol, double* __restrict__ H,
                                                          double* __restrict__ cs, double* __restrict__ sn,
                                                          double* __restrict__ g, double* __restrict__ hist,
                                                          double normr) {
   normalize_givens_body<VT>(n, pin, np, w, v, i, W, hcol, H, cs, sn, g, hist, normr);
}
template <typename VT>
__global__ void __launch_bounds__(256) k_normalize_givens_b(long long n, int np, int i, int W,
                                                            const GmDev* __restrict__ d) {
   const GmDev& a = d[blockIdx.y];
   normalize_givens_body<VT>(n, a.pnorm, np, a.w, (VT*)a.vnext, i, W, a.hbuf, a.H, a.cs, a.sn, a.g, a.hist,
                             a.normr);
}

// w -= sum_j h[j] V_j, summed over j in order.
// With pnorm: pnorm[block] = this block's sum of the updated w^2.
template <typename VT>
__device__ __forceinline__ void multiaxpy_body(long long n, int m, const double* __restrict__ h,
                                               const VT* const* __restrict__ V, double* __restrict__ w,
                                               double* __restrict__ pnorm) {
   __shared__ double red[8];
   double nn = 0.0;
   for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
      double s = 0.0;
      int j = 0;
      // eight basis values in flight
      for (; j + 7 < m; j += 8) {
         double v[8];
#pragma unroll
         for (int u = 0; u < 8; u++) v[u] = (double)V[j + u][i];
#pragma unroll
         for (int u = 0; u < 8; u++) s += h[j + u] * v[u];
      }
      if (j < m) {
         // the last partial batch, all of its loads in flight together
         double v[7];
#pragma unroll
         for (int u = 0; u < 7; u++) {
            const bool ok = j + u < m;
            const double x = (double)V[ok ? j + u : j][i];   // valid vector, no branch
            v[u] = ok ? x : 0.0;
         }
#pragma unroll
         for (int u = 0; u < 7; u++) s += h[j + u < m ? j + u : j] * v[u];   // v[u] = 0 past m
      }
      const double wn = w[i] - s;
      w[i] = wn;
      nn += wn * wn;
   }
   if (pnorm) {
      nn = block_sum(nn, red);
      if (threadIdx.x == 0) pnorm[blockIdx.x] = nn;
   }
}
template <typename VT>
__global__ void __launch_bounds__(256) k_multiaxpy(long long n, int m, const double* __restrict__ h,
                                                   const VT* const* __restrict__ V, double* __restrict__ w,
                                                   double* __restrict__ pnorm) {
   multiaxpy_body<VT>(n, m, h, V, w, pnorm);
}
template <typename VT>
__global__ void __launch_bounds__(256) k_multiaxpy_b(long long n, int m, const GmDev* __restrict__ d, int norm) {
   const GmDev& a = d[blockIdx.y];
   multiaxpy_body<VT>(n, m, a.htmp, (const VT* const*)a.V, a.w, norm ? a.pnorm : nullptr);
}

// x = sum_j y[j] * V[j], in order of j.
template <typename VT>
__global__ void k_combine(long long n, int m, const double* __restrict__ y, const VT* const* __restrict__ V,
                          double* __restrict__ x) {
   for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
      double s = 0.0;
      for (int j0 = 0; j0 < m; j0 += 8) {
         // eight basis values in flight, summed in order of j
         double v[8];
#pragma unroll
         for (int u = 0; u < 8; u++) {
            const bool ok = j0 + u < m;
            const double x = (double)V[ok ? j0 + u : j0][i];   // valid vector, no branch
            v[u] = ok ? x : 0.0;
         }
#pragma unroll
         for (int u = 0; u < 8; u++) s += y[j0 + u < m ? j0 + u : j0] * v[u];   // v[u] = 0 past m
      }
      x[i] = s;
   }
}

__global__ void k_axpby(long long n, double a, const double* __restrict__ x, double b, const double* __restrict__ y,
                        double* __restrict__ z) {
   for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
      z[i] = a * x[i] + b * y[i];
}

// rhs[out[e]] = sum_i q_i / |x_i - c_e| over charges in order
// (SurfaceOperator_makeRHS, SurfaceOperator.c:291-321).
__global__ void k_rhs(int ne, const double* __restrict__ cent, const int* __restrict__ out, int nq,
                      const double* __restrict__ qpos, const double* __restrict__ q, double* __restrict__ rhs) {
   int e = blockIdx.x * blockDim.x + threadIdx.x;
   if (e >= ne) return;
   const double c0 = cent[3 * e], c1 = cent[3 * e + 1], c2 = cent[3 * e + 2];
   double s = 0.0;
   for (int i = 0; i < nq; i++) {
      double qi = q[i];
      if (qi == 0.0) continue;
      double d0 = qpos[3 * i] - c0, d1 = qpos[3 * i + 1] - c1, d2 = qpos[3 * i + 2] - c2;
      s += qi / sqrt(d0 * d0 + d1 * d1 + d2 * d2);
   }
   rhs[out[e]] += s;
}

// Many charge vectors (columns of q, nq x ncol) at once: the distances are
// formed once per (element, charge); each column sums in charge order with
// zero charges skipped, exactly as k_rhs, so column c equals k_rhs on q_c.
constexpr int RHS_NB = 8;
__global__ void k_rhs_batch(int ne, const double* __restrict__ cent, const int* __restrict__ out, int nq, int ncol,
                            const double* __restrict__ qpos, const double* __restrict__ q, long long ldr,
                            double* __restrict__ rhs) {
   for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += gridDim.x * blockDim.x) {
      const double c0 = cent[3 * e], c1 = cent[3 * e + 1], c2 = cent[3 * e + 2];
      for (int col0 = 0; col0 < ncol; col0 += RHS_NB) {
         double s[RHS_NB];
#pragma unroll
         for (int c = 0; c < RHS_NB; c++) s[c] = 0.0;
         for (int i = 0; i < nq; i++) {
            const double d0 = qpos[3 * i] - c0, d1 = qpos[3 * i + 1] - c1, d2 = qpos[3 * i + 2] - c2;
            const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
#pragma unroll
            for (int c = 0; c < RHS_NB; c++) {
               if (col0 + c < ncol) {
                  const double qi = q[(long long)(col0 + c) * nq + i];
                  if (qi != 0.0) s[c] += qi / r;
               }
            }
         }
#pragma unroll
         for (int c = 0; c < RHS_NB; c++)
            if (col0 + c < ncol) rhs[(long long)(col0 + c) * ldr + out[e]] += s[c];
      }
   }
}

// phi[j] += sum over panels of (a_e * slp - b_e * dlp) - (self ? 4 pi b_e : 0)
// with Laplace integrals at the charge (direct branch of
// SurfaceOperator_calculateReactionPotentials, SurfaceOperator.c:364-396).
// For own panels a = sol[dphi], b = sol[phi], self check on; for child
// panels a = -(eps_c/eps) sol[dphi_c], b = -sol[phi_c], self check off.
__global__ void __launch_bounds__(256) k_potentials(int ne, const PanelGeom* __restrict__ panels,
                                                    const int* __restrict__ pidx, const double* __restrict__ coef,
                                                    const int* __restrict__ self_ok,
                                                    const double* __restrict__ qpos, double* __restrict__ phi) {
   __shared__ double red[8];
   const int j = blockIdx.x;
   const double x[3] = {qpos[3 * j], qpos[3 * j + 1], qpos[3 * j + 2]};
   double s = 0.0;
   for (int e = threadIdx.x; e < ne; e += blockDim.x) {
      const PanelGeom& P = panels[pidx[e]];
      double sl, dl;
      laplace_integrals(x, P, sl, dl);
      double a = coef[2 * e], b = coef[2 * e + 1];
      double v = a * sl - b * dl;
      if (self_ok[e] && x[0] == P.c[0] && x[1] == P.c[1] && x[2] == P.c[2]) v -= 4.0 * M_PI * b;
      s += v;
   }
   s = block_sum(s, red);
   if (threadIdx.x == 0) phi[j] += s;
}

// Qualocation operator (Solv_ecf_multiply_qual, GMRES.c:159-181), around
// the transposed double-layer tree apply in tree order:
//   gather: xa[k] = area_t[k] * x[perm[k]]
//   finish: b = (y_t[k] - diag_t[k] xa[k]) * (1/(4 pi eps_in))
//               + (c0 + diag_t[k] / (4 pi eps_in)) xa[k],  out[perm[k]] = b
// with c0 = -eps_out / ((eps_out - eps_in) eps_in).
template <typename XT>
__global__ void k_qual_gather(int n, const int* __restrict__ perm, const double* __restrict__ area_t,
                              const XT* __restrict__ x, double* __restrict__ xa) {
   for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
      xa[k] = area_t[k] * (double)x[perm[k]];
}

__global__ void k_qual_finish(int n, const int* __restrict__ perm, const double* __restrict__ diag_t,
                              const double* __restrict__ xa, const double* __restrict__ yt, double inv4pe,
                              double fourpe, double c0, double* __restrict__ out) {
   for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
      const double d = diag_t[k], a = xa[k];
      double b = yt[k];
      b -= d * a;
      b *= inv4pe;
      b += (c0 + (d) / fourpe) * a;
      out[perm[k]] = b;
   }
}

// Diagonal preconditioner solve, x = b / p (x may alias b).
__global__ void k_diag_div(int n, const double* __restrict__ p, const double* b, double* x) {
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = b[i] / p[i];
}

// Qualocation right-hand side (Charge_makerhs_ecf_qual, Charge.c:129-142):
// rhs[i] = sum_j -dlp(charge j, panel i) q_j / (4 pi eps_in), in charge order.
__global__ void __launch_bounds__(256) k_qual_rhs(int n, const PanelGeom* __restrict__ panels, int nq,
                                                  const double* __restrict__ qpos, const double* __restrict__ q,
                                                  double fourpe, double* __restrict__ rhs) {
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      const PanelGeom& P = panels[i];
      double r = 0.0;
      for (int j = 0; j < nq; j++) {
         const double x[3] = {qpos[3 * j], qpos[3 * j + 1], qpos[3 * j + 2]};
         double sl, dl;
         laplace_integrals(x, P, sl, dl);
         r += -dl * q[j] / fourpe;
      }
      rhs[i] = r;
   }
}

// k_potentials for ncol solution columns: coef holds (a, b) per element and
// column (column-major, 2*ne per column); the integrals are formed once per
// (charge, panel) and each column is summed and reduced as k_potentials.
constexpr int POT_NB = 8;
__global__ void __launch_bounds__(256) k_potentials_batch(int ne, int ncol, const PanelGeom* __restrict__ panels,
                                                          const int* __restrict__ pidx,
                                                          const double* __restrict__ coef,
                                                          const int* __restrict__ self_ok,
                                                          const double* __restrict__ qpos, int nq,
                                                          double* __restrict__ phi) {
   __shared__ double red[8];
   const int j = blockIdx.x;
   const double x[3] = {qpos[3 * j], qpos[3 * j + 1], qpos[3 * j + 2]};
   for (int col0 = 0; col0 < ncol; col0 += POT_NB) {
      double s[POT_NB];
#pragma unroll
      for (int c = 0; c < POT_NB; c++) s[c] = 0.0;
      for (int e = threadIdx.x; e < ne; e += blockDim.x) {
         const PanelGeom& P = panels[pidx[e]];
         double sl, dl;
         laplace_integrals(x, P, sl, dl);
         const bool self = self_ok[e] && x[0] == P.c[0] && x[1] == P.c[1] && x[2] == P.c[2];
#pragma unroll
         for (int c = 0; c < POT_NB; c++) {
            if (col0 + c < ncol) {
               const double* cf = coef + (long long)(col0 + c) * 2 * ne;
               const double a = cf[2 * e], b = cf[2 * e + 1];
               double v = a * sl - b * dl;
               if (self) v -= 4.0 * M_PI * b;
               s[c] += v;
            }
         }
      }
#pragma unroll
      for (int c = 0; c < POT_NB; c++) {
         if (col0 + c < ncol) {   // uniform over the block
            const double t = block_sum(s[c], red);
            if (threadIdx.x == 0) phi[(long long)(col0 + c) * nq + j] += t;
         }
      }
   }
}

// k_qual_rhs for ncol charge columns (q nq x ncol), columns of rhs ldr apart
__global__ void __launch_bounds__(256) k_qual_rhs_batch(int n, const PanelGeom* __restrict__ panels, int nq, int ncol,
                                                        const double* __restrict__ qpos, const double* __restrict__ q,
                                                        double fourpe, long long ldr, double* __restrict__ rhs) {
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      const PanelGeom& P = panels[i];
      for (int col0 = 0; col0 < ncol; col0 += RHS_NB) {
         double r[RHS_NB];
#pragma unroll
         for (int c = 0; c < RHS_NB; c++) r[c] = 0.0;
         for (int j = 0; j < nq; j++) {
            const double x[3] = {qpos[3 * j], qpos[3 * j + 1], qpos[3 * j + 2]};
            double sl, dl;
            laplace_integrals(x, P, sl, dl);
#pragma unroll
            for (int c = 0; c < RHS_NB; c++)
               if (col0 + c < ncol) r[c] += -dl * q[(long long)(col0 + c) * nq + j] / fourpe;
         }
#pragma unroll
         for (int c = 0; c < RHS_NB; c++)
            if (col0 + c < ncol) rhs[(long long)(col0 + c) * ldr + i] = r[c];
      }
   }
}

// Transposed collection: per element e (a panel of a charged region) the
// sums over charges of q_j slp(j, e) and q_j dlp(j, e), with the 4 pi
// self term of k_potentials counted when the charge sits at the centroid.
__global__ void __launch_bounds__(256) k_collect_t(int ne, const PanelGeom* __restrict__ panels,
                                                   const int* __restrict__ pidx, const int* __restrict__ self_ok,
                                                   int nq, const double* __restrict__ qpos,
                                                   const double* __restrict__ q, double* __restrict__ ssl,
                                                   double* __restrict__ sdl) {
   for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += gridDim.x * blockDim.x) {
      const PanelGeom& P = panels[pidx[e]];
      double a = 0.0, b = 0.0;
      for (int j = 0; j < nq; j++) {
         const double x[3] = {qpos[3 * j], qpos[3 * j + 1], qpos[3 * j + 2]};
         double sl, dl;
         laplace_integrals(x, P, sl, dl);
         a += q[j] * sl;
         b += q[j] * dl;
         if (self_ok[e] && x[0] == P.c[0] && x[1] == P.c[1] && x[2] == P.c[2]) b += 4.0 * M_PI * q[j];
      }
      ssl[e] = a;
      sdl[e] = b;
   }
}

// Transposed right-hand side: g[j] += sum_e lam[out[e]] / |c_e - x_j|.
__global__ void __launch_bounds__(256) k_rhs_t(int ne, const double* __restrict__ cent, const int* __restrict__ out,
                                               const double* __restrict__ lam, const double* __restrict__ qpos,
                                               double* __restrict__ g) {
   __shared__ double red[8];
   const int j = blockIdx.x;
   const double x0 = qpos[3 * j], x1 = qpos[3 * j + 1], x2 = qpos[3 * j + 2];
   double s = 0.0;
   for (int e = threadIdx.x; e < ne; e += blockDim.x) {
      const double d0 = x0 - cent[3 * e], d1 = x1 - cent[3 * e + 1], d2 = x2 - cent[3 * e + 2];
      s += lam[out[e]] / sqrt(d0 * d0 + d1 * d1 + d2 * d2);
   }
   s = block_sum(s, red);
   if (threadIdx.x == 0) g[j] += s;
}

// diagonal of the near-field blocks: diag[p0 + i] from leaf rows i < p
__global__ void k_leaf_diag(int nleaf, const LeafDesc* __restrict__ leaves, const double* __restrict__ D,
                            double* __restrict__ ds, double* __restrict__ dd) {
   int l = blockIdx.x;
   if (l >= nleaf) return;
   const LeafDesc L = leaves[l];
   for (int i = threadIdx.x; i < L.p && i < L.m; i += blockDim.x) {
      ds[L.q0 + i] = D[L.d_off + (long long)i * 2 * L.p + i];
      dd[L.q0 + i] = D[L.d_off + (long long)i * 2 * L.p + L.p + i];
   }
}

// GPU batch of panel integrals for tests and tools: out (npt x npan).
__global__ void k_integrate_all(int kernel, double kappa, const PanelGeom* __restrict__ panels, int npan,
                                const double* __restrict__ pts, int npt, double* __restrict__ slp,
                                double* __restrict__ dlp) {
   long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
   if (e >= (long long)npan * npt) return;
   int i = (int)(e / npan), j = (int)(e - (long long)i * npan);
   double s, d;
   panel_integrals(kernel, kappa, pts + 3 * (size_t)i, panels[j], s, d);
   slp[e] = s;
   dlp[e] = d;
}

// LDS read probe: every lane reads W bytes (8 or 16) per load from a 32 KiB
// LDS array, the 64 lanes of a wave one contiguous 64 W bytes (no bank
// conflicts), eight loads in flight, niter rounds; integer three-input adds
// (two or one vector ALU operations per load) keep the ALUs ahead of the
// LDS; the sum is written only if it takes an impossible value, so the
// loads stay.  Bytes read = blocks * 256 * 8 * niter * W.
template <int W>
__global__ void __launch_bounds__(256) k_lds_probe(int niter, float* out) {
   using V = typename std::conditional<W == 16, uint4, uint2>::type;
   constexpr int NV = 32768 / W;
   __shared__ V buf[NV];
   for (int i = threadIdx.x; i < NV; i += 256) {
      V v;
      v.x = (unsigned)i;
      v.y = 1u;
      if constexpr (W == 16) {
         v.z = 2u;
         v.w = 3u;
      }
      buf[i] = v;
   }
   __syncthreads();
   unsigned a0 = 0u, a1 = 0u;
   int idx = threadIdx.x;
   for (int k = 0; k < niter; k++) {
      V v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = buf[(idx + 256 * u) & (NV - 1)];
#pragma unroll
      for (int u = 0; u < 8; u++) {
         a0 = a0 + v[u].x + v[u].y;
         if constexpr (W == 16) a1 = a1 + v[u].z + v[u].w;
      }
      idx = (idx + 64) & (NV - 1);
   }
   if (a0 + a1 == 1u) out[blockIdx.x] = (float)a0;
}

// Plain copy (bandwidth probe and counter calibration).
__global__ void k_copy(long long n4, const double4* __restrict__ a, double4* __restrict__ b) {
   for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x)
      b[i] = a[i];
}

// ---------------------------------------------------------------------------
static inline int nblocks(long long n, int per = 256, int cap = 65535) {
   long long b = (n + per - 1) / per;
   if (b < 1) b = 1;
   if (b > cap) b = cap;
   return (int)b;
}

template <typename XT>
void launch_join(int n, const int* pi, const int* di, const double* sp, const double* sd, const XT* x, double* xs,
                 double* xd, hipStream_t s) {
   if (n > 0) hipLaunchKernelGGL((k_join<XT>), dim3(nblocks(n)), dim3(256), 0, s, n, pi, di, sp, sd, x, xs, xd);
}
template void launch_join<double>(int, const int*, const int*, const double*, const double*, const double*, double*,
                                  double*, hipStream_t);
template void launch_join<float>(int, const int*, const int*, const double*, const double*, const float*, double*,
                                 double*, hipStream_t);
void launch_split(int n, const int* oi, const double* corr, const double* yt, const double* xd, double* y,
                  hipStream_t s) {
   if (n > 0) hipLaunchKernelGGL(k_split, dim3(nblocks(n)), dim3(256), 0, s, n, oi, corr, yt, xd, y);
}
void launch_precond(int nb, const int* ia, const int* ib, const double* M, const double* b, double* x, hipStream_t s) {
   if (nb > 0) hipLaunchKernelGGL(k_precond, dim3(nblocks(nb)), dim3(256), 0, s, nb, ia, ib, M, b, x);
}
void launch_dot_partial(long long n, int np, const double* a, const double* b, double* partial, hipStream_t s) {
   hipLaunchKernelGGL(k_dot_partial, dim3(np), dim3(256), 0, s, n, a, b, partial);
}
void launch_lds_probe(int width, int nblk, int niter, float* out, hipStream_t s) {
   if (width == 16)
      hipLaunchKernelGGL(k_lds_probe<16>, dim3(nblk), dim3(256), 0, s, niter, out);
   else
      hipLaunchKernelGGL(k_lds_probe<8>, dim3(nblk), dim3(256), 0, s, niter, out);
}
void launch_copy(long long n4, const void* a, void* b, hipStream_t s) {
   if (n4 > 0)
      hipLaunchKernelGGL(k_copy, dim3(nblocks(n4, 256, 8192)), dim3(256), 0, s, n4, (const double4*)a, (double4*)b);
}
template <typename VT>
void launch_multidot(long long n, int m, int nch, const VT* const* V, const double* w, double* partial,
                     hipStream_t s) {
   if (m > 0)
      hipLaunchKernelGGL((k_multidot<VT>), dim3(nch, (m + MD_COLS - 1) / MD_COLS), dim3(256), 0, s, n, m, nch, V, w,
                         partial);
}
void launch_multidot_finish(int m, int nch, const double* partial, double* h, double* hacc, int add, hipStream_t s) {
   if (m > 0) hipLaunchKernelGGL(k_multidot_finish, dim3(m), dim3(64), 0, s, m, nch, partial, h, hacc, add);
}
int multiaxpy_blocks(long long n) { return nblocks(n, 256, 4096); }
template <typename VT>
void launch_multiaxpy(long long n, int m, const double* h, const VT* const* V, double* w, double* pnorm,
                      hipStream_t s) {
   if (m > 0)
      hipLaunchKernelGGL((k_multiaxpy<VT>), dim3(multiaxpy_blocks(n)), dim3(256), 0, s, n, m, h, V, w, pnorm);
}
template <typename VT>
void launch_normalize(long long n, int np, const double* pin, double* hout, const double* w, VT* v, hipStream_t s) {
   hipLaunchKernelGGL((k_normalize<VT>), dim3(np), dim3(256), 0, s, n, pin, np, hout, w, v);
}
template <typename VT>
void launch_normalize_givens(long long n, int np, const double* pin, const double* w, VT* v, int i, int W,
                             double* hcol, double* H, double* cs, double* sn, double* g, double* hist, double normr,
                             hipStream_t s) {
   // LDS: the column (i + 2) and the i previous rotations; i + 2 <= W <= 2041
   const size_t lds = sizeof(double) * (size_t)(3 * i + 2);
   hipLaunchKernelGGL((k_normalize_givens<VT>), dim3(nblocks(n, 256, 512)), dim3(256), lds, s, n, pin, np, w, v, i,
                      W, hcol, H, cs, sn, g, hist, normr);
}
// The lockstep GMRES step for ncol columns (d: their device descriptors),
// each launch the single-column launch with one grid slice per column.
template <typename XT>
void launch_join_b(int n, int ncol, const int* pi, const int* di, const double* sp, const double* sd,
                   const XT* const* x, double* xs, double* xd, long long ld, hipStream_t s) {
   if (n > 0 && ncol > 0)
      hipLaunchKernelGGL((k_join_b<XT>), dim3(nblocks(n), ncol), dim3(256), 0, s, n, pi, di, sp, sd, x, xs, xd, ld);
}
template void launch_join_b<double>(int, int, const int*, const int*, const double*, const double*,
                                    const double* const*, double*, double*, long long, hipStream_t);
template void launch_join_b<float>(int, int, const int*, const int*, const double*, const double*,
                                   const float* const*, double*, double*, long long, hipStream_t);
void launch_split_b(int n, int ncol, const int* oi, const double* corr, const double* yt, const double* xd,
                    long long ld, double* const* y, hipStream_t s) {
   if (n > 0 && ncol > 0)
      hipLaunchKernelGGL(k_split_b, dim3(nblocks(n), ncol), dim3(256), 0, s, n, oi, corr, yt, xd, ld, y);
}
void launch_precond_b(int nb, int ncol, const int* ia, const int* ib, const double* M, const GmDev* d,
                      hipStream_t s) {
   if (nb > 0 && ncol > 0) hipLaunchKernelGGL(k_precond_b, dim3(nblocks(nb), ncol), dim3(256), 0, s, nb, ia, ib, M, d);
}
template <typename VT>
void launch_cgs2_b(long long n, int m, int nch, int ncol, const GmDev* d, hipStream_t s) {
   if (m <= 0 || ncol <= 0) return;
   for (int pass = 0; pass < 2; pass++) {
      hipLaunchKernelGGL((k_multidot_b<VT>), dim3(nch, (m + MD_COLS - 1) / MD_COLS, ncol), dim3(256), 0, s, n, m, nch,
                         d);
      hipLaunchKernelGGL(k_multidot_finish_b, dim3(m, ncol), dim3(64), 0, s, m, nch, d, pass);
      hipLaunchKernelGGL((k_multiaxpy_b<VT>), dim3(multiaxpy_blocks(n), ncol), dim3(256), 0, s, n, m, d, pass);
   }
}
template <typename VT>
void launch_normalize_givens_b(long long n, int np, int i, int W, int ncol, const GmDev* d, hipStream_t s) {
   const size_t lds = sizeof(double) * (size_t)(3 * i + 2);
   if (ncol > 0)
      hipLaunchKernelGGL((k_normalize_givens_b<VT>), dim3(nblocks(n, 256, 512), ncol), dim3(256), lds, s, n, np, i, W,
                         d);
}
template void launch_cgs2_b<double>(long long, int, int, int, const GmDev*, hipStream_t);
template void launch_cgs2_b<float>(long long, int, int, int, const GmDev*, hipStream_t);
template void launch_normalize_givens_b<double>(long long, int, int, int, int, const GmDev*, hipStream_t);
template void launch_normalize_givens_b<float>(long long, int, int, int, int, const GmDev*, hipStream_t);

template <typename VT>
void launch_combine(long long n, int m, const double* y, const VT* const* V, double* x, hipStream_t s) {
   hipLaunchKernelGGL((k_combine<VT>), dim3(nblocks(n, 256, 4096)), dim3(256), 0, s, n, m, y, V, x);
}
template <typename XT>
void launch_qual_gather(int n, const int* perm, const double* area_t, const XT* x, double* xa, hipStream_t s) {
   if (n > 0) hipLaunchKernelGGL((k_qual_gather<XT>), dim3(nblocks(n)), dim3(256), 0, s, n, perm, area_t, x, xa);
}
template void launch_qual_gather<double>(int, const int*, const double*, const double*, double*, hipStream_t);
template void launch_qual_gather<float>(int, const int*, const double*, const float*, double*, hipStream_t);
void launch_qual_finish(int n, const int* perm, const double* diag_t, const double* xa, const double* yt,
                        double inv4pe, double fourpe, double c0, double* out, hipStream_t s) {
   if (n > 0)
      hipLaunchKernelGGL(k_qual_finish, dim3(nblocks(n)), dim3(256), 0, s, n, perm, diag_t, xa, yt, inv4pe, fourpe, c0,
                         out);
}
void launch_diag_div(int n, const double* p, const double* b, double* x, hipStream_t s) {
   if (n > 0) hipLaunchKernelGGL(k_diag_div, dim3(nblocks(n)), dim3(256), 0, s, n, p, b, x);
}
void launch_qual_rhs(int n, const PanelGeom* panels, int nq, const double* qpos, const double* q, double fourpe,
                     double* rhs, hipStream_t s) {
   if (n > 0) hipLaunchKernelGGL(k_qual_rhs, dim3(nblocks(n)), dim3(256), 0, s, n, panels, nq, qpos, q, fourpe, rhs);
}

void launch_rhs_batch(int ne, const double* cent, const int* out, int nq, int ncol, const double* qpos,
                      const double* q, long long ldr, double* rhs, hipStream_t s) {
   if (ne > 0 && ncol > 0)
      hipLaunchKernelGGL(k_rhs_batch, dim3(nblocks(ne)), dim3(256), 0, s, ne, cent, out, nq, ncol, qpos, q, ldr, rhs);
}
void launch_potentials_batch(int nq, int ne, int ncol, const PanelGeom* panels, const int* pidx, const double* coef,
                             const int* self_ok, const double* qpos, double* phi, hipStream_t s) {
   if (nq > 0 && ne > 0 && ncol > 0)
      hipLaunchKernelGGL(k_potentials_batch, dim3(nq), dim3(256), 0, s, ne, ncol, panels, pidx, coef, self_ok, qpos,
                         nq, phi);
}
void launch_qual_rhs_batch(int n, const PanelGeom* panels, int nq, int ncol, const double* qpos, const double* q,
                           double fourpe, long long ldr, double* rhs, hipStream_t s) {
   if (n > 0 && ncol > 0)
      hipLaunchKernelGGL(k_qual_rhs_batch, dim3(nblocks(n)), dim3(256), 0, s, n, panels, nq, ncol, qpos, q, fourpe,
                         ldr, rhs);
}

template <typename XT>
void launch_gather_out(int n, const int* out_idx, const XT* y, double* yh, hipStream_t s) {
   if (n > 0) hipLaunchKernelGGL((k_gather_out<XT>), dim3(nblocks(n)), dim3(256), 0, s, n, out_idx, y, yh);
}
template void launch_gather_out<double>(int, const int*, const double*, double*, hipStream_t);
template void launch_gather_out<float>(int, const int*, const float*, double*, hipStream_t);
void launch_scatter_adj(int n, const int* pi, const int* di, const double* sp, const double* sd, const double* corr,
                        const double* yh, const double* vs, const double* vd, double* x, hipStream_t s) {
   if (n > 0)
      hipLaunchKernelGGL(k_scatter_adj, dim3(nblocks(n)), dim3(256), 0, s, n, pi, di, sp, sd, corr, yh, vs, vd, x);
}
void launch_precond_t(int nb, const int* ia, const int* ib, const double* M, const double* b, double* x,
                      hipStream_t s) {
   if (nb > 0) hipLaunchKernelGGL(k_precond_t, dim3(nblocks(nb)), dim3(256), 0, s, nb, ia, ib, M, b, x);
}
void launch_collect_t(int ne, const PanelGeom* panels, const int* pidx, const int* self_ok, int nq, const double* qpos,
                      const double* q, double* ssl, double* sdl, hipStream_t s) {
   if (ne > 0)
      hipLaunchKernelGGL(k_collect_t, dim3(nblocks(ne)), dim3(256), 0, s, ne, panels, pidx, self_ok, nq, qpos, q, ssl,
                         sdl);
}
void launch_rhs_t(int nq, int ne, const double* cent, const int* out, const double* lam, const double* qpos, double* g,
                  hipStream_t s) {
   if (nq > 0 && ne > 0) hipLaunchKernelGGL(k_rhs_t, dim3(nq), dim3(256), 0, s, ne, cent, out, lam, qpos, g);
}

#define FSVD_KRYLOV_INST(VT)                                                                                        \
   template void launch_multidot<VT>(long long, int, int, const VT* const*, const double*, double*, hipStream_t); \
   template void launch_multiaxpy<VT>(long long, int, const double*, const VT* const*, double*, double*,          \
                                      hipStream_t);                                                              \
   template void launch_normalize_givens<VT>(long long, int, const double*, const double*, VT*, int, int,        \
                                             double*, double*, double*, double*, double*, double*, double,        \
                                             hipStream_t);                                                        \
   template void launch_normalize<VT>(long long, int, const double*, double*, const double*, VT*, hipStream_t);  \
   template void launch_combine<VT>(long long, int, const double*, const VT* const*, double*, hipStream_t);
FSVD_KRYLOV_INST(double)
FSVD_KRYLOV_INST(float)
void launch_axpby(long long n, double a, const double* x, double b, const double* y, double* z, hipStream_t s) {
   hipLaunchKernelGGL(k_axpby, dim3(nblocks(n, 256, 4096)), dim3(256), 0, s, n, a, x, b, y, z);
}
void launch_rhs(int ne, const double* cent, const int* out, int nq, const double* qpos, const double* q, double* rhs,
                hipStream_t s) {
   if (ne > 0) hipLaunchKernelGGL(k_rhs, dim3(nblocks(ne)), dim3(256), 0, s, ne, cent, out, nq, qpos, q, rhs);
}
void launch_potentials(int nq, int ne, const PanelGeom* panels, const int* pidx, const double* coef,
                       const int* self_ok, const double* qpos, double* phi, hipStream_t s) {
   if (nq > 0 && ne > 0)
      hipLaunchKernelGGL(k_potentials, dim3(nq), dim3(256), 0, s, ne, panels, pidx, coef, self_ok, qpos, phi);
}
void launch_leaf_diag(int nleaf, const LeafDesc* leaves, const double* D, double* ds, double* dd, hipStream_t s) {
   if (nleaf > 0) hipLaunchKernelGGL(k_leaf_diag, dim3(nleaf), dim3(64), 0, s, nleaf, leaves, D, ds, dd);
}
void launch_integrate_all(int kernel, double kappa, const PanelGeom* panels, int npan, const double* pts, int npt,
                          double* slp, double* dlp, hipStream_t s) {
   long long n = (long long)npan * npt;
   if (n > 0)
      hipLaunchKernelGGL(k_integrate_all, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, kernel, kappa, panels,
                         npan, pts, npt, slp, dlp);
}

}  // namespace fsvd
