// Region-operator glue, block preconditioner, Krylov vector kernels, the
// charge right-hand side and the reaction potentials at the charges.  All
// FP64.
#include <hip/hip_runtime.h>

#include "device.h"
#include "integrals.h"

namespace fsvd {

// Joined potentials of one region in its tree order (reference
// SurfaceOperator_joinpotentials, SurfaceOperator.c:51-65):
// xd = s_phi * x[phi_idx] (double-layer input), xs = s_dphi * x[dphi_idx].
__global__ void k_join(int n, const int* __restrict__ phi_idx, const int* __restrict__ dphi_idx,
                       const double* __restrict__ s_phi, const double* __restrict__ s_dphi, const double* __restrict__ x,
                       double* __restrict__ xs, double* __restrict__ xd) {
   int k = blockIdx.x * blockDim.x + threadIdx.x;
   if (k < n) {
      xd[k] = s_phi[k] * x[phi_idx[k]];
      xs[k] = s_dphi[k] * x[dphi_idx[k]];
   }
}

// Split back to the global result with the -4 pi jump on child panels
// (SurfaceOperator.c:141-147 and splitpotentials/joinglobal, :78-107).
__global__ void k_split(int n, const int* __restrict__ out_idx, const double* __restrict__ corr,
                        const double* __restrict__ yt, const double* __restrict__ xd, double* __restrict__ y) {
   int k = blockIdx.x * blockDim.x + threadIdx.x;
   if (k < n) y[out_idx[k]] = yt[k] + corr[k] * xd[k];
}

// x = P^{-1} b for disjoint 2x2 blocks (factored on the host); in place safe.
__global__ void k_precond(int nb, const int* __restrict__ ia, const int* __restrict__ ib,
                          const double* __restrict__ M, const double* __restrict__ b, double* __restrict__ x) {
   int k = blockIdx.x * blockDim.x + threadIdx.x;
   if (k < nb) {
      int i = ia[k], j = ib[k];
      double bi = b[i], bj = b[j];
      const double* m = M + 4 * (size_t)k;
      x[i] = m[0] * bi + m[1] * bj;
      x[j] = m[2] * bi + m[3] * bj;
   }
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

// v = w / sqrt(sum(pin)); hout[0] = the norm.
__global__ void __launch_bounds__(256) k_normalize(long long n, const double* __restrict__ pin, int np,
                                                   double* __restrict__ hout, const double* __restrict__ w,
                                                   double* __restrict__ v) {
   __shared__ double hs;
   if (threadIdx.x == 0) hs = sqrt(sum_partials(pin, np));
   __syncthreads();
   const double h = hs;
   if (blockIdx.x == 0 && threadIdx.x == 0) hout[0] = h;
   const double inv = 1.0 / h;
   for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
      v[i] = w[i] * inv;
}

// Classical Gram-Schmidt pass over m basis vectors (two passes make CGS2).
// Block (b, jg) owns the contiguous chunk b of the vectors and the columns
// 4jg..4jg+3: partial[j*nch + b] = chunk sum of V_j . w.
__global__ void __launch_bounds__(256) k_multidot(long long n, int m, int nch, const double* const* __restrict__ V,
                                                  const double* __restrict__ w, double* __restrict__ partial) {
   __shared__ double red[8];
   const long long chunk = (n + nch - 1) / nch;
   const long long i0 = blockIdx.x * chunk, i1 = i0 + chunk < n ? i0 + chunk : n;
   const int j0 = blockIdx.y * 4;
   const int nj = m - j0 < 4 ? m - j0 : 4;
   const double* v0 = V[j0];
   const double* v1 = nj > 1 ? V[j0 + 1] : v0;
   const double* v2 = nj > 2 ? V[j0 + 2] : v0;
   const double* v3 = nj > 3 ? V[j0 + 3] : v0;
   double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
   for (long long i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
      const double wi = w[i];
      s0 += v0[i] * wi;
      s1 += v1[i] * wi;
      s2 += v2[i] * wi;
      s3 += v3[i] * wi;
   }
   s0 = block_sum(s0, red);
   s1 = block_sum(s1, red);
   s2 = block_sum(s2, red);
   s3 = block_sum(s3, red);
   if (threadIdx.x == 0) {
      partial[(long long)j0 * nch + blockIdx.x] = s0;
      if (nj > 1) partial[(long long)(j0 + 1) * nch + blockIdx.x] = s1;
      if (nj > 2) partial[(long long)(j0 + 2) * nch + blockIdx.x] = s2;
      if (nj > 3) partial[(long long)(j0 + 3) * nch + blockIdx.x] = s3;
   }
}

// h[j] = sum of column j's nch partials (one wave per column, lanes over
// the partials, then a fixed-order wave reduction); hacc[j] = h[j] (first
// pass) or hacc[j] + h[j] (second pass).
__global__ void __launch_bounds__(64) k_multidot_finish(int m, int nch, const double* __restrict__ partial,
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

// One Arnoldi column's Givens step on the device (reference GMRES.c:3-22,
// 495-510): the previous rotations applied to H[:, i] = hcol, the new
// rotation from (H[i,i], H[i+1,i]), g updated, |g[i+1]|/normr recorded.
// One thread; H is column-major with leading dimension W.
__global__ void k_arnoldi_givens(int i, int W, const double* __restrict__ hcol, double* __restrict__ H,
                                 double* __restrict__ cs, double* __restrict__ sn, double* __restrict__ g,
                                 double* __restrict__ hist, double normr) {
   if (threadIdx.x != 0 || blockIdx.x != 0) return;
   double* Hi = H + (long long)i * W;
   for (int j = 0; j <= i + 1; j++) Hi[j] = hcol[j];
   for (int k = 0; k < i; k++) {
      const double a = Hi[k], b = Hi[k + 1];
      Hi[k] = a * cs[k] - b * sn[k];
      Hi[k + 1] = cs[k] * b + sn[k] * a;
   }
   double c = 1.0, s = 0.0;
   if (Hi[i + 1] != 0.0) {
      const double r1 = fabs(Hi[i]), r2 = fabs(Hi[i + 1]);
      const double L = sqrt(r1 * r1 + r2 * r2);
      c = Hi[i] / L;
      s = -Hi[i + 1] / L;
   }
   cs[i] = c;
   sn[i] = s;
   {
      const double a = Hi[i], b = Hi[i + 1];
      Hi[i] = a * c - b * s;
      Hi[i + 1] = c * b + s * a;
   }
   const double gi = g[i];
   g[i] = gi * c;   // g[i+1] was zero
   g[i + 1] = s * gi;
   hist[i] = fabs(g[i + 1]) / normr;
}

// w -= sum_j h[j] V_j, summed over j in order.
__global__ void k_multiaxpy(long long n, int m, const double* __restrict__ h, const double* const* __restrict__ V,
                            double* __restrict__ w) {
   for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
      double s = 0.0;
      for (int j = 0; j < m; j++) s += h[j] * V[j][i];
      w[i] -= s;
   }
}

// x = sum_j y[j] * V[j]  (basis vectors at stride ld), in order of j.
__global__ void k_combine(long long n, int m, const double* __restrict__ y, const double* const* __restrict__ V,
                          double* __restrict__ x) {
   for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
      double s = 0.0;
      for (int j = 0; j < m; j++) s += y[j] * V[j][i];
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

void launch_join(int n, const int* pi, const int* di, const double* sp, const double* sd, const double* x, double* xs,
                 double* xd, hipStream_t s) {
   if (n > 0) hipLaunchKernelGGL(k_join, dim3(nblocks(n)), dim3(256), 0, s, n, pi, di, sp, sd, x, xs, xd);
}
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
void launch_copy(long long n4, const void* a, void* b, hipStream_t s) {
   if (n4 > 0)
      hipLaunchKernelGGL(k_copy, dim3(nblocks(n4, 256, 8192)), dim3(256), 0, s, n4, (const double4*)a, (double4*)b);
}
void launch_multidot(long long n, int m, int nch, const double* const* V, const double* w, double* partial,
                     hipStream_t s) {
   if (m > 0) hipLaunchKernelGGL(k_multidot, dim3(nch, (m + 3) / 4), dim3(256), 0, s, n, m, nch, V, w, partial);
}
void launch_multidot_finish(int m, int nch, const double* partial, double* h, double* hacc, int add, hipStream_t s) {
   if (m > 0) hipLaunchKernelGGL(k_multidot_finish, dim3(m), dim3(64), 0, s, m, nch, partial, h, hacc, add);
}
void launch_arnoldi_givens(int i, int W, const double* hcol, double* H, double* cs, double* sn, double* g,
                           double* hist, double normr, hipStream_t s) {
   hipLaunchKernelGGL(k_arnoldi_givens, dim3(1), dim3(64), 0, s, i, W, hcol, H, cs, sn, g, hist, normr);
}
void launch_multiaxpy(long long n, int m, const double* h, const double* const* V, double* w, hipStream_t s) {
   if (m > 0) hipLaunchKernelGGL(k_multiaxpy, dim3(nblocks(n, 256, 4096)), dim3(256), 0, s, n, m, h, V, w);
}
void launch_normalize(long long n, int np, const double* pin, double* hout, const double* w, double* v,
                      hipStream_t s) {
   hipLaunchKernelGGL(k_normalize, dim3(np), dim3(256), 0, s, n, pin, np, hout, w, v);
}
void launch_combine(long long n, int m, const double* y, const double* const* V, double* x, hipStream_t s) {
   hipLaunchKernelGGL(k_combine, dim3(nblocks(n, 256, 4096)), dim3(256), 0, s, n, m, y, V, x);
}
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
