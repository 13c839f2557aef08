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

// One modified Gram-Schmidt step fused with the next reduction:
//   h = sum(pin); hout[0] = h; w -= h*v; pout = partials of <w, vnext>
// (vnext == nullptr: partials of <w, w>).
__global__ void __launch_bounds__(256) k_mgs_step(long long n, const double* __restrict__ pin, int np,
                                                  double* __restrict__ hout, double* __restrict__ w,
                                                  const double* __restrict__ v, const double* __restrict__ vnext,
                                                  double* __restrict__ pout) {
   __shared__ double red[8];
   __shared__ double hs;
   if (threadIdx.x == 0) hs = sum_partials(pin, np);
   __syncthreads();
   const double h = hs;
   if (blockIdx.x == 0 && threadIdx.x == 0) hout[0] = h;
   double s = 0.0;
   for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
      double wi = w[i] - h * v[i];
      w[i] = wi;
      double o = vnext ? vnext[i] : wi;
      s += wi * o;
   }
   s = block_sum(s, red);
   if (threadIdx.x == 0) pout[blockIdx.x] = s;
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
void launch_mgs_step(long long n, int np, const double* pin, double* hout, double* w, const double* v,
                     const double* vnext, double* pout, hipStream_t s) {
   hipLaunchKernelGGL(k_mgs_step, dim3(np), dim3(256), 0, s, n, pin, np, hout, w, v, vnext, pout);
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
