/*
 * Oracle (test infrastructure only): dense linear algebra and DFTs, FP64.
 */
#include "oracle.h"
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Moore-Penrose pseudoinverse with MATLAB's default tolerance
 * max(m,n)*s_max*eps (Matrix_pseudoinverse, Matrix.c:467-541).  The
 * reference calls LAPACK dgesvd; here the SVD is a one-sided Jacobi
 * iteration on X^T.  X is m x n row-major, XI is n x m row-major. */
void orc_pinv(const double* X, int m, int n, double* XI) {
   /* A = X^T (n x m), columns are the rows of X */
   double* A = (double*)malloc(sizeof(double) * (size_t)n * m);
   double* V = (double*)calloc((size_t)m * m, sizeof(double));
   for (int i = 0; i < m; i++)
      for (int j = 0; j < n; j++) A[(size_t)j * m + i] = X[(size_t)i * n + j];
   for (int i = 0; i < m; i++) V[(size_t)i * m + i] = 1.0;
   for (int sweep = 0; sweep < 60; sweep++) {
      double off = 0.0;
      for (int p = 0; p < m - 1; p++)
         for (int q = p + 1; q < m; q++) {
            double al = 0, be = 0, ga = 0;
            for (int i = 0; i < n; i++) {
               double ap = A[(size_t)i * m + p], aq = A[(size_t)i * m + q];
               al += ap * ap;
               be += aq * aq;
               ga += ap * aq;
            }
            if (ga == 0.0) continue;
            double rel = fabs(ga) / sqrt(al * be);
            if (rel > off) off = rel;
            if (rel < 1e-15) continue;
            double zeta = (be - al) / (2.0 * ga);
            double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
            for (int i = 0; i < n; i++) {
               double ap = A[(size_t)i * m + p], aq = A[(size_t)i * m + q];
               A[(size_t)i * m + p] = c * ap - s * aq;
               A[(size_t)i * m + q] = s * ap + c * aq;
            }
            for (int i = 0; i < m; i++) {
               double vp = V[(size_t)i * m + p], vq = V[(size_t)i * m + q];
               V[(size_t)i * m + p] = c * vp - s * vq;
               V[(size_t)i * m + q] = s * vp + c * vq;
            }
         }
      if (off < 1e-15) break;
   }
   double* sig = (double*)malloc(sizeof(double) * m);
   double smax = 0.0;
   for (int j = 0; j < m; j++) {
      double s = 0;
      for (int i = 0; i < n; i++) s += A[(size_t)i * m + j] * A[(size_t)i * m + j];
      sig[j] = sqrt(s);
      if (sig[j] > smax) smax = sig[j];
   }
   double tol = (double)(m > n ? m : n) * smax * DBL_EPSILON;
   /* X^T = A = U S V^T with U = A / S  ->  X = V S U^T  ->  X^+ = U S^+ V^T */
   memset(XI, 0, sizeof(double) * (size_t)n * m);
   for (int j = 0; j < m; j++) {
      if (!(sig[j] > tol)) continue;
      double inv2 = 1.0 / (sig[j] * sig[j]);  /* u_j / s_j = a_j / s_j^2 */
      for (int i = 0; i < n; i++) {
         double a = A[(size_t)i * m + j] * inv2;
         for (int k = 0; k < m; k++) XI[(size_t)i * m + k] += a * V[(size_t)k * m + j];
      }
   }
   free(sig);
   free(A);
   free(V);
}

/* Column basis by pivoted modified Gram-Schmidt with two
 * re-orthogonalisation passes and rank cut ||col|| < eps*||A||_F
 * (Matrix_columnbasis_pivotedmgs, Matrix.c:215-307).  A is rows x cols
 * row-major; Q (rows x cols, leading `rank` columns valid) receives the
 * basis; returns the rank. */
int orc_colbasis_mgs(const double* M, int rows, int cols, double eps, double* Q) {
   double* A = (double*)malloc(sizeof(double) * (size_t)rows * cols);
   double* Rk = (double*)malloc(sizeof(double) * (size_t)cols);
   memcpy(A, M, sizeof(double) * (size_t)rows * cols);
   memset(Q, 0, sizeof(double) * (size_t)rows * cols);
   double fro = 0.0;
   for (int i = 0; i < cols; i++) {
      double d = 0.0;
      for (int j = 0; j < rows; j++) d += M[(size_t)j * cols + i] * M[(size_t)j * cols + i];
      fro += d;
   }
   fro = sqrt(fro);
   int rank = 0;
   for (int k = 0; k < cols; k++) {
      double maxn = 0.0;
      int best = 0;
      for (int l = k; l < cols; l++) {
         double cn = 0.0;
         for (int i = 0; i < rows; i++) cn += A[(size_t)i * cols + l] * A[(size_t)i * cols + l];
         cn = sqrt(cn);
         if (cn > maxn) {
            maxn = cn;
            best = l;
         }
      }
      if (k >= 1 && maxn < eps * fro) {
         rank = k;
         break;
      }
      for (int i = 0; i < rows; i++) {
         double t = A[(size_t)i * cols + best];
         A[(size_t)i * cols + best] = A[(size_t)i * cols + k];
         A[(size_t)i * cols + k] = t;
      }
      for (int i = 0; i < rows; i++) Q[(size_t)i * cols + k] = A[(size_t)i * cols + k] / maxn;
      if (k < cols - 1) {
         for (int pass = 0; pass < 2; pass++) {
            for (int j = k + 1; j < cols; j++) Rk[j] = 0.0;
            for (int i = 0; i < rows; i++) {
               double q = Q[(size_t)i * cols + k];
               for (int j = k + 1; j < cols; j++) Rk[j] += q * A[(size_t)i * cols + j];
            }
            for (int i = 0; i < rows; i++) {
               double q = Q[(size_t)i * cols + k];
               for (int j = k + 1; j < cols; j++) A[(size_t)i * cols + j] -= q * Rk[j];
            }
         }
      }
      if (k == cols - 1) rank = cols;
   }
   free(A);
   free(Rk);
   return rank;
}

/* Direct separable DFTs with FFTW's conventions: forward r2c e^{-i},
 * backward c2r e^{+i}, both unnormalised; N x N x (N/2+1) half spectrum
 * (the transforms FFT.c:66-150 obtain from FFTW). */
static void twiddles(int N, double* cs, double* sn) {
   for (int m = 0; m < N; m++) {
      cs[m] = cos(2.0 * M_PI * m / N);
      sn[m] = sin(2.0 * M_PI * m / N);
   }
}

void orc_dft_r2c(int N, const double* in, double* out) {
   int H = N / 2 + 1;
   double* cs = (double*)malloc(sizeof(double) * N);
   double* sn = (double*)malloc(sizeof(double) * N);
   twiddles(N, cs, sn);
   double* t1 = (double*)calloc((size_t)N * N * H * 2, sizeof(double));
   double* t2 = (double*)calloc((size_t)N * N * H * 2, sizeof(double));
   for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
         for (int c = 0; c < H; c++) {
            double re = 0, im = 0;
            for (int k = 0; k < N; k++) {
               int m = (c * k) % N;
               double x = in[((size_t)i * N + j) * N + k];
               re += x * cs[m];
               im -= x * sn[m];
            }
            size_t o = (((size_t)i * N + j) * H + c) * 2;
            t1[o] = re;
            t1[o + 1] = im;
         }
   for (int i = 0; i < N; i++)
      for (int b = 0; b < N; b++)
         for (int c = 0; c < H; c++) {
            double re = 0, im = 0;
            for (int j = 0; j < N; j++) {
               int m = (b * j) % N;
               size_t o = (((size_t)i * N + j) * H + c) * 2;
               re += t1[o] * cs[m] + t1[o + 1] * sn[m];
               im += t1[o + 1] * cs[m] - t1[o] * sn[m];
            }
            size_t o = (((size_t)i * N + b) * H + c) * 2;
            t2[o] = re;
            t2[o + 1] = im;
         }
   for (int a = 0; a < N; a++)
      for (int b = 0; b < N; b++)
         for (int c = 0; c < H; c++) {
            double re = 0, im = 0;
            for (int i = 0; i < N; i++) {
               int m = (a * i) % N;
               size_t o = (((size_t)i * N + b) * H + c) * 2;
               re += t2[o] * cs[m] + t2[o + 1] * sn[m];
               im += t2[o + 1] * cs[m] - t2[o] * sn[m];
            }
            size_t o = (((size_t)a * N + b) * H + c) * 2;
            out[o] = re;
            out[o + 1] = im;
         }
   free(t1);
   free(t2);
   free(cs);
   free(sn);
}

void orc_dft_c2r(int N, const double* in, double* out) {
   int H = N / 2 + 1;
   double* cs = (double*)malloc(sizeof(double) * N);
   double* sn = (double*)malloc(sizeof(double) * N);
   twiddles(N, cs, sn);
   double* t1 = (double*)calloc((size_t)N * N * H * 2, sizeof(double));
   double* t2 = (double*)calloc((size_t)N * N * H * 2, sizeof(double));
   for (int i = 0; i < N; i++)
      for (int b = 0; b < N; b++)
         for (int c = 0; c < H; c++) {
            double re = 0, im = 0;
            for (int a = 0; a < N; a++) {
               int m = (a * i) % N;
               size_t o = (((size_t)a * N + b) * H + c) * 2;
               re += in[o] * cs[m] - in[o + 1] * sn[m];
               im += in[o + 1] * cs[m] + in[o] * sn[m];
            }
            size_t o = (((size_t)i * N + b) * H + c) * 2;
            t1[o] = re;
            t1[o + 1] = im;
         }
   for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
         for (int c = 0; c < H; c++) {
            double re = 0, im = 0;
            for (int b = 0; b < N; b++) {
               int m = (b * j) % N;
               size_t o = (((size_t)i * N + b) * H + c) * 2;
               re += t1[o] * cs[m] - t1[o + 1] * sn[m];
               im += t1[o + 1] * cs[m] + t1[o] * sn[m];
            }
            size_t o = (((size_t)i * N + j) * H + c) * 2;
            t2[o] = re;
            t2[o + 1] = im;
         }
   for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
         for (int k = 0; k < N; k++) {
            size_t o0 = (((size_t)i * N + j) * H) * 2;
            double v = t2[o0];
            for (int c = 1; c < H; c++) {
               int m = (c * k) % N;
               size_t o = o0 + 2 * (size_t)c;
               double term = t2[o] * cs[m] - t2[o + 1] * sn[m];
               v += (2 * c == N) ? term : 2.0 * term;
            }
            out[((size_t)i * N + j) * N + k] = v;
         }
   free(t1);
   free(t2);
   free(cs);
   free(sn);
}
