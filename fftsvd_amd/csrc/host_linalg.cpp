// Small dense host kernels used once per tree level: the equivalent-density
// pseudoinverse and the translation spectra.
#include <cfloat>
#include <cmath>
#include <complex>
#include <cstring>
#include <thread>
#include <vector>

#include "host_linalg.h"

namespace fsvd {

// Moore-Penrose pseudoinverse (MATLAB tolerance max(m,n)*s_max*eps), as the
// reference's Matrix_pseudoinverse (Matrix.c:467-541) defines it; the SVD is
// a one-sided Jacobi iteration on X^T instead of LAPACK dgesvd.
void pinv(const double* X, int m, int n, double* XI) {
   std::vector<double> A((size_t)n * m), V((size_t)m * m, 0.0);
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
            double rel = std::fabs(ga) / std::sqrt(al * be);
            if (rel > off) off = rel;
            if (rel < 1e-15) continue;
            double zeta = (be - al) / (2.0 * ga);
            double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
            double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
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
   std::vector<double> sig(m);
   double smax = 0.0;
   for (int j = 0; j < m; j++) {
      double s = 0;
      for (int i = 0; i < n; i++) s += A[(size_t)i * m + j] * A[(size_t)i * m + j];
      sig[j] = std::sqrt(s);
      if (sig[j] > smax) smax = sig[j];
   }
   double tol = (double)(m > n ? m : n) * smax * DBL_EPSILON;
   std::memset(XI, 0, sizeof(double) * (size_t)n * m);
   for (int j = 0; j < m; j++) {
      if (!(sig[j] > tol)) continue;
      double inv2 = 1.0 / (sig[j] * sig[j]);
      for (int i = 0; i < n; i++) {
         double a = A[(size_t)i * m + j] * inv2;
         for (int k = 0; k < m; k++) XI[(size_t)i * m + k] += a * V[(size_t)k * m + j];
      }
   }
}

// Real-to-complex 3D DFT (forward sign, unnormalised) of an N^3 grid into
// N x N x (N/2+1); separable direct sums in FP64.
void dft3_r2c(int N, const double* in, std::complex<double>* out) {
   const int H = N / 2 + 1;
   std::vector<std::complex<double>> w(N), t1((size_t)N * N * H), t2((size_t)N * N * H);
   for (int m = 0; m < N; m++) w[m] = std::polar(1.0, -2.0 * M_PI * m / N);
   for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
         for (int c = 0; c < H; c++) {
            std::complex<double> s = 0;
            for (int k = 0; k < N; k++) s += in[((size_t)i * N + j) * N + k] * w[(c * k) % N];
            t1[((size_t)i * N + j) * H + c] = s;
         }
   for (int i = 0; i < N; i++)
      for (int b = 0; b < N; b++)
         for (int c = 0; c < H; c++) {
            std::complex<double> s = 0;
            for (int j = 0; j < N; j++) s += t1[((size_t)i * N + j) * H + c] * w[(b * j) % N];
            t2[((size_t)i * N + b) * H + c] = s;
         }
   for (int a = 0; a < N; a++)
      for (int b = 0; b < N; b++)
         for (int c = 0; c < H; c++) {
            std::complex<double> s = 0;
            for (int i = 0; i < N; i++) s += t2[((size_t)i * N + b) * H + c] * w[(a * i) % N];
            out[((size_t)a * N + b) * H + c] = s;
         }
}

// Translation spectra of one level: for each cube offset d in [-3,3]^3
// outside [-1,1]^3, the DFT of G(h (d (n-1) - l)) over the (2n-1)^3
// wrapped grid (reference FFT_calcDiagonalTranslationOperator,
// FFT.c:10-62).  out: 343 slots x pad complex (inner slots left zero).
void translation_spectra(int n, double h, int kernel, double kappa, std::complex<double>* out) {
   const int N = 2 * n - 1, H = N / 2 + 1;
   const size_t pad = (size_t)N * N * H;
   std::vector<int> offs;
   for (int x = -3; x <= 3; x++)
      for (int y = -3; y <= 3; y++)
         for (int z = -3; z <= 3; z++)
            if (!(x >= -1 && x <= 1 && y >= -1 && y <= 1 && z >= -1 && z <= 1)) offs.push_back((x + 3) * 49 + (y + 3) * 7 + (z + 3));
   std::memset((void*)out, 0, sizeof(std::complex<double>) * pad * 343);
   unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
   std::vector<std::thread> pool;
   for (unsigned t = 0; t < nth; t++)
      pool.emplace_back([&, t]() {
         std::vector<double> data((size_t)N * N * N);
         for (size_t q = t; q < offs.size(); q += nth) {
            int o = offs[q];
            int dx = o / 49 - 3, dy = (o / 7) % 7 - 3, dz = o % 7 - 3;
            double sh[3] = {(double)dx * (n - 1), (double)dy * (n - 1), (double)dz * (n - 1)};
            for (int i = 0; i < N; i++) {
               int li = (i > N / 2) ? N - i : -i;
               for (int j = 0; j < N; j++) {
                  int lj = (j > N / 2) ? N - j : -j;
                  for (int k = 0; k < N; k++) {
                     int lk = (k > N / 2) ? N - k : -k;
                     double p0 = h * (sh[0] - li), p1 = h * (sh[1] - lj), p2 = h * (sh[2] - lk);
                     double r = std::sqrt(p0 * p0 + p1 * p1 + p2 * p2);
                     double g = (kernel == 1)   ? std::exp(-kappa * r) / r
                                : (kernel == 2) ? -std::exp(-kappa * r)   // d/dkappa of e^{-kr}/r
                                                : (r == 0.0 ? 0.0 : 1.0 / r);
                     data[((size_t)i * N + j) * N + k] = g;
                  }
               }
            }
            dft3_r2c(N, data.data(), out + (size_t)o * pad);
         }
      });
   for (auto& th : pool) th.join();
}

}  // namespace fsvd
