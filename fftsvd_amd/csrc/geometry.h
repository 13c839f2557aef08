// Flat-panel geometry (FP64), one record per panel, for the host (operator
// set-up) and the device (the allocation-free integral entry point).
// Frames, edge data and FASTCAP moments follow the reference's
// FlatPanel_allocate / FlatPanel_moments (FlatPanel.c:5-179,210-298); the
// arithmetic is kept in the reference's operation order so that centroids,
// frames and thresholds agree bit for bit with a CPU evaluation (host
// code is compiled without FMA contraction).
#pragma once

#include <hip/hip_runtime.h>

#include <math.h>

#include "internal.h"

namespace fsvd {

namespace geom {
__host__ __device__ inline double dot3(const double* a, const double* b) {
#pragma clang fp contract(off)
   return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__host__ __device__ inline void sub3(double* r, const double* a, const double* b) {
#pragma clang fp contract(off)
   r[0] = a[0] - b[0];
   r[1] = a[1] - b[1];
   r[2] = a[2] - b[2];
}
__host__ __device__ inline void cross3(double* r, const double* a, const double* b) {
#pragma clang fp contract(off)
   r[0] = a[1] * b[2] - a[2] * b[1];
   r[1] = a[2] * b[0] - a[0] * b[2];
   r[2] = a[0] * b[1] - a[1] * b[0];
}
__host__ __device__ inline void normalize3(double* a) {
#pragma clang fp contract(off)
   double s = 1.0 / sqrt(dot3(a, a));
   a[0] *= s;
   a[1] *= s;
   a[2] *= s;
}
__host__ __device__ inline double dist3(const double* a, const double* b) {
#pragma clang fp contract(off)
   return sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
}

// Moments about the centroid up to order 4 by edge integration
// (FASTCAP's scheme, FlatPanel.c:210-298).  mom[1..15] = I_mn * CS.
__host__ __device__ inline void moments(PanelGeom* p) {
#pragma clang fp contract(off)
   const double CS[16] = {0.0, 1.0, 1.0, 1.5, 1.5, 3.75, 1.0, 3.0, 1.5, 7.5, 1.5, 1.5, 3.75, 1.5, 7.5, 3.75};
   double X[3][7], Y[3][7], I[5][5];
   for (int a = 0; a < 5; a++)
      for (int b = 0; b < 5; b++) I[a][b] = 0.0;
   for (int i = 0; i < 3; i++) {
      X[i][1] = p->pv[3 * i + 0];
      Y[i][1] = p->pv[3 * i + 1];
      for (int j = 2; j <= 6; j++) {
         X[i][j] = X[i][j - 1] * X[i][1];
         Y[i][j] = Y[i][j - 1] * Y[i][1];
      }
   }
   I[0][0] = p->area;
   for (int s = 0; s < 3; s++) {
      const double *xa = X[s], *ya = Y[s], *xb = X[(s + 1) % 3], *yb = Y[(s + 1) % 3];
      double dx = xb[1] - xa[1], dy = yb[1] - ya[1];
      if (fabs(dx) >= fabs(dy)) {
         double m = dy / dx;
         for (unsigned M = 2; M <= 4; M++) {
            unsigned M1 = M + 1, M2 = M + 2;
            double si = ((xb[M1] * yb[1]) - (xa[M1] * ya[1])) / M1 + m * (xa[M2] - xb[M2]) / (M1 * M2);
            I[M][0] += si;
            for (unsigned N = 1; N <= M; N++) {
               unsigned N1 = N + 1, MN1 = M - N + 1;
               si = (xb[MN1] * yb[N1] - xa[MN1] * ya[N1]) / (MN1 * N1) - (m * N * si) / MN1;
               I[M - N][N] += si;
            }
         }
      } else {
         double m = dx / dy;
         for (unsigned M = 2; M <= 4; M++) {
            unsigned M1 = M + 1, M2 = M + 2;
            double si = (m / (M1 * M2)) * (yb[M2] - ya[M2]);
            I[0][M] += si;
            for (unsigned N = 1; N <= M; N++) {
               unsigned MN1 = M - N + 1, MN2 = MN1 + 1;
               si = m * ((xb[N] * yb[MN2] - xa[N] * ya[MN2]) / (MN1 * MN2) - (N * si / MN1));
               I[N][M - N] += si;
            }
         }
      }
   }
   int k = 0;
   p->mom[0] = 0.0;
   for (int M = 0; M <= 4; M++)
      for (int N = 0; N <= 4 - M; N++) {
         k++;
         p->mom[k] = I[M][N] * CS[k];
      }
}
// The panel centroid (the collocation point), exactly as the record holds it.
__host__ __device__ inline void panel_centroid_hd(const double* v0, const double* v1, const double* v2, double* c) {
#pragma clang fp contract(off)
   for (int d = 0; d < 3; d++) {
      double s = v0[d] + v1[d];
      s = s + v2[d];
      c[d] = s * (1.0 / 3.0);
   }
}
__host__ __device__ inline void panel_geometry_hd(const double* v0, const double* v1, const double* v2, PanelGeom* p) {
#pragma clang fp contract(off)
   double* raw = reinterpret_cast<double*>(p);
   for (int i = 0; i < (int)(sizeof(PanelGeom) / sizeof(double)); i++) raw[i] = 0.0;
   panel_centroid_hd(v0, v1, v2, p->c);
   p->el[0] = dist3(v1, v0);
   p->el[1] = dist3(v2, v1);
   p->el[2] = dist3(v0, v2);

   double* ax0 = p->ax;
   double* ax1 = p->ax + 3;
   double* ax2 = p->ax + 6;
   sub3(ax0, v2, v0);
   sub3(ax1, v1, v0);
   cross3(ax2, ax0, ax1);
   p->area = 0.5 * sqrt(dot3(ax2, ax2));
   normalize3(ax2);
   normalize3(ax0);
   cross3(ax1, ax2, ax0);

   double e[3][3];
   sub3(e[0], v1, v0);
   normalize3(e[0]);
   sub3(e[1], v2, v1);
   normalize3(e[1]);
   sub3(e[2], v0, v2);
   normalize3(e[2]);
   double* n0 = p->axn;
   double* n1 = p->axn + 3;
   double* n2 = p->axn + 6;
   for (int d = 0; d < 3; d++) {
      n0[d] = e[0][d];
      n1[d] = e[1][d];
   }
   if (fabs(dot3(e[2], n0)) < fabs(dot3(n1, n0))) {
      for (int d = 0; d < 3; d++) n1[d] = e[2][d] * -1.0;
   }
   double alpha = dot3(n1, n0);
   for (int d = 0; d < 3; d++) e[0][d] *= -alpha;
   for (int d = 0; d < 3; d++) n1[d] = n1[d] + e[0][d];
   normalize3(n1);
   cross3(n2, n0, n1);

   const double* V[3] = {v0, v1, v2};
   for (int i = 0; i < 3; i++) {
      double r[3];
      sub3(r, V[i], p->c);
      for (int a = 0; a < 3; a++) {
         p->pv[3 * i + a] = p->ax[3 * a + 0] * r[0] + p->ax[3 * a + 1] * r[1] + p->ax[3 * a + 2] * r[2];
         p->pvn[3 * i + a] = p->axn[3 * a + 0] * r[0] + p->axn[3 * a + 1] * r[1] + p->axn[3 * a + 2] * r[2];
      }
   }
   for (int i = 0; i < 3; i++) {
      int nx = (i + 1) % 3, pr = (i + 2) % 3;
      double t[3];
      sub3(t, p->pvn + 3 * nx, p->pvn + 3 * i);
      double* en = p->en + 3 * i;
      en[0] = -t[1];
      en[1] = t[0];
      en[2] = t[2];
      p->erhs[i] = dot3(p->pvn + 3 * i, en);
      p->eov[i] = dot3(p->pvn + 3 * pr, en) - p->erhs[i];
   }
   for (int i = 0; i < 3; i++) {
      int nx = (i + 1) % 3;
      p->cC[i] = (p->pv[3 * nx + 0] - p->pv[3 * i + 0]) / p->el[i];
      p->cS[i] = (p->pv[3 * nx + 1] - p->pv[3 * i + 1]) / p->el[i];
   }
   p->maxd = p->el[0];
   if (p->el[1] > p->maxd) p->maxd = p->el[1];
   if (p->el[2] > p->maxd) p->maxd = p->el[2];
   moments(p);
}

}  // namespace geom

}  // namespace fsvd
