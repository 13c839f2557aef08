/*
 * Oracle (test infrastructure only): flat-panel geometry and the Laplace /
 * Yukawa single- and double-layer collocation integrals, FP64.
 */
#include "oracle.h"
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void sub3(double* r, const double* a, const double* b) { r[0] = a[0] - b[0]; r[1] = a[1] - b[1]; r[2] = a[2] - b[2]; }
static inline void cross3(double* r, const double* a, const double* b) {
   r[0] = a[1] * b[2] - a[2] * b[1];
   r[1] = a[2] * b[0] - a[0] * b[2];
   r[2] = a[0] * b[1] - a[1] * b[0];
}
static inline void scale3(double* a, double s) { a[0] *= s; a[1] *= s; a[2] *= s; }
static inline double len3(const double* a) { return sqrt(dot3(a, a)); }
static inline void normalize3(double* a) { scale3(a, 1.0 / len3(a)); }
static inline double dist3(const double* a, const double* b) {
   return sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
}

/* FASTCAP panel moments about the centroid, up to 4th order
 * (FlatPanel.c:210-298).  Output mom[1..15] = I_mn * CS (1-based). */
static void panel_moments(orc_panel* p) {
   static const double CS[16] = {0.0, 1.0, 1.0, 1.5, 1.5, 3.75, 1.0, 3.0, 1.5, 7.5, 1.5, 1.5, 3.75, 1.5, 7.5, 3.75};
   double X[3][7], Y[3][7], I[5][5];
   memset(I, 0, sizeof(I));
   for (int i = 0; i < 3; i++) {
      X[i][1] = p->pv[i][0];
      Y[i][1] = p->pv[i][1];
      for (int j = 2; j <= 6; j++) {
         X[i][j] = X[i][j - 1] * X[i][1];
         Y[i][j] = Y[i][j - 1] * Y[i][1];
      }
   }
   I[0][0] = p->area;
   for (int s = 0; s < 3; s++) {
      const double* xa = X[s];
      const double* ya = Y[s];
      const double* xb = X[(s + 1) % 3];
      const double* yb = Y[(s + 1) % 3];
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

/* Panel frames and edge data (FlatPanel.c:5-147). */
void orc_panel_init(orc_panel* p, const double* v0, const double* v1, const double* v2) {
   memset(p, 0, sizeof(*p));
   memcpy(p->v[0], v0, 3 * sizeof(double));
   memcpy(p->v[1], v1, 3 * sizeof(double));
   memcpy(p->v[2], v2, 3 * sizeof(double));
   for (int d = 0; d < 3; d++) {
      p->c[d] = v0[d] + v1[d];
      p->c[d] = p->c[d] + v2[d];
      p->c[d] *= 1.0 / 3.0;
   }
   p->el[0] = dist3(v1, v0);
   p->el[1] = dist3(v2, v1);
   p->el[2] = dist3(v0, v2);

   sub3(p->ax[0], v2, v0);
   sub3(p->ax[1], v1, v0);
   cross3(p->ax[2], p->ax[0], p->ax[1]);
   p->area = 0.5 * len3(p->ax[2]);
   normalize3(p->ax[2]);
   normalize3(p->ax[0]);
   cross3(p->ax[1], p->ax[2], p->ax[0]);

   double e[3][3];
   sub3(e[0], v1, v0); normalize3(e[0]);
   sub3(e[1], v2, v1); normalize3(e[1]);
   sub3(e[2], v0, v2); normalize3(e[2]);
   memcpy(p->axn[0], e[0], sizeof(e[0]));
   memcpy(p->axn[1], e[1], sizeof(e[1]));
   if (fabs(dot3(e[2], p->axn[0])) < fabs(dot3(p->axn[1], p->axn[0]))) {
      memcpy(p->axn[1], e[2], sizeof(e[2]));
      scale3(p->axn[1], -1.0);
   }
   double alpha = dot3(p->axn[1], p->axn[0]);
   scale3(e[0], -alpha);
   for (int d = 0; d < 3; d++) p->axn[1][d] = p->axn[1][d] + e[0][d];
   normalize3(p->axn[1]);
   cross3(p->axn[2], p->axn[0], p->axn[1]);

   for (int i = 0; i < 3; i++) {
      double r[3];
      sub3(r, p->v[i], p->c);
      for (int a = 0; a < 3; a++) {
         p->pv[i][a] = p->ax[a][0] * r[0] + p->ax[a][1] * r[1] + p->ax[a][2] * r[2];
         p->pvn[i][a] = p->axn[a][0] * r[0] + p->axn[a][1] * r[1] + p->axn[a][2] * r[2];
      }
   }
   for (int i = 0; i < 3; i++) {
      int next = (i + 1) % 3, prev = (i + 2) % 3;
      double t[3];
      sub3(t, p->pvn[next], p->pvn[i]);
      p->en[i][0] = -t[1];
      p->en[i][1] = t[0];
      p->en[i][2] = t[2];
      p->erhs[i] = dot3(p->pvn[i], p->en[i]);
      p->eov[i] = dot3(p->pvn[prev], p->en[i]) - p->erhs[i];
   }
   for (int i = 0; i < 3; i++) {
      int next = (i + 1) % 3;
      p->cC[i] = (p->pv[next][0] - p->pv[i][0]) / p->el[i];
      p->cS[i] = (p->pv[next][1] - p->pv[i][1]) / p->el[i];
   }
   p->maxd = p->el[0];
   if (p->el[1] > p->maxd) p->maxd = p->el[1];
   if (p->el[2] > p->maxd) p->maxd = p->el[2];
   p->mind = p->el[0];
   if (p->el[1] < p->mind) p->mind = p->el[1];
   if (p->el[2] < p->mind) p->mind = p->el[2];
   panel_moments(p);
}

/* Truncated decimal constants of the far-field moment expansion
 * (FlatIntegration.c:3-10), kept literally for bitwise agreement. */
#define K_FIVE3 1.666666666667
#define K_SEVEN3 2.3333333333333
#define K_ONE6 0.16666666666667
#define K_ONE3 0.3333333333333

/* Laplace 1/r single and double layer over a flat triangle
 * (FlatIntegration_oneoverr, FlatIntegration.c:295-429). */
void orc_laplace(const double* x, const orc_panel* p, double* slp, double* dlp) {
   double r0[3];
   sub3(r0, x, p->c);
   double xn = dot3(p->ax[0], r0), yn = dot3(p->ax[1], r0), zn = dot3(p->ax[2], r0);
   double xsq = xn * xn, ysq = yn * yn, zsq = zn * zn;
   double rsq = xsq + ysq + zsq;
   double diagsq = p->maxd * p->maxd;
   if (rsq > 9.0 * diagsq) {
      const double* s = p->mom;
      double r2i = 1.0 / rsq, ri = sqrt(r2i), r3i = r2i * ri, r5i = r3i * r2i;
      double zr2i = zn * r2i;
      double ss1 = s[1] * ri;
      double ss3 = -(s[3] + s[10]) * r3i;
      double ss5 = (xsq * s[10] + (xn * yn * s[7]) + ysq * s[3]) * r5i;
      double fs = ss1 + K_ONE3 * ss3 + ss5;
      double fdsum = ss1 + ss3 + 5.0 * ss5;
      double fd = zr2i * fdsum;
      if (rsq < 36.0 * diagsq) {
         double s914 = s[9] + s[14], s813 = s[8] + s[13], s411 = s[4] + s[11];
         double s512 = s[5] + s[12], s1215 = s[12] + s[15];
         double r7i = r5i * r2i, r9i = r7i * r2i;
         double t5 = (-xn * s813 - yn * s411 + 0.1 * (s512 + s1215)) * r5i;
         double t7 = (K_FIVE3 * ((xn * xsq * s[13] + yn * ysq * s[4]) + 3.0 * xn * yn * (xn * s[11] + yn * s[8])) -
                      xsq * s1215 - ysq * s512 - xn * yn * s914) * r7i;
         double t9 = (7.0 * (K_ONE6 * (xsq * xsq * s[15] + ysq * ysq * s[5]) + xsq * ysq * s[12]) +
                      K_SEVEN3 * xn * yn * (xsq * s[14] + ysq * s[9])) * r9i;
         fs += t5 + t7 + t9;
         fdsum = 5.0 * t5 + 7.0 * t7 + 9.0 * t9;
         fd += zr2i * fdsum;
      }
      *slp = fs;
      *dlp = fd;
      return;
   }
   double za = fabs(zn);
   double xm[3], ym[3], fe[3], r[3];
   int ok = 1;
   for (int i = 0; i < 3; i++) {
      double xc = xn - p->pv[i][0], yc = yn - p->pv[i][1], zc = zn - p->pv[i][2];
      xm[i] = xc;
      ym[i] = yc;
      fe[i] = xc * xc + zc * zc;
      r[i] = sqrt(yc * yc + fe[i]);
      if (r[i] < 1.005 * za) ok = 0;
   }
   double fs = 0.0, fd = 0.0;
   for (int i = 0; i < 3; i++) {
      int nx = (i + 1) % 3;
      double v = xm[i] * p->cS[i] - ym[i] * p->cC[i];
      double arg = (r[i] + r[nx] - p->el[i]) / (r[i] + r[nx] + p->el[i]);
      if (arg > 0.0) fs -= v * log(arg);
      double s1, c1, s2, c2;
      if (ok) {
         s1 = v * r[i];
         c1 = za * (xm[i] * p->cC[i] + ym[i] * p->cS[i]);
         s2 = v * r[nx];
         c2 = za * (xm[nx] * p->cC[i] + ym[nx] * p->cS[i]);
      } else {
         s1 = (fe[i] * p->cS[i]) - (xm[i] * ym[i] * p->cC[i]);
         c1 = za * r[i] * p->cC[i];
         s2 = (fe[nx] * p->cS[i]) - (xm[nx] * ym[nx] * p->cC[i]);
         c2 = za * r[nx] * p->cC[i];
      }
      double s12 = (s1 * c2) - (s2 * c1);
      double c12 = (c1 * c2) + (s1 * s2);
      fd += atan2(s12, c12);
   }
   if (fd < 0.0) fd += 2.0 * M_PI;
   if (zn < 0.0) fd = -fd;
   fs -= zn * fd;
   if (x[0] == p->c[0] && x[1] == p->c[1] && x[2] == p->c[2]) fd = 2.0 * M_PI;
   *slp = fs;
   *dlp = fd;
}

/* 4-point Gauss-Legendre rule on [0,1], ascending nodes, weights summing
 * to one: the rule QuadratureRule_generatepoints builds by Golub-Welsch
 * (QuadratureRule.c:79-125), written here in closed form. */
void orc_gl4(double* x, double* w) {
   double a = sqrt(3.0 / 7.0 - 2.0 / 7.0 * sqrt(6.0 / 5.0));
   double b = sqrt(3.0 / 7.0 + 2.0 / 7.0 * sqrt(6.0 / 5.0));
   double wa = (18.0 + sqrt(30.0)) / 36.0, wb = (18.0 - sqrt(30.0)) / 36.0;
   x[0] = -b / 2 + 0.5; w[0] = wb / 2;
   x[1] = -a / 2 + 0.5; w[1] = wa / 2;
   x[2] = a / 2 + 0.5;  w[2] = wa / 2;
   x[3] = b / 2 + 0.5;  w[3] = wb / 2;
}

/* Radial line integral from the foot of the perpendicular, with the
 * angular range split geometrically (calcp_LogIntDoLine,
 * FlatIntegration.c:473-635); HELMHOLTZ kernel only, both layers. */
static void yuk_line(double z, double x, double y1, double y2, double kappa, int nq, const double* gx,
                     const double* gw, double* sl, double* dl) {
   double az = z > 0.0 ? z : -z;
   if (x < 0) x = -x;
   double start = y1, end;
   if (x < y1)
      end = (y2 > 10.0 * y1) ? 10.0 * y1 : y2;
   else
      end = (x > y2) ? y2 : x;
   double sz = z > 0.0 ? 1.0 : (z < 0.0 ? -1.0 : 0.0);
   double ekz = exp(-kappa * az);
   double s = 0.0, d = 0.0;
   while (start < y2) {
      double ph1 = atan2(start, x), ph2 = atan2(end, x);
      double dph = ph2 - ph1;
      for (int i = 0; i < nq; i++) {
         double th = ph1 + (ph2 - ph1) * gx[i];
         double q = x * (1.0 / cos(th));
         double R = sqrt(q * q + z * z);
         double ekr = exp(-kappa * R);
         s = s - gw[i] * dph * (ekr - ekz) / kappa;
         d = d - gw[i] * dph * (sz * ekz - (z / R) * ekr);
      }
      start = end;
      if (end == y2)
         end = 10.0 * y2;
      else if (end * 10.0 > y2)
         end = y2;
      else
         end = 10.0 * end;
   }
   *sl = s;
   *dl = d;
}

/* The radial line integral alone, for the schedule tests. */
void orc_yukawa_line(double z, double x, double y1, double y2, double kappa, int nq, const double* gx,
                     const double* gw, double* sl, double* dl) {
   yuk_line(z, x, y1, y2, kappa, nq, gx, gw, sl, dl);
}

/* One edge's contribution, observation point projected to the origin
 * (calcp_LogInt, FlatIntegration.c:637-799).  The in-plane rotation the
 * reference applies when the edge-origin plane is not already the xy plane
 * is a rigid motion of points that already lie in the panel plane; it does
 * not change the integral and is omitted. */
static void yuk_edge(const double* a, const double* b, const double* pnt, double kappa, int nq, const double* gx,
                     const double* gw, double* sl, double* dl) {
   double u[2] = {a[0] - pnt[0], a[1] - pnt[1]};
   double w[2] = {b[0] - pnt[0], b[1] - pnt[1]};
   double z = pnt[2] - a[2];
   double dv[2] = {u[0] - w[0], u[1] - w[1]};
   double L = sqrt(dv[0] * dv[0] + dv[1] * dv[1]);
   double c = dv[0] * (1.0 / L), s = dv[1] * (1.0 / L);
   double ux = s * u[0] - c * u[1], uy = c * u[0] + s * u[1];
   double wy = c * w[0] + s * w[1];
   if (fabs(ux) < 10 * DBL_EPSILON) {
      *sl = 0.0;
      *dl = 0.0;
      return;
   }
   double y1 = wy, y2 = uy;
   if (y2 < 0 && y1 < 0) {
      y1 = -y1;
      y2 = -y2;
   }
   if (y2 < y1) {
      double t = y1;
      y1 = y2;
      y2 = t;
   }
   double st = (y1 >= 0) ? y1 : 0;
   yuk_line(z, ux, st, y2, kappa, nq, gx, gw, sl, dl);
   if (y1 < 0.0) {
      double s2, d2;
      yuk_line(z, ux, 0.0, -y1, kappa, nq, gx, gw, &s2, &d2);
      *sl += s2;
      *dl += d2;
   }
}

/* d/dkappa of yuk_line's sums: the same pieces and nodes with the integrands
 * differentiated, -(e^{-kR} - e^{-k|z|})/k -> (R e^{-kR} - |z| e^{-k|z|})/k
 * + (e^{-kR} - e^{-k|z|})/k^2 and -(sgn z e^{-k|z|} - (z/R) e^{-kR}) ->
 * z (e^{-k|z|} - e^{-kR}). */
static void yuk_line_dk(double z, double x, double y1, double y2, double kappa, double* sl, double* dl) {
   double gx[4], gw[4];
   orc_gl4(gx, gw);
   double az = z > 0.0 ? z : -z;
   if (x < 0) x = -x;
   double start = y1, end;
   if (x < y1)
      end = (y2 > 10.0 * y1) ? 10.0 * y1 : y2;
   else
      end = (x > y2) ? y2 : x;
   double ekz = exp(-kappa * az);
   double s = 0.0, d = 0.0;
   while (start < y2) {
      double ph1 = atan2(start, x), ph2 = atan2(end, x);
      double dph = ph2 - ph1;
      for (int i = 0; i < 4; i++) {
         double th = ph1 + (ph2 - ph1) * gx[i];
         double q = x * (1.0 / cos(th));
         double R = sqrt(q * q + z * z);
         double ekr = exp(-kappa * R);
         double wd = gw[i] * dph;
         s = s + wd * ((R * ekr - az * ekz) / kappa + (ekr - ekz) / (kappa * kappa));
         d = d + wd * z * (ekz - ekr);
      }
      start = end;
      if (end == y2)
         end = 10.0 * y2;
      else if (end * 10.0 > y2)
         end = y2;
      else
         end = 10.0 * end;
   }
   *sl = s;
   *dl = d;
}

/* d/dkappa of orc_yukawa's 4-point scheme: edges, splits and signs as there;
 * the on-plane double layer (a constant) has derivative zero. */
void orc_yukawa_dk(const double* x, const orc_panel* p, double kappa, double* dslp, double* ddlp) {
   double r0[3], pnt[3];
   sub3(r0, x, p->c);
   pnt[0] = dot3(p->axn[0], r0);
   pnt[1] = dot3(p->axn[1], r0);
   pnt[2] = dot3(p->axn[2], r0);
   double S = 0.0, D = 0.0;
   for (int j = 0; j < 3; j++) {
      int nx = (j + 1) % 3;
      double d = dot3(pnt, p->en[j]) - p->erhs[j];
      double sg = ((d > 0.0 && p->eov[j] > 0.0) || (d < 0.0 && p->eov[j] < 0.0)) ? 1.0 : -1.0;
      const double* a = p->pvn[j];
      const double* b = p->pvn[nx];
      double u[2] = {a[0] - pnt[0], a[1] - pnt[1]};
      double w[2] = {b[0] - pnt[0], b[1] - pnt[1]};
      double z = pnt[2] - a[2];
      double dv[2] = {u[0] - w[0], u[1] - w[1]};
      double L = sqrt(dv[0] * dv[0] + dv[1] * dv[1]);
      double c = dv[0] * (1.0 / L), s = dv[1] * (1.0 / L);
      double ux = s * u[0] - c * u[1], uy = c * u[0] + s * u[1];
      double wy = c * w[0] + s * w[1];
      double sl = 0.0, dl = 0.0;
      if (!(fabs(ux) < 10 * DBL_EPSILON)) {
         double y1 = wy, y2 = uy;
         if (y2 < 0 && y1 < 0) {
            y1 = -y1;
            y2 = -y2;
         }
         if (y2 < y1) {
            double t = y1;
            y1 = y2;
            y2 = t;
         }
         yuk_line_dk(z, ux, (y1 >= 0) ? y1 : 0, y2, kappa, &sl, &dl);
         if (y1 < 0.0) {
            double s2, d2;
            yuk_line_dk(z, ux, 0.0, -y1, kappa, &s2, &d2);
            sl += s2;
            dl += d2;
         }
      }
      S += sg * sl;
      D += sg * dl;
   }
   if (fabs(pnt[2]) < 10 * FLT_EPSILON) D = 0.0;
   *dslp = S;
   *ddlp = D;
}

/* Yukawa e^{-kr}/r single and double layer (FlatIntegration_ekroverr_numerical
 * FlatIntegration.c:1035-1049 -> calcp, FlatIntegration.c:882-949). */
void orc_yukawa(const double* x, const orc_panel* p, double kappa, double* slp, double* dlp) {
   if (kappa < 1e-6) {
      orc_laplace(x, p, slp, dlp);
      return;
   }
   double gx[4], gw[4];
   orc_gl4(gx, gw);
   orc_yukawa_rule(x, p, kappa, 4, gx, gw, slp, dlp);
}

/* The same scheme with an arbitrary angular rule (nodes on [0,1], weights
 * summing to one); the reference fixes the 4-point Gauss-Legendre rule. */
void orc_yukawa_rule(const double* x, const orc_panel* p, double kappa, int nq, const double* gx, const double* gw,
                     double* slp, double* dlp) {
   double r0[3], pnt[3];
   sub3(r0, x, p->c);
   pnt[0] = dot3(p->axn[0], r0);
   pnt[1] = dot3(p->axn[1], r0);
   pnt[2] = dot3(p->axn[2], r0);
   double S = 0.0, D = 0.0;
   int left = 0;
   for (int j = 0; j < 3; j++) {
      int nx = (j + 1) % 3;
      double d = dot3(pnt, p->en[j]) - p->erhs[j];
      double sg;
      if ((d > 0.0 && p->eov[j] > 0.0) || (d < 0.0 && p->eov[j] < 0.0)) {
         left++;
         sg = 1.0;
      } else
         sg = -1.0;
      double sl, dl;
      yuk_edge(p->pvn[j], p->pvn[nx], pnt, kappa, nq, gx, gw, &sl, &dl);
      S += sg * sl;
      D += sg * dl;
   }
   if (fabs(pnt[2]) < 10 * FLT_EPSILON) D = (left == 3) ? 2.0 * M_PI : 0.0;
   *slp = S;
   *dlp = D;
}

void orc_integrate(int kernel, double kappa, const double* x, const orc_panel* p, double* slp, double* dlp) {
   if (kernel == ORC_YUKAWA)
      orc_yukawa(x, p, kappa, slp, dlp);
   else
      orc_laplace(x, p, slp, dlp);
}

/* Point kernels (GreensFunction.c:74-80,115-118). */
double orc_green(int kernel, double kappa, const double* dest, const double* src) {
   double r = dist3(dest, src);
   if (kernel == ORC_YUKAWA) return exp(-kappa * r) / r;
   if (r == 0.0) return 0.0;
   return 1.0 / r;
}

void orc_panels_from_verts(const double* v9, int n, orc_panel* out) {
   for (int i = 0; i < n; i++) orc_panel_init(&out[i], v9 + 9 * i, v9 + 9 * i + 3, v9 + 9 * i + 6);
}

size_t orc_panel_size(void) { return sizeof(orc_panel); }

void orc_integrate_batch(int kernel, double kappa, const double* v9, int npan, const double* pts, int npt,
                         double* slp, double* dlp) {
   orc_panel* P = (orc_panel*)malloc(sizeof(orc_panel) * (size_t)npan);
   orc_panels_from_verts(v9, npan, P);
   for (int i = 0; i < npt; i++)
      for (int j = 0; j < npan; j++)
         orc_integrate(kernel, kappa, pts + 3 * i, &P[j], &slp[(size_t)i * npan + j], &dlp[(size_t)i * npan + j]);
   free(P);
}

void orc_panel_geom(const double* v9, double* out) {
   orc_panel p;
   orc_panel_init(&p, v9, v9 + 3, v9 + 6);
   out[0] = p.c[0]; out[1] = p.c[1]; out[2] = p.c[2];
   out[3] = p.area; out[4] = p.maxd; out[5] = p.mind;
   for (int i = 0; i < 15; i++) out[6 + i] = p.mom[i + 1];
   for (int a = 0; a < 3; a++)
      for (int d = 0; d < 3; d++) {
         out[21 + 3 * a + d] = p.ax[a][d];
         out[30 + 3 * a + d] = p.axn[a][d];
      }
}

/* MSMS-style surface files (VertFace.c:26-202): vertex lines "x y z ...",
 * face lines "i j k genus" 1-based; flip reads "i k j"; faces of zero
 * Heron area are dropped; accessible surfaces keep genus-1 faces only. */
int orc_read_vertface(const char* vertfile, const char* facefile, int flip, int accessible, double** verts9) {
   FILE* fv = fopen(vertfile, "r");
   FILE* ff = fopen(facefile, "r");
   if (!fv || !ff) {
      if (fv) fclose(fv);
      if (ff) fclose(ff);
      return -1;
   }
   char line[128];
   int nv = 0, nf = 0;
   while (fgets(line, 128, fv)) nv++;
   rewind(fv);
   double* V = (double*)calloc((size_t)nv * 3 + 3, sizeof(double));
   for (int i = 0; i < nv; i++) {
      if (!fgets(line, 128, fv)) break;
      sscanf(line, "%lf %lf %lf", &V[3 * i], &V[3 * i + 1], &V[3 * i + 2]);
   }
   while (fgets(line, 128, ff)) nf++;
   rewind(ff);
   unsigned* F = (unsigned*)calloc((size_t)nf * 4 + 4, sizeof(unsigned));
   for (int i = 0; i < nf; i++) {
      if (!fgets(line, 128, ff)) break;
      unsigned a = 0, b = 0, c = 0, g = 0;
      sscanf(line, "%u %u %u %u", &a, &b, &c, &g);
      F[4 * i] = a - 1;
      F[4 * i + 1] = flip ? c - 1 : b - 1;
      F[4 * i + 2] = flip ? b - 1 : c - 1;
      F[4 * i + 3] = g;
   }
   fclose(fv);
   fclose(ff);
   double* out = (double*)malloc(sizeof(double) * 9 * (size_t)(nf > 0 ? nf : 1));
   int kept = 0;
   for (int f = 0; f < nf; f++) {
      const double* A = &V[3 * F[4 * f]];
      const double* B = &V[3 * F[4 * f + 1]];
      const double* C = &V[3 * F[4 * f + 2]];
      double d1 = dist3(A, B), d2 = dist3(B, C), d3 = dist3(C, A);
      double s = 0.5 * (d1 + d2 + d3);
      double area = sqrt(s * (s - d1) * (s - d2) * (s - d3));
      if (!(area > 0.0)) continue;
      if (accessible && F[4 * f + 3] != 1) continue;
      memcpy(out + 9 * kept, A, 3 * sizeof(double));
      memcpy(out + 9 * kept + 3, B, 3 * sizeof(double));
      memcpy(out + 9 * kept + 6, C, 3 * sizeof(double));
      kept++;
   }
   free(V);
   free(F);
   *verts9 = out;
   return kept;
}

void orc_free(void* p) { free(p); }

void orc_yukawa_rule_v9(const double* x, const double* v9, double kappa, int nq, const double* gx, const double* gw,
                        double* slp, double* dlp) {
   orc_panel p;
   orc_panel_init(&p, v9, v9 + 3, v9 + 6);
   orc_yukawa_rule(x, &p, kappa, nq, gx, gw, slp, dlp);
}

void orc_yukawa_dk_batch(double kappa, const double* v9, int npan, const double* pts, int npt, double* ds, double* dd) {
   orc_panel* P = (orc_panel*)malloc(sizeof(orc_panel) * (size_t)(npan > 0 ? npan : 1));
   orc_panels_from_verts(v9, npan, P);
   for (int i = 0; i < npt; i++)
      for (int j = 0; j < npan; j++)
         orc_yukawa_dk(pts + 3 * i, &P[j], kappa, &ds[(size_t)i * npan + j], &dd[(size_t)i * npan + j]);
   free(P);
}
