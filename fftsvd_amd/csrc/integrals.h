// Device evaluation of the flat-triangle collocation integrals, FP64.
//
//   Laplace 1/r single and double layer: closed form (Hess-Smith edge sums
//   and the atan2 solid angle) near the panel, FASTCAP moment expansion when
//   r^2 > 9 diag^2 (reference FlatIntegration_oneoverr,
//   FlatIntegration.c:295-429).
//   Yukawa e^{-kr}/r single and double layer: per edge, radial integrals in
//   closed form and the angle by a 4-point Gauss-Legendre rule on
//   geometrically growing sub-intervals (reference calcp and helpers,
//   FlatIntegration.c:473-949,1035-1049).  Both layers are produced in one
//   pass, which the reference does in four separate passes.
//
// Contraction into FMA is disabled in these functions so that the geometric
// branch decisions (moment regime, on-plane tests, exact self-point match)
// see the same rounded values as a plain CPU evaluation.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.h"

namespace fsvd {

__device__ __forceinline__ void laplace_integrals(const double* __restrict__ x, const PanelGeom& p, double& slp,
                                                  double& dlp) {
#pragma clang fp contract(off)
   double r0 = x[0] - p.c[0], r1 = x[1] - p.c[1], r2 = x[2] - p.c[2];
   double xn = p.ax[0] * r0 + p.ax[1] * r1 + p.ax[2] * r2;
   double yn = p.ax[3] * r0 + p.ax[4] * r1 + p.ax[5] * r2;
   double zn = p.ax[6] * r0 + p.ax[7] * r1 + p.ax[8] * r2;
   double xsq = xn * xn, ysq = yn * yn, zsq = zn * zn;
   double rsq = xsq + ysq + zsq;
   double diagsq = p.maxd * p.maxd;
   if (rsq > 9.0 * diagsq) {
      const double* s = p.mom;
      double r2i = 1.0 / rsq, ri = sqrt(r2i), r3i = r2i * ri, r5i = r3i * r2i;
      double zr2i = zn * r2i;
      double ss1 = s[1] * ri;
      double ss3 = -(s[3] + s[10]) * r3i;
      double ss5 = (xsq * s[10] + (xn * yn * s[7]) + ysq * s[3]) * r5i;
      double fs = ss1 + 0.3333333333333 * ss3 + ss5;
      double fdsum = ss1 + ss3 + 5.0 * ss5;
      double fd = zr2i * fdsum;
      if (rsq < 36.0 * diagsq) {
         double s914 = s[9] + s[14], s813 = s[8] + s[13], s411 = s[4] + s[11];
         double s512 = s[5] + s[12], s1215 = s[12] + s[15];
         double r7i = r5i * r2i, r9i = r7i * r2i;
         double t5 = (-xn * s813 - yn * s411 + 0.1 * (s512 + s1215)) * r5i;
         double t7 = (1.666666666667 * ((xn * xsq * s[13] + yn * ysq * s[4]) + 3.0 * xn * yn * (xn * s[11] + yn * s[8])) -
                      xsq * s1215 - ysq * s512 - xn * yn * s914) * r7i;
         double t9 = (7.0 * (0.16666666666667 * (xsq * xsq * s[15] + ysq * ysq * s[5]) + xsq * ysq * s[12]) +
                      2.3333333333333 * xn * yn * (xsq * s[14] + ysq * s[9])) * r9i;
         fs += t5 + t7 + t9;
         fdsum = 5.0 * t5 + 7.0 * t7 + 9.0 * t9;
         fd += zr2i * fdsum;
      }
      slp = fs;
      dlp = fd;
      return;
   }
   double za = fabs(zn);
   double xm[3], ym[3], fe[3], r[3];
   bool ok = true;
#pragma unroll
   for (int i = 0; i < 3; i++) {
      double xc = xn - p.pv[3 * i], yc = yn - p.pv[3 * i + 1], zc = zn - p.pv[3 * i + 2];
      xm[i] = xc;
      ym[i] = yc;
      fe[i] = xc * xc + zc * zc;
      r[i] = sqrt(yc * yc + fe[i]);
      if (r[i] < 1.005 * za) ok = false;
   }
   double fs = 0.0, fd = 0.0;
#pragma unroll
   for (int i = 0; i < 3; i++) {
      int nx = (i == 2) ? 0 : i + 1;
      double cC = p.cC[i], cS = p.cS[i];
      double v = xm[i] * cS - ym[i] * cC;
      double arg = (r[i] + r[nx] - p.el[i]) / (r[i] + r[nx] + p.el[i]);
      if (arg > 0.0) fs -= v * log(arg);
      double s1, c1, s2, c2;
      if (ok) {
         s1 = v * r[i];
         c1 = za * (xm[i] * cC + ym[i] * cS);
         s2 = v * r[nx];
         c2 = za * (xm[nx] * cC + ym[nx] * cS);
      } else {
         s1 = (fe[i] * cS) - (xm[i] * ym[i] * cC);
         c1 = za * r[i] * cC;
         s2 = (fe[nx] * cS) - (xm[nx] * ym[nx] * cC);
         c2 = za * r[nx] * cC;
      }
      double s12 = (s1 * c2) - (s2 * c1);
      double c12 = (c1 * c2) + (s1 * s2);
      fd += atan2(s12, c12);
   }
   if (fd < 0.0) fd += 2.0 * M_PI;
   if (zn < 0.0) fd = -fd;
   fs -= zn * fd;
   if (x[0] == p.c[0] && x[1] == p.c[1] && x[2] == p.c[2]) fd = 2.0 * M_PI;
   slp = fs;
   dlp = fd;
}

// Gauss-Legendre 4-point rule on [0,1] (closed form of the Golub-Welsch
// rule the reference builds, QuadratureRule.c:79-125).
__device__ __forceinline__ void gl4_rule(double* gx, double* gw) {
   const double a = 0.33998104358485626480, b = 0.86113631159405257522;
   const double wa = 0.65214515486254614263, wb = 0.34785484513745385737;
   gx[0] = -b / 2 + 0.5;
   gw[0] = wb / 2;
   gx[1] = -a / 2 + 0.5;
   gw[1] = wa / 2;
   gx[2] = a / 2 + 0.5;
   gw[2] = wa / 2;
   gx[3] = b / 2 + 0.5;
   gw[3] = wb / 2;
}

// Radial integrals along one edge, split at the perpendicular foot; the
// angle range grows geometrically by 10x (calcp_LogIntDoLine).
__device__ __forceinline__ void yukawa_line(double z, double x, double y1, double y2, double kappa, double& sl,
                                            double& dl) {
#pragma clang fp contract(off)
   double gx[4], gw[4];
   gl4_rule(gx, gw);
   double az = z > 0.0 ? z : -z;
   if (x < 0) x = -x;
   double start = y1, end;
   if (x < y1)
      end = (y2 > 10.0 * y1) ? 10.0 * y1 : y2;
   else
      end = (x > y2) ? y2 : x;
   double sz = z > 0.0 ? 1.0 : (z < 0.0 ? -1.0 : 0.0);
   double ekz = exp(-kappa * az);
   double ik = 1.0 / kappa;
   double s = 0.0, d = 0.0;
   int guard = 0;
   while (start < y2 && guard < 64) {
      guard++;
      double ph1 = atan2(start, x), ph2 = atan2(end, x);
      double dph = ph2 - ph1;
#pragma unroll
      for (int i = 0; i < 4; i++) {
         double th = ph1 + (ph2 - ph1) * gx[i];
         double q = x * (1.0 / cos(th));
         double R = sqrt(q * q + z * z);
         double ekr = exp(-kappa * R);
         double wd = gw[i] * dph;
         s = s - wd * (ekr - ekz) * ik;
         d = d - wd * (sz * ekz - (z / R) * ekr);
      }
      start = end;
      if (end == y2)
         end = 10.0 * y2;
      else if (end * 10.0 > y2)
         end = y2;
      else
         end = 10.0 * end;
   }
   sl = s;
   dl = d;
}

__device__ __forceinline__ void yukawa_integrals(const double* __restrict__ x, const PanelGeom& p, double kappa,
                                                 double& slp, double& dlp) {
#pragma clang fp contract(off)
   if (kappa < 1e-6) {
      laplace_integrals(x, p, slp, dlp);
      return;
   }
   double r0 = x[0] - p.c[0], r1 = x[1] - p.c[1], r2 = x[2] - p.c[2];
   double pn[3];
   pn[0] = p.axn[0] * r0 + p.axn[1] * r1 + p.axn[2] * r2;
   pn[1] = p.axn[3] * r0 + p.axn[4] * r1 + p.axn[5] * r2;
   pn[2] = p.axn[6] * r0 + p.axn[7] * r1 + p.axn[8] * r2;
   double S = 0.0, D = 0.0;
   int left = 0;
   for (int j = 0; j < 3; j++) {
      int nx = (j == 2) ? 0 : j + 1;
      const double* en = p.en + 3 * j;
      double dd = (pn[0] * en[0] + pn[1] * en[1] + pn[2] * en[2]) - p.erhs[j];
      double sg;
      if ((dd > 0.0 && p.eov[j] > 0.0) || (dd < 0.0 && p.eov[j] < 0.0)) {
         left++;
         sg = 1.0;
      } else {
         sg = -1.0;
      }
      // edge from vertex j to vertex nx, observation point projected to the origin
      const double* a = p.pvn + 3 * j;
      const double* b = p.pvn + 3 * nx;
      double u0 = a[0] - pn[0], u1 = a[1] - pn[1];
      double w0 = b[0] - pn[0], w1 = b[1] - pn[1];
      double z = pn[2] - a[2];
      double dv0 = u0 - w0, dv1 = u1 - w1;
      double L = sqrt(dv0 * dv0 + dv1 * dv1);
      double il = 1.0 / L;
      double c = dv0 * il, s = dv1 * il;
      double ux = s * u0 - c * u1, uy = c * u0 + s * u1;
      double wy = c * w0 + s * w1;
      double sl = 0.0, dl = 0.0;
      if (!(fabs(ux) < 10 * 2.220446049250313e-16)) {
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
         yukawa_line(z, ux, st, y2, kappa, sl, dl);
         if (y1 < 0.0) {
            double s2, d2;
            yukawa_line(z, ux, 0.0, -y1, kappa, s2, d2);
            sl += s2;
            dl += d2;
         }
      }
      S += sg * sl;
      D += sg * dl;
   }
   if (fabs(pn[2]) < 10 * 1.1920928955078125e-07) D = (left == 3) ? 2.0 * M_PI : 0.0;
   slp = S;
   dlp = D;
}

// kappa-derivative of yukawa_line's two sums: the same pieces and nodes,
// the integrands differentiated:
//   -(e^{-kR} - e^{-k|z|})/k          -> (R e^{-kR} - |z| e^{-k|z|})/k + (e^{-kR} - e^{-k|z|})/k^2
//   -(sgn z e^{-k|z|} - (z/R) e^{-kR}) -> z (e^{-k|z|} - e^{-kR})
__device__ __forceinline__ void yukawa_line_dk(double z, double x, double y1, double y2, double kappa, double& sl,
                                               double& dl) {
#pragma clang fp contract(off)
   double gx[4], gw[4];
   gl4_rule(gx, gw);
   double az = z > 0.0 ? z : -z;
   if (x < 0) x = -x;
   double start = y1, end;
   if (x < y1)
      end = (y2 > 10.0 * y1) ? 10.0 * y1 : y2;
   else
      end = (x > y2) ? y2 : x;
   double ekz = exp(-kappa * az);
   double ik = 1.0 / kappa;
   double ik2 = 1.0 / (kappa * kappa);
   double s = 0.0, d = 0.0;
   int guard = 0;
   while (start < y2 && guard < 64) {
      guard++;
      double ph1 = atan2(start, x), ph2 = atan2(end, x);
      double dph = ph2 - ph1;
#pragma unroll
      for (int i = 0; i < 4; i++) {
         double th = ph1 + (ph2 - ph1) * gx[i];
         double q = x * (1.0 / cos(th));
         double R = sqrt(q * q + z * z);
         double ekr = exp(-kappa * R);
         double wd = gw[i] * dph;
         s = s + wd * ((R * ekr - az * ekz) * ik + (ekr - ekz) * ik2);
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
   sl = s;
   dl = d;
}

// kappa-derivative of yukawa_integrals (the 4-point scheme differentiated
// term by term; the on-plane double layer, a constant, has derivative 0).
__device__ __forceinline__ void yukawa_dk_integrals(const double* __restrict__ x, const PanelGeom& p, double kappa,
                                                    double& dslp, double& ddlp) {
#pragma clang fp contract(off)
   double r0 = x[0] - p.c[0], r1 = x[1] - p.c[1], r2 = x[2] - p.c[2];
   double pn[3];
   pn[0] = p.axn[0] * r0 + p.axn[1] * r1 + p.axn[2] * r2;
   pn[1] = p.axn[3] * r0 + p.axn[4] * r1 + p.axn[5] * r2;
   pn[2] = p.axn[6] * r0 + p.axn[7] * r1 + p.axn[8] * r2;
   double S = 0.0, D = 0.0;
   for (int j = 0; j < 3; j++) {
      int nx = (j == 2) ? 0 : j + 1;
      const double* en = p.en + 3 * j;
      double dd = (pn[0] * en[0] + pn[1] * en[1] + pn[2] * en[2]) - p.erhs[j];
      double sg = ((dd > 0.0 && p.eov[j] > 0.0) || (dd < 0.0 && p.eov[j] < 0.0)) ? 1.0 : -1.0;
      const double* a = p.pvn + 3 * j;
      const double* b = p.pvn + 3 * nx;
      double u0 = a[0] - pn[0], u1 = a[1] - pn[1];
      double w0 = b[0] - pn[0], w1 = b[1] - pn[1];
      double z = pn[2] - a[2];
      double dv0 = u0 - w0, dv1 = u1 - w1;
      double L = sqrt(dv0 * dv0 + dv1 * dv1);
      double il = 1.0 / L;
      double c = dv0 * il, s = dv1 * il;
      double ux = s * u0 - c * u1, uy = c * u0 + s * u1;
      double wy = c * w0 + s * w1;
      double sl = 0.0, dl = 0.0;
      if (!(fabs(ux) < 10 * 2.220446049250313e-16)) {
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
         yukawa_line_dk(z, ux, st, y2, kappa, sl, dl);
         if (y1 < 0.0) {
            double s2, d2;
            yukawa_line_dk(z, ux, 0.0, -y1, kappa, s2, d2);
            sl += s2;
            dl += d2;
         }
      }
      S += sg * sl;
      D += sg * dl;
   }
   if (fabs(pn[2]) < 10 * 1.1920928955078125e-07) D = 0.0;
   dslp = S;
   ddlp = D;
}

__device__ __forceinline__ void panel_integrals(int kernel, double kappa, const double* x, const PanelGeom& p,
                                                double& slp, double& dlp) {
   if (kernel == YUKAWA)
      yukawa_integrals(x, p, kappa, slp, dlp);
   else if (kernel == YUKAWA_DKAPPA)
      yukawa_dk_integrals(x, p, kappa, slp, dlp);
   else
      laplace_integrals(x, p, slp, dlp);
}

// Point kernels (GreensFunction.c:74-80,115-118): 1/r (0 at r = 0), e^{-kr}/r,
// and the kappa-derivative of the latter, -e^{-kr}.
__device__ __forceinline__ double green(int kernel, double kappa, const double* a, const double* b) {
   double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
   double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
   if (kernel == YUKAWA) return exp(-kappa * r) / r;
   if (kernel == YUKAWA_DKAPPA) return -exp(-kappa * r);   // d/dkappa of e^{-kr}/r
   if (r == 0.0) return 0.0;
   return 1.0 / r;
}

}  // namespace fsvd
