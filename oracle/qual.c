/*
 * Oracle: qualocation (second-kind, apparent-surface-charge) operator.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  FP64, single-threaded.
 *
 * One Laplace tree over all dielectric panels whose bases are sampled from
 * the double layer only (Tree_allocate(..., DOUBLE_LAYER_INT),
 * FFTSVDpbeAPI.c:1832; solveBEMproblem, FFTSVDpbeAPI.c:1989).  The unknown
 * is one value per panel.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

orc_qual* orc_qual_create(const double* verts9, int npan, const double* qpos, const double* q, int nq,
                          double idiel, double odiel, int maxpanels, int gridpoints, double svdeps) {
   orc_qual* o = (orc_qual*)calloc(1, sizeof(orc_qual));
   o->npan = npan;
   o->nq = nq;
   o->idiel = idiel;
   o->odiel = odiel;
   o->panels = (orc_panel*)calloc(npan > 0 ? npan : 1, sizeof(orc_panel));
   orc_panels_from_verts(verts9, npan, o->panels);
   o->cent = (double*)malloc(sizeof(double) * 3 * (npan > 0 ? npan : 1));
   for (int i = 0; i < npan; i++)
      for (int d = 0; d < 3; d++) o->cent[3 * i + d] = o->panels[i].c[d];
   o->qpos = (double*)malloc(sizeof(double) * 3 * (nq > 0 ? nq : 1));
   o->q = (double*)malloc(sizeof(double) * (nq > 0 ? nq : 1));
   if (nq) {
      memcpy(o->qpos, qpos, sizeof(double) * 3 * nq);
      memcpy(o->q, q, sizeof(double) * nq);
   }
   o->tree = orc_tree_build(o->panels, npan, o->cent, npan, ORC_LAPLACE, 0.0, maxpanels, gridpoints, svdeps);
   orc_tree_set_double_only(o->tree);
   orc_tree_fill(o->tree);
   /* Tree_extractdiagonal(diag, tree, DOUBLE_LAYER_INT) (Tree.c:861-875;
      Cube_extractdiagonal, Cube.c:2349-2390): leaf D_double[i][i] */
   o->diag = (double*)calloc(npan > 0 ? npan : 1, sizeof(double));
   for (int i = 0; i < o->tree->ncubes; i++) {
      orc_cube* c = &o->tree->cubes[i];
      if (!c->leaf) continue;
      for (int k = 0; k < c->npan; k++) o->diag[c->pan[k]] = c->Dd[(size_t)k * c->dcols + k];
   }
   return o;
}

void orc_qual_free(orc_qual* o) {
   if (!o) return;
   orc_tree_free(o->tree);
   orc_tree_free(o->m1m3);
   free(o->panels);
   free(o->cent);
   free(o->diag);
   free(o->qpos);
   free(o->q);
   free(o);
}

/* Solv_ecf_multiply_qual (GMRES.c:159-181): xa = area .* x, b = A_d^T xa,
 * the tree's diagonal replaced by the qualocation diagonal. */
void orc_qual_matvec(orc_qual* o, const double* x, double* b) {
   const int n = o->npan;
   double* xa = (double*)calloc(n > 0 ? n : 1, sizeof(double));
   for (int i = 0; i < n; i++) xa[i] = o->panels[i].area * x[i];
   orc_tree_apply_transpose(o->tree, 1, xa, b);
   for (int i = 0; i < n; i++) {
      b[i] -= o->diag[i] * xa[i];
      b[i] *= 1.0 / (4.0 * M_PI * o->idiel);
      b[i] += (-o->odiel / ((o->odiel - o->idiel) * o->idiel) + (o->diag[i]) / (4.0 * M_PI * o->idiel)) * xa[i];
   }
   free(xa);
}

/* Diagonal preconditioner (Preconditioner_fill_diagonal_solv_ecf_qual,
 * Preconditioner.c:49-60); its solve is x = b / p. */
void orc_qual_precond_diag(orc_qual* o, double* p) {
   for (int i = 0; i < o->npan; i++)
      p[i] = (-o->odiel / ((o->odiel - o->idiel) * o->idiel) + (o->diag[i]) / (4.0 * M_PI * o->idiel)) *
             o->panels[i].area;
}

/* QualocationOperator_makeRHS (QualocationOperator.c:33-53): through the
 * M1M3 tree, RHS = -A_d^T q / (4 pi eps_in); otherwise
 * Charge_makerhs_ecf_qual (Charge.c:129-142). */
void orc_qual_rhs(orc_qual* o, double* rhs) {
   if (o->m1m3) {
      orc_tree_apply_transpose(o->m1m3, 1, o->q, rhs);
      for (int i = 0; i < o->npan; i++) rhs[i] *= -1.0 / (4.0 * M_PI * o->idiel);
      return;
   }
   for (int i = 0; i < o->npan; i++) {
      rhs[i] = 0.0;
      for (int j = 0; j < o->nq; j++) {
         double sl, dl;
         orc_laplace(&o->qpos[3 * j], &o->panels[i], &sl, &dl);
         rhs[i] += -dl * o->q[j] / (4.0 * M_PI * o->idiel);
      }
   }
}

static double qdot(const double* a, const double* b, int n) {
   double s = 0;
   for (int i = 0; i < n; i++) s += a[i] * b[i];
   return s;
}

static void qgivens(double a, double b, double* c, double* s) {
   if (b == 0.0) {
      *c = 1.0;
      *s = 0.0;
   } else {
      double r1 = fabs(a), r2 = fabs(b);
      double L = sqrt(r1 * r1 + r2 * r2);
      *c = a / L;
      *s = -b / L;
   }
}

static void qrot(double c, double s, double* a, double* b) {
   double a1 = *a, b1 = *b;
   *a = a1 * c - b1 * s;
   *b = c * b1 + s * a1;
}

/* Left-preconditioned full GMRES with MGS and Givens rotations on the
 * qualocation operator (GMRES_QualocationOperator, GMRES.c:541-640, and
 * GMRES_solv_ecf_qual, GMRES.c:358-444, which differ only in the early exit
 * for ||P^-1 b|| < tol). */
int orc_qual_gmres(orc_qual* o, const double* rhs, double* sol, double tol, int maxit, int early_exit, double* hist) {
   const int n = o->npan;
   double* p = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
   orc_qual_precond_diag(o, p);
   double* r = (double*)calloc(n > 0 ? n : 1, sizeof(double));
   for (int k = 0; k < n; k++) r[k] = rhs[k] / p[k];
   double normr = sqrt(qdot(r, r, n));
   if (early_exit && normr < tol) {
      memset(sol, 0, sizeof(double) * n);
      free(r);
      free(p);
      return 0;
   }
   double* c = (double*)calloc(maxit + 1, sizeof(double));
   double* s = (double*)calloc(maxit + 1, sizeof(double));
   double* g = (double*)calloc(maxit + 1, sizeof(double));
   double* y = (double*)calloc(maxit + 1, sizeof(double));
   double* H = (double*)calloc((size_t)(maxit + 1) * (maxit + 1), sizeof(double));
   double** bv = (double**)calloc(maxit + 2, sizeof(double*));
   double* P = (double*)calloc(n > 0 ? n : 1, sizeof(double));
   g[0] = sqrt(qdot(r, r, n));
   bv[0] = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
   for (int k = 0; k < n; k++) bv[0][k] = r[k] * (1.0 / g[0]);
   int i;
   const int W = maxit + 1;
   for (i = 0; i < maxit; i++) {
      orc_qual_matvec(o, bv[i], P);
      for (int k = 0; k < n; k++) P[k] = P[k] / p[k];
      for (int j = 0; j <= i; j++) {
         H[(size_t)i * W + j] = qdot(P, bv[j], n);
         double h = H[(size_t)i * W + j];
         for (int k = 0; k < n; k++) P[k] -= h * bv[j][k];
      }
      H[(size_t)i * W + i + 1] = sqrt(qdot(P, P, n));
      bv[i + 1] = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
      double inv = 1.0 / H[(size_t)i * W + i + 1];
      for (int k = 0; k < n; k++) bv[i + 1][k] = P[k] * inv;
      for (int k = 0; k < i; k++) qrot(c[k], s[k], &H[(size_t)i * W + k], &H[(size_t)i * W + k + 1]);
      qgivens(H[(size_t)i * W + i], H[(size_t)i * W + i + 1], &c[i], &s[i]);
      qrot(c[i], s[i], &H[(size_t)i * W + i], &H[(size_t)i * W + i + 1]);
      g[i + 1] = 0.0;
      qrot(c[i], s[i], &g[i], &g[i + 1]);
      double res = fabs(g[i + 1]) / normr;
      if (hist) hist[i] = res;
      if (res < tol) break;
   }
   if (i == maxit) i = maxit - 1;
   for (int k = 0; k <= i; k++) y[k] = g[k];
   for (int k = i; k >= 0; k--) {
      y[k] /= H[(size_t)k * W + k];
      for (int j = k - 1; j >= 0; j--) y[j] -= H[(size_t)k * W + j] * y[k];
   }
   memset(sol, 0, sizeof(double) * n);
   for (int j = 0; j <= i; j++)
      for (int k = 0; k < n; k++) sol[k] += y[j] * bv[j][k];
   for (int j = 0; j <= i + 1; j++) free(bv[j]);
   free(bv); free(r); free(P); free(p); free(c); free(s); free(g); free(y); free(H);
   return i + 1;
}

/* Direct reaction potentials (QualocationOperator_calculateReactionPotentials,
 * QualocationOperator.c:55-70): phi[c] = sum_i sol[i] * slp(c, panel i). */
void orc_qual_potentials(orc_qual* o, const double* sol, double* phi) {
   memset(phi, 0, sizeof(double) * o->nq);
   if (o->m1m3) {
      /* Tree_multiply(phi, M1M3, sol, SINGLE_LAYER_INT) */
      double* z = (double*)calloc(o->npan > 0 ? o->npan : 1, sizeof(double));
      orc_tree_apply_both(o->m1m3, sol, z, phi);
      free(z);
      return;
   }
   for (int i = 0; i < o->npan; i++)
      for (int c = 0; c < o->nq; c++) {
         double sl, dl;
         orc_laplace(&o->qpos[3 * c], &o->panels[i], &sl, &dl);
         phi[c] += sol[i] * sl;
      }
}

orc_tree* orc_qual_tree(orc_qual* o) { return o->tree; }

void orc_qual_m1m3_build(orc_qual* o) {
   if (o->m1m3 || o->nq == 0) return;
   o->m1m3 = orc_tree_build(o->panels, o->npan, o->qpos, o->nq, ORC_LAPLACE, 0.0, o->tree->maxpanels,
                            o->tree->gridpoints, o->tree->svdeps);
   orc_tree_fill(o->m1m3);
}

orc_tree* orc_qual_m1m3_tree(orc_qual* o) { return o->m1m3; }
