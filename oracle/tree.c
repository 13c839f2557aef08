/*
 * Oracle (test infrastructure only): the FFT-SVD compressed operator of one
 * tree -- octree, interaction lists, fill and apply -- in FP64.
 */
#include "oracle.h"
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static const double SPHERE49[49][3] = {
#include "sphere49.inc"
};

#define NQ 49

static int add_cube(orc_tree* t) {
   if (t->ncubes == t->cap) {
      t->cap = t->cap ? 2 * t->cap : 64;
      t->cubes = (orc_cube*)realloc(t->cubes, sizeof(orc_cube) * (size_t)t->cap);
   }
   memset(&t->cubes[t->ncubes], 0, sizeof(orc_cube));
   return t->ncubes++;
}

static inline int side(double v, double c, int half) { return half == 0 ? (v < c) : (v >= c); }

/* Recursive centroid split (Cube_allocate, Cube.c:352-484). */
static int build_cube(orc_tree* t, const int* pan, int npan, const int* pt, int npt, int level, const int idx[3],
                      const double b0[3], const double b1[3], int parent) {
   int id = add_cube(t);
   orc_cube* c = &t->cubes[id];
   c->level = level;
   memcpy(c->idx, idx, sizeof(int) * 3);
   memcpy(c->b0, b0, sizeof(double) * 3);
   memcpy(c->b1, b1, sizeof(double) * 3);
   for (int d = 0; d < 3; d++) {
      c->ctr[d] = b0[d] + b1[d];
      c->ctr[d] *= 0.5;
   }
   c->parent = parent;
   for (int k = 0; k < 8; k++) c->child[k] = -1;
   c->npan = npan;
   c->npt = npt;
   c->pan = (int*)malloc(sizeof(int) * (npan > 0 ? npan : 1));
   c->pt = (int*)malloc(sizeof(int) * (npt > 0 ? npt : 1));
   memcpy(c->pan, pan, sizeof(int) * npan);
   memcpy(c->pt, pt, sizeof(int) * npt);
   if (level == t->depth - 1) {
      c->leaf = 1;
      return id;
   }
   double ctr[3];
   memcpy(ctr, c->ctr, sizeof(ctr));
   int* cp = (int*)malloc(sizeof(int) * (npan > 0 ? npan : 1));
   int* cq = (int*)malloc(sizeof(int) * (npt > 0 ? npt : 1));
   for (int cx = 0; cx <= 1; cx++)
      for (int cy = 0; cy <= 1; cy++)
         for (int cz = 0; cz <= 1; cz++) {
            int ncp = 0, ncq = 0;
            for (int i = 0; i < npan; i++) {
               const double* g = t->panels[pan[i]].c;
               if (side(g[0], ctr[0], cx) && side(g[1], ctr[1], cy) && side(g[2], ctr[2], cz)) cp[ncp++] = pan[i];
            }
            for (int i = 0; i < npt; i++) {
               const double* g = &t->pts[3 * pt[i]];
               if (side(g[0], ctr[0], cx) && side(g[1], ctr[1], cy) && side(g[2], ctr[2], cz)) cq[ncq++] = pt[i];
            }
            if (ncp == 0 && ncq == 0) continue;
            double cb0[3], cb1[3];
            memcpy(cb0, b0, sizeof(cb0));
            memcpy(cb1, ctr, sizeof(cb1));
            int sel[3] = {cx, cy, cz};
            for (int d = 0; d < 3; d++)
               if (sel[d] == 1) {
                  double diff = ctr[d] - cb0[d];
                  cb0[d] += diff;
                  cb1[d] += diff;
               }
            int cidx[3] = {2 * idx[0] + cx, 2 * idx[1] + cy, 2 * idx[2] + cz};
            int ch = build_cube(t, cp, ncp, cq, ncq, level + 1, cidx, cb0, cb1, id);
            t->cubes[id].child[cx * 4 + cy * 2 + cz] = ch;
         }
   free(cp);
   free(cq);
   return id;
}

static void free_cubes(orc_tree* t) {
   for (int i = 0; i < t->ncubes; i++) {
      orc_cube* c = &t->cubes[i];
      free(c->pan); free(c->pt); free(c->loc); free(c->itr);
      free(c->Ds); free(c->Dd); free(c->VT); free(c->U);
      free(c->PVs); free(c->PVd); free(c->UTI);
      free(c->spec_src); free(c->spec_acc);
   }
   t->ncubes = 0;
}

/* Local (|d|<=1) and interacting lists (Cube_lists, Cube.c:600-713). */
static void make_lists(orc_tree* t) {
   for (int id = 0; id < t->ncubes; id++) {
      orc_cube* c = &t->cubes[id];
      c->nloc = 0;
      c->nint = 0;
      if (c->parent < 0) continue;
      orc_cube* P = &t->cubes[c->parent];
      int cap = 8 + 8 * P->nloc;
      c->loc = (int*)malloc(sizeof(int) * cap);
      c->itr = (int*)malloc(sizeof(int) * cap);
      for (int k = 0; k < 8; k++)
         if (P->child[k] >= 0 && P->child[k] != id) c->loc[c->nloc++] = P->child[k];
      for (int i = 0; i < P->nloc; i++) {
         orc_cube* L = &t->cubes[P->loc[i]];
         for (int k = 0; k < 8; k++) {
            int o = L->child[k];
            if (o < 0) continue;
            orc_cube* O = &t->cubes[o];
            if (abs(O->idx[0] - c->idx[0]) <= 1 && abs(O->idx[1] - c->idx[1]) <= 1 && abs(O->idx[2] - c->idx[2]) <= 1)
               c->loc[c->nloc++] = o;
         }
      }
      for (int i = 0; i < P->nloc; i++) {
         orc_cube* L = &t->cubes[P->loc[i]];
         for (int k = 0; k < 8; k++) {
            int o = L->child[k];
            if (o < 0) continue;
            orc_cube* O = &t->cubes[o];
            if (abs(O->idx[0] - c->idx[0]) > 1 || abs(O->idx[1] - c->idx[1]) > 1 || abs(O->idx[2] - c->idx[2]) > 1)
               c->itr[c->nint++] = o;
         }
      }
   }
}

/* Tree_allocate (Tree.c:12-178) + Tree_lists (Tree.c:250-252). */
orc_tree* orc_tree_build(const orc_panel* panels, int npan, const double* pts, int npt, int kernel, double kappa,
                         int maxpanels, int gridpoints, double svdeps) {
   orc_tree* t = (orc_tree*)calloc(1, sizeof(orc_tree));
   t->kernel = kernel;
   t->kappa = kappa;
   t->npan = npan;
   t->npt = npt;
   t->panels = panels;
   t->pts = (double*)malloc(sizeof(double) * 3 * (npt > 0 ? npt : 1));
   memcpy(t->pts, pts, sizeof(double) * 3 * npt);
   t->maxpanels = maxpanels;
   t->gridpoints = gridpoints;
   t->svdeps = svdeps;
   int* pan = (int*)malloc(sizeof(int) * (npan > 0 ? npan : 1));
   int* pt = (int*)malloc(sizeof(int) * (npt > 0 ? npt : 1));
   for (int i = 0; i < npan; i++) pan[i] = i;
   for (int i = 0; i < npt; i++) pt[i] = i;
   double mn[3] = {1000000.0, 1000000.0, 1000000.0}, mx[3] = {-1000000.0, -1000000.0, -1000000.0};
   for (int i = 0; i < npan; i++)
      for (int d = 0; d < 3; d++) {
         if (panels[i].c[d] < mn[d]) mn[d] = panels[i].c[d];
         if (panels[i].c[d] > mx[d]) mx[d] = panels[i].c[d];
      }
   for (int i = 0; i < npt; i++)
      for (int d = 0; d < 3; d++) {
         if (pts[3 * i + d] < mn[d]) mn[d] = pts[3 * i + d];
         if (pts[3 * i + d] > mx[d]) mx[d] = pts[3 * i + d];
      }
   double ctr[3], ext = 0.0;
   for (int d = 0; d < 3; d++) ctr[d] = (mn[d] + mx[d]) * 0.5;
   for (int d = 0; d < 3; d++)
      if ((mx[d] - mn[d]) > ext) ext = mx[d] - mn[d];
   double b0[3], b1[3];
   for (int d = 0; d < 3; d++) {
      b0[d] = ctr[d] - ext * 0.5;
      b1[d] = ctr[d] + ext * 0.5;
   }
   int zero[3] = {0, 0, 0};
   for (t->depth = 2;; t->depth++) {
      build_cube(t, pan, npan, pt, npt, 0, zero, b0, b1, -1);
      if (npan == 0 || npt == 0) break;
      int nleaf = 0, within = 0;
      for (int i = 0; i < t->ncubes; i++)
         if (t->cubes[i].leaf) {
            nleaf++;
            if (t->cubes[i].npan <= maxpanels) within++;
         }
      if ((double)within / (double)nleaf > 0.90) break;
      free_cubes(t);
   }
   free(pan);
   free(pt);
   t->gpl = (int*)calloc(t->depth, sizeof(int));
   t->gpl[t->depth - 1] = gridpoints;
   if (t->depth >= 2) t->gpl[t->depth - 2] = gridpoints;
   for (int i = t->depth - 3, p = 1; i >= 0; i--, p++) t->gpl[i] = gridpoints + p;
   make_lists(t);
   return t;
}

static inline int gp3_of(const orc_tree* t, int level) { int n = t->gpl[level]; return n * n * n; }
static inline int pad_of(const orc_tree* t, int level) {
   int N = 2 * t->gpl[level] - 1;
   return N * N * (N / 2 + 1);
}

/* Dense near-field blocks of a leaf (Cube_fill_D, Cube.c:715-793). */
static void fill_D(orc_tree* t, orc_cube* c) {
   c->drows = c->npt;
   c->dcols = c->npan;
   for (int i = 0; i < c->nloc; i++) c->dcols += t->cubes[c->loc[i]].npan;
   if (c->drows == 0 || c->dcols == 0) return;
   c->Ds = (double*)malloc(sizeof(double) * (size_t)c->drows * c->dcols);
   c->Dd = (double*)malloc(sizeof(double) * (size_t)c->drows * c->dcols);
   int col = 0;
   for (int j = 0; j < c->npan; j++, col++)
      for (int r = 0; r < c->npt; r++)
         orc_integrate(t->kernel, t->kappa, &t->pts[3 * c->pt[r]], &t->panels[c->pan[j]],
                       &c->Ds[(size_t)r * c->dcols + col], &c->Dd[(size_t)r * c->dcols + col]);
   for (int i = 0; i < c->nloc; i++) {
      orc_cube* L = &t->cubes[c->loc[i]];
      for (int j = 0; j < L->npan; j++, col++)
         for (int r = 0; r < c->npt; r++)
            orc_integrate(t->kernel, t->kappa, &t->pts[3 * c->pt[r]], &t->panels[L->pan[j]],
                          &c->Ds[(size_t)r * c->dcols + col], &c->Dd[(size_t)r * c->dcols + col]);
   }
}

/* Per-level pinv of the sphere -> grid Green's matrix
 * (Tree_fill_pinv_Fitting, Tree.c:295-369; EquivDensity_Q2S_points,
 * EquivDensity.c:461-483). */
static void fill_pinv(orc_tree* t) {
   double root = t->cubes[0].b1[0] - t->cubes[0].b0[0];
   t->pinv = (double**)calloc(t->depth, sizeof(double*));
   for (int d = 0; d < t->depth; d++) {
      int n = t->gpl[d], gp3 = n * n * n;
      double box = root / pow(2.0, d);
      double h = box / (n - 1);
      double* Q = (double*)malloc(sizeof(double) * NQ * gp3);
      double ctr = 0.5 * box;
      double sc = box * 1.5 * 1;
      for (int q = 0; q < NQ; q++) {
         double mq[3];
         for (int a = 0; a < 3; a++) mq[a] = SPHERE49[q][a] * sc + ctr;
         int g = 0;
         for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++)
               for (int k = 0; k < n; k++, g++) {
                  double p[3] = {i * h, j * h, k * h};
                  Q[(size_t)q * gp3 + g] = orc_green(t->kernel, t->kappa, p, mq);
               }
      }
      t->pinv[d] = (double*)malloc(sizeof(double) * gp3 * NQ);
      orc_pinv(Q, NQ, gp3, t->pinv[d]);
      free(Q);
   }
}

/* Translation spectra per level and offset (Tree_fill_T, Tree.c:387-445;
 * FFT_calcDiagonalTranslationOperator, FFT.c:10-62). */
static void fill_T(orc_tree* t) {
   double root = t->cubes[0].b1[0] - t->cubes[0].b0[0];
   t->T = (double**)calloc((size_t)t->depth * 343, sizeof(double*));
   for (int d = 0; d < t->depth; d++) {
      int n = t->gpl[d], N = 2 * n - 1, pad = pad_of(t, d);
      double h = root / pow(2.0, d) / (n - 1);
      double* data = (double*)malloc(sizeof(double) * N * N * N);
      for (int x = -3; x <= 3; x++)
         for (int y = -3; y <= 3; y++)
            for (int z = -3; z <= 3; z++) {
               if (x >= -1 && x <= 1 && y >= -1 && y <= 1 && z >= -1 && z <= 1) continue;
               double sh[3] = {(double)x * (n - 1), (double)y * (n - 1), (double)z * (n - 1)};
               for (int i = 0; i < N; i++) {
                  int li = (i > N / 2) ? N - i : -i;
                  for (int j = 0; j < N; j++) {
                     int lj = (j > N / 2) ? N - j : -j;
                     for (int k = 0; k < N; k++) {
                        int lk = (k > N / 2) ? N - k : -k;
                        double p[3] = {h * (sh[0] - li), h * (sh[1] - lj), h * (sh[2] - lk)};
                        double zero[3] = {0, 0, 0};
                        data[((size_t)i * N + j) * N + k] = orc_green(t->kernel, t->kappa, p, zero);
                     }
                  }
               }
               double* T = (double*)malloc(sizeof(double) * 2 * pad);
               orc_dft_r2c(N, data, T);
               t->T[(size_t)d * 343 + (x + 3) * 49 + (y + 3) * 7 + (z + 3)] = T;
            }
      free(data);
   }
}

static void rowbasis(const double* R, int rows, int cols, double eps, double** VT, int* rank) {
   /* transpose -> column basis -> transpose (Matrix_rowbasis_pivotedmgs, Matrix.c:200-211) */
   double* A = (double*)malloc(sizeof(double) * (size_t)rows * cols);
   for (int i = 0; i < rows; i++)
      for (int j = 0; j < cols; j++) A[(size_t)j * rows + i] = R[(size_t)i * cols + j];
   double* Q = (double*)malloc(sizeof(double) * (size_t)rows * cols);
   int r = orc_colbasis_mgs(A, cols, rows, eps, Q);
   double* V = (double*)malloc(sizeof(double) * (size_t)(r > 0 ? r : 1) * cols);
   for (int i = 0; i < r; i++)
      for (int j = 0; j < cols; j++) V[(size_t)i * cols + j] = Q[(size_t)j * rows + i];
   free(A);
   free(Q);
   *VT = V;
   *rank = r;
}

static void colbasis(const double* C, int rows, int cols, double eps, double** U, int* rank) {
   double* Q = (double*)malloc(sizeof(double) * (size_t)rows * cols);
   int r = orc_colbasis_mgs(C, rows, cols, eps, Q);
   double* W = (double*)malloc(sizeof(double) * (size_t)rows * (r > 0 ? r : 1));
   for (int i = 0; i < rows; i++)
      for (int j = 0; j < r; j++) W[(size_t)i * r + j] = Q[(size_t)i * cols + j];
   free(Q);
   *U = W;
   *rank = r;
}

/* Sampled row/column bases (Cube_fill_U_VT, Cube.c:1104-1412). */
static void fill_UVT(orc_tree* t, orc_cube* c) {
   int* wpt = (int*)malloc(sizeof(int) * (c->nint > 0 ? c->nint : 1));
   int* wpn = (int*)malloc(sizeof(int) * (c->nint > 0 ? c->nint : 1));
   int npts = 0, npns = 0;
   for (int i = 0; i < c->nint; i++) {
      if (t->cubes[c->itr[i]].npt > 0) wpt[npts++] = c->itr[i];
      if (t->cubes[c->itr[i]].npan > 0) wpn[npns++] = c->itr[i];
   }
   /* a DOUBLE_LAYER_INT tree samples one layer: npts*pass rows of
      double-layer integrals and the rank test against half of those */
   const int nl = t->double_only ? 1 : 2;
   if (c->npan > 0 && npts > 0) {
      for (int pass = 1;; pass++) {
         int rows = nl * npts * pass;
         double* R = (double*)malloc(sizeof(double) * (size_t)rows * c->npan);
         for (int p = 0; p < pass; p++)
            for (int r = 0; r < npts; r++) {
               orc_cube* P = &t->cubes[wpt[r]];
               int pi = P->pt[(unsigned)p * (unsigned)P->npt / (unsigned)pass];
               for (int j = 0; j < c->npan; j++) {
                  double sl, dl;
                  orc_integrate(t->kernel, t->kappa, &t->pts[3 * pi], &t->panels[c->pan[j]], &sl, &dl);
                  if (t->double_only) {
                     R[(size_t)(p * npts + r) * c->npan + j] = dl;
                  } else {
                     R[(size_t)(p * npts + r) * c->npan + j] = sl;
                     R[(size_t)(p * npts + r + npts * pass) * c->npan + j] = dl;
                  }
               }
            }
         double* VT;
         int rank;
         rowbasis(R, rows, c->npan, t->svdeps, &VT, &rank);
         free(R);
         if (rank < (nl * pass * npts) / 2) {
            c->VT = VT;
            c->rowrank = rank;
            break;
         }
         free(VT);
      }
   }
   if (c->npt > 0 && npns > 0) {
      for (int pass = 1;; pass++) {
         int cols = nl * npns * pass;
         double* C = (double*)malloc(sizeof(double) * (size_t)c->npt * cols);
         for (int p = 0; p < pass; p++)
            for (int q = 0; q < npns; q++) {
               orc_cube* P = &t->cubes[wpn[q]];
               int pj = P->pan[(unsigned)p * (unsigned)P->npan / (unsigned)pass];
               for (int r = 0; r < c->npt; r++) {
                  double sl, dl;
                  orc_integrate(t->kernel, t->kappa, &t->pts[3 * c->pt[r]], &t->panels[pj], &sl, &dl);
                  if (t->double_only) {
                     C[(size_t)r * cols + p * npns + q] = dl;
                  } else {
                     C[(size_t)r * cols + p * npns + q] = sl;
                     C[(size_t)r * cols + p * npns + q + npns * pass] = dl;
                  }
               }
            }
         double* U;
         int rank;
         colbasis(C, c->npt, cols, t->svdeps, &U, &rank);
         free(C);
         if (rank < (nl * pass * npns) / 2) {
            c->U = U;
            c->colrank = rank;
            break;
         }
         free(U);
      }
   }
   free(wpt);
   free(wpn);
}

/* Equivalent-density projection / interpolation and the products
 * PV = P VT^T, UTI = U^T I (Cube_fill_P_I, Cube.c:795-1002;
 * Cube_fill_PV_UTI, Cube.c:1414-1522; EquivDensity_Q2S_*,
 * EquivDensity.c:461-535). */
static void fill_PV_UTI(orc_tree* t, orc_cube* c) {
   int gp3 = gp3_of(t, c->level);
   double box = c->b1[0] - c->b0[0];
   double sc = box * 1.5 * 1;
   const double* pinv = t->pinv[c->level];
   if (c->npan > 0 && c->rowrank > 0) {
      double* Qs = (double*)malloc(sizeof(double) * NQ * c->npan);
      double* Qd = (double*)malloc(sizeof(double) * NQ * c->npan);
      for (int q = 0; q < NQ; q++) {
         double mq[3];
         for (int a = 0; a < 3; a++) mq[a] = SPHERE49[q][a] * sc + c->ctr[a];
         for (int j = 0; j < c->npan; j++)
            orc_integrate(t->kernel, t->kappa, mq, &t->panels[c->pan[j]], &Qs[q * c->npan + j], &Qd[q * c->npan + j]);
      }
      double* Ps = (double*)calloc((size_t)gp3 * c->npan, sizeof(double));
      double* Pd = (double*)calloc((size_t)gp3 * c->npan, sizeof(double));
      for (int g = 0; g < gp3; g++)
         for (int q = 0; q < NQ; q++) {
            double a = pinv[(size_t)g * NQ + q];
            for (int j = 0; j < c->npan; j++) {
               Ps[(size_t)g * c->npan + j] += a * Qs[q * c->npan + j];
               Pd[(size_t)g * c->npan + j] += a * Qd[q * c->npan + j];
            }
         }
      c->PVs = (double*)calloc((size_t)gp3 * c->rowrank, sizeof(double));
      c->PVd = (double*)calloc((size_t)gp3 * c->rowrank, sizeof(double));
      for (int g = 0; g < gp3; g++)
         for (int j = 0; j < c->npan; j++) {
            double as = Ps[(size_t)g * c->npan + j], ad = Pd[(size_t)g * c->npan + j];
            for (int k = 0; k < c->rowrank; k++) {
               c->PVs[(size_t)g * c->rowrank + k] += as * c->VT[(size_t)k * c->npan + j];
               c->PVd[(size_t)g * c->rowrank + k] += ad * c->VT[(size_t)k * c->npan + j];
            }
         }
      free(Qs); free(Qd); free(Ps); free(Pd);
   }
   if (c->npt > 0 && c->colrank > 0) {
      double* Qp = (double*)malloc(sizeof(double) * NQ * c->npt);
      for (int q = 0; q < NQ; q++) {
         double mq[3];
         for (int a = 0; a < 3; a++) mq[a] = SPHERE49[q][a] * sc + c->ctr[a];
         for (int j = 0; j < c->npt; j++) Qp[q * c->npt + j] = orc_green(t->kernel, t->kappa, &t->pts[3 * c->pt[j]], mq);
      }
      /* I (npt x gp3) = (pinv * Qp)^T */
      double* I = (double*)calloc((size_t)gp3 * c->npt, sizeof(double));
      for (int g = 0; g < gp3; g++)
         for (int q = 0; q < NQ; q++) {
            double a = pinv[(size_t)g * NQ + q];
            for (int j = 0; j < c->npt; j++) I[(size_t)j * gp3 + g] += a * Qp[q * c->npt + j];
         }
      c->UTI = (double*)calloc((size_t)c->colrank * gp3, sizeof(double));
      for (int k = 0; k < c->colrank; k++)
         for (int j = 0; j < c->npt; j++) {
            double u = c->U[(size_t)j * c->colrank + k];
            for (int g = 0; g < gp3; g++) c->UTI[(size_t)k * gp3 + g] += u * I[(size_t)j * gp3 + g];
         }
      free(Qp);
      free(I);
   }
}

void orc_tree_set_double_only(orc_tree* t) { t->double_only = 1; }

/* Tree_fill (Tree.c:513-565): D, pinv, T, U/VT, PV/UTI. */
void orc_tree_fill(orc_tree* t) {
   for (int i = 0; i < t->ncubes; i++)
      if (t->cubes[i].leaf) fill_D(t, &t->cubes[i]);
   fill_pinv(t);
   fill_T(t);
   for (int i = 0; i < t->ncubes; i++)
      if (t->cubes[i].level >= 2) fill_UVT(t, &t->cubes[i]);
   for (int i = 0; i < t->ncubes; i++)
      if (t->cubes[i].level >= 2 && t->cubes[i].nint > 0) fill_PV_UTI(t, &t->cubes[i]);
   for (int i = 0; i < t->ncubes; i++) {
      orc_cube* c = &t->cubes[i];
      if (c->nint > 0) {
         int pad = pad_of(t, c->level);
         c->spec_src = (double*)calloc((size_t)2 * pad, sizeof(double));
         c->spec_acc = (double*)calloc((size_t)2 * pad, sizeof(double));
      }
   }
}

void orc_tree_free(orc_tree* t) {
   if (!t) return;
   free(t->dS);
   free(t->dD);
   free_cubes(t);
   free(t->cubes);
   if (t->pinv)
      for (int d = 0; d < t->depth; d++) free(t->pinv[d]);
   free(t->pinv);
   if (t->T)
      for (int i = 0; i < t->depth * 343; i++) free(t->T[i]);
   free(t->T);
   free(t->gpl);
   free(t->pts);
   free(t);
}

/* Zero-pad the n^3 grid into (2n-1)^3 at offset n/2 and transform
 * (FFT_forwardGridTransform, FFT.c:66-113). */
static void grid_forward(int n, const double* g, double* spec) {
   int N = 2 * n - 1, off = n / 2;
   double* P = (double*)calloc((size_t)N * N * N, sizeof(double));
   for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++)
         for (int k = 0; k < n; k++) P[((size_t)(i + off) * N + (j + off)) * N + (k + off)] = g[(i * n + j) * n + k];
   orc_dft_r2c(N, P, spec);
   free(P);
}

/* Inverse transform, crop at n/2 and scale by 1/N^3
 * (FFT_backwardGridTransform, FFT.c:117-150). */
static void grid_backward(int n, const double* spec, double* g) {
   int N = 2 * n - 1, off = n / 2;
   double* P = (double*)malloc(sizeof(double) * (size_t)N * N * N);
   orc_dft_c2r(N, spec, P);
   for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++)
         for (int k = 0; k < n; k++)
            g[(i * n + j) * n + k] = P[((size_t)(i + off) * N + (j + off)) * N + (k + off)] / (N * N * N);
   free(P);
}

static const double* T_of(orc_tree* t, int level, const orc_cube* a, const orc_cube* b) {
   int dx = a->idx[0] - b->idx[0], dy = a->idx[1] - b->idx[1], dz = a->idx[2] - b->idx[2];
   return t->T[(size_t)level * 343 + (dx + 3) * 49 + (dy + 3) * 7 + (dz + 3)];
}

static void cmac(double* acc, const double* T, const double* S, int pad) {
   for (int k = 0; k < pad; k++) {
      double tr = T[2 * k], ti = T[2 * k + 1], sr = S[2 * k], si = S[2 * k + 1];
      acc[2 * k] += tr * sr - ti * si;
      acc[2 * k + 1] += tr * si + ti * sr;
   }
}

/* Tree_multiplyboth (Tree.c:597-683) with Cube_clear_tempvectors,
 * Cube_multiplyboth_FFT_PV_VT, Cube_multiply_T, Cube_multiply_U_UTI_IFFT
 * and Cube_multiply_D (Cube.c:2304-2329,1986-2060,2062-2111,2156-2225,
 * 1730-1798).  y is overwritten. */
void orc_tree_apply_both(orc_tree* t, const double* xs, const double* xd, double* y) {
   memset(y, 0, sizeof(double) * t->npt);
   if (t->dS) {
      for (int i = 0; i < t->npt; i++) {
         double a = 0.0;
         for (int j = 0; j < t->npan; j++)
            a += t->dS[(size_t)i * t->npan + j] * xs[j] + t->dD[(size_t)i * t->npan + j] * xd[j];
         y[i] = a;
      }
      return;
   }
   for (int i = 0; i < t->ncubes; i++) {
      orc_cube* c = &t->cubes[i];
      if (c->nint > 0) {
         int pad = pad_of(t, c->level);
         memset(c->spec_src, 0, sizeof(double) * 2 * pad);
         memset(c->spec_acc, 0, sizeof(double) * 2 * pad);
      }
   }
   for (int i = 0; i < t->ncubes; i++) {
      orc_cube* c = &t->cubes[i];
      if (!(c->nint > 0 && c->npan > 0 && c->rowrank > 0)) continue;
      int gp3 = gp3_of(t, c->level), r = c->rowrank;
      double* ts = (double*)calloc(r, sizeof(double));
      double* td = (double*)calloc(r, sizeof(double));
      double* g = (double*)calloc(gp3, sizeof(double));
      for (int k = 0; k < r; k++)
         for (int j = 0; j < c->npan; j++) {
            ts[k] += c->VT[(size_t)k * c->npan + j] * xs[c->pan[j]];
            td[k] += c->VT[(size_t)k * c->npan + j] * xd[c->pan[j]];
         }
      for (int q = 0; q < gp3; q++)
         for (int k = 0; k < r; k++) g[q] += c->PVs[(size_t)q * r + k] * ts[k];
      for (int q = 0; q < gp3; q++)
         for (int k = 0; k < r; k++) g[q] += c->PVd[(size_t)q * r + k] * td[k];
      grid_forward(t->gpl[c->level], g, c->spec_src);
      free(ts); free(td); free(g);
   }
   for (int i = 0; i < t->ncubes; i++) {
      orc_cube* c = &t->cubes[i];
      if (!(c->nint > 0 && c->npt > 0)) continue;
      int pad = pad_of(t, c->level);
      for (int k = 0; k < c->nint; k++) {
         orc_cube* s = &t->cubes[c->itr[k]];
         if (s->npan == 0) continue;
         cmac(c->spec_acc, T_of(t, c->level, c, s), s->spec_src, pad);
      }
   }
   for (int i = 0; i < t->ncubes; i++) {
      orc_cube* c = &t->cubes[i];
      if (!(c->nint > 0 && c->npt > 0 && c->colrank > 0)) continue;
      int gp3 = gp3_of(t, c->level), r = c->colrank;
      double* g = (double*)malloc(sizeof(double) * gp3);
      double* w = (double*)calloc(r, sizeof(double));
      grid_backward(t->gpl[c->level], c->spec_acc, g);
      for (int k = 0; k < r; k++)
         for (int q = 0; q < gp3; q++) w[k] += c->UTI[(size_t)k * gp3 + q] * g[q];
      for (int j = 0; j < c->npt; j++) {
         double s = 0.0;
         for (int k = 0; k < r; k++) s += c->U[(size_t)j * r + k] * w[k];
         y[c->pt[j]] += s;
      }
      free(g);
      free(w);
   }
   for (int i = 0; i < t->ncubes; i++) {
      orc_cube* c = &t->cubes[i];
      if (!(c->leaf && c->drows > 0 && c->dcols > 0)) continue;
      double* sx = (double*)malloc(sizeof(double) * c->dcols);
      double* sd = (double*)malloc(sizeof(double) * c->dcols);
      int col = 0;
      for (int j = 0; j < c->npan; j++, col++) {
         sx[col] = xs[c->pan[j]];
         sd[col] = xd[c->pan[j]];
      }
      for (int l = 0; l < c->nloc; l++) {
         orc_cube* L = &t->cubes[c->loc[l]];
         for (int j = 0; j < L->npan; j++, col++) {
            sx[col] = xs[L->pan[j]];
            sd[col] = xd[L->pan[j]];
         }
      }
      for (int r = 0; r < c->drows; r++) {
         double a = 0.0, b = 0.0;
         for (int j = 0; j < c->dcols; j++) a += c->Ds[(size_t)r * c->dcols + j] * sx[j];
         for (int j = 0; j < c->dcols; j++) b += c->Dd[(size_t)r * c->dcols + j] * sd[j];
         y[c->pt[r]] += a;
         y[c->pt[r]] += b;
      }
      free(sx);
      free(sd);
   }
}

/* Exact transpose of one layer (Tree_multiply_transpose, Tree.c:769-832;
 * Cube_multiply_FFT_UTIT_UT, Cube_multiply_T_transpose,
 * Cube_multiply_V_PVT_IFFT, Cube_multiply_DT, Cube.c:1942-1984,2113-2154,
 * 2227-2273,1800-1846).  y (npan) is overwritten. */
void orc_tree_apply_transpose(orc_tree* t, int layer, const double* x, double* y) {
   memset(y, 0, sizeof(double) * t->npan);
   for (int i = 0; i < t->ncubes; i++) {
      orc_cube* c = &t->cubes[i];
      if (c->nint > 0) {
         int pad = pad_of(t, c->level);
         memset(c->spec_src, 0, sizeof(double) * 2 * pad);
         memset(c->spec_acc, 0, sizeof(double) * 2 * pad);
      }
   }
   for (int i = 0; i < t->ncubes; i++) {
      orc_cube* c = &t->cubes[i];
      if (!(c->nint > 0 && c->npt > 0 && c->colrank > 0)) continue;
      int gp3 = gp3_of(t, c->level), r = c->colrank;
      double* u = (double*)calloc(r, sizeof(double));
      double* g = (double*)calloc(gp3, sizeof(double));
      for (int j = 0; j < c->npt; j++)
         for (int k = 0; k < r; k++) u[k] += c->U[(size_t)j * r + k] * x[c->pt[j]];
      for (int k = 0; k < r; k++)
         for (int q = 0; q < gp3; q++) g[q] += c->UTI[(size_t)k * gp3 + q] * u[k];
      grid_forward(t->gpl[c->level], g, c->spec_src);
      free(u);
      free(g);
   }
   for (int i = 0; i < t->ncubes; i++) {
      orc_cube* c = &t->cubes[i];
      if (!(c->nint > 0 && c->npan > 0)) continue;
      int pad = pad_of(t, c->level);
      for (int k = 0; k < c->nint; k++) {
         orc_cube* s = &t->cubes[c->itr[k]];
         if (s->npt == 0) continue;
         cmac(c->spec_acc, T_of(t, c->level, c, s), s->spec_src, pad);
      }
   }
   for (int i = 0; i < t->ncubes; i++) {
      orc_cube* c = &t->cubes[i];
      if (!(c->nint > 0 && c->npan > 0 && c->rowrank > 0)) continue;
      int gp3 = gp3_of(t, c->level), r = c->rowrank;
      const double* PV = layer == 0 ? c->PVs : c->PVd;
      double* g = (double*)malloc(sizeof(double) * gp3);
      double* w = (double*)calloc(r, sizeof(double));
      grid_backward(t->gpl[c->level], c->spec_acc, g);
      for (int q = 0; q < gp3; q++)
         for (int k = 0; k < r; k++) w[k] += PV[(size_t)q * r + k] * g[q];
      for (int k = 0; k < r; k++)
         for (int j = 0; j < c->npan; j++) y[c->pan[j]] += c->VT[(size_t)k * c->npan + j] * w[k];
      free(g);
      free(w);
   }
   for (int i = 0; i < t->ncubes; i++) {
      orc_cube* c = &t->cubes[i];
      if (!(c->leaf && c->drows > 0 && c->dcols > 0)) continue;
      const double* D = layer == 0 ? c->Ds : c->Dd;
      double* sb = (double*)calloc(c->dcols, sizeof(double));
      for (int r = 0; r < c->drows; r++) {
         double xr = x[c->pt[r]];
         for (int j = 0; j < c->dcols; j++) sb[j] += D[(size_t)r * c->dcols + j] * xr;
      }
      int col = 0;
      for (int j = 0; j < c->npan; j++) y[c->pan[j]] += sb[col++];
      for (int l = 0; l < c->nloc; l++) {
         orc_cube* L = &t->cubes[c->loc[l]];
         for (int j = 0; j < L->npan; j++) y[L->pan[j]] += sb[col++];
      }
      free(sb);
   }
}

/* ---------------- accessors for tests ---------------- */
int orc_tree_depth(orc_tree* t) { return t->depth; }
int orc_tree_ncubes(orc_tree* t) { return t->ncubes; }
void orc_tree_gpl(orc_tree* t, int* g) { memcpy(g, t->gpl, sizeof(int) * t->depth); }

void orc_tree_cube_info(orc_tree* t, int id, int* info) {
   orc_cube* c = &t->cubes[id];
   info[0] = c->level; info[1] = c->idx[0]; info[2] = c->idx[1]; info[3] = c->idx[2];
   info[4] = c->leaf; info[5] = c->npan; info[6] = c->npt; info[7] = c->nloc; info[8] = c->nint;
   info[9] = c->rowrank; info[10] = c->colrank; info[11] = c->parent; info[12] = c->dcols;
}

void orc_tree_cube_lists(orc_tree* t, int id, int* pan, int* pt, int* loc, int* itr) {
   orc_cube* c = &t->cubes[id];
   if (pan) memcpy(pan, c->pan, sizeof(int) * c->npan);
   if (pt) memcpy(pt, c->pt, sizeof(int) * c->npt);
   if (loc) memcpy(loc, c->loc, sizeof(int) * c->nloc);
   if (itr) memcpy(itr, c->itr, sizeof(int) * c->nint);
}

void orc_tree_cube_mats(orc_tree* t, int id, double* Ds, double* Dd, double* VT, double* U, double* PVs,
                        double* PVd, double* UTI) {
   orc_cube* c = &t->cubes[id];
   size_t nd = (size_t)c->drows * c->dcols;
   int gp3 = gp3_of(t, c->level);
   if (Ds && c->Ds) memcpy(Ds, c->Ds, sizeof(double) * nd);
   if (Dd && c->Dd) memcpy(Dd, c->Dd, sizeof(double) * nd);
   if (VT && c->VT) memcpy(VT, c->VT, sizeof(double) * (size_t)c->rowrank * c->npan);
   if (U && c->U) memcpy(U, c->U, sizeof(double) * (size_t)c->npt * c->colrank);
   if (PVs && c->PVs) memcpy(PVs, c->PVs, sizeof(double) * (size_t)gp3 * c->rowrank);
   if (PVd && c->PVd) memcpy(PVd, c->PVd, sizeof(double) * (size_t)gp3 * c->rowrank);
   if (UTI && c->UTI) memcpy(UTI, c->UTI, sizeof(double) * (size_t)gp3 * c->colrank);
}

static double* dup(const double* s, size_t n) {
   if (!s || n == 0) return NULL;
   double* d = (double*)malloc(sizeof(double) * n);
   memcpy(d, s, sizeof(double) * n);
   return d;
}

void orc_tree_set_cube_mats(orc_tree* t, int id, int rowrank, int colrank, const double* Ds, const double* Dd,
                            const double* VT, const double* U, const double* PVs, const double* PVd,
                            const double* UTI) {
   orc_cube* c = &t->cubes[id];
   int gp3 = gp3_of(t, c->level);
   size_t nd = (size_t)c->drows * c->dcols;
   if (Ds) { free(c->Ds); c->Ds = dup(Ds, nd); }
   if (Dd) { free(c->Dd); c->Dd = dup(Dd, nd); }
   free(c->VT); free(c->U); free(c->PVs); free(c->PVd); free(c->UTI);
   c->rowrank = rowrank;
   c->colrank = colrank;
   c->VT = dup(VT, (size_t)rowrank * c->npan);
   c->U = dup(U, (size_t)c->npt * colrank);
   c->PVs = dup(PVs, (size_t)gp3 * rowrank);
   c->PVd = dup(PVd, (size_t)gp3 * rowrank);
   c->UTI = dup(UTI, (size_t)gp3 * colrank);
}

void orc_tree_pinv(orc_tree* t, int level, double* out) {
   int g = gp3_of(t, level);
   memcpy(out, t->pinv[level], sizeof(double) * g * NQ);
}

void orc_tree_T(orc_tree* t, int level, int off, double* out) {
   const double* T = t->T[(size_t)level * 343 + off];
   if (T) memcpy(out, T, sizeof(double) * 2 * pad_of(t, level));
}

void orc_tree_dense(orc_tree* t, double* S, double* D) {
   for (int i = 0; i < t->npt; i++)
      for (int j = 0; j < t->npan; j++)
         orc_integrate(t->kernel, t->kappa, &t->pts[3 * i], &t->panels[j], &S[(size_t)i * t->npan + j],
                       &D[(size_t)i * t->npan + j]);
}
