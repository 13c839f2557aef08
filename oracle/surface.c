/*
 * Oracle (test infrastructure only): nested region operator, block
 * preconditioner, GMRES, right-hand side and reaction potentials, FP64.
 */
#include "oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static orc_region* region_new(void) { return (orc_region*)calloc(1, sizeof(orc_region)); }

/* SurfaceOperator_initStartIndices (SurfaceOperator.c:39-47): pre-order. */
static void init_start(orc_region* r, int* index) {
   for (int i = 0; i < r->nchild; i++) {
      orc_region* c = r->child[i];
      c->startphi = *index;
      c->startdphi = *index + c->mynpan;
      *index += 2 * c->mynpan;
      init_start(c, index);
   }
}

/* One dielectric surface with its charges inside: salt
 * (FFTSVDpbeAPI.c:1381-1554) or no-salt (FFTSVDpbeAPI.c:1557-1730)
 * formulation with no cavities and no Stern layer. */
orc_problem* orc_problem_single(const double* verts9, int npan, const double* qpos, const double* q, int ncharge,
                                double eps_in, double eps_out, double kappa, int maxpanels, int gridpoints,
                                double svdeps) {
   orc_problem* pr = (orc_problem*)calloc(1, sizeof(orc_problem));
   pr->ntotal = npan;
   pr->panels = (orc_panel*)malloc(sizeof(orc_panel) * (size_t)npan);
   orc_panels_from_verts(verts9, npan, pr->panels);
   double* cent = (double*)malloc(sizeof(double) * 3 * (size_t)npan);
   for (int i = 0; i < npan; i++) memcpy(&cent[3 * i], pr->panels[i].c, 3 * sizeof(double));

   orc_region* root = region_new();
   root->kernel = kappa > 0.0 ? ORC_YUKAWA : ORC_LAPLACE;
   root->eps = eps_out;
   root->kappa = kappa > 0.0 ? kappa : 0.0;
   root->mynpan = 0;
   root->tree = orc_tree_build(pr->panels, npan, cent, npan, root->kernel, root->kappa, maxpanels, gridpoints, svdeps);
   orc_tree_fill(root->tree);
   root->nchild = 1;
   root->child = (orc_region**)calloc(1, sizeof(orc_region*));
   root->resint = (double*)calloc(npan, sizeof(double));
   root->resext = (double*)calloc(npan, sizeof(double));

   orc_region* in = region_new();
   in->kernel = ORC_LAPLACE;
   in->eps = eps_in;
   in->kappa = 0.0;
   in->mynpan = npan;
   in->parent = root;
   in->tree = orc_tree_build(pr->panels, npan, cent, npan, ORC_LAPLACE, 0.0, maxpanels, gridpoints, svdeps);
   orc_tree_fill(in->tree);
   in->resint = (double*)calloc(npan, sizeof(double));
   in->resext = (double*)calloc(npan, sizeof(double));
   in->ncharge = ncharge;
   in->qpos = (double*)malloc(sizeof(double) * 3 * (ncharge > 0 ? ncharge : 1));
   in->q = (double*)malloc(sizeof(double) * (ncharge > 0 ? ncharge : 1));
   memcpy(in->qpos, qpos, sizeof(double) * 3 * ncharge);
   memcpy(in->q, q, sizeof(double) * ncharge);
   root->child[0] = in;
   pr->root = root;
   int index = 0;
   init_start(root, &index);
   free(cent);
   return pr;
}

/* Leaf region of the nested operator over panels[p0 .. p0+ntree) with the
 * first `mynpan` of them its own. */
static orc_region* make_region(orc_panel* panels, int p0, int ntree, int mynpan, int kernel, double eps,
                               double kappa, int maxpanels, int gridpoints, double svdeps) {
   orc_region* r = region_new();
   r->kernel = kernel;
   r->eps = eps;
   r->kappa = kappa;
   r->mynpan = mynpan;
   double* cent = (double*)malloc(sizeof(double) * 3 * (size_t)(ntree > 0 ? ntree : 1));
   for (int i = 0; i < ntree; i++) memcpy(&cent[3 * i], panels[p0 + i].c, 3 * sizeof(double));
   r->tree = orc_tree_build(panels + p0, ntree, cent, ntree, kernel, kappa, maxpanels, gridpoints, svdeps);
   orc_tree_fill(r->tree);
   r->resint = (double*)calloc(mynpan > 0 ? mynpan : 1, sizeof(double));
   r->resext = (double*)calloc(mynpan > 0 ? mynpan : 1, sizeof(double));
   free(cent);
   return r;
}

static void add_child(orc_region* parent, orc_region* child) {
   parent->child = (orc_region**)realloc(parent->child, sizeof(orc_region*) * (size_t)(parent->nchild + 1));
   parent->child[parent->nchild++] = child;
   child->parent = parent;
}

/* Stern-layer formulation with one salt (ion-exclusion) surface enclosing
 * one dielectric surface and its charges, no cavities
 * (FFTSVDpbeAPI.c:1064-1377): the exterior region (Yukawa, eps_out, kappa)
 * over the salt panels; inside the salt surface and outside the dielectric
 * (Laplace, eps_out) over the salt panels followed by the dielectric
 * panels; inside the dielectric (Laplace, eps_in) over the dielectric
 * panels, holding the charges.  Unknowns follow the pre-order start
 * indices: salt phi, salt dphi, dielectric phi, dielectric dphi. */
orc_problem* orc_problem_stern(const double* salt9, int nsalt, const double* diel9, int ndiel, const double* qpos,
                               const double* q, int ncharge, double eps_in, double eps_out, double kappa,
                               int maxpanels, int gridpoints, double svdeps) {
   orc_problem* pr = (orc_problem*)calloc(1, sizeof(orc_problem));
   const int n = nsalt + ndiel;
   pr->ntotal = n;
   pr->panels = (orc_panel*)malloc(sizeof(orc_panel) * (size_t)n);
   orc_panels_from_verts(salt9, nsalt, pr->panels);
   orc_panels_from_verts(diel9, ndiel, pr->panels + nsalt);
   orc_region* root = make_region(pr->panels, 0, nsalt, 0, ORC_YUKAWA, eps_out, kappa, maxpanels, gridpoints,
                                  svdeps);
   orc_region* salt = make_region(pr->panels, 0, n, nsalt, ORC_LAPLACE, eps_out, 0.0, maxpanels, gridpoints,
                                  svdeps);
   orc_region* in = make_region(pr->panels, nsalt, ndiel, ndiel, ORC_LAPLACE, eps_in, 0.0, maxpanels, gridpoints,
                                svdeps);
   in->ncharge = ncharge;
   in->qpos = (double*)malloc(sizeof(double) * 3 * (ncharge > 0 ? ncharge : 1));
   in->q = (double*)malloc(sizeof(double) * (ncharge > 0 ? ncharge : 1));
   memcpy(in->qpos, qpos, sizeof(double) * 3 * ncharge);
   memcpy(in->q, q, sizeof(double) * ncharge);
   add_child(root, salt);
   add_child(salt, in);
   pr->root = root;
   int index = 0;
   init_start(root, &index);
   return pr;
}

/* Region whose tree runs over the panels ids[0 .. nids) of `all` (own
 * surface first, then the surfaces of its children), gathered into a
 * panel array the region owns (the concatenated Panel arrays
 * generateSurfaceOperator passes to Tree_allocate). */
static orc_region* gather_region(const orc_panel* all, const int* ids, int nids, int mynpan, int kernel, double eps,
                                 double kappa, int maxpanels, int gridpoints, double svdeps) {
   orc_panel* store = (orc_panel*)malloc(sizeof(orc_panel) * (size_t)(nids > 0 ? nids : 1));
   for (int i = 0; i < nids; i++) store[i] = all[ids[i]];
   orc_region* r = make_region(store, 0, nids, mynpan, kernel, eps, kappa, maxpanels, gridpoints, svdeps);
   r->pan_store = store;
   return r;
}

static void set_charges(orc_region* r, const double* qpos, const double* q, int ncharge) {
   r->ncharge = ncharge;
   r->qpos = (double*)malloc(sizeof(double) * 3 * (ncharge > 0 ? ncharge : 1));
   r->q = (double*)malloc(sizeof(double) * (ncharge > 0 ? ncharge : 1));
   memcpy(r->qpos, qpos, sizeof(double) * 3 * ncharge);
   memcpy(r->q, q, sizeof(double) * ncharge);
}

/* Appends the global indices of surface k's panels to ids. */
static int push_surface(int* ids, int n, const int* start, int k) {
   for (int p = start[k]; p < start[k + 1]; p++) ids[n++] = p;
   return n;
}

/* generateSurfaceOperator (FFTSVDpbeAPI.c:1064-1730), every branch:
 * Stern (1064-1377): exterior Yukawa region over all salt panels; per salt
 *   surface a Laplace region (eps_out) over its panels and its
 *   dielectrics'; per dielectric a Laplace region (eps_in, the charges)
 *   over its panels and its cavities'; per dielectric cavity a Laplace
 *   region (eps_out) over its panels and its salt cavities'; per salt
 *   cavity a Yukawa region (eps_out) over its own panels.
 * Salt without Stern (1381-1554) and no salt (1557-1730): the exterior
 *   region (Yukawa with salt, else Laplace; eps_out) over all dielectric
 *   panels; per dielectric a Laplace region (eps_in, the charges) over its
 *   panels and its cavities'; per cavity a region of the exterior kernel
 *   over its own panels.
 * Children are taken in surface order, so the pre-order start indices
 * (SurfaceOperator_initStartIndices) follow the reference's. */
orc_problem* orc_problem_layout(const double* verts9, const int* counts, int nsalt, int ndiel, int ncav, int nscav,
                                const int* diel_parent, const int* cav_parent, const int* scav_parent,
                                const double* qpos, const double* q, int ncharge, double eps_in, double eps_out,
                                double kappa, int stern, int maxpanels, int gridpoints, double svdeps) {
   const int nsurf = nsalt + ndiel + ncav + nscav;
   int* start = (int*)calloc((size_t)nsurf + 1, sizeof(int));
   for (int k = 0; k < nsurf; k++) start[k + 1] = start[k] + counts[k];
   const int nall = start[nsurf];
   orc_problem* pr = (orc_problem*)calloc(1, sizeof(orc_problem));
   pr->panels = (orc_panel*)malloc(sizeof(orc_panel) * (size_t)(nall > 0 ? nall : 1));
   orc_panels_from_verts(verts9, nall, pr->panels);
   int* ids = (int*)malloc(sizeof(int) * (size_t)(nall > 0 ? nall : 1));
   const int S0 = 0, D0 = nsalt, C0 = nsalt + ndiel, L0 = nsalt + ndiel + ncav;   /* surface index bases */
   const int yuk = kappa > 0.0;
   orc_region* root;
   if (stern) {
      int n = 0;
      for (int i = 0; i < nsalt; i++) n = push_surface(ids, n, start, S0 + i);
      root = gather_region(pr->panels, ids, n, 0, ORC_YUKAWA, eps_out, kappa, maxpanels, gridpoints, svdeps);
      for (int i = 0; i < nsalt; i++) {
         n = push_surface(ids, 0, start, S0 + i);
         for (int j = 0; j < ndiel; j++)
            if (diel_parent[j] == i) n = push_surface(ids, n, start, D0 + j);
         orc_region* rs = gather_region(pr->panels, ids, n, counts[S0 + i], ORC_LAPLACE, eps_out, 0.0, maxpanels,
                                        gridpoints, svdeps);
         add_child(root, rs);
         for (int j = 0; j < ndiel; j++) {
            if (diel_parent[j] != i) continue;
            n = push_surface(ids, 0, start, D0 + j);
            for (int k = 0; k < ncav; k++)
               if (cav_parent[k] == j) n = push_surface(ids, n, start, C0 + k);
            orc_region* rd = gather_region(pr->panels, ids, n, counts[D0 + j], ORC_LAPLACE, eps_in, 0.0, maxpanels,
                                           gridpoints, svdeps);
            set_charges(rd, qpos, q, ncharge);
            add_child(rs, rd);
            for (int k = 0; k < ncav; k++) {
               if (cav_parent[k] != j) continue;
               n = push_surface(ids, 0, start, C0 + k);
               for (int l = 0; l < nscav; l++)
                  if (scav_parent[l] == k) n = push_surface(ids, n, start, L0 + l);
               orc_region* rc = gather_region(pr->panels, ids, n, counts[C0 + k], ORC_LAPLACE, eps_out, 0.0,
                                              maxpanels, gridpoints, svdeps);
               add_child(rd, rc);
               for (int l = 0; l < nscav; l++) {
                  if (scav_parent[l] != k) continue;
                  n = push_surface(ids, 0, start, L0 + l);
                  add_child(rc, gather_region(pr->panels, ids, n, counts[L0 + l], ORC_YUKAWA, eps_out, kappa,
                                              maxpanels, gridpoints, svdeps));
               }
            }
         }
      }
   } else {
      const int outer = yuk ? ORC_YUKAWA : ORC_LAPLACE;
      const double ok = yuk ? kappa : 0.0;
      int n = 0;
      for (int j = 0; j < ndiel; j++) n = push_surface(ids, n, start, D0 + j);
      root = gather_region(pr->panels, ids, n, 0, outer, eps_out, ok, maxpanels, gridpoints, svdeps);
      for (int j = 0; j < ndiel; j++) {
         n = push_surface(ids, 0, start, D0 + j);
         for (int k = 0; k < ncav; k++)
            if (cav_parent[k] == j) n = push_surface(ids, n, start, C0 + k);
         orc_region* rd = gather_region(pr->panels, ids, n, counts[D0 + j], ORC_LAPLACE, eps_in, 0.0, maxpanels,
                                        gridpoints, svdeps);
         set_charges(rd, qpos, q, ncharge);
         add_child(root, rd);
         for (int k = 0; k < ncav; k++) {
            if (cav_parent[k] != j) continue;
            n = push_surface(ids, 0, start, C0 + k);
            add_child(rd, gather_region(pr->panels, ids, n, counts[C0 + k], outer, eps_out, ok, maxpanels,
                                        gridpoints, svdeps));
         }
      }
   }
   pr->root = root;
   int index = 0;
   init_start(root, &index);
   pr->ntotal = index / 2;
   free(ids);
   free(start);
   return pr;
}

static void region_free(orc_region* r) {
   if (!r) return;
   for (int i = 0; i < r->nchild; i++) region_free(r->child[i]);
   free(r->child);
   orc_tree_free(r->tree);
   orc_tree_free(r->m3);
   free(r->pan_store);
   free(r->qpos); free(r->q); free(r->resint); free(r->resext);
   free(r);
}

static void region_make_dense(orc_region* r) {
   orc_tree* t = r->tree;
   if (!t->dS) {
      t->dS = (double*)malloc(sizeof(double) * (size_t)t->npt * t->npan + 1);
      t->dD = (double*)malloc(sizeof(double) * (size_t)t->npt * t->npan + 1);
      orc_tree_dense(t, t->dS, t->dD);
   }
   for (int c = 0; c < r->nchild; c++) region_make_dense(r->child[c]);
}

void orc_problem_make_dense(orc_problem* pr) { region_make_dense(pr->root); }

void orc_problem_free(orc_problem* pr) {
   region_free(pr->root);
   free(pr->panels);
   free(pr->precond);
   free(pr->precond_phi);
   free(pr->precond_dphi);
   free(pr);
}

/* Region number `which` in pre-order (0 the root, then each child before
 * its siblings' subtrees: the order generateSurfaceOperator creates them),
 * or NULL. */
static orc_region* region_preorder(orc_region* r, int* which) {
   if ((*which)-- == 0) return r;
   for (int c = 0; c < r->nchild; c++) {
      orc_region* f = region_preorder(r->child[c], which);
      if (f) return f;
   }
   return NULL;
}

int orc_problem_ntotal(orc_problem* pr) { return pr->ntotal; }

orc_tree* orc_problem_tree(orc_problem* pr, int which) {
   orc_region* r = region_preorder(pr->root, &which);
   return r ? r->tree : NULL;
}

/* SurfaceOperator_multiply (SurfaceOperator.c:114-158) with
 * joinpotentials / splitpotentials (SurfaceOperator.c:51-87). */
static void region_multiply(orc_region* r, const double* src) {
   int n = r->tree->npan;
   double* phi = (double*)calloc(n, sizeof(double));
   double* dphi = (double*)calloc(n, sizeof(double));
   double* dest = (double*)calloc(n, sizeof(double));
   int cur = 0;
   for (int i = 0; i < r->mynpan; i++) {
      phi[cur + i] = src[r->startphi + i];
      dphi[cur + i] = -1.0 * src[r->startdphi + i];
   }
   cur += r->mynpan;
   for (int c = 0; c < r->nchild; c++) {
      orc_region* ch = r->child[c];
      for (int i = 0; i < ch->mynpan; i++) {
         phi[cur + i] = -1.0 * src[ch->startphi + i];
         dphi[cur + i] = (ch->eps / r->eps) * src[ch->startdphi + i];
      }
      cur += ch->mynpan;
   }
   orc_tree_apply_both(r->tree, dphi, phi, dest);
   for (int i = r->mynpan; i < n; i++) dest[i] += -4.0 * M_PI * phi[i];
   memcpy(r->resint, dest, sizeof(double) * r->mynpan);
   cur = r->mynpan;
   for (int c = 0; c < r->nchild; c++) {
      memcpy(r->child[c]->resext, dest + cur, sizeof(double) * r->child[c]->mynpan);
      cur += r->child[c]->mynpan;
   }
   free(phi); free(dphi); free(dest);
   for (int c = 0; c < r->nchild; c++) region_multiply(r->child[c], src);
}

static void region_join(orc_region* r, double* v) {
   memcpy(v + r->startphi, r->resint, sizeof(double) * r->mynpan);
   memcpy(v + r->startdphi, r->resext, sizeof(double) * r->mynpan);
   for (int c = 0; c < r->nchild; c++) region_join(r->child[c], v);
}

static void region_clear(orc_region* r) {
   memset(r->resint, 0, sizeof(double) * r->mynpan);
   memset(r->resext, 0, sizeof(double) * r->mynpan);
   for (int c = 0; c < r->nchild; c++) region_clear(r->child[c]);
}

/* SurfaceOperator_topmultiply (SurfaceOperator.c:90-107). */
void orc_so_matvec(orc_problem* pr, const double* x, double* y) {
   region_clear(pr->root);
   region_multiply(pr->root, x);
   region_join(pr->root, y);
}

/* SurfaceOperator_makeRHS (SurfaceOperator.c:291-321); rhs is accumulated. */
static void region_rhs(orc_region* r, double* rhs) {
   for (int i = 0; i < r->ncharge; i++) {
      double qi = r->q[i];
      if (qi == 0.0) continue;
      const double* x = &r->qpos[3 * i];
      for (int j = 0; j < r->mynpan; j++) {
         const double* c = r->tree->panels[j].c;
         double d = sqrt((x[0] - c[0]) * (x[0] - c[0]) + (x[1] - c[1]) * (x[1] - c[1]) + (x[2] - c[2]) * (x[2] - c[2]));
         rhs[r->startphi + j] += qi / d;
      }
      for (int k = 0; k < r->nchild; k++) {
         orc_region* ch = r->child[k];
         for (int j = 0; j < ch->mynpan; j++) {
            const double* c = ch->tree->panels[j].c;
            double d = sqrt((x[0] - c[0]) * (x[0] - c[0]) + (x[1] - c[1]) * (x[1] - c[1]) + (x[2] - c[2]) * (x[2] - c[2]));
            rhs[ch->startdphi + j] += qi / d;
         }
      }
   }
   for (int k = 0; k < r->nchild; k++) region_rhs(r->child[k], rhs);
}

void orc_rhs(orc_problem* pr, double* rhs) {
   memset(rhs, 0, sizeof(double) * 2 * pr->ntotal);
   region_rhs(pr->root, rhs);
}

static void leaf_diag(orc_tree* t, double* ds, double* dd) {
   for (int i = 0; i < t->ncubes; i++) {
      orc_cube* c = &t->cubes[i];
      if (!c->leaf) continue;
      for (int k = 0; k < c->npan; k++) {
         ds[c->pan[k]] = c->Ds[(size_t)k * c->dcols + k];
         dd[c->pan[k]] = c->Dd[(size_t)k * c->dcols + k];
      }
   }
}

/* Per-panel 2x2 blocks [[Dd, -Ds], [Dd, eps_r*Ds]] from the region's own
 * tree diagonal (SurfaceOperator_generatePreconditioner,
 * SurfaceOperator.c:170-197; Cube_extractdiagonal, Cube.c:2349-2390).  The
 * reference factors the assembled 2N x 2N matrix with SuperLU; the blocks
 * are disjoint so the factorisation is the per-block inverse stored here. */
static void region_precond(orc_problem* pr, orc_region* r) {
   if (r->mynpan > 0) {
      double rel = r->parent ? r->eps / r->parent->eps : 1.0;
      int n = r->tree->npan;
      double* ds = (double*)calloc(n, sizeof(double));
      double* dd = (double*)calloc(n, sizeof(double));
      leaf_diag(r->tree, ds, dd);
      for (int i = 0; i < r->mynpan; i++) {
         int b = pr->nblocks++;
         double a11 = dd[i], a12 = -ds[i], a21 = dd[i], a22 = rel * ds[i];
         double det = a11 * a22 - a12 * a21;
         pr->precond[4 * b + 0] = a22 / det;
         pr->precond[4 * b + 1] = -a12 / det;
         pr->precond[4 * b + 2] = -a21 / det;
         pr->precond[4 * b + 3] = a11 / det;
         pr->precond_phi[b] = r->startphi + i;
         pr->precond_dphi[b] = r->startdphi + i;
      }
      free(ds);
      free(dd);
   }
   for (int c = 0; c < r->nchild; c++) region_precond(pr, r->child[c]);
}

void orc_precond_build(orc_problem* pr) {
   free(pr->precond); free(pr->precond_phi); free(pr->precond_dphi);
   pr->precond = (double*)calloc(4 * (size_t)pr->ntotal, sizeof(double));
   pr->precond_phi = (int*)calloc(pr->ntotal, sizeof(int));
   pr->precond_dphi = (int*)calloc(pr->ntotal, sizeof(int));
   pr->nblocks = 0;
   region_precond(pr, pr->root);
}

void orc_precond_solve(orc_problem* pr, const double* b, double* x) {
   int n2 = 2 * pr->ntotal;
   double* t = (double*)malloc(sizeof(double) * n2);
   memcpy(t, b, sizeof(double) * n2);
   for (int k = 0; k < pr->nblocks; k++) {
      int i = pr->precond_phi[k], j = pr->precond_dphi[k];
      const double* M = &pr->precond[4 * k];
      x[i] = M[0] * t[i] + M[1] * t[j];
      x[j] = M[2] * t[i] + M[3] * t[j];
   }
   free(t);
}

static double vdot(const double* a, const double* b, int n) {
   double s = 0;
   for (int i = 0; i < n; i++) s += a[i] * b[i];
   return s;
}

static void givens(double a, double b, double* c, double* s) {
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

static void grot(double c, double s, double* a, double* b) {
   double a1 = *a, b1 = *b;
   *a = a1 * c - b1 * s;
   *b = c * b1 + s * a1;
}

/* Left-preconditioned full GMRES with MGS and Givens rotations
 * (GMRES_SurfaceOperator, GMRES.c:446-539; givens/givensrotate,
 * GMRES.c:3-22).  Returns the iteration count (number of Arnoldi steps);
 * resid_hist (maxit entries, optional) receives |g_{i+1}|/||P^-1 b||. */
int orc_gmres(orc_problem* pr, const double* rhs, double* sol, double tol, int maxit, double* hist) {
   int n = 2 * pr->ntotal;
   double* r = (double*)calloc(n, sizeof(double));
   orc_precond_solve(pr, rhs, r);
   double normr = sqrt(vdot(r, r, n));
   if (normr < tol) {
      memset(sol, 0, sizeof(double) * n);
      free(r);
      return 0;
   }
   double* c = (double*)calloc(maxit + 1, sizeof(double));
   double* s = (double*)calloc(maxit + 1, sizeof(double));
   double* g = (double*)calloc(maxit + 1, sizeof(double));
   double* y = (double*)calloc(maxit + 1, sizeof(double));
   double* H = (double*)calloc((size_t)(maxit + 1) * (maxit + 1), sizeof(double));
   double** bv = (double**)calloc(maxit + 2, sizeof(double*));
   double* P = (double*)calloc(n, sizeof(double));
   g[0] = sqrt(vdot(r, r, n));
   bv[0] = (double*)malloc(sizeof(double) * n);
   for (int k = 0; k < n; k++) bv[0][k] = r[k] * (1.0 / g[0]);
   int i;
   int W = maxit + 1;
   for (i = 0; i < maxit; i++) {
      orc_so_matvec(pr, bv[i], P);
      orc_precond_solve(pr, P, P);
      for (int j = 0; j <= i; j++) {
         H[(size_t)i * W + j] = vdot(P, bv[j], n);
         double h = H[(size_t)i * W + j];
         for (int k = 0; k < n; k++) P[k] -= h * bv[j][k];
      }
      H[(size_t)i * W + i + 1] = sqrt(vdot(P, P, n));
      bv[i + 1] = (double*)malloc(sizeof(double) * n);
      double inv = 1.0 / H[(size_t)i * W + i + 1];
      for (int k = 0; k < n; k++) bv[i + 1][k] = P[k] * inv;
      for (int k = 0; k < i; k++) grot(c[k], s[k], &H[(size_t)i * W + k], &H[(size_t)i * W + k + 1]);
      givens(H[(size_t)i * W + i], H[(size_t)i * W + i + 1], &c[i], &s[i]);
      grot(c[i], s[i], &H[(size_t)i * W + i], &H[(size_t)i * W + i + 1]);
      g[i + 1] = 0.0;
      grot(c[i], s[i], &g[i], &g[i + 1]);
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
   free(bv); free(r); free(P); free(c); free(s); free(g); free(y); free(H);
   return i + 1;
}

/* Direct reaction potentials at the charges (SurfaceOperator_calculateReactionPotentials
 * direct branch, SurfaceOperator.c:364-396); phi is accumulated per charge. */
static void region_react(orc_region* r, const double* sol, double* phi) {
   if (r->ncharge > 0) {
      for (int j = 0; j < r->ncharge; j++) {
         const double* x = &r->qpos[3 * j];
         for (int i = 0; i < r->mynpan; i++) {
            double sl, dl;
            orc_laplace(x, &r->tree->panels[i], &sl, &dl);
            phi[j] += (sol[r->startdphi + i] * sl - sol[r->startphi + i] * dl);
            const double* c = r->tree->panels[i].c;
            if (x[0] == c[0] && x[1] == c[1] && x[2] == c[2]) phi[j] -= 4.0 * M_PI * sol[r->startphi + i];
         }
      }
      for (int k = 0; k < r->nchild; k++) {
         orc_region* ch = r->child[k];
         for (int i = 0; i < ch->mynpan; i++)
            for (int j = 0; j < r->ncharge; j++) {
               double sl, dl;
               orc_laplace(&r->qpos[3 * j], &ch->tree->panels[i], &sl, &dl);
               phi[j] += (-ch->eps / r->eps * sol[ch->startdphi + i] * sl + sol[ch->startphi + i] * dl);
            }
      }
   }
   for (int k = 0; k < r->nchild; k++) region_react(r->child[k], sol, phi);
}

static int region_max_charges(const orc_region* r) {
   int n = r->ncharge;
   for (int c = 0; c < r->nchild; c++) {
      int m = region_max_charges(r->child[c]);
      if (m > n) n = m;
   }
   return n;
}

void orc_reaction_potentials(orc_problem* pr, const double* sol, double* phi) {
   memset(phi, 0, sizeof(double) * region_max_charges(pr->root));
   region_react(pr->root, sol, phi);
}

/* M3 trees: Tree_allocate(region panels, charges, MAX_PANELS_PER_FINEST_CUBE,
 * POISSON_KERNEL, GRID_SIZE, SVD_ERROR, SINGLE_AND_DOUBLE_LAYER_INT, 0)
 * for every region with charges (FFTSVDpbeAPI.c:1503-1513; the Stern
 * branch, :1256-1266). */
static void region_m3(orc_region* r) {
   if (r->ncharge > 0 && !r->m3) {
      orc_tree* t = r->tree;
      r->m3 = orc_tree_build(t->panels, t->npan, r->qpos, r->ncharge, ORC_LAPLACE, 0.0, t->maxpanels, t->gridpoints,
                             t->svdeps);
      orc_tree_fill(r->m3);
   }
   for (int c = 0; c < r->nchild; c++) region_m3(r->child[c]);
}

void orc_m3_build(orc_problem* pr) { region_m3(pr->root); }

/* SurfaceOperator_calculateReactionPotentials, M3 branch (SurfaceOperator.c:347-363):
 * joinpotentials, Tree_multiplyboth(react, M3, dphi, phi), phi_j = -react_j. */
static void region_m3_react(orc_region* r, const double* src, double* out) {
   if (r->m3) {
      int n = r->tree->npan;
      double* phi = (double*)calloc(n > 0 ? n : 1, sizeof(double));
      double* dphi = (double*)calloc(n > 0 ? n : 1, sizeof(double));
      double* react = (double*)calloc(r->ncharge, sizeof(double));
      int cur = 0;
      for (int i = 0; i < r->mynpan; i++) {
         phi[cur + i] = src[r->startphi + i];
         dphi[cur + i] = -1.0 * src[r->startdphi + i];
      }
      cur += r->mynpan;
      for (int c = 0; c < r->nchild; c++) {
         orc_region* ch = r->child[c];
         for (int i = 0; i < ch->mynpan; i++) {
            phi[cur + i] = -1.0 * src[ch->startphi + i];
            dphi[cur + i] = (ch->eps / r->eps) * src[ch->startdphi + i];
         }
         cur += ch->mynpan;
      }
      orc_tree_apply_both(r->m3, dphi, phi, react);
      for (int j = 0; j < r->ncharge; j++) out[j] = -react[j];
      free(phi); free(dphi); free(react);
   }
   for (int c = 0; c < r->nchild; c++) region_m3_react(r->child[c], src, out);
}

void orc_m3_potentials(orc_problem* pr, const double* sol, double* phi) {
   memset(phi, 0, sizeof(double) * region_max_charges(pr->root));
   region_m3_react(pr->root, sol, phi);
}

orc_tree* orc_problem_m3_tree(orc_problem* pr, int which) {
   orc_region* r = region_preorder(pr->root, &which);
   return r ? r->m3 : NULL;
}
