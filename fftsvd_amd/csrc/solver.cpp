// Region operator, preconditioner, right-hand side, GMRES and reaction
// potentials on the GPU.
#include "solver.h"

#include "geometry.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <exception>
#include <functional>
#include <thread>
#include <type_traits>
#include <stdexcept>

namespace fsvd {

template <typename XT>
void launch_join(int n, const int* pi, const int* di, const double* sp, const double* sd, const XT* x, double* xs,
                 double* xd, hipStream_t s);
void launch_split(int n, const int* oi, const double* corr, const double* yt, const double* xd, double* y,
                  hipStream_t s);
void launch_precond(int nb, const int* ia, const int* ib, const double* M, const double* b, double* x, hipStream_t s);
void launch_dot_partial(long long n, int np, const double* a, const double* b, double* partial, hipStream_t s);
template <typename VT>
void launch_normalize(long long n, int np, const double* pin, double* hout, const double* w, VT* v, hipStream_t s);
template <typename VT>
void launch_combine(long long n, int m, const double* y, const VT* const* V, double* x, hipStream_t s);
template <typename VT>
void launch_multidot(long long n, int m, int nch, const VT* const* V, const double* w, double* partial, hipStream_t s);
void launch_multidot_finish(int m, int nch, const double* partial, double* h, double* hacc, int add, hipStream_t s);
int multiaxpy_blocks(long long n);
template <typename VT>
void launch_multiaxpy(long long n, int m, const double* h, const VT* const* V, double* w, double* pnorm,
                      hipStream_t s);
template <typename VT>
void launch_normalize_givens(long long n, int np, const double* pin, const double* w, VT* v, int i, int W,
                             double* hcol, double* H, double* cs, double* sn, double* g, double* hist, double normr,
                             hipStream_t s);
void launch_axpby(long long n, double a, const double* x, double b, const double* y, double* z, hipStream_t s);
template <typename XT>
void launch_join_b(int n, int ncol, const int* pi, const int* di, const double* sp, const double* sd,
                   const XT* const* x, double* xs, double* xd, long long ld, hipStream_t s);
void launch_split_b(int n, int ncol, const int* oi, const double* corr, const double* yt, const double* xd,
                    long long ld, double* const* y, hipStream_t s);
void launch_precond_b(int nb, int ncol, const int* ia, const int* ib, const double* M, const GmDev* d,
                      hipStream_t s);
template <typename VT>
void launch_cgs2_b(long long n, int m, int nch, int ncol, const GmDev* d, hipStream_t s);
template <typename VT>
void launch_normalize_givens_b(long long n, int np, int i, int W, int ncol, const GmDev* d, hipStream_t s);
void launch_rhs(int ne, const double* cent, const int* out, int nq, const double* qpos, const double* q, double* rhs,
                hipStream_t s);
void launch_potentials(int nq, int ne, const PanelGeom* panels, const int* pidx, const double* coef,
                       const int* self_ok, const double* qpos, double* phi, hipStream_t s);
void launch_leaf_diag(int nleaf, const LeafDesc* leaves, const double* D, double* ds, double* dd, hipStream_t s);
template <typename XT>
void launch_qual_gather(int n, const int* perm, const double* area_t, const XT* x, double* xa, hipStream_t s);
void launch_qual_finish(int n, const int* perm, const double* diag_t, const double* xa, const double* yt,
                        double inv4pe, double fourpe, double c0, double* out, hipStream_t s);
void launch_diag_div(int n, const double* p, const double* b, double* x, hipStream_t s);
template <typename XT>
void launch_gather_out(int n, const int* out_idx, const XT* y, double* yh, hipStream_t s);
void launch_scatter_adj(int n, const int* pi, const int* di, const double* sp, const double* sd, const double* corr,
                        const double* yh, const double* vs, const double* vd, double* x, hipStream_t s);
void launch_precond_t(int nb, const int* ia, const int* ib, const double* M, const double* b, double* x,
                      hipStream_t s);
void launch_collect_t(int ne, const PanelGeom* panels, const int* pidx, const int* self_ok, int nq, const double* qpos,
                      const double* q, double* ssl, double* sdl, hipStream_t s);
void launch_rhs_t(int nq, int ne, const double* cent, const int* out, const double* lam, const double* qpos, double* g,
                  hipStream_t s);
void launch_rhs_batch(int ne, const double* cent, const int* out, int nq, int ncol, const double* qpos,
                      const double* q, long long ldr, double* rhs, hipStream_t s);
void launch_potentials_batch(int nq, int ne, int ncol, const PanelGeom* panels, const int* pidx, const double* coef,
                             const int* self_ok, const double* qpos, double* phi, hipStream_t s);
void launch_qual_rhs_batch(int n, const PanelGeom* panels, int nq, int ncol, const double* qpos, const double* q,
                           double fourpe, long long ldr, double* rhs, hipStream_t s);
void launch_qual_rhs(int n, const PanelGeom* panels, int nq, const double* qpos, const double* q, double fourpe,
                     double* rhs, hipStream_t s);

namespace {
double now_ms() {
   return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
int partial_blocks(long long n) {
   long long b = (n + 2047) / 2048;
   if (b < 1) b = 1;
   if (b > 512) b = 512;
   return (int)b;
}
}  // namespace

template <typename VT>
static VT* basis_vec(Problem& pr, int j);
static void reserve_solver_memory(Problem& pr, bool want32);

Problem::~Problem() {
   regions.clear();
   basis.clear();
   basis32.clear();
   if (h_pinned) (void)hipHostFree(h_pinned);
   gm_cols.clear();
   if (h_pinned_b) (void)hipHostFree(h_pinned_b);
   for (hipEvent_t e : ring_ev)
      if (e) (void)hipEventDestroy(e);
   if (ring_h) (void)hipHostFree(ring_h);
   for (int k = 0; k < MAX_AUX; k++) {
      if (ev_join[k]) (void)hipEventDestroy(ev_join[k]);
      if (aux[k]) (void)hipStreamDestroy(aux[k]);
   }
   if (ev_fork) (void)hipEventDestroy(ev_fork);
   if (stream) (void)hipStreamDestroy(stream);
}

// Region layouts (reference generateSurfaceOperator).
// Salt without Stern and no salt (FFTSVDpbeAPI.c:1381-1730): an exterior
// region over every dielectric surface, one interior region per dielectric
// (own panels followed by its cavities' panels) holding all the charges,
// one region per cavity with the exterior kernel.
// Stern layer (FFTSVDpbeAPI.c:1064-1377): an exterior Yukawa region over
// every salt surface; per salt surface a Laplace region (eps_out) over its
// panels and those of the dielectrics it encloses; per dielectric a Laplace
// region (eps_in) over its panels and its cavities', holding the charges;
// per dielectric cavity a Laplace region (eps_out) over its panels and its
// salt cavities'; per salt cavity a Yukawa region.
// Runs fn(i) for i in [0, n) on up to 16 host threads (contiguous ranges).
template <typename F>
static void host_parallel_for(long long n, F&& fn) {
   const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
   if (n < 4096 || nth == 1) {
      for (long long i = 0; i < n; i++) fn(i);
      return;
   }
   std::vector<std::thread> th;
   for (int t = 0; t < nth; t++)
      th.emplace_back([&, t] {
         const long long a = n * t / nth, b = n * (t + 1) / nth;
         for (long long i = a; i < b; i++) fn(i);
      });
   for (auto& x : th) x.join();
}

static void setup_event(const char* what) {
   const char* e = std::getenv("FFTSVD_TIMING");
   if (e && e[0] == '1') std::fprintf(stderr, "[fftsvd setup] %-28s at %.2f\n", what, now_ms());
}

std::unique_ptr<Problem> build_problem(const SurfaceSpec& spec) {
   double t0 = now_ms();
   setup_event("problem start");
   auto prp = std::make_unique<Problem>();
   Problem& pr = *prp;
   FSVD_CHECK(hipStreamCreate(&pr.stream));
   hipStream_t s = pr.stream;
   pr.eps_in = spec.eps_in;
   pr.qpos = spec.qpos;
   pr.q = spec.q;
   const bool stern = !spec.salt_verts.empty();
   // global panel order: [salt surfaces,] dielectrics, dielectric cavities[,
   // salt cavities].  The host keeps the vertex records and the centroids
   // (the octree's input); the geometry records are built on the device.
   std::vector<double> verts;
   int npan_all = 0;
   auto add_surfaces = [&](const std::vector<std::vector<double>>& surf) {
      std::vector<int> start(surf.size() + 1, npan_all);
      for (size_t i = 0; i < surf.size(); i++) {
         const int n = (int)(surf[i].size() / 9);
         start[i + 1] = start[i] + n;
         const size_t base = (size_t)npan_all;
         verts.insert(verts.end(), surf[i].begin(), surf[i].begin() + 9 * (size_t)n);
         pr.cent.resize(3 * (base + n));
         host_parallel_for(n, [&](long long k) {
            const double* v = &surf[i][9 * (size_t)k];
            geom::panel_centroid_hd(v, v + 3, v + 6, &pr.cent[3 * (base + k)]);
         });
         npan_all += n;
      }
      return start;
   };
   const std::vector<int> sstart = add_surfaces(spec.salt_verts);
   const std::vector<int> dstart = add_surfaces(spec.dielectric_verts);
   const std::vector<int> cstart = add_surfaces(spec.cavity_verts);
   const std::vector<int> scstart = add_surfaces(spec.salt_cavity_verts);
   pr.ntotal = npan_all;
   pr.nvec = 2LL * pr.ntotal;
   DBuf<double> d_verts;
   d_verts.upload(verts, s);
   pr.d_allpanels.alloc(std::max(pr.ntotal, 1));
   launch_panel_records(d_verts.p, nullptr, pr.ntotal, pr.d_allpanels.p, s);
   FSVD_CHECK(hipGetLastError());
   setup_event("panel geometry");
   const int nsalt = (int)spec.salt_verts.size(), nd = (int)spec.dielectric_verts.size();
   const int nc = (int)spec.cavity_verts.size(), nsc = (int)spec.salt_cavity_verts.size();
   if (stern && ((int)spec.dielectric_parent.size() != nd || (int)spec.salt_cavity_parent.size() != nsc))
      throw std::runtime_error("Stern layout: one parent per dielectric and per salt cavity");
   if (!stern && nsc > 0) throw std::runtime_error("salt cavities need the Stern-layer formulation");
   const int outer_kernel = spec.kappa > 0.0 ? YUKAWA : LAPLACE;
   const double outer_kappa = spec.kappa > 0.0 ? spec.kappa : 0.0;

   auto new_region = [&](int parent, int kernel, double eps, double kappa, int p0, int p1) {
      const int ri = (int)pr.regions.size();
      pr.regions.emplace_back();
      Region& R = pr.regions[ri];
      R.kernel = kernel;
      R.eps = eps;
      R.kappa = kappa;
      R.parent = parent;
      R.mynpan = p1 - p0;
      for (int k = p0; k < p1; k++) R.panel_ids.push_back(k);
      if (parent >= 0) pr.regions[parent].child.push_back(ri);
      return ri;
   };
   // appends a region's own panels to every ancestor's tree panels up to
   // (not including) `stop`
   auto add_to_parent = [&](int ri) {
      const int parent = pr.regions[ri].parent;
      if (parent < 0) return;
      for (int k = 0; k < pr.regions[ri].mynpan; k++)
         pr.regions[parent].panel_ids.push_back(pr.regions[ri].panel_ids[k]);
   };
   // dielectric j under parent region `par`, with its cavities (and, in
   // the Stern layout, their salt cavities)
   auto add_dielectric = [&](int par, int j) {
      const int rd = new_region(par, LAPLACE, spec.eps_in, 0.0, dstart[j], dstart[j + 1]);
      pr.regions[rd].has_charges = true;
      if (par >= 0 && stern) add_to_parent(rd);
      for (int c = 0; c < nc; c++) {
         if (spec.cavity_parent[c] != j) continue;
         const int rc = stern ? new_region(rd, LAPLACE, spec.eps_out, 0.0, cstart[c], cstart[c + 1])
                              : new_region(rd, outer_kernel, spec.eps_out, outer_kappa, cstart[c], cstart[c + 1]);
         add_to_parent(rc);
         if (!stern) continue;
         for (int l = 0; l < nsc; l++) {
            if (spec.salt_cavity_parent[l] != c) continue;
            const int rl = new_region(rc, YUKAWA, spec.eps_out, outer_kappa, scstart[l], scstart[l + 1]);
            add_to_parent(rl);
         }
      }
   };
   // regions in pre-order
   if (stern) {
      if (!(spec.kappa > 0.0)) throw std::runtime_error("the Stern-layer formulation needs a salt (kappa > 0)");
      const int root = new_region(-1, YUKAWA, spec.eps_out, outer_kappa, 0, 0);
      for (int k = sstart[0]; k < sstart[nsalt]; k++) pr.regions[root].panel_ids.push_back(k);
      for (int i = 0; i < nsalt; i++) {
         const int ri = new_region(root, LAPLACE, spec.eps_out, 0.0, sstart[i], sstart[i + 1]);
         for (int j = 0; j < nd; j++)
            if (spec.dielectric_parent[j] == i) add_dielectric(ri, j);
      }
   } else {
      const int root = new_region(-1, outer_kernel, spec.eps_out, outer_kappa, 0, 0);
      for (int k = dstart[0]; k < dstart[nd]; k++) pr.regions[root].panel_ids.push_back(k);
      for (int j = 0; j < nd; j++) add_dielectric(root, j);
   }
   // start indices (SurfaceOperator_initStartIndices, SurfaceOperator.c:39-47)
   {
      int index = 0;
      std::vector<int> stack;
      std::function<void(int)> walk = [&](int r) {
         for (int c : pr.regions[r].child) {
            Region& C = pr.regions[c];
            C.startphi = index;
            C.startdphi = index + C.mynpan;
            index += 2 * C.mynpan;
            walk(c);
         }
      };
      walk(0);
      if (index != 2 * pr.ntotal) throw std::runtime_error("SurfaceOperator indexing is WRONG!");
   }
   // join / split maps in tree order (the octree's panel order, known
   // before the operators are built)
   auto join_split_maps = [&](Region& R, const std::vector<int>& pan_perm) {
      int n = (int)R.panel_ids.size();
      std::vector<int> pi(n), di(n), oi(n);
      std::vector<double> sp(n), sd(n), co(n);
      // tree-local index -> (child region, offset)
      std::vector<int> owner(n, -1), local(n, 0);
      for (int l = 0; l < R.mynpan; l++) local[l] = l;
      int cur = R.mynpan;
      for (int c : R.child) {
         for (int j = 0; j < pr.regions[c].mynpan; j++) {
            owner[cur + j] = c;
            local[cur + j] = j;
         }
         cur += pr.regions[c].mynpan;
      }
      for (int k = 0; k < n; k++) {
         int l = pan_perm[k];
         if (owner[l] < 0) {
            pi[k] = R.startphi + l;
            di[k] = R.startdphi + l;
            sp[k] = 1.0;
            sd[k] = -1.0;
            oi[k] = R.startphi + l;
            co[k] = 0.0;
         } else {
            const Region& C = pr.regions[owner[l]];
            int j = local[l];
            pi[k] = C.startphi + j;
            di[k] = C.startdphi + j;
            sp[k] = -1.0;
            sd[k] = C.eps / R.eps;
            oi[k] = C.startdphi + j;
            co[k] = -4.0 * M_PI;
         }
      }
      R.phi_idx.upload(pi, s);
      R.dphi_idx.upload(di, s);
      R.out_idx.upload(oi, s);
      R.s_phi.upload(sp, s);
      R.s_dphi.upload(sd, s);
      R.corr.upload(co, s);
      R.xs.alloc(n + 1);
      R.xd.alloc(n + 1);
      R.yt.alloc(n + 1);
   };
   // operators (share identical trees: same kernel, same panels).  The
   // octree depends only on the panels, so it is built once per distinct
   // panel set; the distinct operators are then built concurrently, one host
   // thread and one stream each.
   {
      std::vector<int> todo;
      for (size_t r = 0; r < pr.regions.size(); r++) {
         Region& R = pr.regions[r];
         for (size_t o = 0; o < r; o++) {
            const Region& O = pr.regions[o];
            if (O.shares_tree_with < 0 && O.kernel == R.kernel && O.kappa == R.kappa && O.panel_ids == R.panel_ids) {
               R.shares_tree_with = (int)o;
               break;
            }
         }
         if (R.shares_tree_with < 0) todo.push_back((int)r);
      }
      const size_t nt = todo.size();
      // operators over the same panel set share its device geometry and its
      // octree (tree_src: the first such operator)
      std::vector<std::vector<double>> cent(nt);
      std::vector<int> tree_src(nt, -1);
      std::vector<Tree> trees(nt);
      std::vector<DBuf<int>> ids(nt);
      for (size_t k = 0; k < nt; k++) {
         Region& R = pr.regions[todo[k]];
         for (size_t o = 0; o < k; o++)
            if (tree_src[o] < 0 && pr.regions[todo[o]].panel_ids == R.panel_ids) {
               tree_src[k] = (int)o;
               break;
            }
         if (tree_src[k] >= 0) continue;
         const size_t n = R.panel_ids.size();
         // the records in the region's panel order, built on the device while
         // the host builds the octree
         ids[k].upload(R.panel_ids, s);
         R.d_panels.alloc(std::max<size_t>(n, 1));
         launch_panel_records(d_verts.p, ids[k].p, (int)n, R.d_panels.p, s);
         FSVD_CHECK(hipGetLastError());
         cent[k].resize(3 * n);
         host_parallel_for((long long)n, [&](long long i) {
            for (int d = 0; d < 3; d++) cent[k][3 * i + d] = pr.cent[3 * (size_t)R.panel_ids[i] + d];
         });
         build_tree(trees[k], cent[k].data(), (int)n, cent[k].data(), (int)n, spec.maxpanels, spec.gridpoints);
         setup_event("panels and tree");
      }
      FSVD_CHECK(hipStreamSynchronize(s));
      setup_event("trees ready");
      std::vector<hipStream_t> streams(nt);
      for (size_t k = 0; k < nt; k++) FSVD_CHECK(hipStreamCreate(&streams[k]));
      std::vector<std::exception_ptr> err(nt);
      std::vector<std::thread> th;
      for (size_t k = 0; k < nt; k++)
         th.emplace_back([&, k]() {
            try {
               Region& R = pr.regions[todo[k]];
               const size_t src = tree_src[k] < 0 ? k : (size_t)tree_src[k];
               const Tree& pre = trees[src];
               const int n = (int)(cent[src].size() / 3);
               R.op = build_tree_op(cent[src].data(), n, pr.regions[todo[src]].d_panels.p, cent[src].data(), n,
                                    R.kernel, R.kappa, spec.maxpanels, spec.gridpoints, spec.svdeps, spec.want32,
                                    streams[k], &pre);
               FSVD_CHECK(hipStreamSynchronize(streams[k]));
            } catch (...) {
               err[k] = std::current_exception();
            }
         });
      // meanwhile the regions' join / split maps, from the prebuilt trees
      std::exception_ptr map_err;
      try {
         for (size_t r = 0; r < pr.regions.size(); r++) {
            Region& R = pr.regions[r];
            const int owner = R.shares_tree_with >= 0 ? R.shares_tree_with : (int)r;
            const size_t k = (size_t)(std::find(todo.begin(), todo.end(), owner) - todo.begin());
            const size_t src = tree_src[k] < 0 ? k : (size_t)tree_src[k];
            join_split_maps(R, trees[src].pan_perm);
         }
      } catch (...) {
         map_err = std::current_exception();
      }
      for (auto& x : th) x.join();
      for (size_t k = 0; k < nt; k++) (void)hipStreamDestroy(streams[k]);
      if (map_err) std::rethrow_exception(map_err);
      for (size_t k = 0; k < nt; k++)
         if (err[k]) std::rethrow_exception(err[k]);
      for (int r : todo)
         if (pr.regions[r].op->tree.pt_perm != pr.regions[r].op->tree.pan_perm)
            throw std::runtime_error("point/panel partitions differ");
   }
   // matvec streams: regions that share an operator share its work buffers,
   // so they run on the same stream; distinct operators round-robin
   {
      std::vector<int> op_lane(pr.regions.size(), -1);
      int next = 0;
      for (size_t r = 0; r < pr.regions.size(); r++) {
         Region& R = pr.regions[r];
         const int owner = R.shares_tree_with >= 0 ? R.shares_tree_with : (int)r;
         if (op_lane[owner] < 0) op_lane[owner] = (next++) % (Problem::MAX_AUX + 1);
         R.lane = op_lane[owner];
         pr.naux = std::max(pr.naux, R.lane);
      }
      FSVD_CHECK(hipEventCreateWithFlags(&pr.ev_fork, hipEventDisableTiming));
      for (int k = 0; k < pr.naux; k++) {
         FSVD_CHECK(hipStreamCreate(&pr.aux[k]));
         FSVD_CHECK(hipEventCreateWithFlags(&pr.ev_join[k], hipEventDisableTiming));
      }
   }
   // M3 trees (accelerateM3; FFTSVDpbeAPI.c:1256-1266,1503-1513): Laplace,
   // both layers, from the region's tree panels to the charges; the region
   // join (joinpotentials, SurfaceOperator.c:51-66) in the M3 tree's order
   if (spec.m3 && !pr.q.empty()) {
      for (Region& R : pr.regions) {
         if (!R.has_charges) continue;
         const int n = (int)R.panel_ids.size();
         std::vector<double> mc(3 * (size_t)n);
         for (int i = 0; i < n; i++)
            for (int d = 0; d < 3; d++) mc[3 * (size_t)i + d] = pr.cent[3 * (size_t)R.panel_ids[i] + d];
         DBuf<int> mids;
         mids.upload(R.panel_ids, s);
         R.m3_panels.alloc(std::max(n, 1));
         launch_panel_records(d_verts.p, mids.p, n, R.m3_panels.p, s);
         FSVD_CHECK(hipGetLastError());
         R.m3 = build_tree_op(mc.data(), n, R.m3_panels.p, pr.qpos.data(), (int)pr.q.size(), LAPLACE, 0.0,
                              spec.maxpanels, spec.gridpoints, spec.svdeps, false, s);
         std::vector<int> lpi(n), ldi(n);
         std::vector<double> lsp(n), lsd(n);
         for (int l = 0; l < R.mynpan; l++) {
            lpi[l] = R.startphi + l;
            ldi[l] = R.startdphi + l;
            lsp[l] = 1.0;
            lsd[l] = -1.0;
         }
         int cur = R.mynpan;
         for (int c : R.child) {
            const Region& C = pr.regions[c];
            for (int j = 0; j < C.mynpan; j++) {
               lpi[cur + j] = C.startphi + j;
               ldi[cur + j] = C.startdphi + j;
               lsp[cur + j] = -1.0;
               lsd[cur + j] = C.eps / R.eps;
            }
            cur += C.mynpan;
         }
         std::vector<int> pi(n), di(n);
         std::vector<double> sp(n), sd(n);
         for (int k = 0; k < n; k++) {
            const int l = R.m3->tree.pan_perm[k];
            pi[k] = lpi[l];
            di[k] = ldi[l];
            sp[k] = lsp[l];
            sd[k] = lsd[l];
         }
         R.m3_phi_idx.upload(pi, s);
         R.m3_dphi_idx.upload(di, s);
         R.m3_s_phi.upload(sp, s);
         R.m3_s_dphi.upload(sd, s);
         R.m3_xs.alloc(n + 1);
         R.m3_xd.alloc(n + 1);
         R.m3_y.alloc(pr.q.size() + 1);
      }
   }
   // preconditioner: 2x2 blocks per owned panel from the region's own near-field diagonal
   // (SurfaceOperator_generatePreconditioner, SurfaceOperator.c:170-197)
   {
      std::vector<int> pa, pb;
      std::vector<double> pm;
      for (size_t r = 0; r < pr.regions.size(); r++) {
         Region& R = pr.regions[r];
         if (R.mynpan == 0) continue;
         TreeOp& op = R.shares_tree_with >= 0 ? *pr.regions[R.shares_tree_with].op : *R.op;
         int n = (int)R.panel_ids.size();
         DBuf<double> dds, ddd;
         dds.alloc(n);
         ddd.alloc(n);
         launch_leaf_diag(op.nleaf, op.leaves.p, op.D64.p, dds.p, ddd.p, s);
         std::vector<double> ds(n), dd(n);
         FSVD_CHECK(hipMemcpyAsync(ds.data(), dds.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
         FSVD_CHECK(hipMemcpyAsync(dd.data(), ddd.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
         FSVD_CHECK(hipStreamSynchronize(s));
         std::vector<double> dsl(n), ddl(n);
         for (int k = 0; k < n; k++) {
            dsl[op.tree.pan_perm[k]] = ds[k];
            ddl[op.tree.pan_perm[k]] = dd[k];
         }
         double rel = R.parent >= 0 ? R.eps / pr.regions[R.parent].eps : 1.0;
         for (int i = 0; i < R.mynpan; i++) {
            double a11 = ddl[i], a12 = -dsl[i], a21 = ddl[i], a22 = rel * dsl[i];
            double det = a11 * a22 - a12 * a21;
            if (!(std::fabs(det) > 0.0) || !std::isfinite(det))
               throw std::runtime_error("singular preconditioner block");
            pa.push_back(R.startphi + i);
            pb.push_back(R.startdphi + i);
            pm.push_back(a22 / det);
            pm.push_back(-a12 / det);
            pm.push_back(-a21 / det);
            pm.push_back(a11 / det);
         }
      }
      pr.nblocks = (int)pa.size();
      pr.pc_a.upload(pa, s);
      pr.pc_b.upload(pb, s);
      pr.pc_m.upload(pm, s);
   }
   pr.d_qpos.upload(pr.qpos, s);
   pr.d_q.upload(pr.q, s);
   long long n2 = pr.nvec;
   pr.w.alloc(n2);
   pr.r.alloc(n2);
   pr.tmp.alloc(n2);
   pr.partial.alloc(2 * 512);
   pr.hbuf.alloc(1024 + 8);
   FSVD_CHECK(hipHostMalloc((void**)&pr.h_pinned, sizeof(double) * 2048));
   reserve_solver_memory(pr, spec.want32);
   FSVD_CHECK(hipStreamSynchronize(s));
   pr.setup_ms = now_ms() - t0;
   setup_event("problem end");
   return prp;
}

// Qualocation operator (generateQualocationOperator(Minimal),
// FFTSVDpbeAPI.c:1741-1839; QualocationOperator.c): one Laplace tree with
// double-layer-only bases over the panels in caller order, the qualocation
// diagonal, and the diagonal preconditioner
// (Preconditioner_fill_diagonal_solv_ecf_qual, Preconditioner.c:49-60).
std::unique_ptr<Problem> build_qual_problem(const QualSpec& spec) {
   double t0 = now_ms();
   if (!(spec.eps_out != spec.eps_in)) throw std::runtime_error("qualocation needs eps_out != eps_in");
   auto prp = std::make_unique<Problem>();
   Problem& pr = *prp;
   FSVD_CHECK(hipStreamCreate(&pr.stream));
   hipStream_t s = pr.stream;
   pr.qual = true;
   pr.eps_in = spec.eps_in;
   pr.eps_out = spec.eps_out;
   pr.qpos = spec.qpos;
   pr.q = spec.q;
   const int n = (int)(spec.verts.size() / 9);
   if (n < 1) throw std::runtime_error("qualocation operator without panels");
   pr.panels.resize(n);
   host_parallel_for(n, [&](long long k) {
      const double* v = &spec.verts[9 * (size_t)k];
      panel_geometry(v, v + 3, v + 6, &pr.panels[k]);
   });
   pr.ntotal = n;
   pr.nvec = n;
   pr.regions.emplace_back();
   Region& R = pr.regions[0];
   R.kernel = LAPLACE;
   R.eps = spec.eps_in;
   R.mynpan = n;
   R.has_charges = !spec.q.empty();
   R.panel_ids.resize(n);
   for (int i = 0; i < n; i++) R.panel_ids[i] = i;
   pr.cent.resize(3 * (size_t)n);
   for (int i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) pr.cent[3 * (size_t)i + d] = pr.panels[i].c[d];
   R.d_panels.upload(pr.panels, s);
   R.op = build_tree_op(pr.cent.data(), n, R.d_panels.p, pr.cent.data(), n, LAPLACE, 0.0, spec.maxpanels,
                        spec.gridpoints, spec.svdeps, spec.want32, s, nullptr, 1);
   TreeOp& op = *R.op;
   if (op.tree.pt_perm != op.tree.pan_perm) throw std::runtime_error("point/panel partitions differ");
   // Tree_extractdiagonal(DOUBLE_LAYER_INT) (Tree.c:861-875): leaf D_double[i][i]
   DBuf<double> dds;
   dds.alloc(n);
   pr.qual_diag_t.alloc(n);
   launch_leaf_diag(op.nleaf, op.leaves.p, op.D64.p, dds.p, pr.qual_diag_t.p, s);
   std::vector<double> diag_t(n), area_t(n), pdiag(n);
   FSVD_CHECK(hipMemcpyAsync(diag_t.data(), pr.qual_diag_t.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
   FSVD_CHECK(hipStreamSynchronize(s));
   const double c0 = -spec.eps_out / ((spec.eps_out - spec.eps_in) * spec.eps_in);
   for (int k = 0; k < n; k++) {
      const int i = op.tree.pan_perm[k];
      area_t[k] = pr.panels[i].area;
      pdiag[i] = (c0 + (diag_t[k]) / (4.0 * M_PI * spec.eps_in)) * pr.panels[i].area;
   }
   for (int i = 0; i < n; i++)
      if (!(std::fabs(pdiag[i]) > 0.0) || !std::isfinite(pdiag[i]))
         throw std::runtime_error("singular qualocation preconditioner entry");
   pr.qual_perm.upload(op.tree.pan_perm, s);
   pr.qual_area_t.upload(area_t, s);
   pr.qual_pdiag.upload(pdiag, s);
   R.xs.alloc(n + 1);
   R.yt.alloc(n + 1);
   // M1M3 (generateQualocationOperatorA1A3, FFTSVDpbeAPI.c:1841-1851):
   // Laplace, both layers, panels -> charges
   if (spec.m3 && !spec.q.empty()) {
      const int nq = (int)spec.q.size();
      R.m3 = build_tree_op(pr.cent.data(), n, R.d_panels.p, spec.qpos.data(), nq, LAPLACE, 0.0, spec.maxpanels,
                           spec.gridpoints, spec.svdeps, false, s);
      R.m3_xs.alloc(std::max(n, nq) + 1);
      R.m3_xd.alloc(std::max(n, nq) + 1);
      R.m3_y.alloc(std::max(n, nq) + 1);
   }
   pr.d_qpos.upload(pr.qpos, s);
   pr.d_q.upload(pr.q, s);
   pr.w.alloc(n);
   pr.r.alloc(n);
   pr.tmp.alloc(n);
   pr.partial.alloc(2 * 512);
   pr.hbuf.alloc(1024 + 8);
   FSVD_CHECK(hipHostMalloc((void**)&pr.h_pinned, sizeof(double) * 2048));
   reserve_solver_memory(pr, spec.want32);
   FSVD_CHECK(hipStreamSynchronize(s));
   pr.setup_ms = now_ms() - t0;
   return prp;
}

// Qualocation matvec: y = A x with A applied through the transposed
// double-layer tree in tree order.
template <typename XT>
static void qual_matvec(Problem& pr, int prec, const XT* x, double* y) {
   hipStream_t s = pr.stream;
   Region& R = pr.regions[0];
   TreeOp& op = *R.op;
   const int n = pr.ntotal;
   if (prec == PREC_F32 && !op.have32) throw std::runtime_error("operator has no FP32 copy");
   launch_qual_gather<XT>(n, pr.qual_perm.p, pr.qual_area_t.p, x, R.xs.p, s);
   tree_apply_transpose(op, prec, 1, R.xs.p, R.yt.p, s);
   const double fourpe = 4.0 * M_PI * pr.eps_in;
   const double c0 = -pr.eps_out / ((pr.eps_out - pr.eps_in) * pr.eps_in);
   launch_qual_finish(n, pr.qual_perm.p, pr.qual_diag_t.p, R.xs.p, R.yt.p, 1.0 / fourpe, fourpe, c0, y, s);
}

// x = A^T y: per region (in order, on the library stream) the transposes
// of k_split, the tree apply (both layers) and k_join, scatter-added into x
// (SurfaceOperator_topmultiply transposed; Tree_multiply_transpose,
// Tree.c:769-832, per layer).
template <typename XT>
static void adjoint_matvec(Problem& pr, int prec, const XT* y, double* x) {
   if (pr.qual) throw std::runtime_error("the adjoint of the qualocation operator is not provided");
   hipStream_t s = pr.stream;
   FSVD_CHECK(hipMemsetAsync(x, 0, sizeof(double) * pr.nvec, s));
   for (Region& R : pr.regions) {
      if (R.panel_ids.empty()) continue;
      const int n = (int)R.panel_ids.size();
      TreeOp& op = R.shares_tree_with >= 0 ? *pr.regions[R.shares_tree_with].op : *R.op;
      if (prec == PREC_F32 && !op.have32) throw std::runtime_error("operator has no FP32 copy");
      launch_gather_out<XT>(n, R.out_idx.p, y, R.yt.p, s);
      tree_apply_transpose_both(op, prec, R.yt.p, R.xs.p, R.xd.p, s);
      launch_scatter_adj(n, R.phi_idx.p, R.dphi_idx.p, R.s_phi.p, R.s_dphi.p, R.corr.p, R.yt.p, R.xs.p, R.xd.p, x, s);
   }
}

// Every region writes a disjoint set of rows of y (the phi rows of its own
// panels and the dphi rows of its children's panels), so regions on
// different streams never touch the same output.
template <typename XT>
static void matvec_impl(Problem& pr, int prec, const XT* x, double* y) {
   if (pr.adjoint) return adjoint_matvec(pr, prec, x, y);
   if (pr.qual) return qual_matvec(pr, prec, x, y);
   hipStream_t s = pr.stream;
   if (pr.naux > 0) {
      FSVD_CHECK(hipEventRecord(pr.ev_fork, s));
      for (int k = 0; k < pr.naux; k++) FSVD_CHECK(hipStreamWaitEvent(pr.aux[k], pr.ev_fork, 0));
   }
   // Per lane, its regions' steps in order: join, the four apply parts
   // (projection+source, translation, target, leaf), split.  The steps are
   // enqueued round-robin over the lanes, so every lane has work queued
   // from the start; regions of one lane (which may share a tree operator
   // and its scratch) stay in order on that lane's stream.
   std::vector<std::vector<int>> lane_regions(pr.naux + 1);
   for (size_t r = 0; r < pr.regions.size(); r++)
      if (!pr.regions[r].panel_ids.empty()) lane_regions[pr.regions[r].lane].push_back((int)r);
   // steps: join, projection+source, translation, target, leaf with the
   // split fused in
   constexpr int NSTEP = 5;
   size_t maxsteps = 0;
   for (auto& lr : lane_regions) maxsteps = std::max(maxsteps, lr.size() * NSTEP);
   for (size_t k = 0; k < maxsteps; k++) {
      for (size_t l = 0; l < lane_regions.size(); l++) {
         if (k >= lane_regions[l].size() * NSTEP) continue;
         Region& R = pr.regions[lane_regions[l][k / NSTEP]];
         const int step = (int)(k % NSTEP), n = (int)R.panel_ids.size();
         hipStream_t rs = R.lane == 0 ? s : pr.aux[R.lane - 1];
         TreeOp& op = R.shares_tree_with >= 0 ? *pr.regions[R.shares_tree_with].op : *R.op;
         if (step == 0) {
            if (prec == PREC_F32 && !op.have32) throw std::runtime_error("operator has no FP32 copy");
            launch_join(n, R.phi_idx.p, R.dphi_idx.p, R.s_phi.p, R.s_dphi.p, x, R.xs.p, R.xd.p, rs);
         } else if (step < 4) {
            tree_apply_part(op, prec, step, R.xs.p, R.xd.p, R.yt.p, rs);
         } else {
            tree_apply_leaf_split(op, prec, R.xs.p, R.xd.p, R.out_idx.p, R.corr.p, y, rs);
         }
      }
   }
   for (int k = 0; k < pr.naux; k++) {
      FSVD_CHECK(hipEventRecord(pr.ev_join[k], pr.aux[k]));
      FSVD_CHECK(hipStreamWaitEvent(s, pr.ev_join[k], 0));
   }
}
void problem_matvec(Problem& pr, int prec, const double* x, double* y) { matvec_impl(pr, prec, x, y); }
void problem_matvec(Problem& pr, int prec, const float* x, double* y) { matvec_impl(pr, prec, x, y); }

void problem_precond(Problem& pr, const double* b, double* x) {
   if (pr.qual) return launch_diag_div(pr.ntotal, pr.qual_pdiag.p, b, x, pr.stream);
   if (pr.adjoint) return launch_precond_t(pr.nblocks, pr.pc_a.p, pr.pc_b.p, pr.pc_m.p, b, x, pr.stream);
   launch_precond(pr.nblocks, pr.pc_a.p, pr.pc_b.p, pr.pc_m.p, b, x, pr.stream);
}

void problem_rhs(Problem& pr, double* rhs) {
   hipStream_t s = pr.stream;
   if (pr.qual && pr.regions[0].m3) {
      // QualocationOperator_makeRHS with M1M3 (QualocationOperator.c:33-41):
      // RHS = A_d^T q * (-1 / (4 pi eps_in)) through the tree
      Region& R = pr.regions[0];
      TreeOp& m = *R.m3;
      const int nq = (int)pr.q.size(), n = pr.ntotal;
      std::vector<double> qt(nq), out(n), rh(n);
      for (int k = 0; k < nq; k++) qt[k] = pr.q[m.tree.pt_perm[k]];
      FSVD_CHECK(hipMemcpyAsync(R.m3_xs.p, qt.data(), sizeof(double) * nq, hipMemcpyHostToDevice, s));
      tree_apply_transpose(m, PREC_F64, 1, R.m3_xs.p, R.m3_y.p, s);
      FSVD_CHECK(hipMemcpyAsync(out.data(), R.m3_y.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
      FSVD_CHECK(hipStreamSynchronize(s));
      for (int k = 0; k < n; k++) rh[m.tree.pan_perm[k]] = out[k] * (-1.0 / (4.0 * M_PI * pr.eps_in));
      FSVD_CHECK(hipMemcpyAsync(rhs, rh.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
      FSVD_CHECK(hipStreamSynchronize(s));
      return;
   }
   if (pr.qual) {
      launch_qual_rhs(pr.ntotal, pr.regions[0].d_panels.p, (int)pr.q.size(), pr.d_qpos.p, pr.d_q.p,
                      4.0 * M_PI * pr.eps_in, rhs, s);
      FSVD_CHECK(hipStreamSynchronize(s));
      return;
   }
   long long n2 = pr.nvec;
   FSVD_CHECK(hipMemsetAsync(rhs, 0, sizeof(double) * n2, s));
   for (Region& R : pr.regions) {
      if (!R.has_charges) continue;
      std::vector<double> cent;
      std::vector<int> out;
      for (int j = 0; j < R.mynpan; j++) {
         for (int d = 0; d < 3; d++) cent.push_back(pr.cent[3 * (size_t)R.panel_ids[j] + d]);
         out.push_back(R.startphi + j);
      }
      for (int c : R.child) {
         const Region& C = pr.regions[c];
         for (int j = 0; j < C.mynpan; j++) {
            for (int d = 0; d < 3; d++) cent.push_back(pr.cent[3 * (size_t)C.panel_ids[j] + d]);
            out.push_back(C.startdphi + j);
         }
      }
      DBuf<double> dc;
      DBuf<int> dout;
      dc.upload(cent, s);
      dout.upload(out, s);
      launch_rhs((int)out.size(), dc.p, dout.p, (int)pr.q.size(), pr.d_qpos.p, pr.d_q.p, rhs, s);
      FSVD_CHECK(hipStreamSynchronize(s));
   }
}

// Right-hand sides of ncol charge vectors at once (q: device, ncharge x ncol
// column-major; rhs: device, nvec x ncol); column c equals problem_rhs with
// the problem's charges replaced by q_c.
void problem_rhs_batch(Problem& pr, int ncol, const double* q, double* rhs) {
   hipStream_t s = pr.stream;
   const int nq = (int)pr.q.size();
   if (pr.qual) {
      launch_qual_rhs_batch(pr.ntotal, pr.regions[0].d_panels.p, nq, ncol, pr.d_qpos.p, q, 4.0 * M_PI * pr.eps_in,
                            pr.nvec, rhs, s);
      FSVD_CHECK(hipStreamSynchronize(s));
      return;
   }
   FSVD_CHECK(hipMemsetAsync(rhs, 0, sizeof(double) * pr.nvec * ncol, s));
   for (Region& R : pr.regions) {
      if (!R.has_charges) continue;
      std::vector<double> cent;
      std::vector<int> out;
      for (int j = 0; j < R.mynpan; j++) {
         for (int d = 0; d < 3; d++) cent.push_back(pr.cent[3 * (size_t)R.panel_ids[j] + d]);
         out.push_back(R.startphi + j);
      }
      for (int c : R.child) {
         const Region& C = pr.regions[c];
         for (int j = 0; j < C.mynpan; j++) {
            for (int d = 0; d < 3; d++) cent.push_back(pr.cent[3 * (size_t)C.panel_ids[j] + d]);
            out.push_back(C.startdphi + j);
         }
      }
      DBuf<double> dc;
      DBuf<int> dout;
      dc.upload(cent, s);
      dout.upload(out, s);
      launch_rhs_batch((int)out.size(), dc.p, dout.p, nq, ncol, pr.d_qpos.p, q, pr.nvec, rhs, s);
      FSVD_CHECK(hipStreamSynchronize(s));
   }
}

// Krylov basis vector j (slab j / BASIS_SLAB, allocated on first use).
template <typename VT>
static VT* basis_vec(Problem& pr, int j) {
   constexpr int S = Problem::BASIS_SLAB;
   const size_t off = (size_t)(j % S) * (size_t)pr.nvec;
   if constexpr (std::is_same<VT, double>::value) {
      while ((int)pr.basis.size() <= j / S) {
         pr.basis.emplace_back(new DBuf<double>());
         pr.basis.back()->alloc((size_t)S * pr.nvec);
      }
      return pr.basis[j / S]->p + off;
   } else {
      while ((int)pr.basis32.size() <= j / S) {
         pr.basis32.emplace_back(new DBuf<float>());
         pr.basis32.back()->alloc((size_t)S * pr.nvec);
      }
      return pr.basis32[j / S]->p + off;
   }
}

// The first basis slab of each precision the problem can solve in, made
// while the problem is built.
static void reserve_solver_memory(Problem& pr, bool want32) {
   basis_vec<double>(pr, 0);
   if (want32) basis_vec<float>(pr, 0);
}

// Device array of the basis pointers v_0..v_{m-1} (grown in steps of 64).
template <typename VT>
static const VT* const* basis_ptrs(Problem& pr, int m) {
   constexpr bool f64 = std::is_same<VT, double>::value;
   int& have = f64 ? pr.basis_ptr_n : pr.basis_ptr_n32;
   if (have < m) {
      const int cap = ((m + 63) / 64) * 64;
      std::vector<const VT*> p(cap);
      for (int j = 0; j < cap; j++) p[j] = basis_vec<VT>(pr, j);
      if constexpr (f64)
         pr.d_basis_ptr.upload(p, pr.stream);
      else
         pr.d_basis_ptr32.upload(p, pr.stream);
      FSVD_CHECK(hipStreamSynchronize(pr.stream));
      have = cap;
   }
   if constexpr (f64)
      return pr.d_basis_ptr.p;
   else
      return pr.d_basis_ptr32.p;
}

// Iterations until the next convergence check of a GMRES run: `check`, or
// fewer when the residual history (r_prev at iteration i_prev, r_last at
// i_last) predicts tol sooner at its recent geometric rate.  The cadence
// only decides how many iterations run past convergence before it is seen
// (they are discarded), never the iterate returned.
static int next_check_gap(double r_prev, int i_prev, double r_last, int i_last, double tol, int check) {
   if (!(r_prev > 0.0) || !(r_last > 0.0) || i_last <= i_prev) return check;
   const double rate = std::pow(r_last / r_prev, 1.0 / (double)(i_last - i_prev));
   if (!(rate < 1.0) || !(rate > 0.0)) return check;
   const double k = std::ceil(std::log(tol / r_last) / std::log(rate));
   if (!(k >= 1.0)) return 1;
   return (int)std::min<double>((double)check, k);
}

// GMRES over the basis type VT (the reference keeps FP64 vectors; the
// FP32-operator solves of the mixed mode keep FP32 basis vectors, with the
// residual refined in FP64 by problem_solve).
template <typename VT>
static SolveInfo gmres_impl(Problem& pr, int prec, const double* rhs, double* sol, double tol, int maxit,
                            std::vector<double>* hist) {
   double t0 = now_ms();
   SolveInfo info;
   hipStream_t s = pr.stream;
   const long long n = pr.nvec;
   const int np = partial_blocks(n);
   double* r = pr.r.p;
   double* w = pr.w.p;
   double* P0 = pr.partial.p;
   if (maxit < 1 || maxit > 2040) throw std::runtime_error("GMRES: maxit must be in 1..2040");
   problem_precond(pr, rhs, r);
   launch_dot_partial(n, np, r, r, P0, s);
   FSVD_CHECK(hipMemcpyAsync(pr.h_pinned, P0, sizeof(double) * np, hipMemcpyDeviceToHost, s));
   FSVD_CHECK(hipStreamSynchronize(s));
   double nr2 = 0.0;
   for (int k = 0; k < np; k++) nr2 += pr.h_pinned[k];
   double normr = std::sqrt(nr2);
   if (pr.gmres_zero_test ? normr < tol : !(normr > 0.0)) {
      FSVD_CHECK(hipMemsetAsync(sol, 0, sizeof(double) * n, s));
      FSVD_CHECK(hipStreamSynchronize(s));
      info.ms = now_ms() - t0;
      return info;
   }
   const int W = maxit + 1;
   // multi-dot chunks: about 2048 elements each, at most 256
   const int nch = (int)std::max<long long>(1, std::min<long long>(256, (n + 2047) / 2048));
   if ((long long)pr.partial_m.n < (long long)(maxit + 1) * nch) pr.partial_m.alloc((size_t)(maxit + 1) * nch);
   if ((long long)pr.htmp.n < maxit + 2) pr.htmp.alloc(maxit + 2);
   if ((long long)pr.hbuf.n < maxit + 2) pr.hbuf.alloc(maxit + 2);
   const int nax = multiaxpy_blocks(n);
   if ((long long)pr.pnorm.n < nax) pr.pnorm.alloc(nax);
   // Givens state on the device: H (W x W, column-major), rotations, g, and
   // the residual history; the host reads the history every CHECK iterations
   // (sooner when its rate predicts tol sooner) and stops at the first
   // iteration below tol, so the iteration count and the solution are those
   // of a per-iteration check.
   constexpr int CHECK = 4;
   if ((long long)pr.gm_H.n < (long long)W * W) pr.gm_H.alloc((size_t)W * W);
   if ((long long)pr.gm_v.n < 4LL * (W + 1)) pr.gm_v.alloc(4 * (size_t)(W + 1));
   double* dH = pr.gm_H.p;
   double* dcs = pr.gm_v.p;
   double* dsn = dcs + (W + 1);
   double* dg = dsn + (W + 1);
   double* dhist = dg + (W + 1);
   {
      std::vector<double> g0(W + 1, 0.0);
      g0[0] = normr;
      FSVD_CHECK(hipMemcpyAsync(dg, g0.data(), sizeof(double) * (W + 1), hipMemcpyHostToDevice, s));
   }
   launch_normalize(n, np, P0, pr.hbuf.p, r, basis_vec<VT>(pr, 0), s);
   int i, last = -1, done = -1, next = CHECK - 1, i_prev = -1;
   double r_prev = 1.0;   // the history is |g_{k+1}| / |r_0|
   for (i = 0; i < maxit; i++) {
      VT* vi = basis_vec<VT>(pr, i);
      VT* vnext_store = basis_vec<VT>(pr, i + 1);
      problem_matvec(pr, prec, vi, w);
      if (prec == PREC_F32)
         info.matvecs32++;
      else
         info.matvecs64++;
      problem_precond(pr, w, w);
      // Arnoldi column by classical Gram-Schmidt with one re-orthogonalisation
      // (CGS2): two passes of h = V^T w, w -= V h over v_0..v_i.  (The DGKS
      // test would ask for the second pass at every step here: the operator
      // is close to the identity, so w is mostly removed by the first pass.)
      // The reference orthogonalises by modified Gram-Schmidt
      // (GMRES.c:470-490); both give the same H up to rounding, CGS2 in 8
      // launches per iteration instead of i+3.
      const VT* const* V = basis_ptrs<VT>(pr, i + 2);
      for (int pass = 0; pass < 2; pass++) {
         launch_multidot<VT>(n, i + 1, nch, V, w, pr.partial_m.p, s);
         launch_multidot_finish(i + 1, nch, pr.partial_m.p, pr.htmp.p, pr.hbuf.p, pass, s);
         launch_multiaxpy<VT>(n, i + 1, pr.htmp.p, V, w, pass ? pr.pnorm.p : nullptr, s);
      }
      // |w|, v_{i+1} = w/|w|, H[i+1,i] and the Givens step of column i
      launch_normalize_givens(n, nax, pr.pnorm.p, w, vnext_store, i, W, pr.hbuf.p, dH, dcs, dsn, dg, dhist, normr,
                              s);
      if (i == next || i == maxit - 1) {
         const int a0 = last + 1;
         FSVD_CHECK(hipMemcpyAsync(pr.h_pinned, dhist + a0, sizeof(double) * (i + 1 - a0), hipMemcpyDeviceToHost, s));
         FSVD_CHECK(hipStreamSynchronize(s));
         for (int k = a0; k <= i; k++) {
            const double res = pr.h_pinned[k - a0];
            if (hist) hist->push_back(res);
            info.final_residual = res;
            if (res < tol) {
               done = k;
               break;
            }
         }
         last = i;
         if (done >= 0) break;
         const double r_last = pr.h_pinned[i - a0];
         next = i + next_check_gap(r_prev, i_prev, r_last, i, tol, CHECK);
         r_prev = r_last;
         i_prev = i;
      }
   }
   if (done >= 0) i = done;
   if (i >= maxit) i = maxit - 1;
   // back substitution on the host over the rotated H and g
   std::vector<double> H((size_t)W * (i + 1)), g(i + 1), y(i + 1);
   FSVD_CHECK(hipMemcpyAsync(H.data(), dH, sizeof(double) * (size_t)W * (i + 1), hipMemcpyDeviceToHost, s));
   FSVD_CHECK(hipMemcpyAsync(g.data(), dg, sizeof(double) * (i + 1), hipMemcpyDeviceToHost, s));
   FSVD_CHECK(hipStreamSynchronize(s));
   for (int k = 0; k <= i; k++) y[k] = g[k];
   for (int k = i; k >= 0; k--) {
      y[k] /= H[(size_t)k * W + k];
      for (int j = k - 1; j >= 0; j--) y[j] -= H[(size_t)k * W + j] * y[k];
   }
   info.iterations = i + 1;
   {
      DBuf<double> dy;
      dy.upload(y.data(), i + 1, s);
      launch_combine<VT>(n, i + 1, dy.p, basis_ptrs<VT>(pr, i + 1), sol, s);
      FSVD_CHECK(hipStreamSynchronize(s));
   }
   info.ms = now_ms() - t0;
   return info;
}

SolveInfo problem_gmres(Problem& pr, int prec, const double* rhs, double* sol, double tol, int maxit,
                        std::vector<double>* hist) {
   if (prec == PREC_F32) return gmres_impl<float>(pr, prec, rhs, sol, tol, maxit, hist);
   return gmres_impl<double>(pr, prec, rhs, sol, tol, maxit, hist);
}

static double precond_norm(Problem& pr, const double* v) {
   const long long n = pr.nvec;
   const int np = partial_blocks(n);
   problem_precond(pr, v, pr.tmp.p);
   launch_dot_partial(n, np, pr.tmp.p, pr.tmp.p, pr.partial.p, pr.stream);
   FSVD_CHECK(hipMemcpyAsync(pr.h_pinned, pr.partial.p, sizeof(double) * np, hipMemcpyDeviceToHost, pr.stream));
   FSVD_CHECK(hipStreamSynchronize(pr.stream));
   double a = 0.0;
   for (int k = 0; k < np; k++) a += pr.h_pinned[k];
   return std::sqrt(a);
}

SolveInfo problem_solve(Problem& pr, int mode, const double* rhs, double* sol, double tol, int maxit,
                        std::vector<double>* hist) {
   if (mode == PREC_F64) return problem_gmres(pr, PREC_F64, rhs, sol, tol, maxit, hist);
   double t0 = now_ms();
   hipStream_t s = pr.stream;
   const long long n = pr.nvec;
   SolveInfo tot = problem_gmres(pr, PREC_F32, rhs, sol, tol, maxit, hist);
   double bn = precond_norm(pr, rhs);
   DBuf<double> res, e;
   res.alloc(n);
   e.alloc(n);
   // up to MAX_REFINE corrections; the residual is evaluated once more
   // after the last, so final_residual is always that of the returned sol
   constexpr int MAX_REFINE = 4;
   for (int it = 0;; it++) {
      problem_matvec(pr, PREC_F64, sol, pr.tmp.p);
      tot.matvecs64++;
      launch_axpby(n, 1.0, rhs, -1.0, pr.tmp.p, res.p, s);
      double rn = precond_norm(pr, res.p);
      tot.final_residual = bn > 0 ? rn / bn : rn;
      if (!(tot.final_residual >= tol) || !(rn > 0.0) || it == MAX_REFINE) break;
      // correction solve on the normalised residual (GMRES's zero-solution
      // test is absolute, GMRES.c:461-465), scaled back afterwards
      launch_axpby(n, 1.0 / rn, res.p, 0.0, res.p, res.p, s);
      double inner = std::min(0.5, 0.5 * tol / tot.final_residual);
      SolveInfo si = problem_gmres(pr, PREC_F32, res.p, e.p, inner, maxit, nullptr);
      tot.iterations += si.iterations;
      tot.matvecs32 += si.matvecs32;
      tot.refinements++;
      launch_axpby(n, 1.0, sol, rn, e.p, sol, s);
   }
   FSVD_CHECK(hipStreamSynchronize(s));
   tot.ms = now_ms() - t0;
   return tot;
}

// ---- many right-hand sides in lockstep ----

// Copies `bytes` from host memory to a device slot of the upload ring, on
// the library stream; returns the device address.  A slot's host side is
// rewritten only after its previous copy has completed; its device side is
// rewritten by a copy on the library stream, which runs after everything
// enqueued before it there (and every aux stream is joined back into it).
static const void* ring_upload(Problem& pr, const void* data, size_t bytes) {
   if (bytes > Problem::RING_BYTES) throw std::runtime_error("batched step descriptors above the upload slot size");
   if (!pr.ring_h) {
      FSVD_CHECK(hipHostMalloc((void**)&pr.ring_h, Problem::RING * Problem::RING_BYTES));
      pr.ring_d.alloc(Problem::RING * Problem::RING_BYTES);
      for (hipEvent_t& e : pr.ring_ev) FSVD_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
   }
   const int k = pr.ring_next;
   pr.ring_next = (k + 1) % Problem::RING;
   FSVD_CHECK(hipEventSynchronize(pr.ring_ev[k]));
   unsigned char* h = pr.ring_h + (size_t)k * Problem::RING_BYTES;
   unsigned char* d = pr.ring_d.p + (size_t)k * Problem::RING_BYTES;
   std::memcpy(h, data, bytes);
   FSVD_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, pr.stream));
   FSVD_CHECK(hipEventRecord(pr.ring_ev[k], pr.stream));
   return d;
}

// y_c = A x_c for the ncol columns (x[c], y[c] device vectors): matvec_impl's
// steps per region, every launch covering all columns.
template <typename XT>
static void matvec_batch(Problem& pr, int prec, int ncol, const XT* const* x, double* const* y) {
   if (pr.adjoint || pr.qual) throw std::runtime_error("batched matvec: surface operators only");
   hipStream_t s = pr.stream;
   for (Region& R : pr.regions) {
      if (R.panel_ids.empty()) continue;
      const size_t need = (size_t)ncol * R.panel_ids.size();
      if (R.xsB.n < need) {
         FSVD_CHECK(hipStreamSynchronize(s));
         R.xsB.alloc(need);
         R.xdB.alloc(need);
         R.ytB.alloc(need);
      }
      TreeOp& op = R.shares_tree_with >= 0 ? *pr.regions[R.shares_tree_with].op : *R.op;
      if (prec == PREC_F32 && !op.have32) throw std::runtime_error("operator has no FP32 copy");
      tree_batch_reserve(op, ncol, s);
   }
   const XT* const* dx = (const XT* const*)ring_upload(pr, x, sizeof(XT*) * ncol);
   double* const* dy = (double* const*)ring_upload(pr, y, sizeof(double*) * ncol);
   if (pr.naux > 0) {
      FSVD_CHECK(hipEventRecord(pr.ev_fork, s));
      for (int k = 0; k < pr.naux; k++) FSVD_CHECK(hipStreamWaitEvent(pr.aux[k], pr.ev_fork, 0));
   }
   std::vector<std::vector<int>> lane_regions(pr.naux + 1);
   for (size_t r = 0; r < pr.regions.size(); r++)
      if (!pr.regions[r].panel_ids.empty()) lane_regions[pr.regions[r].lane].push_back((int)r);
   constexpr int NSTEP = 6;
   static const int parts[4] = {1, 2, 3, 0};
   size_t maxsteps = 0;
   for (auto& lr : lane_regions) maxsteps = std::max(maxsteps, lr.size() * NSTEP);
   for (size_t k = 0; k < maxsteps; k++) {
      for (size_t l = 0; l < lane_regions.size(); l++) {
         if (k >= lane_regions[l].size() * NSTEP) continue;
         Region& R = pr.regions[lane_regions[l][k / NSTEP]];
         const int step = (int)(k % NSTEP), n = (int)R.panel_ids.size();
         hipStream_t rs = R.lane == 0 ? s : pr.aux[R.lane - 1];
         TreeOp& op = R.shares_tree_with >= 0 ? *pr.regions[R.shares_tree_with].op : *R.op;
         if (step == 0) {
            launch_join_b<XT>(n, ncol, R.phi_idx.p, R.dphi_idx.p, R.s_phi.p, R.s_dphi.p, dx, R.xsB.p, R.xdB.p, n, rs);
         } else if (step < 5) {
            tree_apply_part_batch(op, prec, parts[step - 1], ncol, R.xsB.p, R.xdB.p, n, R.ytB.p, n, rs);
         } else {
            launch_split_b(n, ncol, R.out_idx.p, R.corr.p, R.ytB.p, R.xdB.p, n, dy, rs);
         }
      }
   }
   for (int k = 0; k < pr.naux; k++) {
      FSVD_CHECK(hipEventRecord(pr.ev_join[k], pr.aux[k]));
      FSVD_CHECK(hipStreamWaitEvent(s, pr.ev_join[k], 0));
   }
}

void problem_matvec_batch(Problem& pr, int prec, int ncol, const double* x, double* y) {
   std::vector<const double*> xs(ncol);
   std::vector<double*> ys(ncol);
   for (int c = 0; c < ncol; c++) {
      xs[c] = x + (size_t)c * pr.nvec;
      ys[c] = y + (size_t)c * pr.nvec;
   }
   matvec_batch(pr, prec, ncol, xs.data(), ys.data());
}

static std::vector<std::unique_ptr<DBuf<double>>>& slabs_of(GmCol& g, double*) { return g.basis64; }
static std::vector<std::unique_ptr<DBuf<float>>>& slabs_of(GmCol& g, float*) { return g.basis32; }

template <typename VT>
static VT* col_basis(Problem& pr, GmCol& g, int j) {
   constexpr int S = Problem::BASIS_SLAB;
   const size_t off = (size_t)(j % S) * (size_t)pr.nvec;
   auto& slabs = slabs_of(g, (VT*)nullptr);
   while ((int)slabs.size() <= j / S) {
      slabs.emplace_back(new DBuf<VT>());
      slabs.back()->alloc((size_t)S * pr.nvec);
   }
   return slabs[j / S]->p + off;
}

template <typename VT>
static const VT* const* col_basis_ptrs(Problem& pr, GmCol& g, int m) {
   constexpr bool f64 = std::is_same<VT, double>::value;
   int& have = f64 ? g.nptr64 : g.nptr32;
   if (have < m) {
      const int cap = ((m + 63) / 64) * 64;
      std::vector<const VT*> p(cap);
      for (int j = 0; j < cap; j++) p[j] = col_basis<VT>(pr, g, j);
      if constexpr (f64)
         g.ptr64.upload(p, pr.stream);
      else
         g.ptr32.upload(p, pr.stream);
      FSVD_CHECK(hipStreamSynchronize(pr.stream));
      have = cap;
   }
   if constexpr (f64)
      return g.ptr64.p;
   else
      return g.ptr32.p;
}

static void grow(DBuf<double>& b, size_t n) {
   if (b.n < n) b.alloc(n);
}

// gmres_impl for the columns cols[0..ncol) at once: column c solves
// rhs[c] -> sol[c] to cols[c]->tol with gmres_impl's kernels, inputs and
// stopping rule (so with its iterates); the matvec of every column still
// iterating is one batched call.  Adds the column's counts to cols[c]->info.
template <typename VT>
static void gmres_batch_impl(Problem& pr, int prec, int ncol, GmCol* const* cols, const double* const* rhs,
                             double* const* sol, int maxit) {
   hipStream_t s = pr.stream;
   const long long n = pr.nvec;
   const int np = partial_blocks(n);
   if (maxit < 1 || maxit > 2040) throw std::runtime_error("GMRES: maxit must be in 1..2040");
   const int W = maxit + 1;
   const int nch = (int)std::max<long long>(1, std::min<long long>(256, (n + 2047) / 2048));
   const int nax = multiaxpy_blocks(n);
   constexpr int CHECK = 4;
   if (pr.h_pinned_bcols < ncol) {
      if (pr.h_pinned_b) FSVD_CHECK(hipHostFree(pr.h_pinned_b));
      pr.h_pinned_b = nullptr;
      FSVD_CHECK(hipHostMalloc((void**)&pr.h_pinned_b, sizeof(double) * Problem::PIN_COL * ncol));
      pr.h_pinned_bcols = ncol;
   }
   double* pin = pr.h_pinned_b;
   for (int c = 0; c < ncol; c++) {
      GmCol& g = *cols[c];
      grow(g.w, n);
      grow(g.r, n);
      grow(g.partial, np);
      grow(g.partial_m, (size_t)(maxit + 1) * nch);
      grow(g.htmp, maxit + 2);
      grow(g.hbuf, maxit + 2);
      grow(g.pnorm, nax);
      grow(g.H, (size_t)W * W);
      grow(g.v, 4 * (size_t)(W + 1));
      col_basis<VT>(pr, g, 0);
   }
   for (int c = 0; c < ncol; c++) {
      GmCol& g = *cols[c];
      problem_precond(pr, rhs[c], g.r.p);
      launch_dot_partial(n, np, g.r.p, g.r.p, g.partial.p, s);
      FSVD_CHECK(hipMemcpyAsync(pin + (size_t)c * Problem::PIN_COL, g.partial.p, sizeof(double) * np,
                                hipMemcpyDeviceToHost, s));
   }
   FSVD_CHECK(hipStreamSynchronize(s));
   std::vector<double> g0(W + 1, 0.0);
   for (int c = 0; c < ncol; c++) {
      GmCol& g = *cols[c];
      double nr2 = 0.0;
      for (int k = 0; k < np; k++) nr2 += pin[(size_t)c * Problem::PIN_COL + k];
      g.normr = std::sqrt(nr2);
      g.done = -1;
      g.run = !(pr.gmres_zero_test ? g.normr < g.tol : !(g.normr > 0.0));
      if (!g.run) {
         FSVD_CHECK(hipMemsetAsync(sol[c], 0, sizeof(double) * n, s));
         continue;
      }
      double* dg = g.v.p + 2 * (W + 1);
      g0[0] = g.normr;
      FSVD_CHECK(hipMemcpyAsync(dg, g0.data(), sizeof(double) * (W + 1), hipMemcpyHostToDevice, s));
      // the upload reads g0 before it is changed for the next column
      FSVD_CHECK(hipStreamSynchronize(s));
      launch_normalize(n, np, g.partial.p, g.hbuf.p, g.r.p, col_basis<VT>(pr, g, 0), s);
   }
   std::vector<int> act;
   std::vector<const VT*> xin;
   std::vector<double*> yout;
   std::vector<int> started;
   for (int c = 0; c < ncol; c++)
      if (cols[c]->run) started.push_back(c);
   int i, last = -1, next = CHECK - 1;
   std::vector<double> r_prev(ncol, 1.0);   // per column, as gmres_impl
   std::vector<int> i_prev(ncol, -1);
   for (i = 0; i < maxit; i++) {
      act.clear();
      xin.clear();
      yout.clear();
      for (int c = 0; c < ncol; c++)
         if (cols[c]->run) {
            act.push_back(c);
            xin.push_back(col_basis<VT>(pr, *cols[c], i));
            yout.push_back(cols[c]->w.p);
         }
      if (act.empty()) break;
      // allocations and pointer uploads before the batched launches
      std::vector<GmDev> desc(act.size());
      for (size_t k = 0; k < act.size(); k++) {
         GmCol& g = *cols[act[k]];
         GmDev& d = desc[k];
         d.vnext = col_basis<VT>(pr, g, i + 1);
         d.V = col_basis_ptrs<VT>(pr, g, i + 2);
         d.w = g.w.p;
         d.partial_m = g.partial_m.p;
         d.htmp = g.htmp.p;
         d.hbuf = g.hbuf.p;
         d.pnorm = g.pnorm.p;
         d.H = g.H.p;
         d.cs = g.v.p;
         d.sn = d.cs + (W + 1);
         d.g = d.sn + (W + 1);
         d.hist = d.g + (W + 1);
         d.normr = g.normr;
      }
      const int na = (int)act.size();
      matvec_batch(pr, prec, na, xin.data(), yout.data());
      const GmDev* dd = (const GmDev*)ring_upload(pr, desc.data(), sizeof(GmDev) * na);
      for (int c : act) {
         if (prec == PREC_F32)
            cols[c]->info.matvecs32++;
         else
            cols[c]->info.matvecs64++;
      }
      // the per-column steps of gmres_impl, one launch for all columns each:
      // P^-1 w, CGS2 over v_0..v_i, |w|, v_{i+1} and the Givens step
      launch_precond_b(pr.nblocks, na, pr.pc_a.p, pr.pc_b.p, pr.pc_m.p, dd, s);
      launch_cgs2_b<VT>(n, i + 1, nch, na, dd, s);
      launch_normalize_givens_b<VT>(n, nax, i, W, na, dd, s);
      if (i == next || i == maxit - 1) {
         const int a0 = last + 1;
         for (int c : act)
            FSVD_CHECK(hipMemcpyAsync(pin + (size_t)c * Problem::PIN_COL, cols[c]->v.p + 3 * (W + 1) + a0,
                                      sizeof(double) * (i + 1 - a0), hipMemcpyDeviceToHost, s));
         FSVD_CHECK(hipStreamSynchronize(s));
         for (int c : act) {
            GmCol& g = *cols[c];
            for (int k = a0; k <= i; k++) {
               const double res = pin[(size_t)c * Problem::PIN_COL + (k - a0)];
               g.info.final_residual = res;
               if (res < g.tol) {
                  g.done = k;
                  break;
               }
            }
            if (g.done >= 0) g.run = false;
         }
         // the next check at the soonest predicted convergence of the
         // columns still running
         int gap = CHECK;
         for (int c : act) {
            GmCol& g = *cols[c];
            if (!g.run) continue;
            const double r_last = pin[(size_t)c * Problem::PIN_COL + (i - a0)];
            gap = std::min(gap, next_check_gap(r_prev[c], i_prev[c], r_last, i, g.tol, CHECK));
            r_prev[c] = r_last;
            i_prev[c] = i;
         }
         last = i;
         next = i + gap;
      }
   }
   // back substitution and the combination, column by column
   for (int c : started) {
      GmCol& g = *cols[c];
      int e = g.done >= 0 ? g.done : i;
      if (e >= maxit) e = maxit - 1;
      std::vector<double> H((size_t)W * (e + 1)), gg(e + 1), y(e + 1);
      FSVD_CHECK(hipMemcpyAsync(H.data(), g.H.p, sizeof(double) * (size_t)W * (e + 1), hipMemcpyDeviceToHost, s));
      FSVD_CHECK(hipMemcpyAsync(gg.data(), g.v.p + 2 * (W + 1), sizeof(double) * (e + 1), hipMemcpyDeviceToHost, s));
      FSVD_CHECK(hipStreamSynchronize(s));
      for (int k = 0; k <= e; k++) y[k] = gg[k];
      for (int k = e; k >= 0; k--) {
         y[k] /= H[(size_t)k * W + k];
         for (int j = k - 1; j >= 0; j--) y[j] -= H[(size_t)k * W + j] * y[k];
      }
      g.info.iterations += e + 1;
      g.run = false;
      DBuf<double> dy;
      dy.upload(y.data(), e + 1, s);
      launch_combine<VT>(n, e + 1, dy.p, col_basis_ptrs<VT>(pr, g, e + 1), sol[c], s);
      FSVD_CHECK(hipStreamSynchronize(s));
   }
}

// precond_norm with the column's own scratch vector
static double col_precond_norm(Problem& pr, GmCol& g, const double* v) {
   const long long n = pr.nvec;
   const int np = partial_blocks(n);
   problem_precond(pr, v, g.tmp.p);
   launch_dot_partial(n, np, g.tmp.p, g.tmp.p, g.partial.p, pr.stream);
   FSVD_CHECK(hipMemcpyAsync(pr.h_pinned, g.partial.p, sizeof(double) * np, hipMemcpyDeviceToHost, pr.stream));
   FSVD_CHECK(hipStreamSynchronize(pr.stream));
   double a = 0.0;
   for (int k = 0; k < np; k++) a += pr.h_pinned[k];
   return std::sqrt(a);
}

void problem_solve_batch(Problem& pr, int mode, int ncol, const double* rhs, double* sol, double tol, int maxit,
                         std::vector<SolveInfo>& info) {
   if (pr.adjoint || pr.qual) throw std::runtime_error("batched solve: surface operators only");
   const double t0 = now_ms();
   hipStream_t s = pr.stream;
   const long long n = pr.nvec;
   const int np = partial_blocks(n);
   info.assign(ncol, SolveInfo());
   // columns per lockstep group (work vectors and bases per column)
   constexpr int GROUP = 64;
   for (int c0 = 0; c0 < ncol; c0 += GROUP) {
      const int nc = std::min(GROUP, ncol - c0);
      while ((int)pr.gm_cols.size() < nc) pr.gm_cols.emplace_back(new GmCol());
      std::vector<GmCol*> cols(nc);
      std::vector<const double*> b(nc);
      std::vector<double*> x(nc);
      for (int c = 0; c < nc; c++) {
         cols[c] = pr.gm_cols[c].get();
         cols[c]->info = SolveInfo();
         cols[c]->tol = tol;
         grow(cols[c]->partial, np);
         b[c] = rhs + (size_t)(c0 + c) * n;
         x[c] = sol + (size_t)(c0 + c) * n;
      }
      if (mode == PREC_F64) {
         gmres_batch_impl<double>(pr, PREC_F64, nc, cols.data(), b.data(), x.data(), maxit);
      } else {
         gmres_batch_impl<float>(pr, PREC_F32, nc, cols.data(), b.data(), x.data(), maxit);
         for (int c = 0; c < nc; c++) {
            GmCol& g = *cols[c];
            grow(g.res, n);
            grow(g.e, n);
            grow(g.tmp, n);
            g.bn = col_precond_norm(pr, g, b[c]);
            g.refine = true;
         }
         constexpr int MAX_REFINE = 4;   // as problem_solve
         for (int it = 0;; it++) {
            std::vector<int> act;
            std::vector<const double*> xin;
            std::vector<double*> yout;
            for (int c = 0; c < nc; c++)
               if (cols[c]->refine) {
                  act.push_back(c);
                  xin.push_back(x[c]);
                  yout.push_back(cols[c]->tmp.p);
               }
            if (act.empty()) break;
            matvec_batch(pr, PREC_F64, (int)act.size(), xin.data(), yout.data());
            std::vector<GmCol*> rc;
            std::vector<const double*> rb;
            std::vector<double*> re;
            for (int c : act) {
               GmCol& g = *cols[c];
               g.info.matvecs64++;
               launch_axpby(n, 1.0, b[c], -1.0, g.tmp.p, g.res.p, s);
               g.rn = col_precond_norm(pr, g, g.res.p);
               g.info.final_residual = g.bn > 0 ? g.rn / g.bn : g.rn;
               if (!(g.info.final_residual >= tol) || !(g.rn > 0.0) || it == MAX_REFINE) {
                  g.refine = false;
                  continue;
               }
               launch_axpby(n, 1.0 / g.rn, g.res.p, 0.0, g.res.p, g.res.p, s);
               g.tol = std::min(0.5, 0.5 * tol / g.info.final_residual);
               rc.push_back(&g);
               rb.push_back(g.res.p);
               re.push_back(g.e.p);
            }
            if (rc.empty()) break;
            // the correction solves' own counts (the residual of the
            // refined solution is set by the next pass)
            std::vector<SolveInfo> keep(rc.size());
            for (size_t k = 0; k < rc.size(); k++) {
               keep[k] = rc[k]->info;
               rc[k]->info = SolveInfo();
            }
            gmres_batch_impl<float>(pr, PREC_F32, (int)rc.size(), rc.data(), rb.data(), re.data(), maxit);
            for (size_t k = 0; k < rc.size(); k++) {
               GmCol& g = *rc[k];
               SolveInfo tot = keep[k];
               tot.iterations += g.info.iterations;
               tot.matvecs32 += g.info.matvecs32;
               tot.refinements++;
               g.info = tot;
            }
            for (int c : act) {
               GmCol& g = *cols[c];
               if (g.refine) launch_axpby(n, 1.0, x[c], g.rn, g.e.p, x[c], s);
            }
         }
      }
      FSVD_CHECK(hipStreamSynchronize(s));
      for (int c = 0; c < nc; c++) info[c0 + c] = cols[c]->info;
   }
   const double ms = now_ms() - t0;
   for (SolveInfo& si : info) si.ms = ms;
}

void problem_potentials(Problem& pr, const double* sol, double* phi_out, bool use_m3) {
   hipStream_t s = pr.stream;
   const int nq = (int)pr.q.size();
   const PanelGeom* dall = pr.d_allpanels.p;
   DBuf<double> dphi;
   dphi.alloc(std::max(nq, 1));
   FSVD_CHECK(hipMemsetAsync(dphi.p, 0, sizeof(double) * std::max(nq, 1), s));
   if (pr.qual && pr.regions[0].m3 && use_m3) {
      // with M1M3: Tree_multiply(phi, M1M3, sol, SINGLE_LAYER_INT) (QualocationOperator.c:60-61)
      Region& R = pr.regions[0];
      TreeOp& m = *R.m3;
      const int n = pr.ntotal;
      std::vector<double> xt(n), zero(n, 0.0), yt(std::max(nq, 1));
      for (int k = 0; k < n; k++) xt[k] = sol[m.tree.pan_perm[k]];
      FSVD_CHECK(hipMemcpyAsync(R.m3_xs.p, xt.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
      FSVD_CHECK(hipMemcpyAsync(R.m3_xd.p, zero.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
      tree_apply(m, PREC_F64, R.m3_xs.p, R.m3_xd.p, R.m3_y.p, s);
      FSVD_CHECK(hipMemcpyAsync(yt.data(), R.m3_y.p, sizeof(double) * nq, hipMemcpyDeviceToHost, s));
      FSVD_CHECK(hipStreamSynchronize(s));
      for (int k = 0; k < nq; k++) phi_out[m.tree.pt_perm[k]] = yt[k];
      return;
   }
   if (pr.qual) {
      // QualocationOperator_calculateReactionPotentials (QualocationOperator.c:55-70):
      // phi[c] = sum_i sol[i] slp(c, panel i)
      std::vector<int> pidx(pr.ntotal), self(pr.ntotal, 0);
      std::vector<double> coef(2 * (size_t)pr.ntotal, 0.0);
      for (int i = 0; i < pr.ntotal; i++) {
         pidx[i] = i;
         coef[2 * (size_t)i] = sol[i];
      }
      DBuf<int> dp, dsf;
      DBuf<double> dc;
      dp.upload(pidx, s);
      dsf.upload(self, s);
      dc.upload(coef, s);
      launch_potentials(nq, pr.ntotal, pr.regions[0].d_panels.p, dp.p, dc.p, dsf.p, pr.d_qpos.p, dphi.p, s);
   }
   for (Region& R : pr.regions) {
      if (pr.qual || !R.m3 || !use_m3) continue;
      // SurfaceOperator_calculateReactionPotentials, M3 branch (SurfaceOperator.c:347-363):
      // phi_j = -(S dphi + D phi)_j through the M3 tree
      const int n = (int)R.panel_ids.size();
      DBuf<double> dsol;
      dsol.upload(sol, pr.nvec, s);
      launch_join(n, R.m3_phi_idx.p, R.m3_dphi_idx.p, R.m3_s_phi.p, R.m3_s_dphi.p, dsol.p, R.m3_xs.p, R.m3_xd.p, s);
      tree_apply(*R.m3, PREC_F64, R.m3_xs.p, R.m3_xd.p, R.m3_y.p, s);
      std::vector<double> react(nq), acc(std::max(nq, 1));
      FSVD_CHECK(hipMemcpyAsync(react.data(), R.m3_y.p, sizeof(double) * nq, hipMemcpyDeviceToHost, s));
      FSVD_CHECK(hipMemcpyAsync(acc.data(), dphi.p, sizeof(double) * std::max(nq, 1), hipMemcpyDeviceToHost, s));
      FSVD_CHECK(hipStreamSynchronize(s));
      for (int k = 0; k < nq; k++) acc[R.m3->tree.pt_perm[k]] += -react[k];
      FSVD_CHECK(hipMemcpyAsync(dphi.p, acc.data(), sizeof(double) * std::max(nq, 1), hipMemcpyHostToDevice, s));
      FSVD_CHECK(hipStreamSynchronize(s));
   }
   for (Region& R : pr.regions) {
      if (pr.qual || !R.has_charges || (R.m3 && use_m3)) continue;
      std::vector<int> pidx, self;
      std::vector<double> coef;
      for (int i = 0; i < R.mynpan; i++) {
         pidx.push_back(R.panel_ids[i]);
         coef.push_back(sol[R.startdphi + i]);
         coef.push_back(sol[R.startphi + i]);
         self.push_back(1);
      }
      for (int c : R.child) {
         const Region& C = pr.regions[c];
         for (int i = 0; i < C.mynpan; i++) {
            pidx.push_back(C.panel_ids[i]);
            coef.push_back(-C.eps / R.eps * sol[C.startdphi + i]);
            coef.push_back(-sol[C.startphi + i]);
            self.push_back(0);
         }
      }
      DBuf<int> dp, dsf;
      DBuf<double> dc;
      dp.upload(pidx, s);
      dsf.upload(self, s);
      dc.upload(coef, s);
      launch_potentials(nq, (int)pidx.size(), dall, dp.p, dc.p, dsf.p, pr.d_qpos.p, dphi.p, s);
      FSVD_CHECK(hipStreamSynchronize(s));
   }
   if (nq) FSVD_CHECK(hipMemcpyAsync(phi_out, dphi.p, sizeof(double) * nq, hipMemcpyDeviceToHost, s));
   FSVD_CHECK(hipStreamSynchronize(s));
}

// Reaction potentials of ncol solutions at once (sol: host, nvec x ncol;
// phi: host, ncharge x ncol); column c equals problem_potentials on sol_c.
void problem_potentials_batch(Problem& pr, int ncol, const double* sol, double* phi_out) {
   hipStream_t s = pr.stream;
   const int nq = (int)pr.q.size();
   const long long nv = pr.nvec;
   const PanelGeom* dall = pr.d_allpanels.p;
   DBuf<double> dphi;
   dphi.alloc(std::max<long long>((long long)nq * ncol, 1));
   FSVD_CHECK(hipMemsetAsync(dphi.p, 0, sizeof(double) * std::max<long long>((long long)nq * ncol, 1), s));
   auto run = [&](const std::vector<int>& pidx, const std::vector<int>& self, const std::vector<double>& coef,
                  const PanelGeom* panels) {
      DBuf<int> dp, dsf;
      DBuf<double> dc;
      dp.upload(pidx, s);
      dsf.upload(self, s);
      dc.upload(coef, s);
      launch_potentials_batch(nq, (int)pidx.size(), ncol, panels, dp.p, dc.p, dsf.p, pr.d_qpos.p, dphi.p, s);
      FSVD_CHECK(hipStreamSynchronize(s));
   };
   if (pr.qual) {
      const int n = pr.ntotal;
      std::vector<int> pidx(n), self(n, 0);
      std::vector<double> coef(2 * (size_t)n * ncol, 0.0);
      for (int i = 0; i < n; i++) pidx[i] = i;
      for (int c = 0; c < ncol; c++)
         for (int i = 0; i < n; i++) coef[2 * ((size_t)c * n + i)] = sol[(size_t)c * nv + i];
      run(pidx, self, coef, pr.regions[0].d_panels.p);
   } else {
      for (Region& R : pr.regions) {
         if (!R.has_charges) continue;
         std::vector<int> pidx, self;
         for (int i = 0; i < R.mynpan; i++) {
            pidx.push_back(R.panel_ids[i]);
            self.push_back(1);
         }
         for (int c : R.child)
            for (int i = 0; i < pr.regions[c].mynpan; i++) {
               pidx.push_back(pr.regions[c].panel_ids[i]);
               self.push_back(0);
            }
         const size_t ne = pidx.size();
         std::vector<double> coef(2 * ne * ncol);
         for (int col = 0; col < ncol; col++) {
            const double* x = sol + (size_t)col * nv;
            double* cf = coef.data() + 2 * ne * col;
            size_t e = 0;
            for (int i = 0; i < R.mynpan; i++, e++) {
               cf[2 * e] = x[R.startdphi + i];
               cf[2 * e + 1] = x[R.startphi + i];
            }
            for (int c : R.child) {
               const Region& C = pr.regions[c];
               for (int i = 0; i < C.mynpan; i++, e++) {
                  cf[2 * e] = -C.eps / R.eps * x[C.startdphi + i];
                  cf[2 * e + 1] = -x[C.startphi + i];
               }
            }
         }
         run(pidx, self, coef, dall);
      }
   }
   if (nq) FSVD_CHECK(hipMemcpyAsync(phi_out, dphi.p, sizeof(double) * nq * ncol, hipMemcpyDeviceToHost, s));
   FSVD_CHECK(hipStreamSynchronize(s));
}

// A^T x = b by the same GMRES (left preconditioner P^T), FP64 or mixed.
SolveInfo problem_solve_adjoint(Problem& pr, int mode, const double* rhs, double* sol, double tol, int maxit) {
   if (pr.qual) throw std::runtime_error("the adjoint of the qualocation operator is not provided");
   struct Flag {
      Problem& p;
      explicit Flag(Problem& q) : p(q) { p.adjoint = true; }
      ~Flag() { p.adjoint = false; }
   } flag(pr);
   return problem_solve(pr, mode, rhs, sol, tol, maxit, nullptr);
}

// The adjoint of the energy E = 1/2 q^T C x at the forward solution x:
// phi = C x (per charge, host) and lambda = A^-T C^T q (device).
static SolveInfo energy_adjoint(Problem& pr, int mode, const double* sol_host, double tol, int maxit,
                                std::vector<double>& phi, DBuf<double>& dlam) {
   hipStream_t s = pr.stream;
   const int nq = (int)pr.q.size();
   const long long nv = pr.nvec;
   // phi = C x
   phi.assign(std::max(nq, 1), 0.0);
   problem_potentials(pr, sol_host, phi.data(), false);
   // g = C^T q
   std::vector<double> g(nv, 0.0);
   const PanelGeom* dall = pr.d_allpanels.p;
   for (Region& R : pr.regions) {
      if (!R.has_charges) continue;
      std::vector<int> pidx, self;
      for (int i = 0; i < R.mynpan; i++) {
         pidx.push_back(R.panel_ids[i]);
         self.push_back(1);
      }
      for (int c : R.child)
         for (int i = 0; i < pr.regions[c].mynpan; i++) {
            pidx.push_back(pr.regions[c].panel_ids[i]);
            self.push_back(0);
         }
      const int ne = (int)pidx.size();
      DBuf<int> dp, dsf;
      DBuf<double> dsl, ddl;
      dp.upload(pidx, s);
      dsf.upload(self, s);
      dsl.alloc(std::max(ne, 1));
      ddl.alloc(std::max(ne, 1));
      launch_collect_t(ne, dall, dp.p, dsf.p, nq, pr.d_qpos.p, pr.d_q.p, dsl.p, ddl.p, s);
      std::vector<double> ssl(ne), sdl(ne);
      FSVD_CHECK(hipMemcpyAsync(ssl.data(), dsl.p, sizeof(double) * ne, hipMemcpyDeviceToHost, s));
      FSVD_CHECK(hipMemcpyAsync(sdl.data(), ddl.p, sizeof(double) * ne, hipMemcpyDeviceToHost, s));
      FSVD_CHECK(hipStreamSynchronize(s));
      int e = 0;
      for (int i = 0; i < R.mynpan; i++, e++) {
         g[R.startdphi + i] += ssl[e];
         g[R.startphi + i] += -sdl[e];
      }
      for (int c : R.child) {
         const Region& C = pr.regions[c];
         for (int i = 0; i < C.mynpan; i++, e++) {
            g[C.startdphi + i] += -C.eps / R.eps * ssl[e];
            g[C.startphi + i] += sdl[e];
         }
      }
   }
   // lambda = A^-T g
   DBuf<double> dg;
   dg.upload(g, s);
   dlam.alloc(nv);
   return problem_solve_adjoint(pr, mode, dg.p, dlam.p, tol, maxit);
}

SolveInfo problem_grad_charges(Problem& pr, int mode, const double* sol_host, double tol, int maxit,
                               double* grad_host) {
   if (pr.qual) throw std::runtime_error("charge gradients of the qualocation operator are not provided");
   hipStream_t s = pr.stream;
   const int nq = (int)pr.q.size();
   std::vector<double> phi;
   DBuf<double> dlam;
   SolveInfo si = energy_adjoint(pr, mode, sol_host, tol, maxit, phi, dlam);
   // B^T lambda over the right-hand-side rows of each charged region
   DBuf<double> dgq;
   dgq.alloc(std::max(nq, 1));
   FSVD_CHECK(hipMemsetAsync(dgq.p, 0, sizeof(double) * std::max(nq, 1), s));
   for (Region& R : pr.regions) {
      if (!R.has_charges) continue;
      std::vector<double> cent;
      std::vector<int> out;
      for (int j = 0; j < R.mynpan; j++) {
         for (int d = 0; d < 3; d++) cent.push_back(pr.cent[3 * (size_t)R.panel_ids[j] + d]);
         out.push_back(R.startphi + j);
      }
      for (int c : R.child) {
         const Region& C = pr.regions[c];
         for (int j = 0; j < C.mynpan; j++) {
            for (int d = 0; d < 3; d++) cent.push_back(pr.cent[3 * (size_t)C.panel_ids[j] + d]);
            out.push_back(C.startdphi + j);
         }
      }
      DBuf<double> dc;
      DBuf<int> dout;
      dc.upload(cent, s);
      dout.upload(out, s);
      launch_rhs_t(nq, (int)out.size(), dc.p, dout.p, dlam.p, pr.d_qpos.p, dgq.p, s);
      FSVD_CHECK(hipStreamSynchronize(s));
   }
   std::vector<double> bt(std::max(nq, 1));
   if (nq) FSVD_CHECK(hipMemcpyAsync(bt.data(), dgq.p, sizeof(double) * nq, hipMemcpyDeviceToHost, s));
   FSVD_CHECK(hipStreamSynchronize(s));
   for (int j = 0; j < nq; j++) grad_host[j] = 0.5 * (phi[j] + bt[j]);
   return si;
}

// dE/dkappa = -1/2 lambda^T (dA/dkappa) x = -1/2 q^T C A^-1 (dA/dkappa) x:
// kappa is one parameter, so the tangent form needs one forward solve,
// A w = (dA/dkappa) x, and the potentials of w, instead of the adjoint
// solve.  The charges sit in Laplace regions, so only the operators of the
// Yukawa regions depend on kappa; per such region the join of x, the
// kappa-derivative of its compressed operator at fixed ranks and bases
// (build_dkappa / tree_apply_dkappa), and the split into the rows the
// region writes (the -4 pi jump does not depend on kappa).
SolveInfo problem_grad_kappa(Problem& pr, int mode, const double* sol_host, double tol, int maxit,
                             double* dedk) {
   if (pr.qual) throw std::runtime_error("the qualocation operator has no kappa");
   hipStream_t s = pr.stream;
   const long long nv = pr.nvec;
   const int nq = (int)pr.q.size();
   DBuf<double> dx, da, dw;
   dx.upload(sol_host, nv, s);
   da.alloc(nv);
   FSVD_CHECK(hipMemsetAsync(da.p, 0, sizeof(double) * nv, s));
   bool any = false;
   for (Region& R : pr.regions) {
      if (R.kernel != YUKAWA || R.panel_ids.empty()) continue;
      any = true;
      const int n = (int)R.panel_ids.size();
      Region& O = R.shares_tree_with >= 0 ? pr.regions[R.shares_tree_with] : R;
      TreeOp& op = *O.op;
      DBuf<double> zero;
      zero.alloc(n);
      FSVD_CHECK(hipMemsetAsync(zero.p, 0, sizeof(double) * n, s));
      launch_join(n, R.phi_idx.p, R.dphi_idx.p, R.s_phi.p, R.s_dphi.p, dx.p, R.xs.p, R.xd.p, s);
      build_dkappa(op, O.d_panels.p, s);
      tree_apply_dkappa(op, R.xs.p, R.xd.p, R.yt.p, s);
      launch_split(n, R.out_idx.p, zero.p, R.yt.p, R.xd.p, da.p, s);
      FSVD_CHECK(hipStreamSynchronize(s));
   }
   SolveInfo si;
   if (!any) {
      *dedk = 0.0;
      return si;
   }
   dw.alloc(nv);
   si = problem_solve(pr, mode, da.p, dw.p, tol, maxit, nullptr);
   std::vector<double> w(nv), phi(std::max(nq, 1), 0.0);
   FSVD_CHECK(hipMemcpyAsync(w.data(), dw.p, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
   FSVD_CHECK(hipStreamSynchronize(s));
   problem_potentials(pr, w.data(), phi.data(), false);
   double acc = 0.0;
   for (int j = 0; j < nq; j++) acc += pr.q[j] * phi[j];
   *dedk = -0.5 * acc;
   return si;
}

}  // namespace fsvd
