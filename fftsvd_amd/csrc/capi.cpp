// C-ABI of libfftsvd_hip.so: the reference's FFTSVDpbeAPI.h solve path and
// the extended array-based entry points of fftsvd_hip.h.
#include <algorithm>
#include <atomic>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/FFTSVDpbeAPI.h"
#include "../../include/fftsvd_hip.h"
#include "solver.h"

namespace fsvd {
void launch_integrate_all(int kernel, double kappa, const PanelGeom* panels, int npan, const double* pts, int npt,
                          double* slp, double* dlp, hipStream_t s);
}

using namespace fsvd;

struct fftsvd_problem {
   std::unique_ptr<Problem> pr;
};

static thread_local std::string g_err;
static int g_precision = 0;

#define API_TRY try {
#define API_CATCH                                 \
   }                                              \
   catch (const std::exception& e) {              \
      g_err = e.what();                           \
      return -1;                                  \
   }                                              \
   catch (...) {                                  \
      g_err = "unknown error";                    \
      return -1;                                  \
   }                                              \
   return 0;

namespace fsvd {
void launch_copy(long long n4, const void* a, void* b, hipStream_t s);
void launch_lds_probe(int width, int nblk, int niter, float* out, hipStream_t s);
// the message fftsvd_last_error() returns, for entry points in other files
void set_last_error(const char* msg) { g_err = msg; }
}

extern "C" {

/* Parameters (reference FFTSVDpbeAPI.c:18-34) */
real dielectricvertexdensity = 0.0;
real saltvertexdensity = 0.0;
real innerdielectric = 0.0;
real outerdielectric = 0.0;
real ionicstrength = 0.0;
real ionexclusionradius = 0.0;
real proberadius = 0.0;
real errortolerance = 0.0;
char energycalculation = '\0';
unsigned int outputmodpdb = 0;
unsigned int outputfrc = 0;
unsigned int addcoulombic = 0;
unsigned int accelerateM3 = 0;
unsigned int usequalocation = 0;
unsigned int usecurved = 0;
real inftydielectric = 0.0;
real lambda = 0.0;

const char* fftsvd_last_error(void) { return g_err.c_str(); }

int fftsvd_set_precision(int precision) {
   if (precision != 0 && precision != 1) {
      g_err = "precision must be 0 (FP64) or 1 (mixed)";
      return -1;
   }
   g_precision = precision;
   return 0;
}

int fftsvd_copy_bandwidth(long long bytes, int nrep, double* gbs) {
   API_TRY
   if (bytes < 32 || nrep < 1) throw std::runtime_error("copy probe: bytes >= 32 and nrep >= 1 required");
   const long long n4 = bytes / 32;
   DBuf<double> a, b;
   a.alloc(4 * (size_t)n4);
   b.alloc(4 * (size_t)n4);
   hipStream_t s;
   FSVD_CHECK(hipStreamCreate(&s));
   FSVD_CHECK(hipMemsetAsync(a.p, 0, 32 * (size_t)n4, s));
   launch_copy(n4, a.p, b.p, s);
   hipEvent_t e0, e1;
   FSVD_CHECK(hipEventCreate(&e0));
   FSVD_CHECK(hipEventCreate(&e1));
   FSVD_CHECK(hipEventRecord(e0, s));
   for (int r = 0; r < nrep; r++) launch_copy(n4, a.p, b.p, s);
   FSVD_CHECK(hipEventRecord(e1, s));
   FSVD_CHECK(hipEventSynchronize(e1));
   float ms = 0;
   FSVD_CHECK(hipEventElapsedTime(&ms, e0, e1));
   (void)hipEventDestroy(e0);
   (void)hipEventDestroy(e1);
   FSVD_CHECK(hipStreamSynchronize(s));
   (void)hipStreamDestroy(s);
   // bytes read + bytes written per copy
   if (gbs) *gbs = 2.0 * 32.0 * (double)n4 * nrep / (ms * 1e-3) / 1e9;
   API_CATCH
}

int fftsvd_lds_bandwidth(int width, int nrep, double* gbs) {
   API_TRY
   if (nrep < 1 || (width != 8 && width != 16)) throw std::runtime_error("LDS probe: width 8 or 16, nrep >= 1");
   int dev = 0, ncu = 0;
   FSVD_CHECK(hipGetDevice(&dev));
   FSVD_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
   // eight 256-thread blocks per CU (32 KiB of LDS each: five resident per
   // CU at a time), enough rounds for about a millisecond
   const int nblk = 8 * std::max(ncu, 1), niter = 2048;
   DBuf<float> out;
   out.alloc(nblk);
   hipStream_t s;
   FSVD_CHECK(hipStreamCreate(&s));
   launch_lds_probe(width, nblk, niter, out.p, s);
   hipEvent_t e0, e1;
   FSVD_CHECK(hipEventCreate(&e0));
   FSVD_CHECK(hipEventCreate(&e1));
   FSVD_CHECK(hipEventRecord(e0, s));
   for (int r = 0; r < nrep; r++) launch_lds_probe(width, nblk, niter, out.p, s);
   FSVD_CHECK(hipEventRecord(e1, s));
   FSVD_CHECK(hipEventSynchronize(e1));
   float ms = 0;
   FSVD_CHECK(hipEventElapsedTime(&ms, e0, e1));
   (void)hipEventDestroy(e0);
   (void)hipEventDestroy(e1);
   FSVD_CHECK(hipStreamSynchronize(s));
   (void)hipStreamDestroy(s);
   if (gbs) *gbs = (double)nblk * 256.0 * 8.0 * niter * width * nrep / (ms * 1e-3) / 1e9;
   API_CATCH
}

int fftsvd_device_info(char* name, int len, int* num_cu) {
   API_TRY
   int dev = 0;
   FSVD_CHECK(hipGetDevice(&dev));
   hipDeviceProp_t p;
   FSVD_CHECK(hipGetDeviceProperties(&p, dev));
   if (name && len > 0) {
      std::snprintf(name, len, "%s (%s)", p.name, p.gcnArchName);
   }
   if (num_cu) *num_cu = p.multiProcessorCount;
   API_CATCH
}

static SurfaceSpec make_spec(int ndiel, const int* npan_diel, int ncav, const int* npan_cav, const int* cav_parent,
                             const double* verts9, const double* qpos, const double* q, int ncharge, double eps_in,
                             double eps_out, double kappa, int maxpanels, int gridpoints, double svdeps,
                             int want_f32) {
   SurfaceSpec sp;
   size_t off = 0;
   for (int i = 0; i < ndiel; i++) {
      sp.dielectric_verts.emplace_back(verts9 + off, verts9 + off + 9 * (size_t)npan_diel[i]);
      off += 9 * (size_t)npan_diel[i];
   }
   // a cavity parent of 999 marks a surface the mesher rejected: readSRF
   // takes its panels out of the count and generateSurfaceOperator attaches
   // it to no region (FFTSVDpbeAPI.c:1006-1015,1441-1444)
   for (int j = 0; j < ncav; j++) {
      const size_t n9 = 9 * (size_t)npan_cav[j];
      if (cav_parent[j] == 999) {
         off += n9;
         continue;
      }
      if (cav_parent[j] < 0 || cav_parent[j] >= ndiel) throw std::runtime_error("cavity parent out of range");
      sp.cavity_verts.emplace_back(verts9 + off, verts9 + off + n9);
      off += n9;
      sp.cavity_parent.push_back(cav_parent[j]);
   }
   sp.qpos.assign(qpos, qpos + 3 * (size_t)ncharge);
   sp.q.assign(q, q + ncharge);
   sp.eps_in = eps_in;
   sp.eps_out = eps_out;
   sp.kappa = kappa;
   sp.maxpanels = maxpanels;
   sp.gridpoints = gridpoints;
   sp.svdeps = svdeps;
   sp.want32 = want_f32 != 0;
   sp.m3 = accelerateM3 != 0;
   return sp;
}

int fftsvd_problem_create_multi(int ndiel, const int* npan_diel, int ncav, const int* npan_cav, const int* cav_parent,
                                const double* verts9, const double* qpos, const double* q, int ncharge, double eps_in,
                                double eps_out, double kappa, int maxpanels, int gridpoints, double svdeps,
                                int want_f32, fftsvd_problem** out) {
   API_TRY
   if (gridpoints != 3 && gridpoints != 4)
      throw std::runtime_error("3 or 4 grid points per cube edge at the finest levels are supported");
   if (ndiel < 1) throw std::runtime_error("need at least one dielectric surface");
   auto sp = make_spec(ndiel, npan_diel, ncav, npan_cav, cav_parent, verts9, qpos, q, ncharge, eps_in, eps_out, kappa,
                       maxpanels, gridpoints, svdeps, want_f32);
   auto* h = new fftsvd_problem;
   h->pr = build_problem(sp);
   *out = h;
   API_CATCH
}

int fftsvd_problem_create_stern(int nsalt, const int* npan_salt, int ndiel, const int* npan_diel,
                                const int* diel_parent, int ncav, const int* npan_cav, const int* cav_parent,
                                int nscav, const int* npan_scav, const int* scav_parent, const double* verts9,
                                const double* qpos, const double* q, int ncharge, double eps_in, double eps_out,
                                double kappa, int maxpanels, int gridpoints, double svdeps, int want_f32,
                                fftsvd_problem** out) {
   API_TRY
   if (gridpoints != 3 && gridpoints != 4)
      throw std::runtime_error("3 or 4 grid points per cube edge at the finest levels are supported");
   if (nsalt < 1 || ndiel < 1) throw std::runtime_error("need at least one salt and one dielectric surface");
   if (!(kappa > 0.0)) throw std::runtime_error("the Stern-layer formulation needs kappa > 0");
   size_t off = 0;
   SurfaceSpec sp;
   for (int i = 0; i < nsalt; i++) {
      sp.salt_verts.emplace_back(verts9 + off, verts9 + off + 9 * (size_t)npan_salt[i]);
      off += 9 * (size_t)npan_salt[i];
   }
   // dielectrics and dielectric cavities as in the salt-free layout
   SurfaceSpec rest = make_spec(ndiel, npan_diel, ncav, npan_cav, cav_parent, verts9 + off, qpos, q, ncharge, eps_in,
                                eps_out, kappa, maxpanels, gridpoints, svdeps, want_f32);
   for (int i = 0; i < ndiel; i++) off += 9 * (size_t)npan_diel[i];
   for (int j = 0; j < ncav; j++) off += 9 * (size_t)npan_cav[j];
   rest.salt_verts = std::move(sp.salt_verts);
   for (int i = 0; i < ndiel; i++) {
      if (diel_parent[i] < 0 || diel_parent[i] >= nsalt) throw std::runtime_error("dielectric parent out of range");
      rest.dielectric_parent.push_back(diel_parent[i]);
   }
   // dielectric cavities kept by make_spec, renumbered without the dropped ones
   std::vector<int> cav_index(ncav, -1);
   for (int j = 0, kept = 0; j < ncav; j++)
      if (cav_parent[j] != 999) cav_index[j] = kept++;
   for (int l = 0; l < nscav; l++) {
      const size_t n9 = 9 * (size_t)npan_scav[l];
      if (scav_parent[l] == 999) {   // rejected by the mesher (FFTSVDpbeAPI.c:1023-1026)
         off += n9;
         continue;
      }
      if (scav_parent[l] < 0 || scav_parent[l] >= ncav) throw std::runtime_error("salt cavity parent out of range");
      if (cav_index[scav_parent[l]] < 0) throw std::runtime_error("salt cavity inside a dropped dielectric cavity");
      rest.salt_cavity_verts.emplace_back(verts9 + off, verts9 + off + n9);
      off += n9;
      rest.salt_cavity_parent.push_back(cav_index[scav_parent[l]]);
   }
   auto* h = new fftsvd_problem;
   h->pr = build_problem(rest);
   *out = h;
   API_CATCH
}

int fftsvd_problem_create(const double* verts9, int npan, const double* qpos, const double* q, int ncharge,
                          double eps_in, double eps_out, double kappa, int maxpanels, int gridpoints, double svdeps,
                          int want_f32, fftsvd_problem** out) {
   return fftsvd_problem_create_multi(1, &npan, 0, nullptr, nullptr, verts9, qpos, q, ncharge, eps_in, eps_out, kappa,
                                      maxpanels, gridpoints, svdeps, want_f32, out);
}

int fftsvd_qual_create(const double* verts9, int npan, const double* qpos, const double* q, int ncharge,
                       double eps_in, double eps_out, int maxpanels, int gridpoints, double svdeps, int want_f32,
                       fftsvd_problem** out) {
   API_TRY
   if (!out) throw std::runtime_error("null output handle");
   *out = nullptr;
   if (npan < 1 || ncharge < 0) throw std::runtime_error("bad sizes");
   if (gridpoints != 3 && gridpoints != 4)
      throw std::runtime_error("3 or 4 grid points per cube edge at the finest levels are supported");
   QualSpec sp;
   sp.verts.assign(verts9, verts9 + 9 * (size_t)npan);
   if (ncharge) {
      sp.qpos.assign(qpos, qpos + 3 * (size_t)ncharge);
      sp.q.assign(q, q + ncharge);
   }
   sp.eps_in = eps_in;
   sp.eps_out = eps_out;
   sp.maxpanels = maxpanels;
   sp.gridpoints = gridpoints;
   sp.svdeps = svdeps;
   sp.want32 = want_f32 != 0;
   sp.m3 = accelerateM3 != 0;
   auto* h = new fftsvd_problem;
   try {
      h->pr = build_qual_problem(sp);
   } catch (...) {
      delete h;
      throw;
   }
   *out = h;
   API_CATCH
}

int fftsvd_problem_unknowns(fftsvd_problem* p, long long* n) {
   API_TRY
   if (n) *n = p->pr->nvec;
   API_CATCH
}

int fftsvd_problem_destroy(fftsvd_problem* p) {
   API_TRY
   delete p;
   API_CATCH
}

int fftsvd_problem_size(fftsvd_problem* p, int* npanels, int* ncharges) {
   API_TRY
   if (npanels) *npanels = p->pr->ntotal;
   if (ncharges) *ncharges = (int)p->pr->q.size();
   API_CATCH
}

int fftsvd_problem_matvec(fftsvd_problem* p, int precision, const double* x, double* y) {
   API_TRY
   Problem& pr = *p->pr;
   size_t n = (size_t)pr.nvec;
   DBuf<double> dx, dy;
   dx.upload(x, n, pr.stream);
   dy.alloc(n);
   problem_matvec(pr, precision, dx.p, dy.p);
   FSVD_CHECK(hipMemcpyAsync(y, dy.p, sizeof(double) * n, hipMemcpyDeviceToHost, pr.stream));
   FSVD_CHECK(hipStreamSynchronize(pr.stream));
   API_CATCH
}

int fftsvd_problem_precond(fftsvd_problem* p, const double* b, double* x) {
   API_TRY
   Problem& pr = *p->pr;
   size_t n = (size_t)pr.nvec;
   DBuf<double> db, dx;
   db.upload(b, n, pr.stream);
   dx.alloc(n);
   problem_precond(pr, db.p, dx.p);
   FSVD_CHECK(hipMemcpyAsync(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost, pr.stream));
   FSVD_CHECK(hipStreamSynchronize(pr.stream));
   API_CATCH
}

int fftsvd_problem_rhs(fftsvd_problem* p, double* rhs) {
   API_TRY
   Problem& pr = *p->pr;
   size_t n = (size_t)pr.nvec;
   DBuf<double> d;
   d.alloc(n);
   problem_rhs(pr, d.p);
   FSVD_CHECK(hipMemcpyAsync(rhs, d.p, sizeof(double) * n, hipMemcpyDeviceToHost, pr.stream));
   FSVD_CHECK(hipStreamSynchronize(pr.stream));
   API_CATCH
}

int fftsvd_problem_solve(fftsvd_problem* p, int precision, const double* rhs, double* sol, double tol, int maxit,
                         int* iters, double* resid_hist, double* info) {
   API_TRY
   Problem& pr = *p->pr;
   size_t n = (size_t)pr.nvec;
   if (maxit < 1 || maxit > 1000) throw std::runtime_error("maxit must be in [1, 1000]");
   DBuf<double> db, dx;
   db.upload(rhs, n, pr.stream);
   dx.alloc(n);
   std::vector<double> hist;
   SolveInfo si = problem_solve(pr, precision, db.p, dx.p, tol, maxit, &hist);
   FSVD_CHECK(hipMemcpyAsync(sol, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost, pr.stream));
   FSVD_CHECK(hipStreamSynchronize(pr.stream));
   if (iters) *iters = si.iterations;
   if (resid_hist)
      for (size_t k = 0; k < hist.size() && (int)k < maxit; k++) resid_hist[k] = hist[k];
   if (info) {
      info[0] = si.iterations;
      info[1] = si.refinements;
      info[2] = si.final_residual;
      info[3] = si.ms;
      info[4] = si.matvecs64;
      info[5] = si.matvecs32;
   }
   API_CATCH
}

int fftsvd_problems_solve(int nprob, fftsvd_problem* const* probs, int precision, const double* const* rhs,
                          double* const* sol, double tol, int maxit, int* iters, double* info) {
   API_TRY
   if (nprob < 1) throw std::runtime_error("nprob must be positive");
   if (maxit < 1 || maxit > 1000) throw std::runtime_error("maxit must be in [1, 1000]");
   for (int k = 0; k < nprob; k++)
      if (!probs[k] || !probs[k]->pr) throw std::runtime_error("null problem in the batch");
   // largest first, taken by MULTI_THREADS workers, each problem's solve on
   // its own streams, so the problems' kernels share the GPU
   constexpr int MULTI_THREADS = 8;
   std::vector<int> order(nprob);
   for (int k = 0; k < nprob; k++) order[k] = k;
   std::stable_sort(order.begin(), order.end(),
                    [&](int a, int b) { return probs[a]->pr->nvec > probs[b]->pr->nvec; });
   int dev = 0;
   FSVD_CHECK(hipGetDevice(&dev));
   std::atomic<int> next{0};
   std::vector<std::exception_ptr> err(nprob);
   auto work = [&] {
      for (int j; (j = next.fetch_add(1)) < nprob;) {
         const int k = order[j];
         try {
            FSVD_CHECK(hipSetDevice(dev));
            Problem& pr = *probs[k]->pr;
            const size_t n = (size_t)pr.nvec;
            DBuf<double> db, dx;
            db.upload(rhs[k], n, pr.stream);
            dx.alloc(n);
            const SolveInfo si = problem_solve(pr, precision, db.p, dx.p, tol, maxit, nullptr);
            FSVD_CHECK(hipMemcpyAsync(sol[k], dx.p, sizeof(double) * n, hipMemcpyDeviceToHost, pr.stream));
            FSVD_CHECK(hipStreamSynchronize(pr.stream));
            if (iters) iters[k] = si.iterations;
            if (info) {
               double* o = info + 6 * (size_t)k;
               o[0] = si.iterations;
               o[1] = si.refinements;
               o[2] = si.final_residual;
               o[3] = si.ms;
               o[4] = si.matvecs64;
               o[5] = si.matvecs32;
            }
         } catch (...) {
            err[k] = std::current_exception();
         }
      }
   };
   std::vector<std::thread> th;
   for (int t = 1; t < std::min(nprob, MULTI_THREADS); t++) th.emplace_back(work);
   work();
   for (auto& t : th) t.join();
   for (auto& e : err)
      if (e) std::rethrow_exception(e);
   API_CATCH
}

int fftsvd_problem_rhs_batch(fftsvd_problem* p, int ncol, const double* q, double* rhs) {
   API_TRY
   Problem& pr = *p->pr;
   if (ncol < 1) throw std::runtime_error("ncol must be positive");
   const size_t nq = pr.q.size(), nv = (size_t)pr.nvec;
   DBuf<double> dq, dr;
   dq.upload(q, std::max<size_t>(nq * ncol, 1), pr.stream);
   dr.alloc(nv * ncol);
   problem_rhs_batch(pr, ncol, dq.p, dr.p);
   FSVD_CHECK(hipMemcpyAsync(rhs, dr.p, sizeof(double) * nv * ncol, hipMemcpyDeviceToHost, pr.stream));
   FSVD_CHECK(hipStreamSynchronize(pr.stream));
   API_CATCH
}

int fftsvd_problem_solve_batch(fftsvd_problem* p, int precision, int ncol, const double* rhs, double* sol,
                               double tol, int maxit, int* iters, double* info) {
   API_TRY
   Problem& pr = *p->pr;
   if (ncol < 1) throw std::runtime_error("ncol must be positive");
   if (maxit < 1 || maxit > 1000) throw std::runtime_error("maxit must be in [1, 1000]");
   const size_t nv = (size_t)pr.nvec;
   DBuf<double> db, dx;
   db.upload(rhs, nv * ncol, pr.stream);
   dx.alloc(nv * ncol);
   // surface operators: all columns in lockstep over the column-batched
   // matvec; the qualocation operator: column after column
   std::vector<SolveInfo> all;
   if (!pr.qual)
      problem_solve_batch(pr, precision, ncol, db.p, dx.p, tol, maxit, all);
   else
      for (int c = 0; c < ncol; c++) all.push_back(problem_solve(pr, precision, db.p + nv * c, dx.p + nv * c, tol,
                                                                 maxit, nullptr));
   for (int c = 0; c < ncol; c++) {
      const SolveInfo& si = all[c];
      if (iters) iters[c] = si.iterations;
      if (info) {
         double* o = info + 6 * (size_t)c;
         o[0] = si.iterations;
         o[1] = si.refinements;
         o[2] = si.final_residual;
         o[3] = si.ms;
         o[4] = si.matvecs64;
         o[5] = si.matvecs32;
      }
   }
   FSVD_CHECK(hipMemcpyAsync(sol, dx.p, sizeof(double) * nv * ncol, hipMemcpyDeviceToHost, pr.stream));
   FSVD_CHECK(hipStreamSynchronize(pr.stream));
   API_CATCH
}

int fftsvd_problem_matvec_batch(fftsvd_problem* p, int precision, int ncol, const double* x, double* y) {
   API_TRY
   Problem& pr = *p->pr;
   if (pr.qual) throw std::runtime_error("the batched matvec takes surface operators");
   if (ncol < 1) throw std::runtime_error("ncol must be positive");
   const size_t nv = (size_t)pr.nvec;
   DBuf<double> dx, dy;
   dx.upload(x, nv * ncol, pr.stream);
   dy.alloc(nv * ncol);
   problem_matvec_batch(pr, precision, ncol, dx.p, dy.p);
   FSVD_CHECK(hipMemcpyAsync(y, dy.p, sizeof(double) * nv * ncol, hipMemcpyDeviceToHost, pr.stream));
   FSVD_CHECK(hipStreamSynchronize(pr.stream));
   API_CATCH
}

int fftsvd_problem_matvec_transpose(fftsvd_problem* p, int precision, const double* y, double* x) {
   API_TRY
   Problem& pr = *p->pr;
   if (pr.qual) throw std::runtime_error("the adjoint of the qualocation operator is not provided");
   const size_t n = (size_t)pr.nvec;
   DBuf<double> dy, dx;
   dy.upload(y, n, pr.stream);
   dx.alloc(n);
   pr.adjoint = true;
   try {
      problem_matvec(pr, precision, dy.p, dx.p);
   } catch (...) {
      pr.adjoint = false;
      throw;
   }
   pr.adjoint = false;
   FSVD_CHECK(hipMemcpyAsync(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost, pr.stream));
   FSVD_CHECK(hipStreamSynchronize(pr.stream));
   API_CATCH
}

int fftsvd_problem_solve_adjoint(fftsvd_problem* p, int precision, const double* rhs, double* sol, double tol,
                                 int maxit, int* iters, double* info) {
   API_TRY
   Problem& pr = *p->pr;
   if (maxit < 1 || maxit > 1000) throw std::runtime_error("maxit must be in [1, 1000]");
   const size_t n = (size_t)pr.nvec;
   DBuf<double> db, dx;
   db.upload(rhs, n, pr.stream);
   dx.alloc(n);
   SolveInfo si = problem_solve_adjoint(pr, precision, db.p, dx.p, tol, maxit);
   FSVD_CHECK(hipMemcpyAsync(sol, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost, pr.stream));
   FSVD_CHECK(hipStreamSynchronize(pr.stream));
   if (iters) *iters = si.iterations;
   if (info) {
      info[0] = si.iterations;
      info[1] = si.refinements;
      info[2] = si.final_residual;
      info[3] = si.ms;
      info[4] = si.matvecs64;
      info[5] = si.matvecs32;
   }
   API_CATCH
}

int fftsvd_problem_grad_charges(fftsvd_problem* p, int precision, const double* sol, double tol, int maxit,
                                double* grad, int* iters) {
   API_TRY
   if (maxit < 1 || maxit > 1000) throw std::runtime_error("maxit must be in [1, 1000]");
   SolveInfo si = problem_grad_charges(*p->pr, precision, sol, tol, maxit, grad);
   if (iters) *iters = si.iterations;
   API_CATCH
}

int fftsvd_problem_grad_kappa(fftsvd_problem* p, int precision, const double* sol, double tol, int maxit,
                              double* dedk, int* iters) {
   API_TRY
   if (maxit < 1 || maxit > 1000) throw std::runtime_error("maxit must be in [1, 1000]");
   if (!dedk) throw std::runtime_error("null output");
   SolveInfo si = problem_grad_kappa(*p->pr, precision, sol, tol, maxit, dedk);
   if (iters) *iters = si.iterations;
   API_CATCH
}

int fftsvd_problem_potentials_batch(fftsvd_problem* p, int ncol, const double* sol, double* phi) {
   API_TRY
   if (ncol < 1) throw std::runtime_error("ncol must be positive");
   problem_potentials_batch(*p->pr, ncol, sol, phi);
   API_CATCH
}

int fftsvd_problem_potentials(fftsvd_problem* p, const double* sol, double* phi) {
   API_TRY
   problem_potentials(*p->pr, sol, phi);
   API_CATCH
}

// which >= 0: the tree of region `which`; which < 0: the M3 tree of region -which-1
static TreeOp& region_op(Problem& pr, int which) {
   if (which < 0) {
      const int r = -which - 1;
      if (r >= (int)pr.regions.size() || !pr.regions[r].m3) throw std::runtime_error("no M3 tree for that region");
      return *pr.regions[r].m3;
   }
   if (which >= (int)pr.regions.size()) throw std::runtime_error("no such region");
   Region& R = pr.regions[which];
   return R.shares_tree_with >= 0 ? *pr.regions[R.shares_tree_with].op : *R.op;
}

/* out: 0 panels, 1 regions, 2 setup ms, 3 near entries (all trees), 4 far
 * entries, 5 spectra bins, 6 translate complex MACs, 7 max depth,
 * 8 leaves, 9 unique trees, 10 algorithmic bytes per FP64 matvec,
 * 11 algorithmic bytes per FP32 matvec, 12 near bytes per FP64 matvec,
 * 13 near bytes per FP32 matvec, 14 near-field column indices (all trees) */
int fftsvd_set_autotune(int on) {
   API_TRY
   set_autotune(on != 0);
   API_CATCH
}

int fftsvd_translation_params(int* mode, int* budget, int* tile, int* tuned, double* tuned_ms) {
   API_TRY
   bool t = false;
   const PipeParams pp = current_pipe_params(&t);
   if (mode) *mode = pp.mode;
   if (budget) *budget = pp.budget;
   if (tile) *tile = pp.tile;
   if (tuned) *tuned = t ? 1 : 0;
   if (tuned_ms) *tuned_ms = pp.tuned_ms;
   API_CATCH
}

int fftsvd_problem_stats(fftsvd_problem* p, double* out, int nout) {
   API_TRY
   Problem& pr = *p->pr;
   double v[32] = {0};
   v[0] = pr.ntotal;
   v[1] = (double)pr.regions.size();
   v[2] = pr.setup_ms;
   double glue = 0, vec = 0;
   for (Region& R : pr.regions) {
      const int n = (int)R.panel_ids.size();
      if (n == 0) continue;
      glue += 8.0 * (2 * n /*gather*/ + 3 * n /*split*/);
      vec += 8.0 * 3 * n;
      TreeOp* opp = R.shares_tree_with >= 0 ? pr.regions[R.shares_tree_with].op.get() : R.op.get();
      if (!opp) continue;
      TreeOp& op = *opp;
      // every region applies its tree once per matvec
      v[3] += op.stats.near_entries;
      v[4] += op.stats.far_entries;
      v[5] += op.stats.spectra_bins;
      v[6] += op.stats.translate_macs;
      v[7] = std::max<double>(v[7], op.stats.depth);
      v[8] += op.stats.nleaf;
      v[9] += 1;
      v[14] += (double)op.h_dcol.size();
      for (const FarLevel& lv : op.levels) {
         v[15] += (double)lv.UT64.n;
         v[16] += (double)lv.VT64.n;
         v[17] += (double)lv.PV64.n;
         v[18] += (double)lv.UTI64.n;
         v[19] += (double)lv.T64.n / 2;
         v[23] += (double)lv.nsrc * lv.pad;
         v[24] += (double)lv.ntgt * lv.pad;
      }
      v[20] += op.plan.nublk;
      if (R.shares_tree_with < 0)
         for (int k = 0; k < 7; k++) v[25 + k] += op.stats.phase_ms[k];
      v[21] += op.plan.nchunk;
      v[22] += (double)n;
   }
   // algorithmic bytes: stored operator entries once, spectra written and
   // read once, vectors (2 gathers + 1 write per point per tree) and glue
   v[10] = 8.0 * (v[3] + v[4]) + 16.0 * 2 * v[5] + vec + glue;
   v[11] = 4.0 * (v[3] + v[4]) + 8.0 * 2 * v[5] + vec + glue;
   v[12] = 8.0 * v[3];
   v[13] = 4.0 * v[3];
   for (int i = 0; i < nout && i < 32; i++) out[i] = v[i];
   API_CATCH
}

int fftsvd_problem_bench_matvec(fftsvd_problem* p, int precision, int nrep, double* ms_total, double* ms_parts) {
   API_TRY
   Problem& pr = *p->pr;
   hipStream_t s = pr.stream;
   size_t n = (size_t)pr.nvec;
   std::vector<double> hx(n);
   std::mt19937_64 rng(12345);
   std::uniform_real_distribution<double> U(-1.0, 1.0);
   for (auto& v : hx) v = U(rng);
   DBuf<double> dx, dy;
   dx.upload(hx, s);
   dy.alloc(n);
   for (int w = 0; w < 2; w++) problem_matvec(pr, precision, dx.p, dy.p);
   hipEvent_t e0, e1;
   FSVD_CHECK(hipEventCreate(&e0));
   FSVD_CHECK(hipEventCreate(&e1));
   FSVD_CHECK(hipEventRecord(e0, s));
   for (int r = 0; r < nrep; r++) problem_matvec(pr, precision, dx.p, dy.p);
   FSVD_CHECK(hipEventRecord(e1, s));
   FSVD_CHECK(hipEventSynchronize(e1));
   float ms = 0;
   FSVD_CHECK(hipEventElapsedTime(&ms, e0, e1));
   if (ms_total) *ms_total = ms;
   if (ms_parts) {
      // per class: sum over trees of one pass, averaged over nrep passes
      for (int part = 0; part < 4; part++) {
         FSVD_CHECK(hipEventRecord(e0, s));
         for (int r = 0; r < nrep; r++)
            for (Region& R : pr.regions) {
               if ((int)R.panel_ids.size() == 0) continue;
               TreeOp& op = R.shares_tree_with >= 0 ? *pr.regions[R.shares_tree_with].op : *R.op;
               tree_apply_part(op, precision, part, R.xs.p, R.xd.p, R.yt.p, s);
            }
         FSVD_CHECK(hipEventRecord(e1, s));
         FSVD_CHECK(hipEventSynchronize(e1));
         float m = 0;
         FSVD_CHECK(hipEventElapsedTime(&m, e0, e1));
         ms_parts[part] = m / nrep;
      }
   }
   (void)hipEventDestroy(e0);
   (void)hipEventDestroy(e1);
   API_CATCH
}

int fftsvd_problem_bench_matvec_batch(fftsvd_problem* p, int precision, int ncol, int nrep, double* ms_total,
                                      double* ms_parts) {
   API_TRY
   Problem& pr = *p->pr;
   if (pr.qual) throw std::runtime_error("the batched matvec takes surface operators");
   if (ncol < 1 || nrep < 1) throw std::runtime_error("ncol and nrep must be positive");
   hipStream_t s = pr.stream;
   const size_t n = (size_t)pr.nvec;
   std::vector<double> hx(n * ncol);
   std::mt19937_64 rng(12345);
   std::uniform_real_distribution<double> U(-1.0, 1.0);
   for (auto& v : hx) v = U(rng);
   DBuf<double> dx, dy;
   dx.upload(hx, s);
   dy.alloc(n * ncol);
   for (int w = 0; w < 2; w++) problem_matvec_batch(pr, precision, ncol, dx.p, dy.p);
   hipEvent_t e0, e1;
   FSVD_CHECK(hipEventCreate(&e0));
   FSVD_CHECK(hipEventCreate(&e1));
   FSVD_CHECK(hipEventRecord(e0, s));
   for (int r = 0; r < nrep; r++) problem_matvec_batch(pr, precision, ncol, dx.p, dy.p);
   FSVD_CHECK(hipEventRecord(e1, s));
   FSVD_CHECK(hipEventSynchronize(e1));
   float ms = 0;
   FSVD_CHECK(hipEventElapsedTime(&ms, e0, e1));
   if (ms_total) *ms_total = ms;
   if (ms_parts) {
      for (int part = 0; part < 4; part++) {
         FSVD_CHECK(hipEventRecord(e0, s));
         for (int r = 0; r < nrep; r++)
            for (Region& R : pr.regions) {
               const int np = (int)R.panel_ids.size();
               if (np == 0) continue;
               TreeOp& op = R.shares_tree_with >= 0 ? *pr.regions[R.shares_tree_with].op : *R.op;
               tree_apply_part_batch(op, precision, part, ncol, R.xsB.p, R.xdB.p, np, R.ytB.p, np, s);
            }
         FSVD_CHECK(hipEventRecord(e1, s));
         FSVD_CHECK(hipEventSynchronize(e1));
         float m = 0;
         FSVD_CHECK(hipEventElapsedTime(&m, e0, e1));
         ms_parts[part] = m / nrep;
      }
   }
   (void)hipEventDestroy(e0);
   (void)hipEventDestroy(e1);
   API_CATCH
}

int fftsvd_tree_info(fftsvd_problem* p, int which, int* depth, int* ncubes, int* npanels) {
   API_TRY
   TreeOp& op = region_op(*p->pr, which);
   if (depth) *depth = op.tree.depth;
   if (ncubes) *ncubes = (int)op.tree.cubes.size();
   if (npanels) *npanels = op.tree.npan;
   API_CATCH
}

int fftsvd_tree_cube_info(fftsvd_problem* p, int which, int cube, int* info) {
   API_TRY
   TreeOp& op = region_op(*p->pr, which);
   const Cube& c = op.tree.cubes.at(cube);
   int dcols = 0;
   if (op.cube_leaf[cube] >= 0) dcols = op.h_leaves[op.cube_leaf[cube]].m;
   int v[13] = {c.level, c.idx[0], c.idx[1], c.idx[2], c.leaf, c.np, c.nq, (int)c.loc.size(), (int)c.itr.size(),
                c.rowrank, c.colrank, c.parent, dcols};
   std::memcpy(info, v, sizeof(v));
   API_CATCH
}

int fftsvd_tree_cube_panels(fftsvd_problem* p, int which, int cube, int* pan, int* dcols) {
   API_TRY
   TreeOp& op = region_op(*p->pr, which);
   const Cube& c = op.tree.cubes.at(cube);
   if (pan)
      for (int k = 0; k < c.np; k++) pan[k] = op.tree.pan_perm[c.p0 + k];
   if (dcols && op.cube_leaf[cube] >= 0) {
      const LeafDesc& L = op.h_leaves[op.cube_leaf[cube]];
      for (int k = 0; k < L.m; k++) dcols[k] = op.tree.pan_perm[op.h_dcol[L.col_off + k]];
   }
   API_CATCH
}

int fftsvd_tree_cube_points(fftsvd_problem* p, int which, int cube, int* pt) {
   API_TRY
   TreeOp& op = region_op(*p->pr, which);
   const Cube& c = op.tree.cubes.at(cube);
   for (int k = 0; k < c.nq; k++) pt[k] = op.tree.pt_perm[c.q0 + k];
   API_CATCH
}

int fftsvd_tree_cube_mats(fftsvd_problem* p, int which, int cube, double* VT, double* U, double* PVs, double* PVd,
                          double* UTI, double* Ds, double* Dd) {
   API_TRY
   Problem& pr = *p->pr;
   TreeOp& op = region_op(pr, which);
   const Cube& c = op.tree.cubes.at(cube);
   hipStream_t s = pr.stream;
   auto get = [&](const double* src, size_t n) {
      std::vector<double> h(n);
      if (n) FSVD_CHECK(hipMemcpyAsync(h.data(), src, sizeof(double) * n, hipMemcpyDeviceToHost, s));
      FSVD_CHECK(hipStreamSynchronize(s));
      return h;
   };
   if (c.level >= 2) {
      FarLevel& lv = op.levels[c.level - 2];
      int gp3 = lv.gp3;
      if (op.cube_src_slot[cube] >= 0) {
         const SrcDesc& sd = lv.h_src[op.cube_src_slot[cube]];
         auto vt = get(lv.VT64.p + sd.vt_off, (size_t)sd.r * sd.np);
         auto pv = get(lv.PV64.p + sd.pv_off, 2 * (size_t)sd.r * gp3);
         if (VT) std::memcpy(VT, vt.data(), sizeof(double) * vt.size());
         for (int g = 0; g < gp3; g++)
            for (int k = 0; k < sd.r; k++) {
               if (PVs) PVs[(size_t)g * sd.r + k] = pv[(size_t)k * gp3 + g];
               if (PVd) PVd[(size_t)g * sd.r + k] = pv[(size_t)sd.r * gp3 + (size_t)k * gp3 + g];
            }
      }
      if (op.cube_tgt_slot[cube] >= 0) {
         const TgtDesc& td = lv.h_tgt[op.cube_tgt_slot[cube]];
         auto ut = get(lv.UT64.p + td.u_off, (size_t)td.r * td.nq);
         auto uti = get(lv.UTI64.p + td.uti_off, (size_t)td.r * gp3);
         if (U)
            for (int i = 0; i < td.nq; i++)
               for (int k = 0; k < td.r; k++) U[(size_t)i * td.r + k] = ut[(size_t)k * td.nq + i];
         if (UTI) std::memcpy(UTI, uti.data(), sizeof(double) * uti.size());
      }
   }
   if (op.cube_leaf[cube] >= 0) {
      const LeafDesc& L = op.h_leaves[op.cube_leaf[cube]];
      auto blk = get(op.D64.p + L.d_off, 2 * (size_t)L.p * L.m);
      for (int r = 0; r < L.p; r++)
         for (int k = 0; k < L.m; k++) {
            if (Ds) Ds[(size_t)r * L.m + k] = blk[(size_t)k * 2 * L.p + r];
            if (Dd) Dd[(size_t)r * L.m + k] = blk[(size_t)k * 2 * L.p + L.p + r];
         }
   }
   API_CATCH
}

int fftsvd_tree_apply(fftsvd_problem* p, int which, int precision, const double* xs, const double* xd, double* y) {
   API_TRY
   Problem& pr = *p->pr;
   TreeOp& op = region_op(pr, which);
   int n = op.tree.npan;
   std::vector<double> a(n), b(n), c(n);
   for (int k = 0; k < n; k++) {
      a[k] = xs[op.tree.pan_perm[k]];
      b[k] = xd[op.tree.pan_perm[k]];
   }
   DBuf<double> da, db, dc;
   da.upload(a, pr.stream);
   db.upload(b, pr.stream);
   dc.alloc(op.tree.npt);
   tree_apply(op, precision, da.p, db.p, dc.p, pr.stream);
   c.resize(op.tree.npt);
   FSVD_CHECK(hipMemcpyAsync(c.data(), dc.p, sizeof(double) * c.size(), hipMemcpyDeviceToHost, pr.stream));
   FSVD_CHECK(hipStreamSynchronize(pr.stream));
   for (int k = 0; k < op.tree.npt; k++) y[op.tree.pt_perm[k]] = c[k];
   API_CATCH
}

int fftsvd_tree_apply_transpose(fftsvd_problem* p, int which, int precision, int layer, const double* x, double* y) {
   API_TRY
   Problem& pr = *p->pr;
   TreeOp& op = region_op(pr, which);
   std::vector<double> a(op.tree.npt), c(op.tree.npan);
   for (int k = 0; k < op.tree.npt; k++) a[k] = x[op.tree.pt_perm[k]];
   DBuf<double> da, dc;
   da.upload(a, pr.stream);
   dc.alloc(op.tree.npan);
   tree_apply_transpose(op, precision, layer, da.p, dc.p, pr.stream);
   FSVD_CHECK(hipMemcpyAsync(c.data(), dc.p, sizeof(double) * c.size(), hipMemcpyDeviceToHost, pr.stream));
   FSVD_CHECK(hipStreamSynchronize(pr.stream));
   for (int k = 0; k < op.tree.npan; k++) y[op.tree.pan_perm[k]] = c[k];
   API_CATCH
}

int fftsvd_integrate(int kernel, double kappa, const double* verts9, int npan, const double* pts, int npt,
                     double* slp, double* dlp) {
   API_TRY
   if (kernel < 0 || kernel > 2) throw std::runtime_error("kernel must be 0 (Laplace), 1 (Yukawa) or 2 (d/dkappa)");
   std::vector<PanelGeom> g(npan);
   for (int i = 0; i < npan; i++) panel_geometry(verts9 + 9 * (size_t)i, verts9 + 9 * (size_t)i + 3, verts9 + 9 * (size_t)i + 6, &g[i]);
   hipStream_t s;
   FSVD_CHECK(hipStreamCreate(&s));
   {
      DBuf<PanelGeom> dg;
      DBuf<double> dp, ds, dd;
      dg.upload(g, s);
      dp.upload(pts, 3 * (size_t)npt, s);
      ds.alloc((size_t)npan * npt);
      dd.alloc((size_t)npan * npt);
      launch_integrate_all(kernel, kappa, dg.p, npan, dp.p, npt, ds.p, dd.p, s);
      FSVD_CHECK(hipMemcpyAsync(slp, ds.p, sizeof(double) * npan * npt, hipMemcpyDeviceToHost, s));
      FSVD_CHECK(hipMemcpyAsync(dlp, dd.p, sizeof(double) * npan * npt, hipMemcpyDeviceToHost, s));
      FSVD_CHECK(hipStreamSynchronize(s));
   }
   (void)hipStreamDestroy(s);
   API_CATCH
}

/* ------------------------------------------------------------------------ */
/* Reference solve-path API (FFTSVDpbeAPI.h)                                  */

void error(char* message, ...) {
   va_list ap;
   va_start(ap, message);
   printf("ERROR: ");
   vprintf(message, ap);
   printf("\n");
   va_end(ap);
   exit(-1);
}

void warning(char* message, ...) {
   va_list ap;
   va_start(ap, message);
   printf("WARNING: ");
   vprintf(message, ap);
   printf("\n");
   va_end(ap);
}

static Panel make_panel(const double* a, const double* b, const double* c) {
   PanelGeom g;
   panel_geometry(a, b, c, &g);
   Panel p = (Panel)calloc(1, sizeof(_Panel));
   p->centroid = (Vector3D)calloc(1, sizeof(_Vector3D));
   p->normal = (Vector3D)calloc(1, sizeof(_Vector3D));
   p->centroid->x = g.c[0];
   p->centroid->y = g.c[1];
   p->centroid->z = g.c[2];
   p->normal->x = g.ax[6];
   p->normal->y = g.ax[7];
   p->normal->z = g.ax[8];
   p->area = g.area;
   p->maxedgelength = g.maxd;
   std::memcpy(p->vertex[0], a, 3 * sizeof(double));
   std::memcpy(p->vertex[1], b, 3 * sizeof(double));
   std::memcpy(p->vertex[2], c, 3 * sizeof(double));
   return p;
}

// .vert/.face reader (reference getFlatPanels + VertFace_*, FFTSVDpbeAPI.c:676-712, VertFace.c:26-202)
static void get_flat_panels(const char* base, Panel** panels, unsigned int* numpanels, int accessible, int cavity) {
   char vn[1100], fn[1100];
   std::snprintf(vn, sizeof(vn), "%s.vert", base);
   std::snprintf(fn, sizeof(fn), "%s.face", base);
   FILE* fv = fopen(vn, "r");
   if (!fv) error((char*)"Could not open vert file %s: %s", vn, strerror(errno));
   FILE* ff = fopen(fn, "r");
   if (!ff) error((char*)"Could not open face file %s: %s", fn, strerror(errno));
   char line[128];
   std::vector<double> V;
   while (fgets(line, 128, fv)) {
      double x = 0, y = 0, z = 0;
      sscanf(line, "%lf %lf %lf", &x, &y, &z);
      V.push_back(x);
      V.push_back(y);
      V.push_back(z);
   }
   std::vector<unsigned> F;
   const bool flip = !(cavity && !usequalocation);
   while (fgets(line, 128, ff)) {
      unsigned a = 0, b = 0, c = 0, g = 0;
      sscanf(line, "%u %u %u %u", &a, &b, &c, &g);
      F.push_back(a - 1);
      F.push_back(flip ? c - 1 : b - 1);
      F.push_back(flip ? b - 1 : c - 1);
      F.push_back(g);
   }
   fclose(fv);
   fclose(ff);
   std::vector<Panel> out;
   size_t nv = V.size() / 3;
   for (size_t f = 0; f < F.size() / 4; f++) {
      unsigned ia = F[4 * f], ib = F[4 * f + 1], ic = F[4 * f + 2];
      if (ia >= nv || ib >= nv || ic >= nv) error((char*)"face %u references a missing vertex", (unsigned)f + 1);
      const double *A = &V[3 * (size_t)ia], *B = &V[3 * (size_t)ib], *C = &V[3 * (size_t)ic];
      auto dist = [](const double* u, const double* v) {
         return std::sqrt((u[0] - v[0]) * (u[0] - v[0]) + (u[1] - v[1]) * (u[1] - v[1]) + (u[2] - v[2]) * (u[2] - v[2]));
      };
      double d1 = dist(A, B), d2 = dist(B, C), d3 = dist(C, A);
      double sp = 0.5 * (d1 + d2 + d3);
      double area = std::sqrt(sp * (sp - d1) * (sp - d2) * (sp - d3));
      if (!(area > 0.0)) continue;
      if (accessible && F[4 * f + 3] != 1) continue;
      out.push_back(make_panel(A, B, C));
   }
   *numpanels = (unsigned)out.size();
   *panels = (Panel*)calloc(out.size() + 1, sizeof(Panel));
   for (size_t k = 0; k < out.size(); k++) (*panels)[k] = out[k];
}

static std::vector<std::string> split_names(const char* line) {
   std::vector<std::string> v;
   std::string cur;
   for (const char* p = line; *p; p++) {
      if (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t') {
         if (!cur.empty()) v.push_back(cur);
         cur.clear();
      } else {
         cur.push_back(*p);
      }
   }
   if (!cur.empty()) v.push_back(cur);
   return v;
}

void readSRF(const char* filename, Panel*** saltpanels, unsigned int** numsaltpanels, unsigned int* numsalts,
             Panel*** dielectricpanels, unsigned int** numdielectricpanels, unsigned int* numdielectrics,
             unsigned int** dielectricparent, Panel*** dielectriccavitypanels,
             unsigned int** numdielectriccavitypanels, unsigned int* numdielectriccavities,
             unsigned int** dielectriccavityparent, Panel*** saltcavitypanels, unsigned int** numsaltcavitypanels,
             unsigned int* numsaltcavities, unsigned int** saltcavityparent, unsigned int* numtotalpanels) {
   FILE* f = fopen(filename, "r");
   if (!f) error((char*)"Could not open surface file %s: %s", filename, strerror(errno));
   std::string dir(filename);
   size_t slash = dir.rfind('/');
   dir = (slash == std::string::npos) ? std::string() : dir.substr(0, slash + 1);
   char line[1024];
   auto next_line = [&]() -> std::string {
      if (!fgets(line, sizeof(line), f)) return std::string();
      return std::string(line);
   };
   std::string st = next_line(), dt = next_line();
   char salttype = st.empty() ? 'f' : st[0], dieltype = dt.empty() ? 'f' : dt[0];
   if (salttype != 'f' || dieltype != 'f') error((char*)"curved panels are not supported (flat 'f' surfaces only)");
   *numtotalpanels = 0;
   auto load = [&](const std::string& names, Panel*** panels, unsigned int** counts, unsigned int* num, int accessible,
                   int cavity) {
      auto v = split_names(names.c_str());
      *num = (unsigned)v.size();
      if (v.empty()) return;
      *panels = (Panel**)calloc(v.size(), sizeof(Panel*));
      *counts = (unsigned int*)calloc(v.size(), sizeof(unsigned int));
      for (size_t i = 0; i < v.size(); i++) {
         std::string path = v[i][0] == '/' ? v[i] : dir + v[i];
         get_flat_panels(path.c_str(), &(*panels)[i], &(*counts)[i], accessible, cavity);
         *numtotalpanels += (*counts)[i];
      }
   };
   load(next_line(), saltpanels, numsaltpanels, numsalts, 1, 0);
   load(next_line(), dielectricpanels, numdielectricpanels, numdielectrics, 0, 0);
   load(next_line(), dielectriccavitypanels, numdielectriccavitypanels, numdielectriccavities, 0, 1);
   load(next_line(), saltcavitypanels, numsaltcavitypanels, numsaltcavities, 1, 1);
   auto parents = [&](unsigned n, unsigned int** out, const unsigned int* counts, bool drop999) {
      std::string l = next_line();
      auto v = split_names(l.c_str());
      if (n == 0) return;
      *out = (unsigned int*)calloc(n, sizeof(unsigned int));
      for (unsigned i = 0; i < n; i++) {
         (*out)[i] = i < v.size() ? (unsigned)atoi(v[i].c_str()) : 0;
         if (drop999 && (*out)[i] == 999) *numtotalpanels -= counts[i];
      }
   };
   parents(*numdielectrics, dielectricparent, *numdielectricpanels, false);
   parents(*numdielectriccavities, dielectriccavityparent, *numdielectriccavities ? *numdielectriccavitypanels : nullptr,
           true);
   parents(*numsaltcavities, saltcavityparent, *numsaltcavities ? *numsaltcavitypanels : nullptr, true);
   fclose(f);
}

static SurfaceOperator new_so() { return (SurfaceOperator)calloc(1, sizeof(_SurfaceOperator)); }

void generateSurfaceOperator(SurfaceOperator* pbesurfaceoperator, PDBentry* PDBentries, unsigned int numPDBentries,
                             Panel** saltpanels, unsigned int* numsaltpanels, unsigned int numsalts,
                             Panel** dielectricpanels, unsigned int* numdielectricpanels, unsigned int numdielectrics,
                             unsigned int* dielectricparent, Panel** dielectriccavitypanels,
                             unsigned int* numdielectriccavitypanels, unsigned int numdielectriccavities,
                             unsigned int* dielectriccavityparent, Panel** saltcavitypanels,
                             unsigned int* numsaltcavitypanels, unsigned int numsaltcavities,
                             unsigned int* saltcavityparent, unsigned int numtotalpanels) {
   const double kappa = std::sqrt(ionicstrength) / DEBYE_CONSTANT;
   const bool stern = ionicstrength > 0.0 && ionexclusionradius > 1e-6;
   SurfaceSpec sp;
   auto panel_verts = [](Panel* ps, unsigned np) {
      std::vector<double> v;
      for (unsigned k = 0; k < np; k++)
         for (int a = 0; a < 3; a++)
            for (int d = 0; d < 3; d++) v.push_back(ps[k]->vertex[a][d]);
      return v;
   };
   if (stern) {
      // Stern layer (FFTSVDpbeAPI.c:1064-1377): salt surfaces enclose the
      // dielectrics; salt cavities sit inside dielectric cavities
      if (numsalts == 0) error((char*)"the Stern-layer formulation needs a salt surface");
      for (unsigned i = 0; i < numsalts; i++) sp.salt_verts.push_back(panel_verts(saltpanels[i], numsaltpanels[i]));
      for (unsigned j = 0; j < numdielectrics; j++) {
         if (dielectricparent[j] >= numsalts) error((char*)"dielectric %u has no enclosing salt surface", j);
         sp.dielectric_parent.push_back((int)dielectricparent[j]);
      }
   }
   for (unsigned i = 0; i < numdielectrics; i++) {
      std::vector<double> v;
      for (unsigned k = 0; k < numdielectricpanels[i]; k++)
         for (int a = 0; a < 3; a++)
            for (int d = 0; d < 3; d++) v.push_back(dielectricpanels[i][k]->vertex[a][d]);
      sp.dielectric_verts.push_back(v);
   }
   std::vector<unsigned> cav_used;
   std::vector<int> cav_index(numdielectriccavities, -1);
   for (unsigned j = 0; j < numdielectriccavities; j++) {
      if (dielectriccavityparent[j] >= numdielectrics) continue;  // 999: removed by the mesher
      cav_index[j] = (int)sp.cavity_verts.size();
      sp.cavity_verts.push_back(panel_verts(dielectriccavitypanels[j], numdielectriccavitypanels[j]));
      sp.cavity_parent.push_back((int)dielectriccavityparent[j]);
      cav_used.push_back(j);
   }
   if (stern)
      for (unsigned l = 0; l < numsaltcavities; l++) {
         if (saltcavityparent[l] >= numdielectriccavities || cav_index[saltcavityparent[l]] < 0) continue;
         sp.salt_cavity_verts.push_back(panel_verts(saltcavitypanels[l], numsaltcavitypanels[l]));
         sp.salt_cavity_parent.push_back(cav_index[saltcavityparent[l]]);
      }
   for (unsigned a = 0; a < numPDBentries; a++) {
      sp.qpos.push_back(PDBentries[a].x);
      sp.qpos.push_back(PDBentries[a].y);
      sp.qpos.push_back(PDBentries[a].z);
      sp.q.push_back(toupper(PDBentries[a].chain) != 'X' ? PDBentries[a].charge : 0.0);
   }
   sp.eps_in = innerdielectric;
   sp.eps_out = outerdielectric;
   sp.kappa = ionicstrength > 0.0 ? kappa : 0.0;
   sp.maxpanels = MAX_PANELS_PER_FINEST_CUBE;
   sp.gridpoints = GRID_SIZE;
   sp.svdeps = SVD_ERROR;
   sp.want32 = true;
   sp.m3 = accelerateM3 != 0;
   std::unique_ptr<Problem> pr;
   try {
      pr = build_problem(sp);
   } catch (const std::exception& e) {
      error((char*)"%s", e.what());
   }
   if ((unsigned)2 * pr->ntotal != 2 * numtotalpanels)
      error((char*)"SurfaceOperator indexing is WRONG! (%u != %u)", 2 * pr->ntotal, 2 * numtotalpanels);
   // host mirror of the region tree for callers that walk it
   auto* h = new fftsvd_problem;
   h->pr = std::move(pr);
   Problem& P = *h->pr;
   std::vector<SurfaceOperator> so(P.regions.size());
   for (size_t r = 0; r < P.regions.size(); r++) so[r] = new_so();
   for (size_t r = 0; r < P.regions.size(); r++) {
      Region& R = P.regions[r];
      SurfaceOperator o = so[r];
      o->kernel = R.kernel == YUKAWA ? HELMHOLTZ_KERNEL : POISSON_KERNEL;
      o->epsilon = R.eps;
      o->kappa = R.kappa;
      o->mynumpanels = R.mynpan;
      o->startphi = R.startphi;
      o->startdphi = R.startdphi;
      o->tree = h;
      o->M3 = R.m3 ? (void*)h : nullptr;
      o->parent = R.parent >= 0 ? so[R.parent] : nullptr;
      o->numchildren = (unsigned)R.child.size();
      o->children = (SurfaceOperator*)calloc(R.child.size() + 1, sizeof(SurfaceOperator));
      for (size_t c = 0; c < R.child.size(); c++) o->children[c] = so[R.child[c]];
      o->resultInternal = (Vector)calloc(R.panel_ids.size() + 1, sizeof(real));
      o->resultExternal = (Vector)calloc(R.panel_ids.size() + 1, sizeof(real));
      if (R.has_charges) {
         Charge ch = (Charge)calloc(1, sizeof(_Charge));
         ch->numcharges = numPDBentries;
         ch->globalindexstart = 0;
         ch->points = (Vector3D*)calloc(numPDBentries + 1, sizeof(Vector3D));
         ch->charges = (Vector)calloc(numPDBentries + 1, sizeof(real));
         for (unsigned a = 0; a < numPDBentries; a++) {
            ch->points[a] = (Vector3D)calloc(1, sizeof(_Vector3D));
            ch->points[a]->x = sp.qpos[3 * a];
            ch->points[a]->y = sp.qpos[3 * a + 1];
            ch->points[a]->z = sp.qpos[3 * a + 2];
            ch->charges[a] = sp.q[a];
         }
         o->charges = ch;
      }
   }
   *pbesurfaceoperator = so[0];
}

static Problem& so_problem(SurfaceOperator so) {
   if (!so || !so->tree) error((char*)"surface operator was not generated by this library");
   return *((fftsvd_problem*)so->tree)->pr;
}

void generateSurfaceOperatorPreconditioner(Preconditioner* preconditioner, SurfaceOperator so,
                                           unsigned int numtotalpanels) {
   Problem& pr = so_problem(so);
   Preconditioner P = (Preconditioner)calloc(1, sizeof(_Preconditioner));
   P->rows = P->columns = 2 * numtotalpanels;
   P->nnz = 4 * pr.nblocks;
   P->device = so->tree;
   *preconditioner = P;
}

void generateRHS(Vector* rhs, SurfaceOperator so, unsigned int numtotalpanels) {
   Problem& pr = so_problem(so);
   *rhs = (Vector)calloc(2 * (size_t)numtotalpanels, sizeof(real));
   if (fftsvd_problem_rhs((fftsvd_problem*)so->tree, *rhs)) error((char*)"%s", g_err.c_str());
   (void)pr;
}

void solveSurfaceOperator(Vector* solution, SurfaceOperator so, Preconditioner preconditioner, Vector rhs,
                          unsigned int numtotalpanels) {
   (void)preconditioner;
   so_problem(so);
   *solution = (Vector)calloc(2 * (size_t)numtotalpanels, sizeof(real));
   if (fftsvd_problem_solve((fftsvd_problem*)so->tree, g_precision, rhs, *solution, errortolerance, 1000, nullptr,
                            nullptr, nullptr))
      error((char*)"%s", g_err.c_str());
}

// potentials at the collect entries in kT/e, plus the Coulomb potential of
// the input charges when addcoulombic (collectPotentials,
// FFTSVDpbeAPI.c:1896-1922; collectQualocationPotentials, :1924-1950)
static void collect_into(PDBentry* PDBcollectentries, unsigned int numPDBcollectentries, PDBentry* PDBinputentries,
                         unsigned int numPDBinputentries, Problem& pr, Vector solution) {
   std::vector<double> phi(pr.q.size() + 1, 0.0);
   problem_potentials(pr, solution, phi.data());
   for (unsigned a = 0; a < numPDBcollectentries; a++)
      PDBcollectentries[a].potential = KT_CONVERSION * (a < pr.q.size() ? phi[a] : 0.0) / innerdielectric;
   if (addcoulombic) {
      for (unsigned c = 0; c < numPDBcollectentries; c++)
         for (unsigned a = 0; a < numPDBinputentries; a++)
            if (toupper(PDBinputentries[a].chain) != 'X') {
               double dx = PDBinputentries[a].x - PDBcollectentries[c].x;
               double dy = PDBinputentries[a].y - PDBcollectentries[c].y;
               double dz = PDBinputentries[a].z - PDBcollectentries[c].z;
               double distance = std::sqrt(dx * dx + dy * dy + dz * dz);
               if (distance > 0.0)
                  PDBcollectentries[c].potential +=
                      KT_CONVERSION * PDBinputentries[a].charge / (innerdielectric * distance);
            }
   }
}

void collectPotentials(PDBentry* PDBcollectentries, unsigned int numPDBcollectentries, PDBentry* PDBinputentries,
                       unsigned int numPDBinputentries, SurfaceOperator so, Vector solution) {
   collect_into(PDBcollectentries, numPDBcollectentries, PDBinputentries, numPDBinputentries, so_problem(so),
                solution);
}

// ---- qualocation formulation (FFTSVDpbeAPI.c:1741-1869,1878-1894,1924-1950) ----

void generateQualocationOperatorMinimal(QualocationOperator* qualocationoperator, PDBentry* PDBentries,
                                        unsigned int numPDBentries, Panel** dielectricpanels,
                                        unsigned int* numdielectricpanels, unsigned int numdielectrics,
                                        Panel** dielectriccavitypanels, unsigned int* numdielectriccavitypanels,
                                        unsigned int numdielectriccavities, unsigned int* dielectriccavityparent,
                                        unsigned int numtotalpanels) {
   if (ionicstrength != 0.0) error((char*)"Qualocation method only works without salt!");
   QualSpec sp;
   // all dielectric panels, then the cavities' (parent 999: removed)
   for (unsigned i = 0; i < numdielectrics; i++)
      for (unsigned k = 0; k < numdielectricpanels[i]; k++)
         for (int a = 0; a < 3; a++)
            for (int d = 0; d < 3; d++) sp.verts.push_back(dielectricpanels[i][k]->vertex[a][d]);
   for (unsigned i = 0; i < numdielectriccavities; i++) {
      if (dielectriccavityparent[i] == 999) continue;
      for (unsigned k = 0; k < numdielectriccavitypanels[i]; k++)
         for (int a = 0; a < 3; a++)
            for (int d = 0; d < 3; d++) sp.verts.push_back(dielectriccavitypanels[i][k]->vertex[a][d]);
   }
   if (sp.verts.size() / 9 != numtotalpanels)
      error((char*)"qualocation panels (%u) differ from numtotalpanels (%u)", (unsigned)(sp.verts.size() / 9),
            numtotalpanels);
   for (unsigned a = 0; a < numPDBentries; a++) {
      sp.qpos.push_back(PDBentries[a].x);
      sp.qpos.push_back(PDBentries[a].y);
      sp.qpos.push_back(PDBentries[a].z);
      sp.q.push_back(toupper(PDBentries[a].chain) != 'X' ? PDBentries[a].charge : 0.0);
   }
   sp.eps_in = innerdielectric;
   sp.eps_out = outerdielectric;
   sp.maxpanels = MAX_PANELS_PER_FINEST_CUBE;
   sp.gridpoints = GRID_SIZE;
   sp.svdeps = SVD_ERROR;
   sp.want32 = true;
   sp.m3 = accelerateM3 != 0;   // generateQualocationOperatorA1A3 (FFTSVDpbeAPI.c:1836-1838)
   auto* h = new fftsvd_problem;
   try {
      h->pr = build_qual_problem(sp);
   } catch (const std::exception& e) {
      delete h;
      error((char*)"%s", e.what());
   }
   QualocationOperator o = (QualocationOperator)calloc(1, sizeof(_QualocationOperator));
   o->tree = h;
   o->M1M3 = h->pr->regions[0].m3 ? (void*)h : nullptr;
   o->innerdielectric = innerdielectric;
   o->outerdielectric = outerdielectric;
   Charge ch = (Charge)calloc(1, sizeof(_Charge));
   ch->numcharges = numPDBentries;
   ch->points = (Vector3D*)calloc(numPDBentries + 1, sizeof(Vector3D));
   ch->charges = (Vector)calloc(numPDBentries + 1, sizeof(real));
   for (unsigned a = 0; a < numPDBentries; a++) {
      ch->points[a] = (Vector3D)calloc(1, sizeof(_Vector3D));
      ch->points[a]->x = sp.qpos[3 * a];
      ch->points[a]->y = sp.qpos[3 * a + 1];
      ch->points[a]->z = sp.qpos[3 * a + 2];
      ch->charges[a] = sp.q[a];
   }
   o->charges = ch;
   *qualocationoperator = o;
}

void generateQualocationOperator(QualocationOperator* qualocationoperator, PDBentry* PDBentries,
                                 unsigned int numPDBentries, Panel** dielectricpanels,
                                 unsigned int* numdielectricpanels, unsigned int numdielectrics,
                                 Panel** dielectriccavitypanels, unsigned int* numdielectriccavitypanels,
                                 unsigned int numdielectriccavities, unsigned int* dielectriccavityparent,
                                 unsigned int numtotalpanels) {
   generateQualocationOperatorMinimal(qualocationoperator, PDBentries, numPDBentries, dielectricpanels,
                                      numdielectricpanels, numdielectrics, dielectriccavitypanels,
                                      numdielectriccavitypanels, numdielectriccavities, dielectriccavityparent,
                                      numtotalpanels);
}

static Problem& qo_problem(QualocationOperator qo) {
   if (!qo || !qo->tree) error((char*)"qualocation operator was not generated by this library");
   return *((fftsvd_problem*)qo->tree)->pr;
}

void generateQualocationOperatorPreconditioner(Preconditioner* preconditioner, QualocationOperator qo,
                                               unsigned int numtotalpanels) {
   qo_problem(qo);
   Preconditioner P = (Preconditioner)calloc(1, sizeof(_Preconditioner));
   P->rows = P->columns = numtotalpanels;
   P->nnz = numtotalpanels;
   P->device = qo->tree;
   *preconditioner = P;
}

void generateQualocationRHS(Vector* rhs, QualocationOperator qo, unsigned int numtotalpanels) {
   qo_problem(qo);
   *rhs = (Vector)calloc((size_t)numtotalpanels + 1, sizeof(real));
   if (fftsvd_problem_rhs((fftsvd_problem*)qo->tree, *rhs)) error((char*)"%s", g_err.c_str());
}

void solveQualocationOperator(Vector* solution, QualocationOperator qo, Preconditioner preconditioner, Vector rhs,
                              unsigned int numtotalpanels) {
   (void)preconditioner;
   qo_problem(qo);
   *solution = (Vector)calloc((size_t)numtotalpanels + 1, sizeof(real));
   if (fftsvd_problem_solve((fftsvd_problem*)qo->tree, g_precision, rhs, *solution, errortolerance, 1000, nullptr,
                            nullptr, nullptr))
      error((char*)"%s", g_err.c_str());
}

void collectQualocationPotentials(PDBentry* PDBcollectentries, unsigned int numPDBcollectentries,
                                  PDBentry* PDBinputentries, unsigned int numPDBinputentries, QualocationOperator qo,
                                  Vector solution) {
   collect_into(PDBcollectentries, numPDBcollectentries, PDBinputentries, numPDBinputentries, qo_problem(qo),
                solution);
}

void solveBEMproblem(unsigned int* numpanels, real* paneldata, real* innerdiel, real* outerdiel, real* rhs,
                     real* sol) {
   const unsigned n = *numpanels;
   QualSpec sp;
   sp.verts.assign(paneldata, paneldata + 9 * (size_t)n);
   sp.eps_in = *innerdiel;
   sp.eps_out = *outerdiel;
   sp.maxpanels = 27;
   sp.gridpoints = 3;
   sp.svdeps = 1e-5;
   sp.want32 = g_precision != PREC_F64;
   fftsvd_problem h;
   try {
      h.pr = build_qual_problem(sp);
   } catch (const std::exception& e) {
      error((char*)"%s", e.what());
   }
   h.pr->gmres_zero_test = false;   // GMRES_solv_ecf_qual has no zero-RHS test
   if (fftsvd_problem_solve(&h, g_precision, rhs, sol, 1e-5, 1000, nullptr, nullptr, nullptr))
      error((char*)"%s", g_err.c_str());
}

}  // extern "C"
