// Construction and application of one compressed FFT-SVD operator on the GPU.
//
// Construction (reference Tree_fill, Tree.c:513-565) runs as batched
// device work: near-field integral blocks, sampled interaction matrices,
// pivoted-MGS bases, equivalent-density projections and the small GEMMs
// that fold them into PV / UTI.  The host decides the tree, the sampling
// pattern and the per-level pseudoinverses and translation spectra.
#include <algorithm>
#include <atomic>
#include <climits>
#include <mutex>
#include <exception>
#include <thread>
#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <cstdlib>
#include <functional>
#include <string>
#include <unordered_map>
#include <iterator>
#include <stdexcept>

#include "host_linalg.h"
#include "op.h"

namespace fsvd {

void launch_batched_copy(int n, const void* descs, hipStream_t s);

namespace {

const double SPHERE49[49][3] = {
#include "sphere49.inc"
};
constexpr int NQ = 49;

struct CopyDesc {
   const double* src;
   double* dst;
   long long n;
};

inline int offset_code(const int* a, const int* b) {
   return (a[0] - b[0] + 3) * 49 + (a[1] - b[1] + 3) * 7 + (a[2] - b[2] + 3);
}

// FFTSVD_TIMING=1 prints the host wall time of setup sub-steps to stderr.
static bool timing_on() {
   static const bool on = [] {
      const char* e = std::getenv("FFTSVD_TIMING");
      return e && e[0] == '1';
   }();
   return on;
}
static void timing_mark(const char* what, double& t) {
   const double n = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
   // duration of the step and its end on the process clock
   if (timing_on()) std::fprintf(stderr, "[fftsvd setup] %-28s %8.2f ms  (ends %.2f)\n", what, n - t, n);
   t = n;
}

double now_ms() {
   return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Splits an IntBlock along its larger dimension so no workgroup gets more
// than `cap` integrals.
void push_blocks(std::vector<IntBlock>& out, IntBlock b, int cap = 8192) {
   if ((long long)b.npt * b.npan <= cap || (b.npt <= 1 && b.npan <= 1)) {
      if (b.npt > 0 && b.npan > 0) out.push_back(b);
      return;
   }
   if (b.npt >= b.npan) {
      int step = std::max(1, cap / std::max(1, b.npan));
      for (int i0 = 0; i0 < b.npt; i0 += step) {
         IntBlock c = b;
         c.npt = std::min(step, b.npt - i0);
         c.pt_off = b.pt_off + i0;
         c.out_off = b.out_off + (long long)i0 * b.ld_pt;
         out.push_back(c);
      }
   } else {
      int step = std::max(1, cap / std::max(1, b.npt));
      for (int j0 = 0; j0 < b.npan; j0 += step) {
         IntBlock c = b;
         c.npan = std::min(step, b.npan - j0);
         c.pan_off = b.pan_off + j0;
         c.out_off = b.out_off + (long long)j0 * b.ld_pan;
         out.push_back(c);
      }
   }
}

void run_integrals(const std::vector<IntBlock>& blocks, const std::vector<double>& bpts, const std::vector<int>& bpan,
                   const PanelGeom* d_panels, int kernel, double kappa, double* out, hipStream_t s) {
   if (blocks.empty()) return;
   DBuf<IntBlock> db;
   DBuf<double> dp;
   DBuf<int> dn;
   db.upload(blocks, s);
   dp.upload(bpts, s);
   dn.upload(bpan, s);
   launch_integral_blocks((int)blocks.size(), db.p, dp.p, dn.p, d_panels, kernel, kappa, out, s);
   FSVD_CHECK(hipStreamSynchronize(s));
}

void run_gemms(const std::vector<GemmDesc>& g, hipStream_t s) {
   if (g.empty()) return;
   DBuf<GemmDesc> dg;
   dg.upload(g, s);
   int maxtiles = 0;
   for (auto& d : g) maxtiles = std::max(maxtiles, ((d.M + 31) / 32) * ((d.N + 31) / 32));
   int tiles = std::min(256, std::max(1, maxtiles));
   // batch dimension limited to 65535 per launch
   for (size_t b0 = 0; b0 < g.size(); b0 += 65535) {
      int nb = (int)std::min<size_t>(65535, g.size() - b0);
      launch_gemm_batched(nb, tiles, dg.p + b0, s);
   }
   FSVD_CHECK(hipStreamSynchronize(s));
}

void to_f32(const DBuf<double>& a, DBuf<float>& b, hipStream_t s) {
   b.alloc(a.n);
   launch_f64_to_f32(a.p, b.p, (long long)a.n, s);
}

}  // namespace

std::vector<int> xcd_block_map(const std::vector<double>& work) {
   const int n = (int)work.size();
   double total = 0;
   for (double w : work) total += w;
   std::vector<std::vector<int>> part(8);
   double acc = 0;
   for (int i = 0; i < n; i++) {
      int x = total > 0 ? (int)(8.0 * (acc + 0.5 * work[i]) / total) : (int)(8LL * i / std::max(1, n));
      x = std::min(7, std::max(0, x));
      part[x].push_back(i);
      acc += work[i];
   }
   size_t len = 0;
   for (auto& p : part) len = std::max(len, p.size());
   std::vector<int> map(8 * len, -1);
   for (int x = 0; x < 8; x++)
      for (size_t k = 0; k < part[x].size(); k++) map[8 * k + x] = part[x][k];
   return map;
}


// The slot of an absolute spectrum offset of level L: (so - sbase) / pst,
// an exact division, done as a shift by pst's trailing zeros and a multiply
// by the inverse of its odd part modulo 2^32 (the planners take it for every
// interaction-list entry; a hardware divide per entry dominated them).
struct SlotOf {
   long long sbase;
   int shift;
   uint32_t inv;
   explicit SlotOf(const TrLevel& L) : sbase(L.sbase), shift(0), inv(1) {
      uint32_t odd = (uint32_t)L.pst;
      if (odd == 0) throw std::runtime_error("spectrum stride is zero");
      while (!(odd & 1u)) {
         odd >>= 1;
         shift++;
      }
      uint32_t x = odd;   // Newton's iteration doubles the correct low bits: 3, 6, 12, 24, 48
      for (int k = 0; k < 4; k++) x *= 2u - odd * x;
      inv = x;
   }
   int operator()(int so) const { return (int)(((uint32_t)((long long)so - sbase) >> shift) * inv); }
};

// Tiles of consecutive targets of one level: a tile grows while it has
// fewer than max_tile targets and the union of their sources stays within
// budget; a tile whose union exceeds max_src is halved until it fits.
// emit(k0, k1, union) receives each tile (targets tgt[t0 + k0 .. t0 + k1)
// and the absolute spectrum offsets of its sources in order of first
// appearance).
template <typename F>
static void make_tiles(const TrLevel& L, const std::vector<TgtPlan>& tgt, const std::vector<int>& isrc, int max_tile,
                       int budget, int max_src, F&& emit, int quantum = 1) {
   std::vector<int> stamp(std::max(L.nsrc, 1), -1);
   const SlotOf slot_of(L);
   int epoch = 0;
   std::function<void(int, int)> make = [&](int k0, int k1) {
      const int mark = -2 - (epoch++);
      std::vector<int> so;
      for (int k = k0; k < k1; k++) {
         const TgtPlan& tp = tgt[L.t0 + k];
         for (int e = 0; e < tp.nint; e++) {
            const int sl = slot_of(isrc[tp.int_off + e]);
            if (stamp[sl] != mark) {
               stamp[sl] = mark;
               so.push_back(isrc[tp.int_off + e]);
            }
         }
      }
      if ((int)so.size() > max_src) {
         if (k1 - k0 < 2) throw std::runtime_error("translation tile exceeds its LDS bound");
         const int km = (k0 + k1) / 2;
         make(k0, km);
         make(km, k1);
         return;
      }
      emit(k0, k1, so);
   };
   for (int k0 = 0; k0 < L.nt;) {
      int k1 = k0, un = 0;
      while (k1 < L.nt && k1 - k0 < max_tile) {
         const TgtPlan& tp = tgt[L.t0 + k1];
         int add = 0;
         for (int e = 0; e < tp.nint; e++) add += stamp[slot_of(isrc[tp.int_off + e])] != k0;   // lists hold distinct sources
         if (k1 > k0 && un + add > budget) break;
         for (int e = 0; e < tp.nint; e++) stamp[slot_of(isrc[tp.int_off + e])] = k0;
         un += add;
         k1++;
      }
      // a tile that stops short of the level's end is cut to a multiple of
      // the quantum when it holds more than one
      if (k1 < L.nt && k1 - k0 > quantum) k1 = k0 + (k1 - k0) / quantum * quantum;
      make(k0, k1);
      k0 = k1;
   }
}

// FP64 translation (k_translate<double>): tiles of up to TR_TILE targets, one
// task per tile and 4-bin chunk; each target's words hold the two LDS row
// byte offsets (T code * TR_ROW | local source * TR_ROW << 16) in
// interaction-list order, from a 16-byte aligned start, and every list is
// followed by at least four readable words.
static void build_translate_f64(ApplyPlan& P, const std::vector<TrLevel>& lvls, const std::vector<TgtPlan>& tgt,
                                const std::vector<int>& isrc, const std::vector<int>& itoff, hipStream_t s) {
   std::vector<TrTile> tiles;
   std::vector<GTgt> gtgt;
   std::vector<int> gsrc, gint;
   std::vector<int4> td;
   for (const TrLevel& L : lvls) {
      const int first = (int)tiles.size();
      // a tile's local index of each of its sources, by the source's slot
      std::vector<int> loc(std::max(L.nsrc, 1), -1);
      const SlotOf slot_of(L);
      make_tiles(L, tgt, isrc, TR_TILE, TR_SRC_BUDGET, TR_MAX_SRC, [&](int k0, int k1, const std::vector<int>& so) {
         for (size_t i = 0; i < so.size(); i++) loc[slot_of(so[i])] = (int)i;
         TrTile t;
         t.nt = k1 - k0;
         t.tgt0 = (int)gtgt.size();
         t.nsrc = (int)so.size();
         t.src_off = (int)gsrc.size();
         t.pad = L.pad;
         t.pst = L.pst;
         t.t_off = L.t_off;
         gsrc.insert(gsrc.end(), so.begin(), so.end());
         P.tr_max_src = std::max(P.tr_max_src, t.nsrc);
         for (int k = k0; k < k1; k++) {
            const TgtPlan& tp = tgt[L.t0 + k];
            GTgt g;
            g.acc_off = tp.acc_off;
            while (gint.size() % 4) gint.push_back(0);
            g.int_off = (int)gint.size();
            g.nint = tp.nint;
            for (int e = 0; e < tp.nint; e++)
               gint.push_back((int)((unsigned)(itoff[tp.int_off + e] * TR_ROW) |
                                    ((unsigned)(loc[slot_of(isrc[tp.int_off + e])] * TR_ROW) << 16)));
            gtgt.push_back(g);
         }
         tiles.push_back(t);
      });
      for (int ti = first; ti < (int)tiles.size(); ti++)
         for (int b0 = 0; b0 < L.pad; b0 += TR_BINS_F64) td.push_back(make_int4(ti, 1, b0, 0));
   }
   for (int k = 0; k < 8; k++) gint.push_back(0);   // read-ahead past the last list
   P.ntask_d = (int)td.size();
   P.task_d.upload(td, s);
   P.tiles.upload(tiles, s);
   P.gtgt.upload(gtgt, s);
   P.gsrc.upload(gsrc, s);
   P.gint.upload(gint, s);
   FSVD_CHECK(hipStreamSynchronize(s));
}

// Words of an octet for k_translate_pipe: the interaction lists of up to
// eight targets stepped together (target t's word at step k read by the
// lanes 8t..8t+7), nk4 steps (the longest list rounded up to 8), stored in
// groups of four steps: target t's word of step k at 32 (k / 4) + 4 t + k % 4.
// A half-wave (targets 0-3 or 4-7) reads four 64-byte T rows and four
// source rows per step, each row one quarter of the 256-byte LDS bank row;
// each list is reordered so that at every step the half-wave's four T
// rows, and separately its four source rows, fall in distinct quarters
// where the lists allow (greedy: the target with most words left picks
// first, each taking the word class (T quarter, source quarter) that
// collides least with the picks made so far, ties to the class it holds
// most of).  Finished lists take the null word (zero T row).  Quarters:
// T code c -> c & 3; local source l -> l & 3 (both regions start on a
// quarter boundary).
static void octet_words(const std::vector<std::vector<unsigned>>& lists, int nk4, std::vector<unsigned>& out) {
   out.assign(32 * (size_t)(nk4 / 4), TP_NULL_WORD);
   auto cls_of = [](unsigned w) { return (int)(((w & 0xffffu) / 64 & 3) * 4 + ((w >> 16) / 64 & 3)); };
   std::vector<unsigned> flat;   // a half's words: per target, grouped by class, list order within a class
   for (int h = 0; h < 2; h++) {
      int start[4][17] = {}, left[4][16] = {}, rem[4] = {0, 0, 0, 0};
      size_t base[4] = {0, 0, 0, 0}, tot = 0;
      for (int j = 0; j < 4; j++)
         if (4 * h + j < (int)lists.size()) tot += lists[4 * h + j].size();
      flat.resize(tot);
      size_t b = 0;
      for (int j = 0; j < 4; j++) {
         const int t = 4 * h + j;
         base[j] = b;
         if (t >= (int)lists.size()) continue;
         int* st = start[j];
         for (unsigned w : lists[t]) st[cls_of(w) + 1]++;
         for (int c = 0; c < 16; c++) st[c + 1] += st[c];
         int fill[16];
         for (int c = 0; c < 16; c++) fill[c] = st[c];
         for (unsigned w : lists[t]) flat[b + fill[cls_of(w)]++] = w;
         for (int c = 0; c < 16; c++) left[j][c] = st[c + 1] - st[c];
         rem[j] = (int)lists[t].size();
         b += lists[t].size();
      }
      for (int k = 0; k < nk4; k++) {
         // targets by words left, most first, ties in target order
         int order[4] = {0, 1, 2, 3};
         for (int a = 1; a < 4; a++) {
            const int v = order[a];
            int q = a - 1;
            while (q >= 0 && rem[order[q]] < rem[v]) {
               order[q + 1] = order[q];
               q--;
            }
            order[q + 1] = v;
         }
         // collisions so far in this step, per class: T quarter c >> 2 and
         // source quarter c & 3
         int used[16] = {};
         for (int oi = 0; oi < 4; oi++) {
            const int j = order[oi];
            unsigned w = TP_NULL_WORD;
            if (rem[j] > 0) {
               // least collisions, then most words left, then lowest class:
               // the smallest key among the non-empty classes
               int best_key = INT_MAX;
               for (int c = 0; c < 16; c++) {
                  const int n = left[j][c];
                  const int key = n > 0 ? (used[c] << 20) - (n << 4) + c : INT_MAX;
                  best_key = std::min(best_key, key);
               }
               const int best = best_key & 15;
               w = flat[base[j] + start[j][best] + (start[j][best + 1] - start[j][best] - left[j][best])];
               left[j][best]--;
               rem[j]--;
            }
            const int tq = (int)((w & 0xffffu) / 64 & 3), sq = (int)((w >> 16) / 64 & 3);
            for (int c = 0; c < 4; c++) {
               used[4 * tq + c]++;
               used[4 * c + sq]++;
            }
            out[(size_t)32 * (k / 4) + 4 * (4 * h + j) + (k & 3)] = w;
         }
      }
   }
}

// FP32 translation (k_translate_pipe): tiles of up to TP_TILE targets whose
// source union fits an LDS buffer; per tile the targets sorted by list
// length and cut into octets of eight, longest first (the consumer waves
// claim them in this order); tasks (tile, 8-bin chunk, first octet, octets)
// listed by level, then chunk, then tile, and dealt to the persistent
// blocks as described below.  The tiles are cut level by level; then, on up
// to 16 host threads, each tile's targets are ordered and cut into octets
// (lists per octet), every octet's words are scheduled (octets are
// independent), and the words are concatenated in tile and octet order, so
// the plan does not depend on the thread count.
namespace {
struct PipeTileJob {
   int level, k0, k1;
   std::vector<int> so;        // the tile's sources (absolute spectrum offsets)
   // per octet: the targets' words in list order, then the scheduled words
   std::vector<std::vector<std::vector<unsigned>>> lists;
   std::vector<std::vector<unsigned>> ow;
   // results
   std::vector<TpOct> octs;    // int_off relative to words, acc_idx relative to oacc
   std::vector<long long> oacc;
   size_t nwords = 0;          // the tile's words, alignment included
   size_t wbase = 0;           // its first word in the plan
   int cost = 0;
};
}  // namespace

// Runs fn(i) for i in [0, n) on up to 16 host threads (dynamic claiming);
// the first exception is rethrown.
template <typename F>
static void plan_parallel_for(int n, F&& fn) {
   const int nth = (int)std::max<size_t>(1, std::min<size_t>({16u, std::thread::hardware_concurrency(),
                                                              (size_t)std::max(n, 1)}));
   std::atomic<int> next{0};
   std::exception_ptr err;
   std::mutex em;
   auto work = [&] {
      for (int j; (j = next.fetch_add(1)) < n;) {
         try {
            fn(j);
         } catch (...) {
            std::lock_guard<std::mutex> g(em);
            if (!err) err = std::current_exception();
         }
      }
   };
   std::vector<std::thread> th;
   for (int t = 1; t < nth; t++) th.emplace_back(work);
   work();
   for (auto& x : th) x.join();
   if (err) std::rethrow_exception(err);
}

// A tile's targets sorted by list length (longest first) and cut into
// octets: each octet's accumulator offsets, step count and word lists.
static void pipe_tile_lists(PipeTileJob& J, const TrLevel& L, const std::vector<TgtPlan>& tgt,
                            const std::vector<int>& isrc, const std::vector<int>& itoff, std::vector<int>& loc) {
   const SlotOf slot_of(L);
   for (size_t i = 0; i < J.so.size(); i++) loc[slot_of(J.so[i])] = (int)i;
   const int n = J.k1 - J.k0;
   std::vector<int> order(n);
   for (int k = 0; k < n; k++) order[k] = k;
   std::stable_sort(order.begin(), order.end(),
                    [&](int x, int y) { return tgt[L.t0 + J.k0 + x].nint > tgt[L.t0 + J.k0 + y].nint; });
   for (int o0 = 0; o0 < n; o0 += 8) {
      std::vector<std::vector<unsigned>> lists;
      TpOct o;
      o.acc_idx = (int)J.oacc.size();
      o.pad_ = 0;
      o.int_off = 0;
      int longest = 0;
      for (int t8 = 0; t8 < 8; t8++) {
         if (o0 + t8 >= n) {
            J.oacc.push_back(-1);
            continue;
         }
         const TgtPlan& tp = tgt[L.t0 + J.k0 + order[o0 + t8]];
         std::vector<unsigned> lw(tp.nint);
         for (int e = 0; e < tp.nint; e++)
            lw[e] = (unsigned)(itoff[tp.int_off + e] * 64) |
                    ((unsigned)(loc[slot_of(isrc[tp.int_off + e])] * 64) << 16);
         longest = std::max(longest, tp.nint);
         lists.push_back(std::move(lw));
         J.oacc.push_back(tp.acc_off);
      }
      o.nk = std::max(8, (longest + 7) / 8 * 8);
      J.octs.push_back(o);
      J.lists.push_back(std::move(lists));
   }
   J.ow.resize(J.octs.size());
}

// The offsets of the tile's octet words (each octet from a 16-byte aligned
// start, relative to the tile's) and the modelled cost of a task of this
// tile in steps per consumer wave.
static void pipe_tile_words(PipeTileJob& J) {
   size_t nw = 0;
   for (size_t k = 0; k < J.octs.size(); k++) {
      nw = (nw + 3) / 4 * 4;
      J.octs[k].int_off = (int)nw;
      nw += J.ow[k].size();
   }
   J.nwords = nw;
   std::vector<std::vector<std::vector<unsigned>>>().swap(J.lists);
   double steps = 0;
   for (const TpOct& o : J.octs) steps += o.nk;
   J.cost = (int)(steps / TP_CW + (J.octs.empty() ? 0 : J.octs[0].nk));
}

static void build_translate_pipe(ApplyPlan& P, const std::vector<TrLevel>& lvls, const std::vector<TgtPlan>& tgt,
                                 const std::vector<int>& isrc, const std::vector<int>& itoff, const PipeParams& pp) {
   PipeHost& H = P.pipe_h;
   H = PipeHost();
   std::vector<PipeTileJob> jobs;
   int maxsrc = 1;
   for (int li = 0; li < (int)lvls.size(); li++) maxsrc = std::max(maxsrc, lvls[li].nsrc);
   {
      // the levels' tiles cut concurrently, listed level by level
      std::vector<std::vector<PipeTileJob>> lj(lvls.size());
      plan_parallel_for((int)lvls.size(), [&](int li) {
         make_tiles(
             lvls[li], tgt, isrc, pp.tile, pp.budget, TP_MAX_SRC,
             [&](int k0, int k1, const std::vector<int>& so) {
                PipeTileJob J;
                J.level = li;
                J.k0 = k0;
                J.k1 = k1;
                J.so = so;
                lj[li].push_back(std::move(J));
             },
             TP_QUANTUM);
      });
      for (auto& v : lj)
         for (PipeTileJob& J : v) jobs.push_back(std::move(J));
   }
   {
      plan_parallel_for((int)jobs.size(), [&](int j) {
         thread_local std::vector<int> loc;
         if ((int)loc.size() < maxsrc) loc.assign(maxsrc, 0);
         pipe_tile_lists(jobs[j], lvls[jobs[j].level], tgt, isrc, itoff, loc);
      });
      std::vector<std::pair<int, int>> octets;   // (job, octet)
      for (size_t j = 0; j < jobs.size(); j++)
         for (size_t k = 0; k < jobs[j].octs.size(); k++) octets.push_back({(int)j, (int)k});
      plan_parallel_for((int)octets.size(), [&](int i) {
         PipeTileJob& J = jobs[octets[i].first];
         const int k = octets[i].second;
         octet_words(J.lists[k], J.octs[k].nk, J.ow[k]);
      });
      plan_parallel_for((int)jobs.size(), [&](int j) { pipe_tile_words(jobs[j]); });
   }
   // concatenation in tile order; tasks by level, then chunk, then tile
   size_t j0 = 0, gint_n = 0;
   for (int li = 0; li < (int)lvls.size(); li++) {
      const TrLevel& L = lvls[li];
      std::vector<int4> lt;   // per tile of the level: index, first octet, octets, modelled cost
      for (; j0 < jobs.size() && jobs[j0].level == li; j0++) {
         PipeTileJob& J = jobs[j0];
         TpTile t;
         t.T = nullptr;
         t.spec = nullptr;
         t.acc = nullptr;
         t.src_off = (int)H.gsrc.size();
         t.pad = L.pad;
         t.pst = L.pst;
         if (J.so.empty())
            H.gsrc.push_back(0);   // a readable row for the producers' clamped loads
         else
            H.gsrc.insert(H.gsrc.end(), J.so.begin(), J.so.end());
         t.nsrc = std::max(1, (int)J.so.size());
         gint_n = (gint_n + 3) / 4 * 4;
         J.wbase = gint_n;
         const int wbase = (int)gint_n, abase = (int)H.oacc.size(), oct0 = (int)H.octs.size();
         gint_n += J.nwords;
         H.oacc.insert(H.oacc.end(), J.oacc.begin(), J.oacc.end());
         for (TpOct o : J.octs) {
            o.int_off += wbase;
            o.acc_idx += abase;
            H.octs.push_back(o);
         }
         lt.push_back(make_int4((int)H.tiles.size(), oct0, (int)J.octs.size(), J.cost));
         H.tiles.push_back(t);
         H.t_off.push_back(L.t_off);
      }
      for (int b0 = 0; b0 < L.pad; b0 += 8)
         for (const int4& e : lt) {
            H.tasks.push_back(make_int4(e.x, b0, e.y, e.z));
            H.cost.push_back(e.w + 16.0);   // summing steps plus a fixed share per task
         }
   }
   // every octet's words in place
   H.gint.assign(gint_n, (int)TP_NULL_WORD);
   {
      std::vector<std::pair<int, int>> octets;   // (job, octet)
      for (size_t j = 0; j < jobs.size(); j++)
         for (size_t k = 0; k < jobs[j].octs.size(); k++) octets.push_back({(int)j, (int)k});
      plan_parallel_for((int)octets.size(), [&](int i) {
         const PipeTileJob& J = jobs[octets[i].first];
         const int k = octets[i].second;
         std::copy(J.ow[k].begin(), J.ow[k].end(), (unsigned*)H.gint.data() + J.wbase + J.octs[k].int_off);
      });
   }
   // the producers address the FP32 spectra by 32-bit byte offsets
   for (size_t i = 0; i < H.tiles.size(); i++)
      for (int k = 0; k < H.tiles[i].nsrc; k++)
         if (8.0 * ((double)H.gsrc[H.tiles[i].src_off + k] + H.tiles[i].pst) >= 4294967296.0)
            throw std::runtime_error("FP32 spectra above 4 GiB: the pipelined translation cannot address them");
}

// Tasks to persistent blocks: the list cut into 8 contiguous eighths of
// equal cost, eighth x on the blocks b = x + 8 j of XCD x (blocks are dealt
// to the 8 XCDs round-robin, so one L2 serves the rows that neighbouring
// tiles of an eighth share); within an eighth, tasks longest first to the
// least loaded block; each block then takes its tasks in list order
// (shared T rows in a row).  Returns the tasks in block order.
static std::vector<int4> assign_pipe_blocks(const std::vector<int4>& tasks, const std::vector<double>& cost, int nblk,
                                            std::vector<int2>& range, double* imbalance) {
   const int ntask = (int)tasks.size();
   range.assign(std::max(nblk, 1), make_int2(0, 0));
   std::vector<int4> out;
   if (nblk <= 0) return out;
   const int ngrp = (nblk % 8 == 0) ? 8 : 1;
   double total = 0;
   for (double c : cost) total += c;
   std::vector<int> grp(ntask);
   double acc = 0;
   for (int i = 0; i < ntask; i++) {
      grp[i] = std::min(ngrp - 1, (int)((acc + 0.5 * cost[i]) * ngrp / total));
      acc += cost[i];
   }
   std::vector<std::vector<int>> mine(nblk);
   std::vector<double> load(nblk, 0.0);
   for (int x = 0; x < ngrp; x++) {
      std::vector<int> ids;
      for (int i = 0; i < ntask; i++)
         if (grp[i] == x) ids.push_back(i);
      std::stable_sort(ids.begin(), ids.end(), [&](int u, int v) { return cost[u] > cost[v]; });
      for (int i : ids) {
         int best = -1;
         for (int b = x; b < nblk; b += ngrp)
            if (best < 0 || load[b] < load[best]) best = b;
         mine[best].push_back(i);
         load[best] += cost[i];
      }
   }
   for (int b = 0; b < nblk; b++) {
      std::sort(mine[b].begin(), mine[b].end());
      range[b] = make_int2((int)out.size(), (int)mine[b].size());
      for (int i : mine[b]) out.push_back(tasks[i]);
   }
   if (imbalance) *imbalance = *std::max_element(load.begin(), load.end()) / (total / nblk);
   return out;
}

void make_pipe(PipeDev& d, const std::vector<const TreeOp*>& ops, hipStream_t s, int mode, bool pair) {
   std::vector<TpTile> tiles;
   std::vector<TpOct> octs;
   std::vector<long long> oacc;
   std::vector<int> gsrc, gint;
   std::vector<int4> tasks;
   std::vector<double> cost;
   for (const TreeOp* op : ops) {
      const ApplyPlan& P = op->plan;
      const PipeHost& H = P.pipe_h;
      if (!P.Tsp32.p) throw std::runtime_error("pipelined translation needs the FP32 operator");
      const int tb = (int)tiles.size(), ob = (int)octs.size(), ab = (int)oacc.size(), sb = (int)gsrc.size();
      const int gb = (int)gint.size();
      for (size_t i = 0; i < H.tiles.size(); i++) {
         TpTile t = H.tiles[i];
         t.T = (const float2*)P.Tsp32.p + H.t_off[i];
         t.spec = (const float2*)P.spec.p;
         t.acc = (float2*)P.acc.p;
         t.src_off += sb;
         tiles.push_back(t);
      }
      for (TpOct o : H.octs) {
         o.int_off += gb;
         o.acc_idx += ab;
         octs.push_back(o);
      }
      oacc.insert(oacc.end(), H.oacc.begin(), H.oacc.end());
      gsrc.insert(gsrc.end(), H.gsrc.begin(), H.gsrc.end());
      gint.insert(gint.end(), H.gint.begin(), H.gint.end());
      for (size_t i = 0; i < H.tasks.size(); i++) {
         int4 t = H.tasks[i];
         t.x += tb;
         t.z += ob;
         tasks.push_back(t);
         cost.push_back(H.cost[i]);
      }
   }
   if (pair) {
      // each run of tasks of one level and 8-bin chunk (consecutive, over
      // the level's tiles) becomes the run of its first 4-bin halves, then
      // the run of its second halves, so consecutive tasks keep sharing
      // their T rows
      std::vector<int4> t2;
      std::vector<double> c2;
      for (size_t a = 0; a < tasks.size();) {
         size_t b = a + 1;
         while (b < tasks.size() && tasks[b].y == tasks[a].y && tiles[tasks[b].x].T == tiles[tasks[a].x].T) b++;
         for (int half = 0; half < 2; half++)
            for (size_t i = a; i < b; i++) {
               int4 t = tasks[i];
               t.y += 4 * half;
               if (t.y >= tiles[t.x].pad) continue;
               t2.push_back(t);
               c2.push_back(0.5 * cost[i]);
            }
         a = b;
      }
      tasks.swap(t2);
      cost.swap(c2);
   }
   if (gint.empty()) gint.push_back((int)TP_NULL_WORD);
   if (oacc.empty()) oacc.push_back(-1);
   if (gsrc.empty()) gsrc.push_back(0);
   int ncu = 256;
   {
      int dev = 0, v = 0;
      if (hipGetDevice(&dev) == hipSuccess &&
          hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
         ncu = v;
   }
   const int nblk = std::min((int)tasks.size(), ncu);
   std::vector<int2> range;
   double imb = 1.0;
   std::vector<int4> btasks = assign_pipe_blocks(tasks, cost, nblk, range, &imb);
   if (timing_on() && nblk > 0) {
      long long targets = 0, steps = 0, nsrc = 0;
      for (const TpOct& o : octs) steps += o.nk;
      for (long long a : oacc) targets += a >= 0;
      for (const TpTile& t : tiles) nsrc += t.nsrc;
      std::fprintf(stderr,
                   "[fftsvd setup] pipelined translation (%d tree(s)): %d tiles (%.1f targets, %.1f sources each), "
                   "%d tasks on %d blocks, block cost max/mean %.3f, %.1f steps per octet, %.3f of target slots "
                   "used\n",
                   (int)ops.size(), (int)tiles.size(), (double)targets / std::max<size_t>(tiles.size(), 1),
                   (double)nsrc / std::max<size_t>(tiles.size(), 1), (int)tasks.size(), nblk, imb,
                   (double)steps / std::max<size_t>(octs.size(), 1),
                   (double)targets / std::max<size_t>(8 * octs.size(), 1));
   }
   d.nblk = nblk;
   d.mode = mode;
   if (btasks.empty()) btasks.push_back(make_int4(0, 0, 0, 0));
   d.range.upload(range, s);
   d.tasks.upload(btasks, s);
   d.tiles.upload(tiles.empty() ? std::vector<TpTile>(1) : tiles, s);
   d.octs.upload(octs.empty() ? std::vector<TpOct>(1) : octs, s);
   d.oacc.upload(oacc, s);
   d.gsrc.upload(gsrc, s);
   d.gint.upload(gint, s);
   FSVD_CHECK(hipStreamSynchronize(s));
}

// ---- autotuner of the pipelined translation ----
// Staging mode, tile source budget and tile size of k_translate_pipe.  By
// default the hand-tuned values; with autotuning on (fftsvd_set_autotune,
// or FFTSVD_AUTOTUNE=1), the first FP32 operator of the process times each
// candidate on its own translation (one warm-up and three timed launches,
// HIP events on the setup stream) and the fastest is used from then on.
namespace {
std::mutex g_tune_mu;
bool g_autotune = [] {
   const char* e = std::getenv("FFTSVD_AUTOTUNE");
   return e && e[0] == '1';
}();
bool g_tuned = false;
PipeParams g_params;
}  // namespace

void set_autotune(bool on) {
   std::lock_guard<std::mutex> g(g_tune_mu);
   g_autotune = on;
   g_tuned = false;
   g_params = PipeParams();
}

PipeParams current_pipe_params(bool* tuned) {
   std::lock_guard<std::mutex> g(g_tune_mu);
   if (tuned) *tuned = g_tuned;
   return g_params;
}

static void build_translate_pipe(ApplyPlan& P, const std::vector<TrLevel>& lvls, const std::vector<TgtPlan>& tgt,
                                 const std::vector<int>& isrc, const std::vector<int>& itoff, const PipeParams& pp);

static PipeParams pipe_params(TreeOp& op, hipStream_t s) {
   std::lock_guard<std::mutex> g(g_tune_mu);
   if (!g_autotune || g_tuned) return g_params;
   ApplyPlan& P = op.plan;
   const PipeParams cand[] = {{1, TP_SRC_BUDGET, TP_TILE}, {0, TP_SRC_BUDGET, TP_TILE}, {1, 720, TP_TILE},
                              {1, TP_SRC_BUDGET, 256}};
   hipEvent_t e0, e1;
   FSVD_CHECK(hipEventCreate(&e0));
   FSVD_CHECK(hipEventCreate(&e1));
   PipeParams best = g_params;
   float best_ms = 1e30f;
   for (const PipeParams& c : cand) {
      build_translate_pipe(P, P.tr_lv, P.tr_tgt, P.tr_isrc, P.tr_itoff, c);
      PipeDev d;
      make_pipe(d, {&op}, s, c.mode);
      launch_pipe(d, s);
      FSVD_CHECK(hipEventRecord(e0, s));
      for (int r = 0; r < 3; r++) launch_pipe(d, s);
      FSVD_CHECK(hipEventRecord(e1, s));
      FSVD_CHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      FSVD_CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best_ms) {
         best_ms = ms;
         best = c;
      }
   }
   (void)hipEventDestroy(e0);
   (void)hipEventDestroy(e1);
   best.tuned_ms = best_ms / 3.0f;
   g_params = best;
   g_tuned = true;
   return best;
}

// Concatenates every level's far-field blocks into single buffers with
// absolute offsets (row bases re-laid panel-major for the projection),
// splits source cubes into projection chunks, and attaches to every leaf
// the U blocks of the target cubes that contain it.
static void build_apply_plan(TreeOp& op, bool want32, hipStream_t s) {
   // basis entries per projection chunk (one wavefront), at most
   // PROJ_MAXCOLS panels
   constexpr int CHUNK = 4096, PROJ_MAXCOLS = 512;
   double tm = now_ms();
   ApplyPlan& P = op.plan;
   const Tree& t = op.tree;
   const size_t L = op.levels.size();
   long long vt_n = 0, pv_n = 0, t_n = 0, uti_n = 0, ut_n = 0;
   for (const FarLevel& lv : op.levels) {
      vt_n += (long long)lv.VT64.n;
      pv_n += (long long)lv.PV64.n;
      t_n += 2LL * 343 * tr_pst(lv.pad);
      uti_n += (long long)lv.UTI64.n;
      ut_n += (long long)lv.UT64.n;
   }
   P.VTc64.alloc(std::max<long long>(vt_n, 1));
   P.PV64.alloc(std::max<long long>(pv_n, 1));
   P.Tsp64.alloc(std::max<long long>(t_n, 2));
   // the bins between pad and the padded stride stay zero
   FSVD_CHECK(hipMemsetAsync(P.Tsp64.p, 0, sizeof(double) * P.Tsp64.n, s));
   P.UTI64.alloc(std::max<long long>(uti_n, 1));
   P.UT64.alloc(std::max<long long>(ut_n, 1));

   std::vector<ProjChunk> chunks;
   std::vector<SrcPlan> src;
   std::vector<TgtPlan> tgt;
   std::vector<int> isrc, itoff;
   std::vector<TransDesc> tr;
   std::vector<long long> lv_ut(L);
   std::vector<std::vector<int>> lv_w(L);
   long long vt_b = 0, pv_b = 0, t_b = 0, uti_b = 0, ut_b = 0, spec_b = 0, acc_b = 0;
   int tpart_n = 0, w_n = 0;
   for (size_t li = 0; li < L; li++) {
      const FarLevel& lv = op.levels[li];
      P.max_n = std::max(P.max_n, lv.n);
      for (const SrcDesc& sd : lv.h_src) {
         SrcPlan sp;
         sp.n = lv.n;
         sp.r = sd.r;
         sp.part_off = tpart_n;
         const int cpc = std::max(1, std::min(PROJ_MAXCOLS, CHUNK / sd.r));
         sp.nchunk = 0;
         for (int c0 = 0; c0 < sd.np; c0 += cpc) {
            ProjChunk c;
            c.vt_off = vt_b + sd.vt_off + (long long)c0 * sd.r;
            c.p0 = sd.p0 + c0;
            c.ncol = std::min(cpc, sd.np - c0);
            c.r = sd.r;
            c.out_off = tpart_n;
            tpart_n += 2 * sd.r;
            chunks.push_back(c);
            sp.nchunk++;
            P.proj_maxcols = std::max(P.proj_maxcols, c.ncol);
         }
         sp.pv_off = pv_b + sd.pv_off;
         sp.spec_off = spec_b + (long long)sd.spec_slot * tr_pst(lv.pad);
         src.push_back(sp);
         P.maxr_src = std::max(P.maxr_src, sd.r);
         tr.push_back({lv.VT64.p + sd.vt_off, P.VTc64.p + vt_b + sd.vt_off, sd.r, sd.np});
      }
      lv_ut[li] = ut_b;
      lv_w[li].resize(lv.h_tgt.size());
      for (size_t k = 0; k < lv.h_tgt.size(); k++) {
         const TgtDesc& td = lv.h_tgt[k];
         TgtPlan tp;
         tp.n = lv.n;
         tp.r = td.r;
         tp.pad = lv.pad;
         tp.nint = td.nint;
         tp.int_off = (int)isrc.size();
         tp.w_off = w_n;
         tp.t_off = t_b / 2;
         tp.acc_off = acc_b + (long long)td.spec_slot * tr_pst(lv.pad);
         tp.uti_off = uti_b + td.uti_off;
         for (int e = 0; e < td.nint; e++) {
            long long so = spec_b + (long long)lv.h_isrc[td.int_off + e] * tr_pst(lv.pad);
            if (so > 0x7fffffffLL) throw std::runtime_error("spectrum offsets exceed 32 bits");
            isrc.push_back((int)so);
            itoff.push_back(lv.h_itoff[td.int_off + e]);
         }
         lv_w[li][k] = w_n;
         w_n += td.r;
         tgt.push_back(tp);
      }
      auto cp = [&](double* dst, const DBuf<double>& b) {
         if (b.n) FSVD_CHECK(hipMemcpyAsync(dst, b.p, sizeof(double) * b.n, hipMemcpyDeviceToDevice, s));
      };
      cp(P.PV64.p + pv_b, lv.PV64);
      // T spectra re-laid at the padded row stride (343 rows of pad bins)
      if (lv.T64.n)
         FSVD_CHECK(hipMemcpy2DAsync(P.Tsp64.p + t_b, sizeof(double) * 2 * tr_pst(lv.pad), lv.T64.p,
                                     sizeof(double) * 2 * lv.pad, sizeof(double) * 2 * lv.pad, 343,
                                     hipMemcpyDeviceToDevice, s));
      cp(P.UTI64.p + uti_b, lv.UTI64);
      cp(P.UT64.p + ut_b, lv.UT64);
      vt_b += (long long)lv.VT64.n;
      pv_b += (long long)lv.PV64.n;
      t_b += 2LL * 343 * tr_pst(lv.pad);
      uti_b += (long long)lv.UTI64.n;
      ut_b += (long long)lv.UT64.n;
      spec_b += (long long)lv.nsrc * tr_pst(lv.pad);
      acc_b += (long long)lv.ntgt * tr_pst(lv.pad);
   }
   {
      DBuf<TransDesc> dtr;
      dtr.upload(tr, s);
      launch_transpose_batched<double>((int)tr.size(), dtr.p, s);
      FSVD_CHECK(hipStreamSynchronize(s));
   }
   timing_mark("plan: concatenate + VT^T", tm);
   timing_mark("plan: leaf U blocks start", tm);
   // leaf -> far-field output blocks, finest level first
   std::vector<int> leaf_cube(op.nleaf, -1);
   for (size_t i = 0; i < op.cube_leaf.size(); i++)
      if (op.cube_leaf[i] >= 0) leaf_cube[op.cube_leaf[i]] = (int)i;
   std::vector<LeafU> ublk;
   for (int li = 0; li < op.nleaf; li++) {
      LeafDesc& Ld = op.h_leaves[li];
      Ld.u0 = (int)ublk.size();
      for (int a = leaf_cube[li]; a >= 0 && t.cubes[a].level >= 2; a = t.cubes[a].parent) {
         const int lvl = t.cubes[a].level - 2, slot = op.cube_tgt_slot[a];
         if (lvl >= (int)L || slot < 0) continue;
         const TgtDesc& td = op.levels[lvl].h_tgt[slot];
         LeafU u;
         u.u_off = lv_ut[lvl] + td.u_off + (Ld.q0 - td.q0);
         u.ld = td.nq;
         u.r = td.r;
         u.w_off = lv_w[lvl][slot];
         u.pad_ = 0;
         ublk.push_back(u);
      }
      Ld.nu = (int)ublk.size() - Ld.u0;
   }
   op.leaves.upload(op.h_leaves, s);
   P.nchunk = (int)chunks.size();
   P.nsrc = (int)src.size();
   P.ntgt = (int)tgt.size();
   P.nublk = (int)ublk.size();
   P.chunks.upload(chunks, s);
   timing_mark("plan: leaf U blocks", tm);
   // translation plans over the targets of each level (tree order, so
   // consecutive targets are spatially compact): FP64 tiles for k_translate,
   // FP32 tiles and persistent-block task ranges for k_translate_pipe
   // The FP64 plan is worked out on a host thread of its own, beside the
   // rest of the plan (it reads tl, tgt, isrc and itoff, which stay
   // unchanged, and writes only its own fields of P); joined before return.
   std::vector<TrLevel> tl;
   {
      size_t tbase = 0;
      long long tb = 0, sbase = 0;
      for (size_t li = 0; li < L; li++) {
         const FarLevel& lv = op.levels[li];
         tl.push_back({(int)tbase, (int)lv.h_tgt.size(), lv.pad, tr_pst(lv.pad), lv.nsrc, tb / 2, sbase});
         tbase += lv.h_tgt.size();
         tb += 2LL * 343 * tr_pst(lv.pad);
         sbase += (long long)lv.nsrc * tr_pst(lv.pad);
      }
      if (want32) {
         P.tr_lv = tl;
         P.tr_tgt = tgt;
         P.tr_isrc = isrc;
         P.tr_itoff = itoff;
      }
   }
   int dev = 0;
   FSVD_CHECK(hipGetDevice(&dev));
   std::exception_ptr f64_err;
   struct Joiner {
      std::thread t;
      ~Joiner() {
         if (t.joinable()) t.join();
      }
   } f64_plan;
   f64_plan.t = std::thread([&] {
      try {
         FSVD_CHECK(hipSetDevice(dev));
         build_translate_f64(P, tl, tgt, isrc, itoff, s);
      } catch (...) {
         f64_err = std::current_exception();
      }
   });
   {
      std::vector<double> wl(op.h_leaves.size());
      for (size_t k = 0; k < wl.size(); k++) {
         const LeafDesc& Ld = op.h_leaves[k];
         wl[k] = 2.0 * Ld.p * Ld.m + 64.0;
      }
      std::vector<int> ml = xcd_block_map(wl);
      P.nmap_leaf = (int)ml.size();
      P.map_leaf.upload(ml, s);
   }
   {
      // cubes of four grid points per side (one per wavefront) after the
      // others (one per block): each class is one contiguous launch
      auto big_first = [](auto v) {
         std::stable_partition(v.begin(), v.end(), [](const auto& e) { return e.n != 4; });
         return v;
      };
      const std::vector<SrcPlan> src_o = big_first(src);
      const std::vector<TgtPlan> tgt_o = big_first(tgt);
      P.nsrc_big = (int)std::count_if(src.begin(), src.end(), [](const SrcPlan& e) { return e.n != 4; });
      P.ntgt_big = (int)std::count_if(tgt.begin(), tgt.end(), [](const TgtPlan& e) { return e.n != 4; });
      P.src.upload(src_o, s);
      P.tgt.upload(tgt_o, s);
      FSVD_CHECK(hipStreamSynchronize(s));
   }
   P.ublk.upload(ublk, s);
   timing_mark("plan: tiles + maps", tm);
   P.tpart.alloc(std::max(tpart_n, 1));
   P.spec.alloc(2 * std::max<long long>(spec_b, 1));
   P.acc.alloc(2 * std::max<long long>(acc_b, 1));
   FSVD_CHECK(hipMemsetAsync(P.spec.p, 0, sizeof(double) * P.spec.n, s));
   FSVD_CHECK(hipMemsetAsync(P.acc.p, 0, sizeof(double) * P.acc.n, s));
   P.w.alloc(std::max(w_n, 1));
   if (want32) {
      to_f32(P.VTc64, P.VTc32, s);
      to_f32(P.PV64, P.PV32, s);
      to_f32(P.Tsp64, P.Tsp32, s);
      to_f32(P.UTI64, P.UTI32, s);
      to_f32(P.UT64, P.UT32, s);
      const PipeParams pp = pipe_params(op, s);
      timing_mark("plan: FP32 copies, tuning", tm);
      build_translate_pipe(P, P.tr_lv, P.tr_tgt, P.tr_isrc, P.tr_itoff, pp);
      timing_mark("plan: pipe tiles", tm);
      make_pipe(P.pipe, {&op}, s, pp.mode);
      op.stats.pipe = pp;
      timing_mark("plan: pipelined translation", tm);
   }
   f64_plan.t.join();
   if (f64_err) std::rethrow_exception(f64_err);
   timing_mark("plan: FP64 translation joined", tm);
}

// Equivalent densities of one level (Cube_fill_P_I + Cube_fill_PV_UTI,
// Cube.c:795-1002,1414-1522): P = pinv Q2S(panels -> sphere) (both layers),
// PV^T = VT P^T; I = (pinv Q2S(sphere -> points))^T, UTI = U^T I, for the
// level's source and target cubes cl, with the sphere integrals and point
// kernel of `kernel`; results in the level's PV and UTI layouts.
static void level_pv_uti(const Tree& t, const std::vector<int>& cl, const FarLevel& lv,
                         const std::vector<double>& ptsT, const PanelGeom* d_panels, const double* pinvp, int kernel,
                         double kappa, const double* VT, const double* UT, double* PV, double* UTI, hipStream_t s) {
   const int gp3 = lv.gp3;
   std::vector<IntBlock> blocks;
   std::vector<double> bpts(ptsT);
   std::vector<double> sph;   // sphere points per cube (target side, Green blocks)
   std::vector<GreenBlock> gblocks;
   long long qs = 0, ps = 0, qp = 0, ip = 0;
   struct Aux {
      long long q_off, p_off;
   };
   std::vector<Aux> saux(lv.h_src.size()), taux(lv.h_tgt.size());
   size_t si = 0, ti = 0;
   for (size_t ci = 0; ci < cl.size(); ci++) {
      const Cube& c = t.cubes[cl[ci]];
      double box = c.b1[0] - c.b0[0];
      double sc = box * 1.5 * 1;
      bool is_src = c.np > 0 && c.rowrank > 0, is_tgt = c.nq > 0 && c.colrank > 0;
      if (!is_src && !is_tgt) continue;
      int sbase = (int)(bpts.size() / 3);
      for (int q = 0; q < NQ; q++)
         for (int a = 0; a < 3; a++) bpts.push_back(SPHERE49[q][a] * sc + c.ctr[a]);
      if (is_src) {
         IntBlock b;
         b.npt = NQ;
         b.npan = c.np;
         b.pt_off = sbase;
         b.pan_off = c.p0;
         b.out_off = qs;
         b.ld_pt = c.np;
         b.ld_pan = 1;
         b.layer_stride = (long long)NQ * c.np;
         push_blocks(blocks, b);
         saux[si] = {qs, ps};
         qs += 2LL * NQ * c.np;
         ps += 2LL * gp3 * c.np;
         si++;
      }
      if (is_tgt) {
         GreenBlock g;
         g.ndest = c.nq;
         g.nsrc = NQ;
         g.dest_off = c.q0;
         g.src_off = sbase;
         g.out_off = qp;
         g.ld_dest = 1;
         g.ld_src = c.nq;
         gblocks.push_back(g);
         taux[ti] = {qp, ip};
         qp += (long long)NQ * c.nq;
         ip += (long long)gp3 * c.nq;
         ti++;
      }
   }
   DBuf<double> Qs, Pb, Qpb, Ib;
   Qs.alloc(std::max<long long>(qs, 1));
   Pb.alloc(std::max<long long>(ps, 1));
   Qpb.alloc(std::max<long long>(qp, 1));
   Ib.alloc(std::max<long long>(ip, 1));
   run_integrals(blocks, bpts, t.pan_perm, d_panels, kernel, kappa, Qs.p, s);
   if (!gblocks.empty()) {
      DBuf<GreenBlock> dg;
      DBuf<double> dp;
      dg.upload(gblocks, s);
      dp.upload(bpts, s);
      launch_green_blocks((int)gblocks.size(), dg.p, dp.p, dp.p, kernel, kappa, Qpb.p, s);
      FSVD_CHECK(hipStreamSynchronize(s));
   }
   // P = pinv * Q2S (both layers); I^T = (pinv * Q2S_points)^T
   std::vector<GemmDesc> g1, g2;
   for (size_t k = 0; k < lv.h_src.size(); k++) {
      const SrcDesc& sd = lv.h_src[k];
      for (int layer = 0; layer < 2; layer++) {
         GemmDesc d;
         d.M = gp3;
         d.N = sd.np;
         d.K = NQ;
         d.A = pinvp;
         d.a_rs = NQ;
         d.a_cs = 1;
         d.B = Qs.p + saux[k].q_off + (long long)layer * NQ * sd.np;
         d.b_rs = sd.np;
         d.b_cs = 1;
         d.C = Pb.p + saux[k].p_off + (long long)layer * gp3 * sd.np;
         d.c_rs = sd.np;
         d.c_cs = 1;
         g1.push_back(d);
      }
   }
   for (size_t k = 0; k < lv.h_tgt.size(); k++) {
      const TgtDesc& td = lv.h_tgt[k];
      GemmDesc d;
      d.M = td.nq;
      d.N = gp3;
      d.K = NQ;
      d.A = Qpb.p + taux[k].q_off;
      d.a_rs = 1;
      d.a_cs = td.nq;
      d.B = pinvp;
      d.b_rs = 1;
      d.b_cs = NQ;
      d.C = Ib.p + taux[k].p_off;
      d.c_rs = gp3;
      d.c_cs = 1;
      g1.push_back(d);
   }
   run_gemms(g1, s);
   // PV^T (r x gp3) = VT (r x np) * P^T ; UTI (r x gp3) = U^T (r x nq) * I (nq x gp3)
   for (size_t k = 0; k < lv.h_src.size(); k++) {
      const SrcDesc& sd = lv.h_src[k];
      for (int layer = 0; layer < 2; layer++) {
         GemmDesc d;
         d.M = sd.r;
         d.N = gp3;
         d.K = sd.np;
         d.A = VT + sd.vt_off;
         d.a_rs = sd.np;
         d.a_cs = 1;
         d.B = Pb.p + saux[k].p_off + (long long)layer * gp3 * sd.np;
         d.b_rs = 1;
         d.b_cs = sd.np;
         d.C = PV + sd.pv_off + (long long)layer * sd.r * gp3;
         d.c_rs = gp3;
         d.c_cs = 1;
         g2.push_back(d);
      }
   }
   for (size_t k = 0; k < lv.h_tgt.size(); k++) {
      const TgtDesc& td = lv.h_tgt[k];
      GemmDesc d;
      d.M = td.r;
      d.N = gp3;
      d.K = td.nq;
      d.A = UT + td.u_off;
      d.a_rs = td.nq;
      d.a_cs = 1;
      d.B = Ib.p + taux[k].p_off;
      d.b_rs = gp3;
      d.b_cs = 1;
      d.C = UTI + td.uti_off;
      d.c_rs = gp3;
      d.c_cs = 1;
      g2.push_back(d);
   }
   run_gemms(g2, s);
}

std::unique_ptr<TreeOp> build_tree_op(const double* pan_cent, int npan, const PanelGeom* d_panels,
                                      const double* pts, int npt, int kernel, double kappa, int maxpanels,
                                      int gridpoints, double svdeps, bool want32, hipStream_t s,
                                      const Tree* pre, int layers) {
   if (layers != 1 && layers != 2) throw std::runtime_error("layers must be 1 (double) or 2 (single and double)");
   if (kernel != LAPLACE && kernel != YUKAWA) throw std::runtime_error("operators are built for the Laplace and Yukawa kernels");
   auto opp = std::make_unique<TreeOp>();
   TreeOp& op = *opp;
   op.kernel = kernel;
   op.kappa = kappa;
   Tree& t = op.tree;
   t.kernel = kernel;
   t.kappa = kappa;
   t.svdeps = svdeps;
   double tp = now_ms();
   if (pre) {
      t = *pre;
   } else {
      build_tree(t, pan_cent, npan, pts, npt, maxpanels, gridpoints);
   }
   t.kernel = kernel;
   t.kappa = kappa;
   t.svdeps = svdeps;
   op.stats.phase_ms[0] += now_ms() - tp;
   timing_mark("tree", tp);
   op.stats.depth = t.depth;
   op.stats.ncubes = (int)t.cubes.size();

   std::vector<double> ptsT(3 * (size_t)npt);
   for (int k = 0; k < npt; k++)
      for (int d = 0; d < 3; d++) ptsT[3 * (size_t)k + d] = pts[3 * (size_t)t.pt_perm[k] + d];
   op.pts_tree.upload(ptsT, s);
   op.pan_perm.upload(t.pan_perm, s);
   op.yfar.alloc(std::max(npt, npan) + 1);

   // ---------------- near field (Cube_fill_D, Cube.c:715-793) ----------------
   // on a host thread and stream of its own, beside the far-field levels
   // (it writes only the near-field members of op; its stream is
   // synchronized and the thread joined before the interaction lists)
   int dev = 0;
   FSVD_CHECK(hipGetDevice(&dev));
   std::exception_ptr near_err;
   auto near_field = [&]() {
      hipStream_t s = nullptr;
      try {
         FSVD_CHECK(hipSetDevice(dev));
         FSVD_CHECK(hipStreamCreate(&s));
         double t0 = now_ms();
         long long dsize = 0;
         for (size_t i = 0; i < t.cubes.size(); i++) {
            const Cube& c = t.cubes[i];
            if (!c.leaf || c.nq == 0) continue;
            LeafDesc L;
            L.q0 = c.q0;
            L.p = c.nq;
            L.col_off = (int)op.h_dcol.size();
            for (int k = 0; k < c.np; k++) op.h_dcol.push_back(c.p0 + k);
            for (int l : c.loc) {
               const Cube& o = t.cubes[l];
               for (int k = 0; k < o.np; k++) op.h_dcol.push_back(o.p0 + k);
            }
            L.m = (int)op.h_dcol.size() - L.col_off;
            L.d_off = dsize;
            L.u0 = 0;
            L.nu = 0;
            dsize += 2LL * L.p * L.m;
            op.h_leaves.push_back(L);
            op.maxm = std::max(op.maxm, L.m);
            op.maxp = std::max(op.maxp, L.p);
         }
         op.nleaf = (int)op.h_leaves.size();
         op.stats.nleaf = op.nleaf;
         op.stats.near_entries = dsize;
         op.leaves.upload(op.h_leaves, s);
         op.dcol.upload(op.h_dcol, s);
         op.D64.alloc(std::max<long long>(dsize, 1));
         {
            std::vector<int> bpan(op.h_dcol.size());
            for (size_t k = 0; k < bpan.size(); k++) bpan[k] = t.pan_perm[op.h_dcol[k]];
            std::vector<IntBlock> blocks;
            for (const LeafDesc& L : op.h_leaves) {
               if (L.m == 0) continue;
               IntBlock b;
               b.npt = L.p;
               b.npan = L.m;
               b.pt_off = L.q0;
               b.pan_off = L.col_off;
               b.out_off = L.d_off;
               b.ld_pt = 1;
               b.ld_pan = 2 * L.p;
               b.layer_stride = L.p;
               push_blocks(blocks, b);
            }
            run_integrals(blocks, ptsT, bpan, d_panels, kernel, kappa, op.D64.p, s);
         }
         FSVD_CHECK(hipStreamSynchronize(s));
         op.stats.setup_ms_near = now_ms() - t0;
         timing_mark("near field", t0);
      } catch (...) {
         near_err = std::current_exception();
      }
      if (s) (void)hipStreamDestroy(s);
   };
   struct Joiner {
      std::thread t;
      ~Joiner() {
         if (t.joinable()) t.join();
      }
   } near_thread;
   near_thread.t = std::thread(near_field);
   double t0 = now_ms();

   // ---------------- far field ----------------
   const double root = t.cubes[0].b1[0] - t.cubes[0].b0[0];
   op.h_pinv.resize(t.depth);
   std::vector<int> src_slot(t.cubes.size(), -1), tgt_slot(t.cubes.size(), -1);
   // Levels are independent until the interaction lists: each is built by
   // its own host thread on its own stream (level l -> op.levels[l-2]).
   const int nlev = std::max(0, t.depth - 2);
   op.levels.resize(nlev);
   std::vector<std::array<double, 8>> lev_ph(nlev);
   std::vector<long long> lev_far(nlev, 0);
   auto build_level = [&](int l, hipStream_t s) {
      double* ph = lev_ph[l - 2].data();
      FarLevel lv;
      lv.level = l;
      lv.n = t.gpl[l];
      const int n = lv.n, N = 2 * n - 1, H = N / 2 + 1;
      lv.pad = N * N * H;
      lv.gp3 = n * n * n;
      const int gp3 = lv.gp3, pad = lv.pad;

      double tp = now_ms();
      double tl = tp;
      // pseudoinverse of the sphere -> grid Green's matrix (Tree_fill_pinv_Fitting, Tree.c:295-369)
      {
         double box = root / std::pow(2.0, l);
         double h = box / (n - 1);
         double sc = box * 1.5 * 1, ctr = 0.5 * box;
         std::vector<double> Q((size_t)NQ * gp3);
         for (int q = 0; q < NQ; q++) {
            double mq[3];
            for (int a = 0; a < 3; a++) mq[a] = SPHERE49[q][a] * sc + ctr;
            int g = 0;
            for (int i = 0; i < n; i++)
               for (int j = 0; j < n; j++)
                  for (int k = 0; k < n; k++, g++) {
                     double p0 = i * h - mq[0], p1 = j * h - mq[1], p2 = k * h - mq[2];
                     double r = std::sqrt(p0 * p0 + p1 * p1 + p2 * p2);
                     Q[(size_t)q * gp3 + g] = (kernel == YUKAWA) ? std::exp(-kappa * r) / r : (r == 0.0 ? 0.0 : 1.0 / r);
                  }
         }
         op.h_pinv[l].resize((size_t)gp3 * NQ);
         pinv(Q.data(), NQ, gp3, op.h_pinv[l].data());
      }
      // translation spectra (Tree_fill_T, Tree.c:387-445)
      {
         std::vector<std::complex<double>> T((size_t)343 * pad);
         translation_spectra(n, root / std::pow(2.0, l) / (n - 1), kernel, kappa, T.data());
         lv.T64.upload((const double*)T.data(), 2 * T.size(), s);
      }
      DBuf<double> d_pinv;
      d_pinv.upload(op.h_pinv[l], s);
      ph[2] += now_ms() - tp;
      timing_mark((std::string("L") + std::to_string(l) + " " + "pinv + T").c_str(), tl);
      tp = now_ms();

      // cubes of this level with interactions
      std::vector<int> cl;
      for (size_t i = 0; i < t.cubes.size(); i++)
         if (t.cubes[i].level == l && !t.cubes[i].itr.empty()) cl.push_back((int)i);

      // ---- sampled bases (Cube_fill_U_VT, Cube.c:1104-1412) ----
      struct Prob {
         int cube;
         int isU;       // 0: row basis VT over panels, 1: column basis U over points
         int pass;
         int nsamp;     // interacting cubes with points (VT) / panels (U)
         int rank = -1;
      };
      std::vector<Prob> probs;
      std::vector<std::vector<int>> ipt(cl.size()), ipn(cl.size());
      for (size_t ci = 0; ci < cl.size(); ci++) {
         const Cube& c = t.cubes[cl[ci]];
         for (int o : c.itr) {
            if (t.cubes[o].nq > 0) ipt[ci].push_back(o);
            if (t.cubes[o].np > 0) ipn[ci].push_back(o);
         }
         if (c.np > 0 && !ipt[ci].empty()) probs.push_back({(int)ci, 0, 1, (int)ipt[ci].size()});
         if (c.nq > 0 && !ipn[ci].empty()) probs.push_back({(int)ci, 1, 1, (int)ipn[ci].size()});
      }
      // final Q storage per problem (device) collected across resampling rounds
      std::vector<std::shared_ptr<DBuf<double>>> keep;
      std::vector<const double*> qsrc(probs.size(), nullptr);
      std::vector<int> pending(probs.size());
      for (size_t i = 0; i < probs.size(); i++) pending[i] = (int)i;
      while (!pending.empty()) {
         std::vector<IntBlock> blocks;
         std::vector<MgsDesc> mgs;
         std::vector<double> bpts(ptsT);
         std::vector<int> bpan(t.pan_perm);
         long long asz = 0, qsz = 0;
         for (size_t pi = 0; pi < pending.size(); pi++) {
            Prob& P = probs[pending[pi]];
            const Cube& c = t.cubes[cl[P.cube]];
            const int K = layers * P.nsamp * P.pass;
            const int half = P.nsamp * P.pass;
            MgsDesc m;
            m.rank_slot = (int)pi;
            m.a_off = asz;
            m.q_off = qsz;
            IntBlock b;
            b.out_off = asz;
            b.layer_stride = half;
            if (!P.isU) {
               // rows of R: sampled points of interacting cubes, columns: my panels -> A = R^T (np x K)
               int base = (int)(bpts.size() / 3);
               for (int p = 0; p < P.pass; p++)
                  for (int r = 0; r < P.nsamp; r++) {
                     const Cube& I = t.cubes[ipt[P.cube][r]];
                     int ci = (int)(&I - &t.cubes[0]);
                     int off = t.cube_pt_sorted_off[ci];
                     int idx = t.cube_pt_sorted[off + (unsigned)p * (unsigned)I.nq / (unsigned)P.pass];
                     for (int d = 0; d < 3; d++) bpts.push_back(pts[3 * (size_t)idx + d]);
                  }
               // column-major A (np x K): sample s of layer L at column L*half + s
               b.npt = half;
               b.npan = c.np;
               b.pt_off = base;
               b.pan_off = c.p0;
               b.ld_pt = c.np;
               b.ld_pan = 1;
               b.layer_stride = layers == 2 ? (long long)half * c.np : -1;
               m.rows = c.np;
               m.cols = K;
            } else {
               int base = (int)bpan.size();
               for (int p = 0; p < P.pass; p++)
                  for (int q = 0; q < P.nsamp; q++) {
                     const Cube& I = t.cubes[ipn[P.cube][q]];
                     int ci = (int)(&I - &t.cubes[0]);
                     int off = t.cube_pan_sorted_off[ci];
                     bpan.push_back(t.cube_pan_sorted[off + (unsigned)p * (unsigned)I.np / (unsigned)P.pass]);
                  }
               b.npt = c.nq;
               b.npan = half;
               b.pt_off = c.q0;
               b.pan_off = base;
               b.ld_pt = 1;
               b.ld_pan = c.nq;
               b.layer_stride = layers == 2 ? (long long)half * c.nq : -1;
               m.rows = c.nq;
               m.cols = K;
            }
            push_blocks(blocks, b);
            mgs.push_back(m);
            asz += (long long)m.rows * m.cols;
            qsz += (long long)m.rows * std::min(m.rows, m.cols);
         }
         auto Abuf = std::make_shared<DBuf<double>>();
         auto Qbuf = std::make_shared<DBuf<double>>();
         Abuf->alloc(std::max<long long>(asz, 1));
         Qbuf->alloc(std::max<long long>(qsz, 1));
         run_integrals(blocks, bpts, bpan, d_panels, kernel, kappa, Abuf->p, s);
         FSVD_CHECK(hipStreamSynchronize(s));
         ph[3] += now_ms() - tp;
         tp = now_ms();
         timing_mark((std::string("L") + std::to_string(l) + " " + "sampling").c_str(), tl);
         std::vector<int> ranks;
         run_mgs(mgs, Abuf->p, Qbuf->p, ranks, svdeps, s);
         timing_mark((std::string("L") + std::to_string(l) + " " + "pivoted MGS").c_str(), tl);
         ph[4] += now_ms() - tp;
         tp = now_ms();
         std::vector<int> next;
         for (size_t pi = 0; pi < pending.size(); pi++) {
            Prob& P = probs[pending[pi]];
            int r = ranks[pi];
            if (r < (layers * P.pass * P.nsamp) / 2) {
               P.rank = r;
               qsrc[pending[pi]] = Qbuf->p + mgs[pi].q_off;
            } else {
               P.pass++;
               next.push_back(pending[pi]);
            }
         }
         keep.push_back(Qbuf);
         pending.swap(next);
      }
      ph[3] += now_ms() - tp;
      tp = now_ms();
      // ranks into cubes; compact bases into the level buffers
      std::vector<int> vt_prob(cl.size(), -1), u_prob(cl.size(), -1);
      for (size_t i = 0; i < probs.size(); i++) {
         Cube& c = t.cubes[cl[probs[i].cube]];
         if (!probs[i].isU) {
            c.rowrank = probs[i].rank;
            vt_prob[probs[i].cube] = (int)i;
         } else {
            c.colrank = probs[i].rank;
            u_prob[probs[i].cube] = (int)i;
         }
      }
      long long vts = 0, pvs = 0, uts = 0, utis = 0;
      std::vector<CopyDesc> copies;
      for (size_t ci = 0; ci < cl.size(); ci++) {
         Cube& c = t.cubes[cl[ci]];
         if (c.np > 0 && c.rowrank > 0) {
            SrcDesc sd;
            sd.p0 = c.p0;
            sd.np = c.np;
            sd.r = c.rowrank;
            sd.spec_slot = (int)lv.h_src.size();
            sd.vt_off = vts;
            sd.pv_off = pvs;
            vts += (long long)c.rowrank * c.np;
            pvs += 2LL * c.rowrank * gp3;
            src_slot[cl[ci]] = sd.spec_slot;
            lv.h_src.push_back(sd);
            lv.maxr_src = std::max(lv.maxr_src, c.rowrank);
         }
         if (c.nq > 0 && c.colrank > 0) {
            TgtDesc td;
            td.q0 = c.q0;
            td.nq = c.nq;
            td.r = c.colrank;
            td.spec_slot = (int)lv.h_tgt.size();
            td.u_off = uts;
            td.uti_off = utis;
            uts += (long long)c.colrank * c.nq;
            utis += (long long)c.colrank * gp3;
            tgt_slot[cl[ci]] = td.spec_slot;
            lv.h_tgt.push_back(td);
            lv.maxr_tgt = std::max(lv.maxr_tgt, c.colrank);
         }
      }
      lv.VT64.alloc(std::max<long long>(vts, 1));
      lv.PV64.alloc(std::max<long long>(pvs, 1));
      lv.UT64.alloc(std::max<long long>(uts, 1));
      lv.UTI64.alloc(std::max<long long>(utis, 1));
      for (auto& sd : lv.h_src) {
         (void)sd;
      }
      {
         size_t si = 0, ti = 0;
         for (size_t ci = 0; ci < cl.size(); ci++) {
            Cube& c = t.cubes[cl[ci]];
            if (c.np > 0 && c.rowrank > 0) {
               const SrcDesc& sd = lv.h_src[si++];
               copies.push_back({qsrc[vt_prob[ci]], lv.VT64.p + sd.vt_off, (long long)sd.r * sd.np});
            }
            if (c.nq > 0 && c.colrank > 0) {
               const TgtDesc& td = lv.h_tgt[ti++];
               copies.push_back({qsrc[u_prob[ci]], lv.UT64.p + td.u_off, (long long)td.r * td.nq});
            }
         }
         DBuf<CopyDesc> dc;
         dc.upload(copies, s);
         for (size_t b0 = 0; b0 < copies.size(); b0 += 65535)
            launch_batched_copy((int)std::min<size_t>(65535, copies.size() - b0), dc.p + b0, s);
         FSVD_CHECK(hipStreamSynchronize(s));
      }
      keep.clear();

      timing_mark((std::string("L") + std::to_string(l) + " " + "compaction").c_str(), tl);
      // ---- equivalent densities: PV = P VT^T, UTI = U^T I (Cube.c:795-1002,1414-1522) ----
      level_pv_uti(t, cl, lv, ptsT, d_panels, d_pinv.p, kernel, kappa, lv.VT64.p, lv.UT64.p, lv.PV64.p, lv.UTI64.p, s);
      FSVD_CHECK(hipStreamSynchronize(s));
      ph[5] += now_ms() - tp;
      timing_mark((std::string("L") + std::to_string(l) + " " + "PV, UTI").c_str(), tl);
      lev_far[l - 2] = (long long)(lv.VT64.n + lv.PV64.n + lv.UT64.n + lv.UTI64.n);
      op.levels[l - 2] = std::move(lv);
   };
   {
      std::vector<std::exception_ptr> err(nlev);
      std::vector<std::thread> th;
      for (int l = 2; l < t.depth; l++)
         th.emplace_back([&, l]() {
            hipStream_t ls = nullptr;
            try {
               FSVD_CHECK(hipStreamCreate(&ls));
               build_level(l, ls);
               FSVD_CHECK(hipStreamSynchronize(ls));
            } catch (...) {
               err[l - 2] = std::current_exception();
            }
            if (ls) (void)hipStreamDestroy(ls);
         });
      for (auto& x : th) x.join();
      for (auto& e : err)
         if (e) std::rethrow_exception(e);
      for (int k = 0; k < nlev; k++) {
         for (int q = 0; q < 8; q++) op.stats.phase_ms[q] += lev_ph[k][q];
         op.stats.far_entries += lev_far[k];
      }
   }

   near_thread.t.join();
   if (near_err) std::rethrow_exception(near_err);
   op.stats.phase_ms[1] += op.stats.setup_ms_near;
   op.cube_src_slot = src_slot;
   op.cube_tgt_slot = tgt_slot;
   op.cube_leaf.assign(t.cubes.size(), -1);
   {
      size_t li = 0;
      for (size_t i = 0; i < t.cubes.size(); i++)
         if (t.cubes[i].leaf && t.cubes[i].nq > 0) op.cube_leaf[i] = (int)li++;
   }
   tp = now_ms();
   double tm = tp;
   // interaction lists (Cube_multiply_T, Cube.c:2062-2111) and their transposes
   for (FarLevel& lv : op.levels) {
      for (TgtDesc& td : lv.h_tgt) td.nint = 0;
      size_t ti = 0;
      for (size_t i = 0; i < t.cubes.size(); i++) {
         const Cube& c = t.cubes[i];
         if (c.level != lv.level || tgt_slot[i] < 0) continue;
         TgtDesc& td = lv.h_tgt[ti++];
         td.int_off = (int)lv.h_isrc.size();
         for (int o : c.itr)
            if (src_slot[o] >= 0) {
               lv.h_isrc.push_back(src_slot[o]);
               lv.h_itoff.push_back(offset_code(c.idx, t.cubes[o].idx));
            }
         td.nint = (int)lv.h_isrc.size() - td.int_off;
         op.stats.translate_macs += (long long)td.nint * lv.pad;
      }
      lv.nsrc = (int)lv.h_src.size();
      lv.ntgt = (int)lv.h_tgt.size();
      op.stats.spectra_bins += (long long)(lv.nsrc + lv.ntgt) * lv.pad;
   }
   timing_mark("interaction lists", tm);
   if (want32) {
      to_f32(op.D64, op.D32, s);
      for (FarLevel& lv : op.levels) {
         to_f32(lv.T64, lv.T32, s);
         to_f32(lv.VT64, lv.VT32, s);
         to_f32(lv.PV64, lv.PV32, s);
         to_f32(lv.UT64, lv.UT32, s);
         to_f32(lv.UTI64, lv.UTI32, s);
      }
      op.have32 = true;
   }
   FSVD_CHECK(hipStreamSynchronize(s));
   timing_mark("FP32 copies", tm);
   build_apply_plan(op, want32, s);
   FSVD_CHECK(hipStreamSynchronize(s));
   timing_mark("apply plan", tm);
   op.stats.phase_ms[6] += now_ms() - tp;
   op.stats.setup_ms_far = now_ms() - t0;
   return opp;
}

template <typename T>
static ApplyArgs make_args(TreeOp& op, const double* xs, const double* xd, double* y) {
   const bool f32 = std::is_same<T, float>::value;
   ApplyPlan& P = op.plan;
   ApplyArgs a;
   a.nchunk = P.nchunk;
   a.nsrc = P.nsrc;
   a.ntgt = P.ntgt;
   a.nsrc_big = P.nsrc_big;
   a.ntgt_big = P.ntgt_big;
   a.nleaf = op.nleaf;
   a.proj_maxcols = P.proj_maxcols;
   a.maxr_src = P.maxr_src;
   a.max_n = P.max_n;
   a.maxm = op.maxm;
   a.maxp = op.maxp;
   a.chunks = P.chunks.p;
   a.src = P.src.p;
   a.tgt = P.tgt.p;
   a.ntask_d = P.ntask_d;
   a.tr_max_src = P.tr_max_src;
   a.task_d = P.task_d.p;
   a.nblk_p = P.pipe.nblk;
   a.mode_p = P.pipe.mode;
   a.range_p = P.pipe.range.p;
   a.task_p = P.pipe.tasks.p;
   a.tiles_p = P.pipe.tiles.p;
   a.oct_p = P.pipe.octs.p;
   a.oacc_p = P.pipe.oacc.p;
   a.gsrc_p = P.pipe.gsrc.p;
   a.gint_p = P.pipe.gint.p;
   a.tiles = P.tiles.p;
   a.gtgt = P.gtgt.p;
   a.gsrc = P.gsrc.p;
   a.gint = P.gint.p;
   a.nmap_leaf = P.nmap_leaf;
   a.map_leaf = P.map_leaf.p;
   a.leaves = op.leaves.p;
   a.dcol = op.dcol.p;
   a.ublk = P.ublk.p;
   a.VTc = f32 ? (const void*)P.VTc32.p : (const void*)P.VTc64.p;
   a.PV = f32 ? (const void*)P.PV32.p : (const void*)P.PV64.p;
   a.Tsp = f32 ? (const void*)P.Tsp32.p : (const void*)P.Tsp64.p;
   a.UTI = f32 ? (const void*)P.UTI32.p : (const void*)P.UTI64.p;
   a.UT = f32 ? (const void*)P.UT32.p : (const void*)P.UT64.p;
   a.D = f32 ? (const void*)op.D32.p : (const void*)op.D64.p;
   a.tpart = P.tpart.p;
   a.spec = P.spec.p;
   a.acc = P.acc.p;
   a.w = P.w.p;
   a.xs = xs;
   a.xd = xd;
   a.y = y;
   return a;
}

template <typename T>
static void apply_impl(TreeOp& op, const double* xs, const double* xd, double* y, hipStream_t s) {
   const ApplyArgs a = make_args<T>(op, xs, xd, y);
   for (int stage = 0; stage < 5; stage++) launch_apply_stage<T>(stage, a, s);
}

template <typename T>
static void apply_part_impl(TreeOp& op, int part, const double* xs, const double* xd, double* y, hipStream_t s) {
   const ApplyArgs a = make_args<T>(op, xs, xd, y);
   switch (part) {
      case 0: launch_apply_stage<T>(4, a, s); break;
      case 1:
         launch_apply_stage<T>(0, a, s);
         launch_apply_stage<T>(1, a, s);
         break;
      case 2: launch_apply_stage<T>(2, a, s); break;
      case 3: launch_apply_stage<T>(3, a, s); break;
      default: break;
   }
}

void tree_apply_part(TreeOp& op, int prec, int part, const double* xs, const double* xd, double* y, hipStream_t s) {
   if (prec == PREC_F32)
      apply_part_impl<float>(op, part, xs, xd, y, s);
   else
      apply_part_impl<double>(op, part, xs, xd, y, s);
}

template <typename T>
static void leaf_split_impl(TreeOp& op, const double* xs, const double* xd, const int* out_idx, const double* corr,
                            double* yout, hipStream_t s) {
   ApplyArgs a = make_args<T>(op, xs, xd, nullptr);
   a.sp_out = out_idx;
   a.sp_corr = corr;
   a.sp_y = yout;
   launch_apply_stage<T>(4, a, s);
}

void tree_apply_leaf_split(TreeOp& op, int prec, const double* xs, const double* xd, const int* out_idx,
                           const double* corr, double* yout, hipStream_t s) {
   if (prec == PREC_F32)
      leaf_split_impl<float>(op, xs, xd, out_idx, corr, yout, s);
   else
      leaf_split_impl<double>(op, xs, xd, out_idx, corr, yout, s);
}

// Column copies of the work vectors.  The spectra buffers start zeroed (the
// bins between a spectrum's pad and its padded stride stay zero), as the
// single-column ones.
void tree_batch_reserve(TreeOp& op, int ncol, hipStream_t s) {
   ApplyPlan& P = op.plan;
   if (P.ncolB >= ncol) return;
   FSVD_CHECK(hipStreamSynchronize(s));   // the old buffers may be in use
   P.tpartB.alloc((size_t)ncol * P.tpart.n);
   P.specB.alloc((size_t)ncol * P.spec.n);
   P.accB.alloc((size_t)ncol * P.acc.n);
   P.wB.alloc((size_t)ncol * P.w.n);
   FSVD_CHECK(hipMemsetAsync(P.specB.p, 0, sizeof(double) * P.specB.n, s));
   FSVD_CHECK(hipMemsetAsync(P.accB.p, 0, sizeof(double) * P.accB.n, s));
   P.ncolB = ncol;
}

template <typename T>
static void apply_part_batch_impl(TreeOp& op, int part, int ncol, const double* xs, const double* xd, long long ldx,
                                  double* y, long long ldy, hipStream_t s) {
   ApplyPlan& P = op.plan;
   tree_batch_reserve(op, ncol, s);
   ApplyArgs a = make_args<T>(op, xs, xd, y);
   // strides in elements of each buffer's type; every single-column buffer
   // is sized for FP64, so these strides hold a column of either precision
   a.ncol = ncol;
   a.cs_x = ldx;
   a.cs_y = ldy;
   a.tpart = P.tpartB.p;
   a.cs_tpart = (long long)P.tpart.n;
   a.spec = P.specB.p;
   a.cs_spec = (long long)P.spec.n / 2;
   a.acc = P.accB.p;
   a.cs_acc = (long long)P.acc.n / 2;
   a.w = P.wB.p;
   a.cs_w = (long long)P.w.n;
   a.spec_delta = (long long)((const char*)P.specB.p - (const char*)P.spec.p);
   a.acc_delta = (long long)((const char*)P.accB.p - (const char*)P.acc.p);
   if (std::is_same<T, float>::value && ncol >= 2 && part == 2) {
      if (!P.pipe2.nblk && P.pipe.nblk) make_pipe(P.pipe2, {&op}, s, 0, true);
      a.nblk_p2 = P.pipe2.nblk;
      a.range_p2 = P.pipe2.range.p;
      a.task_p2 = P.pipe2.tasks.p;
      a.tiles_p2 = P.pipe2.tiles.p;
      a.oct_p2 = P.pipe2.octs.p;
      a.oacc_p2 = P.pipe2.oacc.p;
      a.gsrc_p2 = P.pipe2.gsrc.p;
      a.gint_p2 = P.pipe2.gint.p;
   }
   switch (part) {
      case 0: launch_apply_stage<T>(4, a, s); break;
      case 1:
         launch_apply_stage<T>(0, a, s);
         launch_apply_stage<T>(1, a, s);
         break;
      case 2: launch_apply_stage<T>(2, a, s); break;
      case 3: launch_apply_stage<T>(3, a, s); break;
      default: break;
   }
}

void tree_apply_part_batch(TreeOp& op, int prec, int part, int ncol, const double* xs, const double* xd,
                           long long ldx, double* y, long long ldy, hipStream_t s) {
   if (ncol < 1) throw std::runtime_error("column-batched apply: ncol must be positive");
   if (prec == PREC_F32) {
      if (!op.have32) throw std::runtime_error("operator has no FP32 copy");
      apply_part_batch_impl<float>(op, part, ncol, xs, xd, ldx, y, ldy, s);
   } else {
      apply_part_batch_impl<double>(op, part, ncol, xs, xd, ldx, y, ldy, s);
   }
}

void tree_apply(TreeOp& op, int prec, const double* xs, const double* xd, double* y, hipStream_t s) {
   if (prec == PREC_F32) {
      if (!op.have32) throw std::runtime_error("operator has no FP32 copy");
      apply_impl<float>(op, xs, xd, y, s);
   } else {
      apply_impl<double>(op, xs, xd, y, s);
   }
}

// The transposed apply's per-level descriptors and interaction lists
// (each source cube receives from the target cubes of its list), built and
// uploaded on the transposed apply's first use.
static void transpose_lists(TreeOp& op, hipStream_t s) {
   if (op.haveT) return;
   const Tree& t = op.tree;
   for (FarLevel& lv : op.levels) {
      lv.h_tgtT.assign(lv.h_src.size(), TgtDesc{});
      lv.h_isrcT.clear();
      lv.h_itoffT.clear();
      size_t si = 0;
      for (size_t i = 0; i < t.cubes.size(); i++) {
         const Cube& c = t.cubes[i];
         if (c.level != lv.level || op.cube_src_slot[i] < 0) continue;
         TgtDesc& td = lv.h_tgtT[si];
         std::memset(&td, 0, sizeof(td));
         td.spec_slot = op.cube_src_slot[i];
         td.int_off = (int)lv.h_isrcT.size();
         for (int o : c.itr)
            if (op.cube_tgt_slot[o] >= 0) {
               lv.h_isrcT.push_back(op.cube_tgt_slot[o]);
               lv.h_itoffT.push_back(offset_code(c.idx, t.cubes[o].idx));
            }
         td.nint = (int)lv.h_isrcT.size() - td.int_off;
         si++;
      }
      lv.src.upload(lv.h_src, s);
      lv.tgt.upload(lv.h_tgt, s);
      lv.tgtT.upload(lv.h_tgtT, s);
      lv.isrcT.upload(lv.h_isrcT, s);
      lv.itoffT.upload(lv.h_itoffT, s);
   }
   FSVD_CHECK(hipStreamSynchronize(s));
   op.haveT = true;
}

// y2 set: both layers from one pass (single layer into y, double into y2)
template <typename T>
static void applyT_impl(TreeOp& op, int layer, const double* x, double* y, hipStream_t s, double* y2 = nullptr) {
   const bool f32 = std::is_same<T, float>::value;
   transpose_lists(op, s);
   for (FarLevel& lv : op.levels)
      if (!lv.specT.p) {
         lv.specT.alloc(2 * (size_t)std::max(1, lv.ntgt) * lv.pad);
         lv.accT.alloc(2 * (size_t)std::max(1, lv.nsrc) * lv.pad);
      }
   launch_zero(y, op.tree.npan, s);
   if (y2) launch_zero(y2, op.tree.npan, s);
   for (FarLevel& lv : op.levels)
      if (lv.ntgt)
         launch_farT_source<T>(lv.n, lv.ntgt, lv.tgt.p, f32 ? (const T*)lv.UT32.p : (const T*)lv.UT64.p,
                               f32 ? (const T*)lv.UTI32.p : (const T*)lv.UTI64.p, x, lv.specT.p, lv.maxr_tgt, s);
   for (FarLevel& lv : op.levels)
      if (lv.nsrc)
         launch_far_translate<T>(lv.nsrc, lv.tgtT.p, lv.isrcT.p, lv.itoffT.p,
                                 f32 ? (const void*)lv.T32.p : (const void*)lv.T64.p, lv.specT.p, lv.accT.p, lv.pad,
                                 s);
   for (FarLevel& lv : op.levels)
      if (lv.nsrc)
         launch_farT_target<T>(lv.n, lv.nsrc, lv.src.p, f32 ? (const T*)lv.VT32.p : (const T*)lv.VT64.p,
                               f32 ? (const T*)lv.PV32.p : (const T*)lv.PV64.p, lv.accT.p, layer, y, y2, lv.maxr_src,
                               s);
   if (op.nleaf)
      launch_nearT<T>(op.nleaf, op.leaves.p, op.dcol.p, f32 ? (const T*)op.D32.p : (const T*)op.D64.p, x, layer, y, y2,
                      s);
}

void tree_apply_transpose_both(TreeOp& op, int prec, const double* x, double* ys, double* yd, hipStream_t s) {
   if (prec == PREC_F32) {
      if (!op.have32) throw std::runtime_error("operator has no FP32 copy");
      applyT_impl<float>(op, 0, x, ys, s, yd);
   } else {
      applyT_impl<double>(op, 0, x, ys, s, yd);
   }
}

void tree_apply_transpose(TreeOp& op, int prec, int layer, const double* x, double* y, hipStream_t s) {
   if (prec == PREC_F32) {
      if (!op.have32) throw std::runtime_error("operator has no FP32 copy");
      applyT_impl<float>(op, layer, x, y, s);
   } else {
      applyT_impl<double>(op, layer, x, y, s);
   }
}

}  // namespace fsvd

namespace fsvd {

// Derivative of the Moore-Penrose pseudoinverse Ap (n x m) of A (m x n,
// both row-major) along dA, at constant rank (Golub and Pereyra 1973):
// dAp = -Ap dA Ap + Ap Ap^T dA^T (I - A Ap) + (I - Ap A) dA^T Ap^T Ap.
static void pinv_derivative(const double* A, const double* dA, const double* Ap, int m, int n, double* dAp) {
   auto mul = [](const std::vector<double>& X, const std::vector<double>& Y, int r, int k, int c) {
      std::vector<double> Z((size_t)r * c, 0.0);
      for (int i = 0; i < r; i++)
         for (int l = 0; l < k; l++) {
            const double x = X[(size_t)i * k + l];
            if (x == 0.0) continue;
            for (int j = 0; j < c; j++) Z[(size_t)i * c + j] += x * Y[(size_t)l * c + j];
         }
      return Z;
   };
   auto tr = [](const std::vector<double>& X, int r, int c) {
      std::vector<double> Z((size_t)r * c);
      for (int i = 0; i < r; i++)
         for (int j = 0; j < c; j++) Z[(size_t)j * r + i] = X[(size_t)i * c + j];
      return Z;
   };
   const std::vector<double> vA(A, A + (size_t)m * n), vdA(dA, dA + (size_t)m * n), vAp(Ap, Ap + (size_t)n * m);
   const std::vector<double> dAT = tr(vdA, m, n), ApT = tr(vAp, n, m);
   // -Ap dA Ap
   std::vector<double> t1 = mul(mul(vAp, vdA, n, m, n), vAp, n, n, m);
   // Ap Ap^T dA^T (I - A Ap)
   std::vector<double> R = mul(vA, vAp, m, n, m);
   for (int i = 0; i < m; i++)
      for (int j = 0; j < m; j++) R[(size_t)i * m + j] = (i == j ? 1.0 : 0.0) - R[(size_t)i * m + j];
   std::vector<double> t2 = mul(mul(vAp, ApT, n, m, n), mul(dAT, R, n, m, m), n, n, m);
   // (I - Ap A) dA^T Ap^T Ap
   std::vector<double> Pn = mul(vAp, vA, n, m, n);
   for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) Pn[(size_t)i * n + j] = (i == j ? 1.0 : 0.0) - Pn[(size_t)i * n + j];
   std::vector<double> t3 = mul(Pn, mul(dAT, mul(ApT, vAp, m, n, m), n, m, m), n, n, m);
   for (size_t k = 0; k < (size_t)n * m; k++) dAp[k] = -t1[k] + t2[k] + t3[k];
}

void launch_axpby(long long n, double a, const double* x, double b, const double* y, double* z, hipStream_t s);

// The kappa-derivative of the compressed operator as it is applied, with
// the tree, ranks and bases (VT, U) held fixed: the near-field blocks of
// the differentiated integrals; per level P' = pinv' Q2S + pinv Q2S' and
// I' = (pinv' Q2Sp + pinv Q2Sp')^T, so PV' = VT P'^T and UTI' = U^T I'
// (level_pv_uti with pinv' and the Yukawa integrals plus level_pv_uti with
// pinv and the differentiated ones); the spectra of -e^{-kr}.  pinv' of the
// sphere -> grid matrix Q follows from Q' by the pseudoinverse derivative.
void build_dkappa(TreeOp& op, const PanelGeom* d_panels, hipStream_t s) {
   if (op.kernel != YUKAWA) throw std::runtime_error("kappa derivatives exist for Yukawa operators only");
   if (op.dk && op.dk->kappa == op.kappa) return;
   auto dk = std::make_unique<DkData>();
   dk->kappa = op.kappa;
   const double kappa = op.kappa;
   const Tree& t = op.tree;
   const ApplyPlan& P = op.plan;
   std::vector<double> ptsT(op.pts_tree.n);
   FSVD_CHECK(hipMemcpyAsync(ptsT.data(), op.pts_tree.p, sizeof(double) * ptsT.size(), hipMemcpyDeviceToHost, s));
   FSVD_CHECK(hipStreamSynchronize(s));
   // near field, in the layout of D64
   dk->D64.alloc(std::max<size_t>(op.D64.n, 1));
   {
      std::vector<int> bpan(op.h_dcol.size());
      for (size_t k = 0; k < bpan.size(); k++) bpan[k] = t.pan_perm[op.h_dcol[k]];
      std::vector<IntBlock> blocks;
      for (const LeafDesc& L : op.h_leaves) {
         if (L.m == 0) continue;
         IntBlock b;
         b.npt = L.p;
         b.npan = L.m;
         b.pt_off = L.q0;
         b.pan_off = L.col_off;
         b.out_off = L.d_off;
         b.ld_pt = 1;
         b.ld_pan = 2 * L.p;
         b.layer_stride = L.p;
         push_blocks(blocks, b);
      }
      run_integrals(blocks, ptsT, bpan, d_panels, YUKAWA_DKAPPA, kappa, dk->D64.p, s);
   }
   // far field, in the layouts of the plan's PV64, Tsp64 and UTI64
   dk->PV64.alloc(std::max<size_t>(P.PV64.n, 1));
   dk->Tsp64.alloc(std::max<size_t>(P.Tsp64.n, 2));
   FSVD_CHECK(hipMemsetAsync(dk->Tsp64.p, 0, sizeof(double) * dk->Tsp64.n, s));
   dk->UTI64.alloc(std::max<size_t>(P.UTI64.n, 1));
   const double root = t.cubes[0].b1[0] - t.cubes[0].b0[0];
   long long pv_b = 0, t_b = 0, uti_b = 0;
   for (const FarLevel& lv : op.levels) {
      const int l = lv.level, n = lv.n, gp3 = lv.gp3;
      std::vector<int> cl;
      for (size_t i = 0; i < t.cubes.size(); i++)
         if (t.cubes[i].level == l && !t.cubes[i].itr.empty()) cl.push_back((int)i);
      // Q (sphere -> grid, as in the build) and Q'
      const double box = root / std::pow(2.0, l), h = box / (n - 1), sc = box * 1.5 * 1, ctr = 0.5 * box;
      std::vector<double> Q((size_t)NQ * gp3), dQ((size_t)NQ * gp3), dpinv((size_t)gp3 * NQ);
      for (int q = 0; q < NQ; q++) {
         double mq[3];
         for (int a = 0; a < 3; a++) mq[a] = SPHERE49[q][a] * sc + ctr;
         int g = 0;
         for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++)
               for (int k = 0; k < n; k++, g++) {
                  double p0 = i * h - mq[0], p1 = j * h - mq[1], p2 = k * h - mq[2];
                  double r = std::sqrt(p0 * p0 + p1 * p1 + p2 * p2);
                  Q[(size_t)q * gp3 + g] = std::exp(-kappa * r) / r;
                  dQ[(size_t)q * gp3 + g] = -std::exp(-kappa * r);
               }
      }
      pinv_derivative(Q.data(), dQ.data(), op.h_pinv[l].data(), NQ, gp3, dpinv.data());
      DBuf<double> d_pinv, d_dpinv, pva, utia, pvb, utib;
      d_pinv.upload(op.h_pinv[l], s);
      d_dpinv.upload(dpinv, s);
      pva.alloc(std::max<size_t>(lv.PV64.n, 1));
      pvb.alloc(std::max<size_t>(lv.PV64.n, 1));
      utia.alloc(std::max<size_t>(lv.UTI64.n, 1));
      utib.alloc(std::max<size_t>(lv.UTI64.n, 1));
      level_pv_uti(t, cl, lv, ptsT, d_panels, d_dpinv.p, YUKAWA, kappa, lv.VT64.p, lv.UT64.p, pva.p, utia.p, s);
      level_pv_uti(t, cl, lv, ptsT, d_panels, d_pinv.p, YUKAWA_DKAPPA, kappa, lv.VT64.p, lv.UT64.p, pvb.p, utib.p,
                   s);
      if (lv.PV64.n) launch_axpby((long long)lv.PV64.n, 1.0, pva.p, 1.0, pvb.p, dk->PV64.p + pv_b, s);
      if (lv.UTI64.n) launch_axpby((long long)lv.UTI64.n, 1.0, utia.p, 1.0, utib.p, dk->UTI64.p + uti_b, s);
      {
         std::vector<std::complex<double>> T((size_t)343 * lv.pad);
         translation_spectra(n, h, YUKAWA_DKAPPA, kappa, T.data());
         DBuf<double> dT;
         dT.upload((const double*)T.data(), 2 * T.size(), s);
         FSVD_CHECK(hipMemcpy2DAsync(dk->Tsp64.p + t_b, sizeof(double) * 2 * tr_pst(lv.pad), dT.p,
                                     sizeof(double) * 2 * lv.pad, sizeof(double) * 2 * lv.pad, 343,
                                     hipMemcpyDeviceToDevice, s));
         FSVD_CHECK(hipStreamSynchronize(s));
      }
      pv_b += (long long)lv.PV64.n;
      t_b += 2LL * 343 * tr_pst(lv.pad);
      uti_b += (long long)lv.UTI64.n;
   }
   FSVD_CHECK(hipStreamSynchronize(s));
   op.dk = std::move(dk);
}

// The far field U UTI T PV VT is a product of three kappa-dependent
// factors: its derivative is U (UTI' T PV + UTI T' PV + UTI T PV') VT.
// Each term is one apply with one factor swapped (the near field D' in
// each), and the near field alone (the leaf stage with the far-field
// weights zeroed) is taken out twice: y = y1 + y2 + y3 - 2 D'x.
void tree_apply_dkappa(TreeOp& op, const double* xs, const double* xd, double* y, hipStream_t s) {
   if (!op.dk) throw std::runtime_error("build_dkappa first");
   const long long npt = (long long)op.tree.npt;
   DBuf<double> yk, y0;
   yk.alloc(std::max<long long>(3 * npt, 1));
   y0.alloc(std::max<long long>(npt, 1));
   for (int term = 0; term < 3; term++) {
      ApplyArgs a = make_args<double>(op, xs, xd, yk.p + term * npt);
      a.D = op.dk->D64.p;
      if (term == 0) a.UTI = op.dk->UTI64.p;
      if (term == 1) a.Tsp = op.dk->Tsp64.p;
      if (term == 2) a.PV = op.dk->PV64.p;
      for (int stage = 0; stage < 5; stage++) launch_apply_stage<double>(stage, a, s);
   }
   {
      ApplyArgs a = make_args<double>(op, xs, xd, y0.p);
      a.D = op.dk->D64.p;
      FSVD_CHECK(hipMemsetAsync(op.plan.w.p, 0, sizeof(double) * op.plan.w.n, s));
      launch_apply_stage<double>(4, a, s);
   }
   launch_axpby(npt, 1.0, yk.p, 1.0, yk.p + npt, y, s);
   launch_axpby(npt, 1.0, y, 1.0, yk.p + 2 * npt, y, s);
   launch_axpby(npt, 1.0, y, -2.0, y0.p, y, s);
   FSVD_CHECK(hipStreamSynchronize(s));
}

}  // namespace fsvd
