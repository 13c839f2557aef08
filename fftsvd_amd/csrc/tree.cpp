// Octree construction on the host.  The partition decisions follow the
// reference exactly (Tree_allocate, Tree.c:12-178; Cube_allocate,
// Cube.c:352-484; Cube_lists, Cube.c:600-713) because they fix which
// integrals are near-field; the storage is this library's own: cubes in a
// flat array (parents first) and panels/points renumbered into "tree order"
// so that every cube owns contiguous index ranges.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>

#include "internal.h"

namespace fsvd {

namespace {

// fn(a, b) over [0, n) in contiguous chunks on up to 16 host threads, at
// least `grain` items per thread; each item is handled by exactly one call,
// so results do not depend on the split.
template <typename F>
void tree_parallel_for(long long n, F&& fn, long long grain = 16384) {
   int nth = (int)std::min<long long>({16LL, (long long)std::max(1u, std::thread::hardware_concurrency()),
                                        (n + grain - 1) / grain});
   if (nth <= 1) {
      if (n > 0) fn(0LL, n);
      return;
   }
   std::vector<std::thread> th;
   const long long chunk = (n + nth - 1) / nth;
   for (int t = 1; t < nth; t++) {
      const long long a = t * chunk, b = std::min(n, a + chunk);
      if (a < b) th.emplace_back([&fn, a, b] { fn(a, b); });
   }
   fn(0LL, std::min(n, chunk));
   for (auto& x : th) x.join();
}

struct Builder {
   Tree& t;
   const double* P;   // panel centroids, 3 per panel
   const double* pts;
   // temporary per-cube sorted index lists
   std::vector<std::vector<int>> pan, pt;

   explicit Builder(Tree& tt, const double* p, const double* q) : t(tt), P(p), pts(q) {}

   int build(std::vector<int>&& mypan, std::vector<int>&& mypt, int level, const int idx[3], const double b0[3],
             const double b1[3], int parent) {
      int id = (int)t.cubes.size();
      t.cubes.emplace_back();
      pan.emplace_back(std::move(mypan));
      pt.emplace_back(std::move(mypt));
      {
         Cube& c = t.cubes[id];
         c.level = level;
         for (int d = 0; d < 3; d++) {
            c.idx[d] = idx[d];
            c.b0[d] = b0[d];
            c.b1[d] = b1[d];
            double s = b0[d] + b1[d];
            c.ctr[d] = s * 0.5;
         }
         c.parent = parent;
         for (int k = 0; k < 8; k++) c.child[k] = -1;
         c.leaf = (level == t.depth - 1);
      }
      if (t.cubes[id].leaf) return id;
      double ctr[3] = {t.cubes[id].ctr[0], t.cubes[id].ctr[1], t.cubes[id].ctr[2]};
      // child k = 4*cx + 2*cy + cz with c* = 0 for v < ctr, 1 for v >= ctr
      // (a coordinate that compares neither way joins no child)
      std::vector<int> bp[8], bq[8];
      auto child_of = [&](const double* g) {
         int k = 0;
         for (int d = 0; d < 3; d++) {
            if (g[d] >= ctr[d]) k |= 4 >> d;
            else if (!(g[d] < ctr[d])) return -1;
         }
         return k;
      };
      for (int i : pan[id]) {
         int k = child_of(P + 3 * (size_t)i);
         if (k >= 0) bp[k].push_back(i);
      }
      for (int i : pt[id]) {
         int k = child_of(pts + 3 * (size_t)i);
         if (k >= 0) bq[k].push_back(i);
      }
      for (int cx = 0; cx <= 1; cx++)
         for (int cy = 0; cy <= 1; cy++)
            for (int cz = 0; cz <= 1; cz++) {
               std::vector<int>& cp = bp[cx * 4 + cy * 2 + cz];
               std::vector<int>& cq = bq[cx * 4 + cy * 2 + cz];
               if (cp.empty() && cq.empty()) continue;
               double cb0[3] = {b0[0], b0[1], b0[2]}, cb1[3] = {ctr[0], ctr[1], ctr[2]};
               int sel[3] = {cx, cy, cz};
               for (int d = 0; d < 3; d++)
                  if (sel[d]) {
                     double diff = ctr[d] - cb0[d];
                     cb0[d] += diff;
                     cb1[d] += diff;
                  }
               int cidx[3] = {2 * idx[0] + cx, 2 * idx[1] + cy, 2 * idx[2] + cz};
               int ch = build(std::move(cp), std::move(cq), level + 1, cidx, cb0, cb1, id);
               t.cubes[id].child[cx * 4 + cy * 2 + cz] = ch;
            }
      return id;
   }
};

void make_lists(Tree& t) {
   // compact copies of the fields the lists read
   const int nc = (int)t.cubes.size();
   std::vector<int> idx(3 * (size_t)nc), child(8 * (size_t)nc);
   for (int i = 0; i < nc; i++) {
      for (int d = 0; d < 3; d++) idx[3 * (size_t)i + d] = t.cubes[i].idx[d];
      for (int k = 0; k < 8; k++) child[8 * (size_t)i + k] = t.cubes[i].child[k];
   }
   // a cube's lists read its parent's local list: level by level, the
   // cubes of one level in parallel
   int maxlev = 0;
   for (const Cube& c : t.cubes) maxlev = std::max(maxlev, c.level);
   std::vector<std::vector<int>> bylev(maxlev + 1);
   for (int id = 0; id < nc; id++) bylev[t.cubes[id].level].push_back(id);
   for (const std::vector<int>& ids : bylev)
      tree_parallel_for((long long)ids.size(), [&](long long a, long long b) {
         int near[8 * 27], far[8 * 27];
         for (long long r = a; r < b; r++) {
      const int id = ids[r];
      Cube& c = t.cubes[id];
      c.loc.clear();
      c.itr.clear();
      if (c.parent < 0) continue;
      const int* pch = &child[8 * (size_t)c.parent];
      const int x = idx[3 * (size_t)id], y = idx[3 * (size_t)id + 1], z = idx[3 * (size_t)id + 2];
      // siblings, then the parent's neighbours' children in list order:
      // adjacent ones are local, the rest interact
      int nn = 0, nf = 0;
      for (int k = 0; k < 8; k++)
         if (pch[k] >= 0 && pch[k] != id) near[nn++] = pch[k];
      for (int l : t.cubes[c.parent].loc) {
         const int* lch = &child[8 * (size_t)l];
         for (int k = 0; k < 8; k++) {
            const int o = lch[k];
            if (o < 0) continue;
            const int* oi = &idx[3 * (size_t)o];
            if (std::abs(oi[0] - x) <= 1 && std::abs(oi[1] - y) <= 1 && std::abs(oi[2] - z) <= 1)
               near[nn++] = o;
            else
               far[nf++] = o;
         }
      }
      c.loc.assign(near, near + nn);
      c.itr.assign(far, far + nf);
         }
      });
}

// Level-synchronous construction of the same tree on flat arrays.  Every
// panel and point carries its cell bounds and a cell code (3 bits per level,
// child k = 4*cx + 2*cy + cz, the Builder's order); one level deeper is the
// Builder's centroid comparison and bound update per entity, then a stable
// 8-way regroup inside each cell, so each cell's entities stay in ascending
// index order -- exactly the Builder's per-cube lists -- and cells in code
// order are the Builder's depth-first leaf order.  The depth loop adds one
// level per step instead of rebuilding.  Returns false (tree untouched) for
// non-finite coordinates or a depth past MAXLEV; the Builder then decides.
class LevelTree {
 public:
   static constexpr int MAXLEV = 20;

   LevelTree(const double* pan_cent, int npan, const double* pts, int npt) : npan_(npan), npt_(npt) {
      const int ne = npan + npt;
      x_.resize(3 * (size_t)ne);
      if (npan) std::memcpy(x_.data(), pan_cent, sizeof(double) * 3 * (size_t)npan);
      if (npt) std::memcpy(&x_[3 * (size_t)npan], pts, sizeof(double) * 3 * (size_t)npt);
   }

   bool build(Tree& t, const double b0[3], const double b1[3], int maxpanels) {
      const int ne = npan_ + npt_;
      lo_.resize(3 * (size_t)ne);
      hi_.resize(3 * (size_t)ne);
      code_.resize(ne);
      ord_.assign(1, std::vector<int>(ne));
      std::atomic<bool> finite{true};
      tree_parallel_for(ne, [&](long long a, long long b) {
         for (long long e = a; e < b; e++) {
            for (int d = 0; d < 3; d++) {
               if (!std::isfinite(x_[3 * (size_t)e + d])) finite = false;
               lo_[3 * (size_t)e + d] = b0[d];
               hi_[3 * (size_t)e + d] = b1[d];
            }
            code_[e] = 0;
            ord_[0][e] = (int)e;
         }
      });
      if (!finite) return false;
      cells_.assign(1, {});
      cells_[0].push_back({0, 0, npan_, npan_, ne});
      int depth = 2;
      for (;; depth++) {
         if (depth - 1 > MAXLEV) return false;
         while ((int)ord_.size() < depth)
            if (!deepen()) return false;
         if (npan_ == 0 || npt_ == 0) break;
         const std::vector<Cell>& leaves = cells_[depth - 1];
         int within = 0;
         for (const Cell& c : leaves)
            if (c.pb - c.pa <= maxpanels) within++;
         if ((double)within / (double)leaves.size() > 0.90) break;
      }
      emit(t, depth, b0, b1);
      return true;
   }

 private:
   struct Cell {
      uint64_t code;
      int pa, pb;   // panel range in this level's order
      int qa, qb;   // point range (offset by npan)
   };

   // one level deeper for every entity, regroup, list the cells
   bool deepen() {
      const int ne = npan_ + npt_;
      std::atomic<bool> ok{true};
      tree_parallel_for(ne, [&](long long a, long long b) {
         for (long long e = a; e < b; e++) {
            int k = 0;
            for (int d = 0; d < 3; d++) {
               double& lo = lo_[3 * (size_t)e + d];
               double& hi = hi_[3 * (size_t)e + d];
               const double s = lo + hi;
               const double ctr = s * 0.5;
               const double g = x_[3 * (size_t)e + d];
               if (g >= ctr) {
                  const double diff = ctr - lo;
                  lo += diff;
                  hi = ctr + diff;
                  k |= 4 >> d;
               } else if (g < ctr) {
                  hi = ctr;
               } else {
                  ok = false;
                  return;
               }
            }
            code_[e] = code_[e] << 3 | (uint64_t)k;
         }
      });
      if (!ok) return false;
      const std::vector<int>& prev = ord_.back();
      const std::vector<Cell>& pcells = cells_.back();
      const size_t npc = pcells.size();
      std::vector<int> next(ne);
      // the panels, then the points, of each previous cell (a run of one
      // parent code) regrouped by child octant, stably
      auto regroup = [&](int a, int b) {
         int cnt[9] = {0};
         for (int r = a; r < b; r++) cnt[(code_[prev[r]] & 7) + 1]++;
         for (int k = 0; k < 8; k++) cnt[k + 1] += cnt[k];
         for (int r = a; r < b; r++) next[a + cnt[code_[prev[r]] & 7]++] = prev[r];
      };
      // each previous cell's child cells: merge of its distinct panel and
      // point codes (the previous cells are in code order, so the
      // concatenation is in code order)
      std::vector<std::array<Cell, 8>> kids(npc);
      std::vector<int> nkid(npc + 1, 0);
      tree_parallel_for((long long)npc, [&](long long a, long long b) {
         for (long long pc = a; pc < b; pc++) {
            const Cell& P = pcells[pc];
            regroup(P.pa, P.pb);
            regroup(P.qa, P.qb);
            int i = P.pa, j = P.qa, m = 0;
            while (i < P.pb || j < P.qb) {
               uint64_t c;
               if (j >= P.qb) c = code_[next[i]];
               else if (i >= P.pb) c = code_[next[j]];
               else c = std::min(code_[next[i]], code_[next[j]]);
               Cell cell{c, i, i, j, j};
               while (i < P.pb && code_[next[i]] == c) i++;
               while (j < P.qb && code_[next[j]] == c) j++;
               cell.pb = i;
               cell.qb = j;
               kids[pc][m++] = cell;
            }
            nkid[pc + 1] = m;
         }
      }, 1);
      for (size_t pc = 0; pc < npc; pc++) nkid[pc + 1] += nkid[pc];
      std::vector<Cell> cells(nkid[npc]);
      for (size_t pc = 0; pc < npc; pc++)
         std::copy(kids[pc].begin(), kids[pc].begin() + (nkid[pc + 1] - nkid[pc]), cells.begin() + nkid[pc]);
      ord_.push_back(std::move(next));
      cells_.push_back(std::move(cells));
      return true;
   }

   // cubes in the Builder's depth-first order, lists and tree order
   void emit(Tree& t, int depth, const double b0[3], const double b1[3]) {
      t.depth = depth;
      t.cubes.clear();
      const int L = depth - 1;
      std::vector<size_t> cursor(L + 1, 0);
      std::vector<std::pair<int, int>> where;   // per cube: level, cell
      const int zero[3] = {0, 0, 0};
      add_cube(t, 0, zero, b0, b1, -1, L, cursor, where);
      const int nc = (int)t.cubes.size();
      t.pan_perm.resize(npan_);
      t.pt_perm.resize(npt_);
      for (int i = 0; i < npan_; i++) t.pan_perm[i] = ord_[L][i];
      for (int i = 0; i < npt_; i++) t.pt_perm[i] = ord_[L][npan_ + i] - npan_;
      t.cube_pan_sorted_off.assign(nc + 1, 0);
      t.cube_pt_sorted_off.assign(nc + 1, 0);
      for (int id = 0; id < nc; id++) {
         const Cell& c = cells_[where[id].first][where[id].second];
         Cube& cb = t.cubes[id];
         cb.np = c.pb - c.pa;
         cb.nq = c.qb - c.qa;
         if (cb.leaf) {
            cb.p0 = c.pa;
            cb.q0 = c.qa - npan_;
         }
         t.cube_pan_sorted_off[id + 1] = t.cube_pan_sorted_off[id] + cb.np;
         t.cube_pt_sorted_off[id + 1] = t.cube_pt_sorted_off[id] + cb.nq;
      }
      t.cube_pan_sorted.resize(t.cube_pan_sorted_off[nc]);
      t.cube_pt_sorted.resize(t.cube_pt_sorted_off[nc]);
      for (int id = 0; id < nc; id++) {
         const std::vector<int>& o = ord_[where[id].first];
         const Cell& c = cells_[where[id].first][where[id].second];
         std::copy(o.begin() + c.pa, o.begin() + c.pb, t.cube_pan_sorted.begin() + t.cube_pan_sorted_off[id]);
         int* q = t.cube_pt_sorted.data() + t.cube_pt_sorted_off[id];
         for (int r = c.qa; r < c.qb; r++) *q++ = o[r] - npan_;
      }
   }

   // the cube of cell cursor[level] (the next unvisited cell of that level)
   int add_cube(Tree& t, int level, const int idx[3], const double b0[3], const double b1[3], int parent, int L,
                std::vector<size_t>& cursor, std::vector<std::pair<int, int>>& where) {
      const int id = (int)t.cubes.size();
      const int cell = (int)cursor[level]++;
      const uint64_t code = cells_[level][cell].code;
      t.cubes.emplace_back();
      where.push_back({level, cell});
      {
         Cube& c = t.cubes[id];
         c.level = level;
         for (int d = 0; d < 3; d++) {
            c.idx[d] = idx[d];
            c.b0[d] = b0[d];
            c.b1[d] = b1[d];
            double s = b0[d] + b1[d];
            c.ctr[d] = s * 0.5;
         }
         c.parent = parent;
         for (int k = 0; k < 8; k++) c.child[k] = -1;
         c.leaf = (level == L);
      }
      if (level == L) return id;
      const double ctr[3] = {t.cubes[id].ctr[0], t.cubes[id].ctr[1], t.cubes[id].ctr[2]};
      const std::vector<Cell>& next = cells_[level + 1];
      while (cursor[level + 1] < next.size() && next[cursor[level + 1]].code >> 3 == code) {
         const int k = (int)(next[cursor[level + 1]].code & 7);
         const int sel[3] = {(k >> 2) & 1, (k >> 1) & 1, k & 1};
         double cb0[3] = {b0[0], b0[1], b0[2]}, cb1[3] = {ctr[0], ctr[1], ctr[2]};
         for (int d = 0; d < 3; d++)
            if (sel[d]) {
               double diff = ctr[d] - cb0[d];
               cb0[d] += diff;
               cb1[d] += diff;
            }
         const int cidx[3] = {2 * idx[0] + sel[0], 2 * idx[1] + sel[1], 2 * idx[2] + sel[2]};
         const int ch = add_cube(t, level + 1, cidx, cb0, cb1, id, L, cursor, where);
         t.cubes[id].child[k] = ch;
      }
      return id;
   }

   int npan_, npt_;
   std::vector<double> x_, lo_, hi_;
   std::vector<uint64_t> code_;
   std::vector<std::vector<int>> ord_;    // per level: panels then points, cell order
   std::vector<std::vector<Cell>> cells_;
};

}  // namespace

void build_tree(Tree& t, const double* pan_cent, int npan, const double* pts, int npt, int maxpanels,
                int gridpoints) {
   t.npan = npan;
   t.npt = npt;
   t.maxpanels = maxpanels;
   t.gridpoints = gridpoints;
   double mn[3] = {1000000.0, 1000000.0, 1000000.0}, mx[3] = {-1000000.0, -1000000.0, -1000000.0};
   for (int i = 0; i < npan; i++)
      for (int d = 0; d < 3; d++) {
         if (pan_cent[3 * (size_t)i + d] < mn[d]) mn[d] = pan_cent[3 * (size_t)i + d];
         if (pan_cent[3 * (size_t)i + d] > mx[d]) mx[d] = pan_cent[3 * (size_t)i + d];
      }
   for (int i = 0; i < npt; i++)
      for (int d = 0; d < 3; d++) {
         if (pts[3 * (size_t)i + d] < mn[d]) mn[d] = pts[3 * (size_t)i + d];
         if (pts[3 * (size_t)i + d] > mx[d]) mx[d] = pts[3 * (size_t)i + d];
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
   if (LevelTree(pan_cent, npan, pts, npt).build(t, b0, b1, maxpanels)) {
      t.gpl.assign(t.depth, 0);
      t.gpl[t.depth - 1] = gridpoints;
      if (t.depth >= 2) t.gpl[t.depth - 2] = gridpoints;
      for (int i = t.depth - 3, p = 1; i >= 0; i--, p++) t.gpl[i] = gridpoints + p;
      make_lists(t);
      for (int i = (int)t.cubes.size() - 1; i >= 0; i--) {
         Cube& c = t.cubes[i];
         if (c.leaf) continue;
         int p0 = -1, q0 = -1;
         for (int k = 0; k < 8; k++) {
            int ch = c.child[k];
            if (ch < 0) continue;
            if (p0 < 0) p0 = t.cubes[ch].p0;
            if (q0 < 0) q0 = t.cubes[ch].q0;
         }
         c.p0 = p0 < 0 ? 0 : p0;
         c.q0 = q0 < 0 ? 0 : q0;
      }
      return;
   }
   int zero[3] = {0, 0, 0};
   std::vector<int> all_pan(npan), all_pt(npt);
   for (int i = 0; i < npan; i++) all_pan[i] = i;
   for (int i = 0; i < npt; i++) all_pt[i] = i;

   std::vector<std::vector<int>> pan_lists, pt_lists;
   for (t.depth = 2;; t.depth++) {
      t.cubes.clear();
      Builder b(t, pan_cent, pts);
      b.build(std::vector<int>(all_pan), std::vector<int>(all_pt), 0, zero, b0, b1, -1);
      pan_lists.swap(b.pan);
      pt_lists.swap(b.pt);
      if (npan == 0 || npt == 0) break;
      int nleaf = 0, within = 0;
      for (size_t i = 0; i < t.cubes.size(); i++)
         if (t.cubes[i].leaf) {
            nleaf++;
            if ((int)pan_lists[i].size() <= maxpanels) within++;
         }
      if ((double)within / (double)nleaf > 0.90) break;
   }
   t.gpl.assign(t.depth, 0);
   t.gpl[t.depth - 1] = gridpoints;
   if (t.depth >= 2) t.gpl[t.depth - 2] = gridpoints;
   for (int i = t.depth - 3, p = 1; i >= 0; i--, p++) t.gpl[i] = gridpoints + p;
   make_lists(t);

   // tree order: leaves in creation (depth-first) order, each leaf's sorted list
   t.pan_perm.clear();
   t.pt_perm.clear();
   int nc = (int)t.cubes.size();
   for (int i = 0; i < nc; i++) {
      Cube& c = t.cubes[i];
      c.np = (int)pan_lists[i].size();
      c.nq = (int)pt_lists[i].size();
      if (c.leaf) {
         c.p0 = (int)t.pan_perm.size();
         c.q0 = (int)t.pt_perm.size();
         t.pan_perm.insert(t.pan_perm.end(), pan_lists[i].begin(), pan_lists[i].end());
         t.pt_perm.insert(t.pt_perm.end(), pt_lists[i].begin(), pt_lists[i].end());
      }
   }
   // non-leaf ranges start at their first descendant leaf; children have larger ids
   for (int i = nc - 1; i >= 0; i--) {
      Cube& c = t.cubes[i];
      if (c.leaf) continue;
      int p0 = -1, q0 = -1;
      for (int k = 0; k < 8; k++) {
         int ch = c.child[k];
         if (ch < 0) continue;
         if (p0 < 0) p0 = t.cubes[ch].p0;
         if (q0 < 0) q0 = t.cubes[ch].q0;
      }
      c.p0 = p0 < 0 ? 0 : p0;
      c.q0 = q0 < 0 ? 0 : q0;
   }
   t.cube_pan_sorted_off.assign(nc + 1, 0);
   t.cube_pt_sorted_off.assign(nc + 1, 0);
   for (int i = 0; i < nc; i++) {
      t.cube_pan_sorted_off[i + 1] = t.cube_pan_sorted_off[i] + (int)pan_lists[i].size();
      t.cube_pt_sorted_off[i + 1] = t.cube_pt_sorted_off[i] + (int)pt_lists[i].size();
   }
   t.cube_pan_sorted.resize(t.cube_pan_sorted_off[nc]);
   t.cube_pt_sorted.resize(t.cube_pt_sorted_off[nc]);
   for (int i = 0; i < nc; i++) {
      std::copy(pan_lists[i].begin(), pan_lists[i].end(), t.cube_pan_sorted.begin() + t.cube_pan_sorted_off[i]);
      std::copy(pt_lists[i].begin(), pt_lists[i].end(), t.cube_pt_sorted.begin() + t.cube_pt_sorted_off[i]);
   }
}

}  // namespace fsvd
