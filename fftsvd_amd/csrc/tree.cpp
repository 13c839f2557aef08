// Octree construction on the host.  The partition decisions follow the
// reference exactly (Tree_allocate, Tree.c:12-178; Cube_allocate,
// Cube.c:352-484; Cube_lists, Cube.c:600-713) because they fix which
// integrals are near-field; the storage is this library's own: cubes in a
// flat array (parents first) and panels/points renumbered into "tree order"
// so that every cube owns contiguous index ranges.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "internal.h"

namespace fsvd {

namespace {

struct Builder {
   Tree& t;
   const PanelGeom* P;
   const double* pts;
   // temporary per-cube sorted index lists
   std::vector<std::vector<int>> pan, pt;

   explicit Builder(Tree& tt, const PanelGeom* p, const double* q) : t(tt), P(p), pts(q) {}

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
         int k = child_of(P[i].c);
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
   for (size_t id = 0; id < t.cubes.size(); id++) {
      Cube& c = t.cubes[id];
      c.loc.clear();
      c.itr.clear();
      if (c.parent < 0) continue;
      const Cube& P = t.cubes[c.parent];
      for (int k = 0; k < 8; k++)
         if (P.child[k] >= 0 && P.child[k] != (int)id) c.loc.push_back(P.child[k]);
      for (int l : P.loc) {
         const Cube& L = t.cubes[l];
         for (int k = 0; k < 8; k++) {
            int o = L.child[k];
            if (o < 0) continue;
            const Cube& O = t.cubes[o];
            if (std::abs(O.idx[0] - c.idx[0]) <= 1 && std::abs(O.idx[1] - c.idx[1]) <= 1 &&
                std::abs(O.idx[2] - c.idx[2]) <= 1)
               c.loc.push_back(o);
         }
      }
      for (int l : P.loc) {
         const Cube& L = t.cubes[l];
         for (int k = 0; k < 8; k++) {
            int o = L.child[k];
            if (o < 0) continue;
            const Cube& O = t.cubes[o];
            if (std::abs(O.idx[0] - c.idx[0]) > 1 || std::abs(O.idx[1] - c.idx[1]) > 1 ||
                std::abs(O.idx[2] - c.idx[2]) > 1)
               c.itr.push_back(o);
         }
      }
   }
}

}  // namespace

void build_tree(Tree& t, const PanelGeom* panels, int npan, const double* pts, int npt, int maxpanels,
                int gridpoints) {
   t.npan = npan;
   t.npt = npt;
   t.maxpanels = maxpanels;
   t.gridpoints = gridpoints;
   double mn[3] = {1000000.0, 1000000.0, 1000000.0}, mx[3] = {-1000000.0, -1000000.0, -1000000.0};
   for (int i = 0; i < npan; i++)
      for (int d = 0; d < 3; d++) {
         if (panels[i].c[d] < mn[d]) mn[d] = panels[i].c[d];
         if (panels[i].c[d] > mx[d]) mx[d] = panels[i].c[d];
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
   int zero[3] = {0, 0, 0};
   std::vector<int> all_pan(npan), all_pt(npt);
   for (int i = 0; i < npan; i++) all_pan[i] = i;
   for (int i = 0; i < npt; i++) all_pt[i] = i;

   std::vector<std::vector<int>> pan_lists, pt_lists;
   for (t.depth = 2;; t.depth++) {
      t.cubes.clear();
      Builder b(t, panels, pts);
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
