// Solvent-excluded surface mesher (host, C++ threads).
//
// The reference triangulates molecular surfaces by running MSMS from
// meshmaker (meshmaker.c:119-270).  This library builds its own: the SES is
// the zero set of f(x) = dist(x, A) - r_p, where A is the set of probe
// centres (outside every atom sphere grown by the probe radius r_p); the
// distance to A is measured to dense samples of the accessible part of the
// grown spheres.  f is sampled on a regular grid and contoured by marching
// tetrahedra (six Kuhn tetrahedra per cell, so neighbouring cells agree and
// the surface is closed).  Triangles are oriented counter-clockwise seen
// from the solute into the solvent (the MSMS convention: outward on an
// outer surface, into the void on a cavity).  fftsvd_mesh_ses keeps the
// largest connected component; fftsvd_mesh_components returns all of them.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

struct CellGrid {
   double lo[3], cell;
   int dim[3];
   std::vector<int> start, items;
   void build(const std::vector<double>& p, double cellsize, const double* blo, const double* bhi) {
      cell = cellsize;
      for (int d = 0; d < 3; d++) {
         lo[d] = blo[d];
         dim[d] = std::max(1, (int)std::ceil((bhi[d] - blo[d]) / cell) + 1);
      }
      size_t nc = (size_t)dim[0] * dim[1] * dim[2];
      size_t n = p.size() / 3;
      std::vector<int> cnt(nc + 1, 0), which(n);
      for (size_t i = 0; i < n; i++) {
         int c = cell_of(&p[3 * i]);
         which[i] = c;
         cnt[c + 1]++;
      }
      for (size_t c = 0; c < nc; c++) cnt[c + 1] += cnt[c];
      start = cnt;
      items.resize(n);
      std::vector<int> fill(cnt.begin(), cnt.end() - 1);
      for (size_t i = 0; i < n; i++) items[fill[which[i]]++] = (int)i;
   }
   int coord(double v, int d) const {
      int c = (int)std::floor((v - lo[d]) / cell);
      return std::min(std::max(c, 0), dim[d] - 1);
   }
   int cell_of(const double* x) const {
      return (coord(x[2], 2) * dim[1] + coord(x[1], 1)) * dim[0] + coord(x[0], 0);
   }
   template <typename F>
   void visit(const double* x, double radius, F&& f) const {
      int a0 = coord(x[0] - radius, 0), a1 = coord(x[0] + radius, 0);
      int b0 = coord(x[1] - radius, 1), b1 = coord(x[1] + radius, 1);
      int c0 = coord(x[2] - radius, 2), c1 = coord(x[2] + radius, 2);
      for (int k = c0; k <= c1; k++)
         for (int j = b0; j <= b1; j++)
            for (int i = a0; i <= a1; i++) {
               int c = (k * dim[1] + j) * dim[0] + i;
               for (int q = start[c]; q < start[c + 1]; q++) f(items[q]);
            }
   }
};

template <typename F>
void parallel_for(long long n, F&& f) {
   unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
   std::atomic<long long> next(0);
   std::vector<std::thread> th;
   for (unsigned t = 0; t < nt; t++)
      th.emplace_back([&]() {
         for (;;) {
            long long b = next.fetch_add(4096);
            if (b >= n) break;
            long long e = std::min(n, b + 4096);
            for (long long i = b; i < e; i++) f(i);
         }
      });
   for (auto& t : th) t.join();
}

struct UF {
   std::vector<int> p;
   explicit UF(int n) : p(n) {
      for (int i = 0; i < n; i++) p[i] = i;
   }
   int find(int a) {
      while (p[a] != a) a = p[a] = p[p[a]];
      return a;
   }
   void unite(int a, int b) {
      a = find(a);
      b = find(b);
      if (a != b) p[a] = b;
   }
};

// SES of the atoms with a probe of radius `probe` (union_mode false), or
// the surface of the union of the atom spheres grown by `probe` (union_mode
// true: the ion-exclusion surface, which the reference's meshmaker asks of
// MSMS with a vanishing probe, meshmaker.c:105-133).
// comp_faces set: every connected component, largest first (faces and
// vertices of a component contiguous), its face counts in *comp_faces;
// otherwise only the largest component.
int mesh_impl(const double* xyzr, int natoms, double probe, double h, bool union_mode, double** verts_out,
              int* nverts, int** faces_out, int* nfaces, std::vector<int>* comp_faces = nullptr) {
   {
      if (natoms < 1 || !(h > 0) || !(probe > 0)) throw std::runtime_error("bad mesher arguments");
      std::vector<double> ctr(3 * (size_t)natoms);
      std::vector<double> R(natoms);
      double rmax = 0, lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
      for (int i = 0; i < natoms; i++) {
         for (int d = 0; d < 3; d++) {
            ctr[3 * i + d] = xyzr[4 * i + d];
            lo[d] = std::min(lo[d], xyzr[4 * i + d]);
            hi[d] = std::max(hi[d], xyzr[4 * i + d]);
         }
         R[i] = xyzr[4 * i + 3] + probe;
         rmax = std::max(rmax, R[i]);
      }
      const double pad = rmax + 3 * h;
      for (int d = 0; d < 3; d++) {
         lo[d] -= pad;
         hi[d] += pad;
      }
      CellGrid atoms;
      atoms.build(ctr, 2 * rmax, lo, hi);
      // accessible samples on the grown spheres (golden-angle spirals)
      const double delta = std::min(0.25, 0.5 * h);
      std::vector<double> samp;
      if (!union_mode) {
         std::vector<std::vector<double>> part(natoms);
         parallel_for(natoms, [&](long long i) {
            int n = std::max(32, (int)std::ceil(4 * M_PI * R[i] * R[i] / (delta * delta)));
            const double ga = M_PI * (3.0 - std::sqrt(5.0));
            std::vector<double>& out = part[i];
            for (int k = 0; k < n; k++) {
               double z = 1.0 - (2.0 * k + 1.0) / n;
               double rr = std::sqrt(std::max(0.0, 1.0 - z * z));
               double th = ga * k;
               double x[3] = {ctr[3 * i] + R[i] * rr * std::cos(th), ctr[3 * i + 1] + R[i] * rr * std::sin(th),
                              ctr[3 * i + 2] + R[i] * z};
               bool buried = false;
               atoms.visit(x, rmax, [&](int j) {
                  if (buried || j == (int)i) return;
                  double d0 = x[0] - ctr[3 * j], d1 = x[1] - ctr[3 * j + 1], d2 = x[2] - ctr[3 * j + 2];
                  if (d0 * d0 + d1 * d1 + d2 * d2 < R[j] * R[j]) buried = true;
               });
               if (!buried) out.insert(out.end(), x, x + 3);
            }
         });
         for (auto& v : part) samp.insert(samp.end(), v.begin(), v.end());
      }
      const double cut = probe + 2 * h;
      CellGrid sgrid;
      sgrid.build(samp, cut, lo, hi);
      // implicit function on the grid (positive inside the SES)
      int nx = (int)std::ceil((hi[0] - lo[0]) / h) + 1, ny = (int)std::ceil((hi[1] - lo[1]) / h) + 1,
          nz = (int)std::ceil((hi[2] - lo[2]) / h) + 1;
      const long long ng = (long long)nx * ny * nz;
      if (ng > 600000000LL) throw std::runtime_error("mesher grid too large");
      std::vector<float> f(ng);
      parallel_for(ng, [&](long long g) {
         int i = (int)(g % nx), j = (int)((g / nx) % ny), k = (int)(g / ((long long)nx * ny));
         double x[3] = {lo[0] + i * h, lo[1] + j * h, lo[2] + k * h};
         double dsas = 1e300;
         atoms.visit(x, rmax, [&](int a) {
            double d0 = x[0] - ctr[3 * a], d1 = x[1] - ctr[3 * a + 1], d2 = x[2] - ctr[3 * a + 2];
            double d = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2) - R[a];
            dsas = std::min(dsas, d);
         });
         double val;
         if (union_mode) {
            val = -std::max(-10.0, std::min(dsas, 10.0));   // positive inside the grown spheres
         } else if (dsas >= 0) {
            val = -probe - std::min(dsas, 10.0);
         } else {
            double best = cut * cut;
            bool found = false;
            sgrid.visit(x, cut, [&](int s) {
               double d0 = x[0] - samp[3 * s], d1 = x[1] - samp[3 * s + 1], d2 = x[2] - samp[3 * s + 2];
               double d = d0 * d0 + d1 * d1 + d2 * d2;
               if (d < best) {
                  best = d;
                  found = true;
               }
            });
            val = found ? std::sqrt(best) - probe : 2 * h;
         }
         if (val == 0.0) val = 1e-9;
         f[g] = (float)val;
      });
      // marching tetrahedra
      static const int tets[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
      std::unordered_map<long long, int> emap;
      emap.reserve(1 << 20);
      std::vector<double> V;
      std::vector<int> Fc;
      auto gidx = [&](int i, int j, int k) { return ((long long)k * ny + j) * nx + i; };
      auto edge_vertex = [&](const int* a, const int* b) -> int {
         long long ga = gidx(a[0], a[1], a[2]), gb = gidx(b[0], b[1], b[2]);
         long long key = std::min(ga, gb) * 8 + ((std::abs(b[0] - a[0])) | (std::abs(b[1] - a[1]) << 1) |
                                                 (std::abs(b[2] - a[2]) << 2));
         auto it = emap.find(key);
         if (it != emap.end()) return it->second;
         double fa = f[ga], fb = f[gb];
         double t = fa / (fa - fb);
         t = std::min(0.98, std::max(0.02, t));
         int id = (int)(V.size() / 3);
         for (int d = 0; d < 3; d++) V.push_back(lo[d] + h * (a[d] + t * (b[d] - a[d])));
         emap.emplace(key, id);
         return id;
      };
      for (int k = 0; k + 1 < nz; k++)
         for (int j = 0; j + 1 < ny; j++)
            for (int i = 0; i + 1 < nx; i++) {
               // quick reject: all 8 corners on one side
               bool anyin = false, anyout = false;
               for (int c = 0; c < 8; c++) {
                  float v = f[gidx(i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1))];
                  (v > 0 ? anyin : anyout) = true;
               }
               if (!(anyin && anyout)) continue;
               for (int t = 0; t < 6; t++) {
                  int P[4][3];
                  P[0][0] = i;
                  P[0][1] = j;
                  P[0][2] = k;
                  for (int s = 1; s <= 3; s++) {
                     for (int d = 0; d < 3; d++) P[s][d] = P[s - 1][d];
                     P[s][tets[t][s - 1]] += 1;
                  }
                  int in[4], nin = 0, out[4], nout = 0;
                  for (int s = 0; s < 4; s++) {
                     if (f[gidx(P[s][0], P[s][1], P[s][2])] > 0)
                        in[nin++] = s;
                     else
                        out[nout++] = s;
                  }
                  if (nin == 0 || nin == 4) continue;
                  double cin[3] = {0, 0, 0}, cout_[3] = {0, 0, 0};
                  for (int q = 0; q < nin; q++)
                     for (int d = 0; d < 3; d++) cin[d] += P[in[q]][d] / (double)nin;
                  for (int q = 0; q < nout; q++)
                     for (int d = 0; d < 3; d++) cout_[d] += P[out[q]][d] / (double)nout;
                  double od[3] = {cout_[0] - cin[0], cout_[1] - cin[1], cout_[2] - cin[2]};
                  auto emit = [&](int a, int b, int c) {
                     const double *A = &V[3 * (size_t)a], *B = &V[3 * (size_t)b], *C = &V[3 * (size_t)c];
                     double u[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]}, w[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
                     double nrm[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
                     double area2 = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
                     if (!(area2 > 1e-8 * h * h) || a == b || b == c || a == c) return;
                     if (nrm[0] * od[0] + nrm[1] * od[1] + nrm[2] * od[2] < 0) std::swap(b, c);
                     Fc.push_back(a);
                     Fc.push_back(b);
                     Fc.push_back(c);
                  };
                  if (nin == 1 || nin == 3) {
                     int lone = nin == 1 ? in[0] : out[0];
                     int* others = nin == 1 ? out : in;
                     int e0 = edge_vertex(P[lone], P[others[0]]);
                     int e1 = edge_vertex(P[lone], P[others[1]]);
                     int e2 = edge_vertex(P[lone], P[others[2]]);
                     emit(e0, e1, e2);
                  } else {
                     int a0 = edge_vertex(P[in[0]], P[out[0]]);
                     int a1 = edge_vertex(P[in[0]], P[out[1]]);
                     int b0 = edge_vertex(P[in[1]], P[out[0]]);
                     int b1 = edge_vertex(P[in[1]], P[out[1]]);
                     emit(a0, a1, b1);
                     emit(a0, b1, b0);
                  }
               }
            }
      int nv = (int)(V.size() / 3), nf = (int)(Fc.size() / 3);
      if (nf == 0) throw std::runtime_error("empty surface");
      UF uf(nv);
      for (int q = 0; q < nf; q++) {
         uf.unite(Fc[3 * q], Fc[3 * q + 1]);
         uf.unite(Fc[3 * q], Fc[3 * q + 2]);
      }
      std::unordered_map<int, int> compsize, firstface;
      for (int q = 0; q < nf; q++) {
         const int r = uf.find(Fc[3 * q]);
         if (compsize[r]++ == 0) firstface[r] = q;
      }
      // components by face count, largest first (ties by first face)
      std::vector<std::pair<int, int>> order;
      for (auto& kv : compsize) order.push_back({kv.first, kv.second});
      std::sort(order.begin(), order.end(), [&](const std::pair<int, int>& a, const std::pair<int, int>& b) {
         return a.second != b.second ? a.second > b.second : firstface[a.first] < firstface[b.first];
      });
      if (!comp_faces) order.resize(1);
      std::vector<int> remap(nv, -1);
      std::vector<double> Vo;
      std::vector<int> Fo;
      for (const auto& c : order) {
         for (int q = 0; q < nf; q++) {
            if (uf.find(Fc[3 * q]) != c.first) continue;
            for (int s = 0; s < 3; s++) {
               int v = Fc[3 * q + s];
               if (remap[v] < 0) {
                  remap[v] = (int)(Vo.size() / 3);
                  Vo.insert(Vo.end(), &V[3 * (size_t)v], &V[3 * (size_t)v] + 3);
               }
               Fo.push_back(remap[v]);
            }
         }
         if (comp_faces) comp_faces->push_back(c.second);
      }
      *nverts = (int)(Vo.size() / 3);
      *nfaces = (int)(Fo.size() / 3);
      *verts_out = (double*)std::malloc(sizeof(double) * Vo.size());
      *faces_out = (int*)std::malloc(sizeof(int) * Fo.size());
      std::memcpy(*verts_out, Vo.data(), sizeof(double) * Vo.size());
      std::memcpy(*faces_out, Fo.data(), sizeof(int) * Fo.size());
      return 0;
   }
}

}  // namespace

extern "C" {

int fftsvd_mesh_ses(const double* xyzr, int natoms, double probe, double h, double** verts_out, int* nverts,
                    int** faces_out, int* nfaces) {
   try {
      return mesh_impl(xyzr, natoms, probe, h, false, verts_out, nverts, faces_out, nfaces);
   } catch (const std::exception& e) {
      std::fprintf(stderr, "fftsvd_mesh_ses: %s\n", e.what());
      return -1;
   }
}

int fftsvd_mesh_union(const double* xyzr, int natoms, double grow, double h, double** verts_out, int* nverts,
                      int** faces_out, int* nfaces) {
   try {
      return mesh_impl(xyzr, natoms, grow, h, true, verts_out, nverts, faces_out, nfaces);
   } catch (const std::exception& e) {
      std::fprintf(stderr, "fftsvd_mesh_union: %s\n", e.what());
      return -1;
   }
}

int fftsvd_mesh_components(const double* xyzr, int natoms, double probe, double h, int union_mode, double** verts_out,
                           int* nverts, int** faces_out, int* nfaces, int** comp_faces, int* ncomp) {
   try {
      std::vector<int> cf;
      const int rc = mesh_impl(xyzr, natoms, probe, h, union_mode != 0, verts_out, nverts, faces_out, nfaces, &cf);
      *ncomp = (int)cf.size();
      *comp_faces = (int*)std::malloc(sizeof(int) * std::max<size_t>(cf.size(), 1));
      std::memcpy(*comp_faces, cf.data(), sizeof(int) * cf.size());
      return rc;
   } catch (const std::exception& e) {
      std::fprintf(stderr, "fftsvd_mesh_components: %s\n", e.what());
      return -1;
   }
}

void fftsvd_free(void* p) { std::free(p); }

}  // extern "C"
