/*
 * geometry.cpp -- seeded scene construction on the host (see pfg_geometry.h).
 *
 * The reference reads a Gambit .neu platelet mesh (NEUReader, gpuSEM_LB.c:346-424),
 * derives topology and rest area/volume from it (gpuSEM_LB.c:846-1055), draws
 * receptor positions with GSL (gpuSEM_LB.c:732-747) and reads the fibrin network
 * from a text file (fiber.cpp:109-498).  The workloads here are synthetic: a
 * Fibonacci-spiral ellipsoid triangulated by its convex hull, receptors and
 * ligands from the Philox counter RNG, and a random geometric fibrin graph (the
 * construction of network::definelinks, network.cpp:197-218).
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <new>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "pfg.h"
#include "pfg_geometry.h"
#include "pfg_rng.h"

namespace {

/* Uniform double in [0,1) from (seed, tag, i, k): 53 bits of one Philox block. */
double urand(uint64_t seed, uint32_t tag, uint32_t i, uint32_t k) {
  const uint32_t ctr[4] = {i, k, tag, 0x5ce7e5u};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  const pfg_u32x4 r = pfg_philox4x32_10(ctr, key);
  const uint64_t bits = ((uint64_t)(r.v[0] >> 5) << 26) | (uint64_t)(r.v[1] >> 6);
  return (double)bits * (1.0 / 9007199254740992.0);
}

enum : uint32_t {
  TAG_JITTER = 1,
  TAG_RECEPTOR = 2,
  TAG_LIGAND = 3,
  TAG_FIBRIN = 4,
  TAG_FIBRIN_PICK = 5,
  TAG_PLACE = 6
};

struct V3 {
  double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

/* Incremental convex hull of points (all on the hull for points on a sphere).
 * Faces come out wound counter-clockwise seen from outside. */
bool convex_hull(const std::vector<V3> &P, std::vector<int> &tri) {
  const int n = (int)P.size();
  if (n < 4) return false;
  struct Face {
    int v[3];
    V3 nrm;
    double off;
    bool alive;
  };
  std::vector<Face> F;
  std::unordered_map<uint64_t, int> edge;
  auto key = [](int a, int b) { return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b; };
  auto add_face = [&](int a, int b, int c) {
    Face f;
    f.v[0] = a;
    f.v[1] = b;
    f.v[2] = c;
    f.nrm = cross(sub(P[b], P[a]), sub(P[c], P[a]));
    f.off = dot(f.nrm, P[a]);
    f.alive = true;
    F.push_back(f);
    const int id = (int)F.size() - 1;
    edge[key(a, b)] = id;
    edge[key(b, c)] = id;
    edge[key(c, a)] = id;
  };
  /* initial tetrahedron: 0, farthest from 0, farthest from that line, farthest
   * from that plane */
  int i0 = 0, i1 = -1, i2 = -1, i3 = -1;
  double best = -1;
  for (int i = 1; i < n; ++i) {
    const V3 d = sub(P[i], P[i0]);
    if (dot(d, d) > best) { best = dot(d, d); i1 = i; }
  }
  best = -1;
  for (int i = 1; i < n; ++i) {
    if (i == i1) continue;
    const V3 c = cross(sub(P[i1], P[i0]), sub(P[i], P[i0]));
    if (dot(c, c) > best) { best = dot(c, c); i2 = i; }
  }
  const V3 pn = cross(sub(P[i1], P[i0]), sub(P[i2], P[i0]));
  best = -1;
  for (int i = 1; i < n; ++i) {
    if (i == i1 || i == i2) continue;
    const double d = std::fabs(dot(pn, sub(P[i], P[i0])));
    if (d > best) { best = d; i3 = i; }
  }
  if (i1 < 0 || i2 < 0 || i3 < 0 || best <= 0) return false;
  if (dot(pn, sub(P[i3], P[i0])) > 0) std::swap(i1, i2); /* i3 below face 0-1-2 */
  add_face(i0, i1, i2);
  add_face(i0, i3, i1);
  add_face(i1, i3, i2);
  add_face(i2, i3, i0);
  std::vector<char> used(n, 0);
  used[i0] = used[i1] = used[i2] = used[i3] = 1;
  std::vector<int> visible;
  std::vector<std::pair<int, int>> horizon;
  for (int p = 0; p < n; ++p) {
    if (used[p]) continue;
    visible.clear();
    for (int f = 0; f < (int)F.size(); ++f) {
      if (!F[f].alive) continue;
      const double s = dot(F[f].nrm, P[p]) - F[f].off;
      const double scale = std::sqrt(dot(F[f].nrm, F[f].nrm));
      if (s > 1e-12 * scale) visible.push_back(f);
    }
    if (visible.empty()) return false; /* point inside: not on a convex surface */
    std::vector<char> vis(F.size(), 0);
    for (int f : visible) vis[f] = 1;
    horizon.clear();
    for (int f : visible) {
      for (int k = 0; k < 3; ++k) {
        const int a = F[f].v[k], b = F[f].v[(k + 1) % 3];
        auto it = edge.find(key(b, a));
        if (it == edge.end()) return false;
        if (!vis[it->second]) horizon.push_back({a, b});
      }
    }
    for (int f : visible) {
      F[f].alive = false;
      for (int k = 0; k < 3; ++k) {
        const int a = F[f].v[k], b = F[f].v[(k + 1) % 3];
        auto it = edge.find(key(a, b));
        if (it != edge.end() && it->second == f) edge.erase(it);
      }
    }
    for (auto &e : horizon) add_face(e.first, e.second, p);
    used[p] = 1;
  }
  tri.clear();
  for (auto &f : F)
    if (f.alive) {
      tri.push_back(f.v[0]);
      tri.push_back(f.v[1]);
      tri.push_back(f.v[2]);
    }
  return true;
}

}  // namespace

extern "C" {

pfg_status pfg_geom_ellipsoid(int n, float a, float b, float c, float cx, float cy,
                              float cz, float jitter, uint64_t seed, float *x,
                              float *y, float *z, int32_t *tri, int32_t *n_tri) {
  if (n < 4 || !x || !y || !z || !tri || !n_tri || a <= 0 || b <= 0 || c <= 0)
    return PFG_ERR_ARG;
  std::vector<V3> P((size_t)n);
  const double golden = M_PI * (3.0 - std::sqrt(5.0));
  for (int i = 0; i < n; ++i) {
    const double zz = 1.0 - (2.0 * i + 1.0) / n;
    const double r = std::sqrt(std::max(0.0, 1.0 - zz * zz));
    const double ph = golden * i;
    P[(size_t)i] = {r * std::cos(ph), r * std::sin(ph), zz};
  }
  /* the hull of the unit-sphere points depends on n only: computed once per
   * n and reused (a large scene builds thousands of equal platelets) */
  static std::mutex hull_mu;
  static std::map<int, std::vector<int>> hulls;
  std::vector<int> T;
  {
    std::lock_guard<std::mutex> lock(hull_mu);
    auto it = hulls.find(n);
    if (it != hulls.end()) {
      T = it->second;
    } else {
      if (!convex_hull(P, T)) return PFG_ERR_ARG;
      hulls[n] = T;
    }
  }
  if ((int)T.size() != 3 * (2 * n - 4)) return PFG_ERR_ARG;
  for (int i = 0; i < n; ++i) {
    const double jx = jitter * (2.0 * urand(seed, TAG_JITTER, (uint32_t)i, 0) - 1.0);
    const double jy = jitter * (2.0 * urand(seed, TAG_JITTER, (uint32_t)i, 1) - 1.0);
    const double jz = jitter * (2.0 * urand(seed, TAG_JITTER, (uint32_t)i, 2) - 1.0);
    x[i] = (float)(cx + a * P[(size_t)i].x + jx);
    y[i] = (float)(cy + b * P[(size_t)i].y + jy);
    z[i] = (float)(cz + c * P[(size_t)i].z + jz);
  }
  for (size_t k = 0; k < T.size(); ++k) tri[k] = T[k];
  *n_tri = (int32_t)(T.size() / 3);
  return PFG_OK;
}

/* NEUReader (gpuSEM_LB.c:346-424): counts, then "index x y z" per node and
 * "index type nodes n1 n2 n3" per triangle, 1-based.  Unlike the reference,
 * which writes through whatever indices it reads, every index is checked and
 * every node and triangle must appear. */
pfg_status pfg_geom_read_neu(const char *path, int max_nodes, int max_tri, float *x,
                             float *y, float *z, int32_t *tri, int32_t *n_nodes,
                             int32_t *n_tri) {
  if (!path || !n_nodes || !n_tri) return PFG_ERR_ARG;
  FILE *in = std::fopen(path, "r");
  if (!in) return PFG_ERR_IO;
  pfg_status st = PFG_OK;
  int nn = 0, nt = 0;
  if (std::fscanf(in, "%d %d", &nn, &nt) != 2 || nn < 3 || nt < 1) st = PFG_ERR_ARG;
  if (st == PFG_OK) {
    *n_nodes = nn;
    *n_tri = nt;
  }
  if (st == PFG_OK && x) {
    if (!y || !z || !tri || nn > max_nodes || nt > max_tri) st = PFG_ERR_ARG;
    std::vector<char> seen(st == PFG_OK ? (size_t)nn : 0, 0);
    for (int i = 0; i < nn && st == PFG_OK; ++i) {
      int idx;
      double a, b, c;
      if (std::fscanf(in, "%d %lf %lf %lf", &idx, &a, &b, &c) != 4 || idx < 1 || idx > nn ||
          seen[(size_t)idx - 1] || !std::isfinite(a) || !std::isfinite(b) || !std::isfinite(c)) {
        st = PFG_ERR_ARG;
        break;
      }
      seen[(size_t)idx - 1] = 1;
      x[idx - 1] = (float)a;
      y[idx - 1] = (float)b;
      z[idx - 1] = (float)c;
    }
    std::vector<char> tseen(st == PFG_OK ? (size_t)nt : 0, 0);
    for (int t = 0; t < nt && st == PFG_OK; ++t) {
      int idx, type, nodes, v[3];
      if (std::fscanf(in, "%d %d %d %d %d %d", &idx, &type, &nodes, &v[0], &v[1], &v[2]) != 6 ||
          idx < 1 || idx > nt || tseen[(size_t)idx - 1]) {
        st = PFG_ERR_ARG;
        break;
      }
      tseen[(size_t)idx - 1] = 1;
      for (int k = 0; k < 3 && st == PFG_OK; ++k) {
        if (v[k] < 1 || v[k] > nn) st = PFG_ERR_ARG;
        else tri[3 * (idx - 1) + k] = v[k] - 1;
      }
    }
  }
  std::fclose(in);
  return st;
}

/* init_fibrin_network's reader (fiber.cpp:145-217), checked: see the header. */
pfg_status pfg_geom_read_fibrin(const char *path, int max_nodes, int max_links, float *x,
                                float *y, float *z, int32_t *link, float *thickness,
                                int32_t *n_nodes, int32_t *n_links) {
  if (!path || !n_nodes || !n_links) return PFG_ERR_ARG;
  std::ifstream in(path);
  if (!in.is_open()) return PFG_ERR_IO;
  struct Node {
    double p[3];
    std::vector<int> nbr;
    std::vector<double> th;
  };
  std::vector<Node> nodes;
  std::vector<int> index;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream iss(line);
    std::vector<double> dat;
    double v;
    while (iss >> v) dat.push_back(v);
    if (dat.empty()) continue;
    if (dat.size() < 5) return PFG_ERR_ARG;
    const double nc = dat[4];
    if (nc < 0 || nc != std::floor(nc) || dat.size() != 5 + 2 * (size_t)nc) return PFG_ERR_ARG;
    Node nd;
    for (int k = 0; k < 3; ++k) {
      if (!std::isfinite(dat[(size_t)1 + k])) return PFG_ERR_ARG;
      nd.p[k] = dat[(size_t)1 + k];
    }
    for (size_t k = 0; k < (size_t)nc; ++k) {
      const double c = dat[5 + k];
      if (c != std::floor(c)) return PFG_ERR_ARG;
      nd.nbr.push_back((int)c);
      nd.th.push_back(dat[5 + (size_t)nc + k]);
    }
    if (dat[0] != std::floor(dat[0])) return PFG_ERR_ARG;
    index.push_back((int)dat[0]);
    nodes.push_back(std::move(nd));
  }
  const int nn = (int)nodes.size();
  if (nn < 1) return PFG_ERR_ARG;
  std::vector<int> slot((size_t)nn, -1); /* file index - 1 -> line */
  for (int l = 0; l < nn; ++l) {
    const int id = index[(size_t)l];
    if (id < 1 || id > nn || slot[(size_t)id - 1] >= 0) return PFG_ERR_ARG;
    slot[(size_t)id - 1] = l;
  }
  /* symmetric neighbour lists (as multisets, self links aside) */
  std::map<std::pair<int, int>, int> count;
  for (int i = 0; i < nn; ++i)
    for (int c : nodes[(size_t)slot[(size_t)i]].nbr) {
      if (c < 1 || c > nn) return PFG_ERR_ARG;
      if (c - 1 != i) count[{i, c - 1}]++;
    }
  for (const auto &kv : count) {
    auto it = count.find({kv.first.second, kv.first.first});
    if (it == count.end() || it->second != kv.second) return PFG_ERR_ARG;
  }
  int nl = 0;
  for (int i = 0; i < nn; ++i)
    for (int c : nodes[(size_t)slot[(size_t)i]].nbr)
      if (c - 1 > i) ++nl;
  *n_nodes = nn;
  *n_links = nl;
  if (!x) return PFG_OK;
  if (!y || !z || (nl > 0 && (!link || !thickness)) || nn > max_nodes || nl > max_links)
    return PFG_ERR_ARG;
  int l = 0;
  for (int i = 0; i < nn; ++i) {
    const Node &nd = nodes[(size_t)slot[(size_t)i]];
    x[i] = (float)nd.p[0];
    y[i] = (float)nd.p[1];
    z[i] = (float)nd.p[2];
    for (size_t k = 0; k < nd.nbr.size(); ++k)
      if (nd.nbr[k] - 1 > i) {
        link[2 * l] = i;
        link[2 * l + 1] = nd.nbr[k] - 1;
        thickness[l] = (float)nd.th[k];
        ++l;
      }
  }
  return PFG_OK;
}

pfg_status pfg_geom_area_volume(int n_tri, const int32_t *tri, const float *x,
                                const float *y, const float *z, double *tri_S,
                                double *S_all, double *V) {
  if (n_tri < 0 || (n_tri > 0 && (!tri || !x || !y || !z))) return PFG_ERR_ARG;
  double S = 0, Vol = 0;
  for (int t = 0; t < n_tri; ++t) {
    const int i0 = tri[3 * t], i1 = tri[3 * t + 1], i2 = tri[3 * t + 2];
    const V3 p0 = {x[i0], y[i0], z[i0]}, p1 = {x[i1], y[i1], z[i1]},
             p2 = {x[i2], y[i2], z[i2]};
    V3 nn = cross(sub(p1, p0), sub(p2, p0));
    const double len = std::sqrt(dot(nn, nn));
    const double St = 0.5 * len;
    nn = {-nn.x / len, -nn.y / len, -nn.z / len};
    const V3 cc = {(p0.x + p1.x + p2.x) / 3.0, (p0.y + p1.y + p2.y) / 3.0,
                   (p0.z + p1.z + p2.z) / 3.0};
    if (tri_S) tri_S[t] = St;
    S += St;
    Vol += dot(cc, nn) * St / 3.0;
  }
  if (S_all) *S_all = S;
  if (V) *V = Vol;
  return PFG_OK;
}

pfg_status pfg_geom_fibrin(int n, const float lo[3], const float hi[3], float radius,
                           int max_links, uint64_t seed, float *x, float *y, float *z,
                           int32_t *link, float *link_l0, int32_t *n_links) {
  if (n < 0 || !lo || !hi || radius <= 0 || max_links < 0 || !n_links) return PFG_ERR_ARG;
  if (n > 0 && (!x || !y || !z)) return PFG_ERR_ARG;
  if (max_links > 0 && (!link || !link_l0)) return PFG_ERR_ARG;
  for (int i = 0; i < n; ++i) {
    x[i] = (float)(lo[0] + (hi[0] - lo[0]) * urand(seed, TAG_FIBRIN, (uint32_t)i, 0));
    y[i] = (float)(lo[1] + (hi[1] - lo[1]) * urand(seed, TAG_FIBRIN, (uint32_t)i, 1));
    z[i] = (float)(lo[2] + (hi[2] - lo[2]) * urand(seed, TAG_FIBRIN, (uint32_t)i, 2));
  }
  /* pairs within radius via a uniform grid of cell size radius */
  int g[3];
  for (int k = 0; k < 3; ++k)
    g[k] = std::max(1, (int)std::floor((hi[k] - lo[k]) / radius));
  auto gi = [&](float v, int k) {
    int c = (int)std::floor((v - lo[k]) / (hi[k] - lo[k]) * g[k]);
    return std::min(std::max(c, 0), g[k] - 1);
  };
  std::vector<std::vector<int>> cells((size_t)g[0] * g[1] * g[2]);
  for (int i = 0; i < n; ++i)
    cells[(size_t)gi(x[i], 0) + (size_t)g[0] * (gi(y[i], 1) + (size_t)g[1] * gi(z[i], 2))]
        .push_back(i);
  const double r2 = (double)radius * radius;
  std::vector<std::pair<int, int>> pairs;
  for (int i = 0; i < n; ++i) {
    const int cx = gi(x[i], 0), cy = gi(y[i], 1), cz = gi(z[i], 2);
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int ax = cx + dx, ay = cy + dy, az = cz + dz;
          if (ax < 0 || ay < 0 || az < 0 || ax >= g[0] || ay >= g[1] || az >= g[2]) continue;
          for (int j : cells[(size_t)ax + (size_t)g[0] * (ay + (size_t)g[1] * az)]) {
            if (j <= i) continue;
            const double ddx = (double)x[j] - x[i], ddy = (double)y[j] - y[i],
                         ddz = (double)z[j] - z[i];
            if (ddx * ddx + ddy * ddy + ddz * ddz <= r2) pairs.push_back({i, j});
          }
        }
  }
  if ((int)pairs.size() > max_links) {
    std::vector<std::pair<double, size_t>> keyed(pairs.size());
    for (size_t k = 0; k < pairs.size(); ++k)
      keyed[k] = {urand(seed, TAG_FIBRIN_PICK, (uint32_t)pairs[k].first,
                        (uint32_t)pairs[k].second),
                  k};
    std::sort(keyed.begin(), keyed.end());
    std::vector<std::pair<int, int>> kept;
    for (int k = 0; k < max_links; ++k) kept.push_back(pairs[keyed[(size_t)k].second]);
    pairs.swap(kept);
  }
  std::sort(pairs.begin(), pairs.end());
  for (size_t k = 0; k < pairs.size(); ++k) {
    const int i = pairs[k].first, j = pairs[k].second;
    link[2 * k] = i;
    link[2 * k + 1] = j;
    const double ddx = (double)x[j] - x[i], ddy = (double)y[j] - y[i],
                 ddz = (double)z[j] - z[i];
    link_l0[k] = (float)std::sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
  }
  *n_links = (int32_t)pairs.size();
  return PFG_OK;
}

}  // extern "C"

/* ---------------- scenes ---------------- */

struct pfg_scene {
  pfg_params p{};
  int64_t N = 0;
  std::vector<float> x, y, z;
  std::vector<int32_t> kind, body;
  std::vector<int32_t> plt_elem_start, plt_tri_start, tri;
  std::vector<double> tri_S0, plt_S0, plt_V0;
  std::vector<int32_t> edge;
  std::vector<float> edge_l0;
  std::vector<int32_t> rec_tri;
  std::vector<float> rec_r1, rec_r2;
  std::vector<float> lig_x, lig_y, lig_z;
  int fibrin_ligands = 0;
  std::vector<int32_t> link;
  std::vector<float> link_l0;
  std::vector<uint8_t> flags;
  /* initial flow: u_x = couette_u * (z + 0.5) / nz */
  float couette_u = 0.f;
};

namespace {

struct Shape {
  int elements;
  float a, b, c;
};

/* Semi-axes scaled from (4, 4, 1.5) at 256 elements so that the element
 * spacing stays that of the 256-element platelet. */
Shape platelet_shape(int elements) {
  const float s = std::sqrt((float)elements / 256.0f);
  return {elements, 4.0f * s, 4.0f * s, 1.5f * s};
}

/* A platelet surface read from a file: nodes relative to their centroid,
 * already scaled to lattice units. */
struct Mesh {
  std::vector<float> x, y, z;
  std::vector<int32_t> tri;
};

pfg_status load_mesh(const char *path, float scale, Mesh *m, Shape *sh) {
  int32_t nn = 0, nt = 0;
  pfg_status st = pfg_geom_read_neu(path, 0, 0, nullptr, nullptr, nullptr, nullptr, &nn, &nt);
  if (st != PFG_OK) return st;
  m->x.resize((size_t)nn);
  m->y.resize((size_t)nn);
  m->z.resize((size_t)nn);
  m->tri.resize((size_t)3 * nt);
  st = pfg_geom_read_neu(path, nn, nt, m->x.data(), m->y.data(), m->z.data(), m->tri.data(),
                         &nn, &nt);
  if (st != PFG_OK) return st;
  double c[3] = {0, 0, 0};
  for (int i = 0; i < nn; ++i) {
    c[0] += m->x[(size_t)i];
    c[1] += m->y[(size_t)i];
    c[2] += m->z[(size_t)i];
  }
  for (double &v : c) v /= nn;
  if (!(scale > 0)) { /* mean edge 0.75, the spacing of the generated platelets */
    double len = 0;
    for (int t = 0; t < nt; ++t)
      for (int k = 0; k < 3; ++k) {
        const int a = m->tri[(size_t)3 * t + k], b = m->tri[(size_t)3 * t + (k + 1) % 3];
        const double dx = (double)m->x[(size_t)a] - m->x[(size_t)b];
        const double dy = (double)m->y[(size_t)a] - m->y[(size_t)b];
        const double dz = (double)m->z[(size_t)a] - m->z[(size_t)b];
        len += std::sqrt(dx * dx + dy * dy + dz * dz);
      }
    len /= 3.0 * nt;
    if (!(len > 0)) return PFG_ERR_ARG;
    scale = (float)(0.75 / len);
  }
  float ext[3] = {0, 0, 0};
  for (int i = 0; i < nn; ++i) {
    m->x[(size_t)i] = (float)((m->x[(size_t)i] - c[0]) * scale);
    m->y[(size_t)i] = (float)((m->y[(size_t)i] - c[1]) * scale);
    m->z[(size_t)i] = (float)((m->z[(size_t)i] - c[2]) * scale);
    ext[0] = std::max(ext[0], std::fabs(m->x[(size_t)i]));
    ext[1] = std::max(ext[1], std::fabs(m->y[(size_t)i]));
    ext[2] = std::max(ext[2], std::fabs(m->z[(size_t)i]));
  }
  *sh = {nn, ext[0], ext[1], ext[2]};
  return PFG_OK;
}

/* Append one platelet (mesh, rest area/volume, mesh edges, receptors): the
 * file mesh when given, else a generated ellipsoid. */
pfg_status add_platelet(pfg_scene *S, const Shape &sh, const Mesh *mesh, float cx, float cy,
                        float cz, float jitter, uint64_t seed, int receptors, uint32_t pid) {
  const int n = sh.elements;
  std::vector<float> x((size_t)n), y((size_t)n), z((size_t)n);
  std::vector<int32_t> tri;
  int32_t nt = 0;
  if (mesh) {
    for (int i = 0; i < n; ++i) {
      x[(size_t)i] = cx + mesh->x[(size_t)i];
      y[(size_t)i] = cy + mesh->y[(size_t)i];
      z[(size_t)i] = cz + mesh->z[(size_t)i];
    }
    tri = mesh->tri;
    nt = (int32_t)(tri.size() / 3);
  } else {
    tri.resize((size_t)3 * (2 * n - 4));
    pfg_status st = pfg_geom_ellipsoid(n, sh.a, sh.b, sh.c, cx, cy, cz, jitter,
                                       seed + 7919ull * pid, x.data(), y.data(),
                                       z.data(), tri.data(), &nt);
    if (st != PFG_OK) return st;
  }
  const int base = (int)S->x.size();
  const int tbase = (int)(S->tri.size() / 3);
  const int body = (int)S->plt_S0.size();
  for (int i = 0; i < n; ++i) {
    S->x.push_back(x[(size_t)i]);
    S->y.push_back(y[(size_t)i]);
    S->z.push_back(z[(size_t)i]);
    S->kind.push_back(PFG_MARKER_ELEMENT);
    S->body.push_back(body);
  }
  std::vector<double> tS((size_t)nt);
  double Sall = 0, V = 0;
  pfg_geom_area_volume(nt, tri.data(), x.data(), y.data(), z.data(), tS.data(), &Sall, &V);
  for (int t = 0; t < nt; ++t) {
    for (int k = 0; k < 3; ++k) S->tri.push_back(base + tri[(size_t)3 * t + k]);
    S->tri_S0.push_back(tS[(size_t)t]);
  }
  S->plt_S0.push_back(Sall);
  S->plt_V0.push_back(V);
  if (S->plt_elem_start.empty()) S->plt_elem_start.push_back(0);
  if (S->plt_tri_start.empty()) S->plt_tri_start.push_back(0);
  S->plt_elem_start.push_back(base + n);
  S->plt_tri_start.push_back(tbase + nt);
  /* mesh edges (each once, a < b), rest length as built */
  std::vector<std::pair<int, int>> edges;
  for (int t = 0; t < nt; ++t)
    for (int k = 0; k < 3; ++k) {
      int a = tri[(size_t)3 * t + k], b = tri[(size_t)3 * t + (k + 1) % 3];
      if (a > b) std::swap(a, b);
      edges.push_back({a, b});
    }
  std::sort(edges.begin(), edges.end());
  edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
  for (auto &e : edges) {
    const int a = e.first, b = e.second;
    const double dx = (double)x[(size_t)b] - x[(size_t)a], dy = (double)y[(size_t)b] - y[(size_t)a],
                 dz = (double)z[(size_t)b] - z[(size_t)a];
    S->edge.push_back(base + a);
    S->edge.push_back(base + b);
    S->edge_l0.push_back((float)std::sqrt(dx * dx + dy * dy + dz * dz));
  }
  /* receptors: triangle chosen with probability proportional to area,
   * r1 = sqrt(u) so that (1-r1) x0 + r1(1-r2) x1 + r1 r2 x2 is uniform on it */
  std::vector<double> cum((size_t)nt);
  double acc = 0;
  for (int t = 0; t < nt; ++t) {
    acc += tS[(size_t)t];
    cum[(size_t)t] = acc;
  }
  for (int r = 0; r < receptors; ++r) {
    const double u = urand(seed, TAG_RECEPTOR, pid, 3u * r) * acc;
    int t = (int)(std::upper_bound(cum.begin(), cum.end(), u) - cum.begin());
    if (t >= nt) t = nt - 1;
    S->rec_tri.push_back(tbase + t);
    S->rec_r1.push_back((float)std::sqrt(urand(seed, TAG_RECEPTOR, pid, 3u * r + 1)));
    S->rec_r2.push_back((float)urand(seed, TAG_RECEPTOR, pid, 3u * r + 2));
  }
  return PFG_OK;
}

void base_params(pfg_params *p, int nx, int ny, int nz) {
  pfg_params_default(p);
  p->nx = nx;
  p->ny = ny;
  p->nz = nz;
  p->morse_u0 = 1.0f;
  p->morse_rho = 2.0f;
  p->morse_req = 0.8f;
  /* nearest neighbours only (spacing ~0.75): a longer attractive tail
   * crumples a membrane that has no bending rigidity */
  p->morse_cutoff = 1.2f;
  p->morse_taper = 0.0f;
  p->bin_cells = 2;
  /* model force unit -> lattice force: keeps the explicit IB coupling of the
   * Morse stiffness 2 rho^2 U0 = 8 inside its stability range */
  p->ib_force_scale = 1e-2f;
}

void bell_params(pfg_params *p) {
  p->k_on = 1e-3f;
  p->k_off0 = 1e-4f;
  p->bond_k = 0.5f;
  p->bond_l0 = 0.3f;
  p->bond_reach = 1.0f;
  p->bell_xb_kT = 1.0f;
}

/* Seeded non-overlapping centres (bounding spheres of radius rad + gap). */
std::vector<V3> place_centres(int count, float rad, const float lo[3], const float hi[3],
                              float tube_r, float tcy, float tcz, uint64_t seed) {
  std::vector<V3> out;
  const double min_d = 2.0 * rad + 1.0;
  for (uint32_t attempt = 0; (int)out.size() < count && attempt < 2000000u; ++attempt) {
    V3 c = {lo[0] + (hi[0] - lo[0]) * urand(seed, TAG_PLACE, attempt, 0),
            lo[1] + (hi[1] - lo[1]) * urand(seed, TAG_PLACE, attempt, 1),
            lo[2] + (hi[2] - lo[2]) * urand(seed, TAG_PLACE, attempt, 2)};
    if (tube_r > 0) {
      const double ry = c.y - tcy, rz = c.z - tcz;
      if (std::sqrt(ry * ry + rz * rz) > tube_r - rad - 2.0) continue;
    }
    bool ok = true;
    for (const V3 &o : out) {
      const V3 d = sub(c, o);
      if (dot(d, d) < min_d * min_d) {
        ok = false;
        break;
      }
    }
    if (ok) out.push_back(c);
  }
  return out;
}

}  // namespace

extern "C" {

size_t pfg_struct_size(int which) {
  switch (which) {
    case 0: return sizeof(pfg_params);
    case 1: return sizeof(pfg_model);
    case 2: return sizeof(pfg_step_info);
    case 3: return sizeof(pfg_scene_opts);
    case 4: return sizeof(pfg_tuning);
  }
  return 0;
}

void pfg_scene_opts_init(pfg_scene_opts *o, int config) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->config = config;
  o->store = -1;
}

pfg_status pfg_scene_build(const pfg_scene_opts *o, pfg_params *params, pfg_scene **out) {
  if (!o || !params || !out) return PFG_ERR_ARG;
  *out = nullptr;
  pfg_scene *S = new (std::nothrow) pfg_scene();
  if (!S) return PFG_ERR_ALLOC;
  const uint64_t seed = o->seed ? o->seed : 1101ull;
  auto pick = [](int v, int def) { return v != 0 ? v : def; };
  pfg_params &p = S->p;
  pfg_status st = PFG_OK;
  int nplt = 0, elems = 0, recs = 0, ligs = 0, fnodes = 0, flinks = 0;
  float tube_r = 0;
  switch (o->config) {
    case PFG_SCENE_CHANNEL: {
      base_params(&p, pick(o->nx, 64), pick(o->ny, 32), pick(o->nz, 32));
      p.tau = 0.8f;
      p.wall[0] = PFG_WALL_PERIODIC;
      p.wall[1] = PFG_WALL_BOUNCE;
      p.wall[2] = PFG_WALL_BOUNCE;
      p.body_force[0] = 1e-6f;
      nplt = pick(o->n_platelets, 1);
      elems = pick(o->elements, 256);
      recs = pick(o->receptors, -1);
      ligs = pick(o->ligands, -1);
      fnodes = pick(o->fibrin_nodes, -1);
      break;
    }
    case PFG_SCENE_SHEAR:
    case PFG_SCENE_LARGE:
    case PFG_SCENE_BATCH: {
      const bool large = o->config == PFG_SCENE_LARGE;
      const bool batch = o->config == PFG_SCENE_BATCH;
      base_params(&p, pick(o->nx, large ? 1024 : (batch ? 64 : 256)),
                  pick(o->ny, large ? 512 : (batch ? 64 : 128)),
                  pick(o->nz, large ? 512 : (batch ? 64 : 128)));
      p.tau = 0.7f;
      p.wall[0] = PFG_WALL_PERIODIC;
      p.wall[1] = PFG_WALL_PERIODIC;
      p.wall[2] = PFG_WALL_BOUNCE;
      p.wall_u_hi[2][0] = 0.05f;
      S->couette_u = 0.05f;
      if (large) p.store = PFG_STORE_F16;
      bell_params(&p);
      p.detach_rate = 0.05f / (float)p.nz; /* gamma = shear rate * dt */
      nplt = pick(o->n_platelets, large ? 2000 : 1);
      elems = pick(o->elements, large ? 1000 : (batch ? 256 : 2048));
      recs = pick(o->receptors, large ? 64 : (batch ? 32 : 512));
      ligs = pick(o->ligands, large ? 65536 : (batch ? 512 : 4096));
      fnodes = pick(o->fibrin_nodes, large ? 80000 : -1);
      flinks = pick(o->fibrin_links, 120000);
      break;
    }
    case PFG_SCENE_THROMBUS: {
      base_params(&p, pick(o->nx, 512), pick(o->ny, 256), pick(o->nz, 256));
      p.tau = 0.8f;
      p.wall[0] = PFG_WALL_PERIODIC;
      p.wall[1] = PFG_WALL_BOUNCE;
      p.wall[2] = PFG_WALL_BOUNCE;
      p.use_flags = 1;
      tube_r = o->tube_radius > 0 ? o->tube_radius : 120.0f * p.nz / 256.0f;
      /* Poiseuille u_max = g R^2 / (4 nu) = 0.05 */
      p.body_force[0] = 0.05f * 4.0f * ((p.tau - 0.5f) / 3.0f) / (tube_r * tube_r);
      bell_params(&p);
      nplt = pick(o->n_platelets, 500);
      elems = pick(o->elements, 1000);
      recs = pick(o->receptors, 64);
      ligs = pick(o->ligands, 16384);
      fnodes = pick(o->fibrin_nodes, 20000);
      flinks = pick(o->fibrin_links, 30000);
      break;
    }
    default:
      delete S;
      return PFG_ERR_ARG;
  }
  if (o->store >= 0) p.store = o->store;
  p.seed = seed;
  if (nplt < 0) nplt = 0;
  if (recs < 0) recs = 0;
  if (ligs < 0) ligs = 0;
  if (fnodes < 0) fnodes = 0;
  S->N = (int64_t)p.nx * p.ny * p.nz;
  const float tcy = 0.5f * (p.ny - 1), tcz = 0.5f * (p.nz - 1);
  if (p.use_flags) {
    S->flags.assign((size_t)S->N, 0);
    for (int k = 0; k < p.nz; ++k)
      for (int j = 0; j < p.ny; ++j) {
        const float dy = j - tcy, dz = k - tcz;
        const uint8_t solid = (dy * dy + dz * dz > tube_r * tube_r) ? 1 : 0;
        for (int i = 0; i < p.nx; ++i)
          S->flags[(size_t)i + (size_t)p.nx * (j + (size_t)p.ny * k)] = solid;
      }
  }
  /* platelets */
  Mesh mesh;
  Shape sh = platelet_shape(elems);
  if (nplt > 0 && o->mesh_path) {
    st = load_mesh(o->mesh_path, o->mesh_scale, &mesh, &sh);
    if (st != PFG_OK) {
      delete S;
      return st;
    }
  }
  const Mesh *pm = o->mesh_path ? &mesh : nullptr;
  if (nplt > 0) {
    std::vector<V3> centres;
    if (o->config == PFG_SCENE_CHANNEL) {
      centres.push_back({0.5 * p.nx, tcy, tcz});
    } else if (o->config == PFG_SCENE_SHEAR && nplt == 1) {
      /* lowest element half a cell above the ligand wall */
      centres.push_back({0.25 * p.nx, tcy, p.wall_z + sh.c + 0.5});
    } else {
      const float rad = std::max(std::max(sh.a, sh.b), sh.c) + 0.5f;
      const float lo[3] = {0.f, tube_r > 0 ? tcy - tube_r : rad + 1.f,
                           tube_r > 0 ? tcz - tube_r : rad + 1.f};
      const float hi[3] = {(float)p.nx, tube_r > 0 ? tcy + tube_r : p.ny - rad - 1.f,
                           tube_r > 0 ? tcz + tube_r : p.nz - rad - 1.f};
      centres = place_centres(nplt, rad, lo, hi, tube_r, tcy, tcz, seed);
      if ((int)centres.size() < nplt) st = PFG_ERR_ARG;
    }
    for (size_t q = 0; q < centres.size() && st == PFG_OK; ++q)
      st = add_platelet(S, sh, pm, (float)centres[q].x, (float)centres[q].y,
                        (float)centres[q].z, 0.05f, seed, recs, (uint32_t)q);
  }
  if (S->plt_elem_start.empty()) {
    S->plt_elem_start.push_back(0);
    S->plt_tri_start.push_back(0);
  }
  /* static ligands */
  for (int l = 0; l < ligs && st == PFG_OK; ++l) {
    const double u0 = urand(seed, TAG_LIGAND, (uint32_t)l, 0);
    const double u1 = urand(seed, TAG_LIGAND, (uint32_t)l, 1);
    if (tube_r > 0) {
      const double th = 2.0 * M_PI * u1, rr = tube_r - 0.5;
      S->lig_x.push_back((float)(u0 * p.nx));
      S->lig_y.push_back((float)(tcy + rr * std::cos(th)));
      S->lig_z.push_back((float)(tcz + rr * std::sin(th)));
    } else {
      S->lig_x.push_back((float)(u0 * p.nx));
      S->lig_y.push_back((float)(u1 * p.ny));
      S->lig_z.push_back(p.wall_z);
    }
  }
  /* fibrin network in a box inscribed in the flow domain: generated, or read
   * from a file, centred in the box and scaled */
  const bool ffile = o->fibrin_path != nullptr;
  if ((fnodes > 0 || ffile) && st == PFG_OK) {
    float lo[3], hi[3];
    const float half_yz = tube_r > 0 ? tube_r / std::sqrt(2.0f) - 5.0f : 0.4f * p.ny;
    double len = 0.8 * p.nx;
    if (!ffile) {
      const double pairs_per_volume = 0.5 * (double)fnodes * fnodes *
                                      (4.0 / 3.0 * M_PI * 216.0);
      len = pairs_per_volume / std::max(1, flinks) / (4.0 * half_yz * half_yz);
    }
    len = std::min(len, 0.9 * p.nx);
    lo[0] = (float)(0.5 * p.nx - 0.5 * len);
    hi[0] = (float)(0.5 * p.nx + 0.5 * len);
    lo[1] = tcy - half_yz;
    hi[1] = tcy + half_yz;
    const float zc = tube_r > 0 ? tcz : 0.5f * p.nz;
    const float half_z = tube_r > 0 ? half_yz : std::min(half_yz, 0.4f * p.nz);
    lo[2] = zc - half_z;
    hi[2] = zc + half_z;
    std::vector<float> fx, fy, fz, l0;
    std::vector<int32_t> lk;
    int32_t nl = 0;
    if (ffile) {
      int32_t nn = 0;
      st = pfg_geom_read_fibrin(o->fibrin_path, 0, 0, nullptr, nullptr, nullptr, nullptr,
                                nullptr, &nn, &nl);
      if (st == PFG_OK) {
        fnodes = nn;
        fx.resize((size_t)nn);
        fy.resize((size_t)nn);
        fz.resize((size_t)nn);
        lk.resize((size_t)2 * nl);
        l0.resize((size_t)nl);
        std::vector<float> thick((size_t)nl);
        st = pfg_geom_read_fibrin(o->fibrin_path, nn, nl, fx.data(), fy.data(), fz.data(),
                                  lk.data(), thick.data(), &nn, &nl);
      }
      if (st == PFG_OK) {
        float flo[3] = {fx[0], fy[0], fz[0]}, fhi[3] = {fx[0], fy[0], fz[0]};
        for (int i = 0; i < fnodes; ++i) {
          const float q[3] = {fx[(size_t)i], fy[(size_t)i], fz[(size_t)i]};
          for (int k = 0; k < 3; ++k) {
            flo[k] = std::min(flo[k], q[k]);
            fhi[k] = std::max(fhi[k], q[k]);
          }
        }
        double s = o->fibrin_scale;
        if (!(s > 0)) { /* fit into the box */
          s = 1e30;
          for (int k = 0; k < 3; ++k)
            if (fhi[k] > flo[k]) s = std::min(s, (double)(hi[k] - lo[k]) / (fhi[k] - flo[k]));
          if (!(s < 1e30)) s = 1.0;
        }
        for (int i = 0; i < fnodes; ++i) {
          float *q[3] = {&fx[(size_t)i], &fy[(size_t)i], &fz[(size_t)i]};
          for (int k = 0; k < 3; ++k)
            *q[k] = (float)(0.5 * (lo[k] + hi[k]) + s * (*q[k] - 0.5 * (flo[k] + fhi[k])));
        }
        for (int k = 0; k < 3; ++k) { /* the box the network now fills */
          lo[k] = (float)(0.5 * (lo[k] + hi[k]) - 0.5 * s * (fhi[k] - flo[k]));
        }
        for (int l = 0; l < nl; ++l) { /* rest lengths as placed */
          const int a = lk[(size_t)2 * l], b = lk[(size_t)2 * l + 1];
          const double dx = (double)fx[(size_t)b] - fx[(size_t)a];
          const double dy = (double)fy[(size_t)b] - fy[(size_t)a];
          const double dz = (double)fz[(size_t)b] - fz[(size_t)a];
          l0[(size_t)l] = (float)std::sqrt(dx * dx + dy * dy + dz * dz);
        }
      }
    } else {
      fx.resize((size_t)fnodes);
      fy.resize((size_t)fnodes);
      fz.resize((size_t)fnodes);
      lk.resize((size_t)2 * flinks);
      l0.resize((size_t)flinks);
      st = pfg_geom_fibrin(fnodes, lo, hi, 6.0f, flinks, seed, fx.data(), fy.data(),
                           fz.data(), lk.data(), l0.data(), &nl);
    }
    const int base = (int)S->x.size();
    for (int i = 0; i < fnodes && st == PFG_OK; ++i) {
      S->x.push_back(fx[(size_t)i]);
      S->y.push_back(fy[(size_t)i]);
      S->z.push_back(fz[(size_t)i]);
      /* anchor the upstream face of the network */
      S->kind.push_back(fx[(size_t)i] < lo[0] + 4.0f ? PFG_MARKER_FIXED : PFG_MARKER_FIBER);
      S->body.push_back(-1);
    }
    for (int l = 0; l < nl && st == PFG_OK; ++l) {
      S->link.push_back(base + lk[(size_t)2 * l]);
      S->link.push_back(base + lk[(size_t)2 * l + 1]);
      S->link_l0.push_back(l0[(size_t)l]);
    }
    S->fibrin_ligands = 1;
    p.fiber_k = 0.2f;
    p.fiber_law = PFG_FIBER_WLC;
    p.bin_cells = std::max(p.bin_cells, 2);
  }
  if (st != PFG_OK) {
    delete S;
    return st;
  }
  *params = p;
  *out = S;
  return PFG_OK;
}

pfg_status pfg_scene_destroy(pfg_scene *s) {
  if (!s) return PFG_ERR_ARG;
  delete s;
  return PFG_OK;
}

pfg_status pfg_scene_model(const pfg_scene *s, pfg_model *m) {
  if (!s || !m) return PFG_ERR_ARG;
  std::memset(m, 0, sizeof(*m));
  m->n_markers = (int32_t)s->x.size();
  m->x = s->x.data();
  m->y = s->y.data();
  m->z = s->z.data();
  m->kind = s->kind.data();
  m->body = s->body.data();
  m->n_platelets = (int32_t)s->plt_S0.size();
  m->plt_elem_start = s->plt_elem_start.data();
  m->plt_tri_start = s->plt_tri_start.data();
  m->n_tri = (int32_t)(s->tri.size() / 3);
  m->tri = s->tri.data();
  m->tri_S0 = s->tri_S0.data();
  m->plt_S0 = s->plt_S0.data();
  m->plt_V0 = s->plt_V0.data();
  m->n_edges = (int32_t)s->edge_l0.size();
  m->edge = s->edge.data();
  m->edge_l0 = s->edge_l0.data();
  m->n_receptors = (int32_t)s->rec_tri.size();
  m->rec_tri = s->rec_tri.data();
  m->rec_r1 = s->rec_r1.data();
  m->rec_r2 = s->rec_r2.data();
  m->n_wall_ligands = (int32_t)s->lig_x.size();
  m->lig_x = s->lig_x.data();
  m->lig_y = s->lig_y.data();
  m->lig_z = s->lig_z.data();
  m->fibrin_ligands = s->fibrin_ligands;
  m->n_links = (int32_t)s->link_l0.size();
  m->link = s->link.data();
  m->link_l0 = s->link_l0.data();
  return PFG_OK;
}

const uint8_t *pfg_scene_flags(const pfg_scene *s) {
  if (!s || s->flags.empty()) return nullptr;
  return s->flags.data();
}

pfg_status pfg_scene_velocity(const pfg_scene *s, float *ux, float *uy, float *uz) {
  if (!s) return PFG_ERR_ARG;
  const pfg_params &p = s->p;
  for (int k = 0; k < p.nz; ++k) {
    const float u = s->couette_u * (k + 0.5f) / (float)p.nz;
    for (int64_t c = (int64_t)k * p.nx * p.ny; c < (int64_t)(k + 1) * p.nx * p.ny; ++c) {
      if (ux) ux[c] = u;
      if (uy) uy[c] = 0.f;
      if (uz) uz[c] = 0.f;
    }
  }
  return PFG_OK;
}

pfg_status pfg_scene_velocity_z(const pfg_scene *s, float *ux, float *uy, float *uz) {
  if (!s) return PFG_ERR_ARG;
  for (int k = 0; k < s->p.nz; ++k) {
    if (ux) ux[k] = s->couette_u * (k + 0.5f) / (float)s->p.nz;
    if (uy) uy[k] = 0.f;
    if (uz) uz[k] = 0.f;
  }
  return PFG_OK;
}

pfg_status pfg_scene_params(const pfg_scene *s, pfg_params *out) {
  if (!s || !out) return PFG_ERR_ARG;
  *out = s->p;
  return PFG_OK;
}

int64_t pfg_scene_fluid_cells(const pfg_scene *s) {
  if (!s) return 0;
  if (s->flags.empty()) return s->N;
  int64_t c = 0;
  for (uint8_t f : s->flags) c += f == 0;
  return c;
}

}  // extern "C"
