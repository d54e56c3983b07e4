/*
 * cpu_port.cpp -- the coupled step on the host's cores: FP32, OpenMP.
 *
 * TEST INFRASTRUCTURE: the CPU baseline that bench.py times beside the
 * device (and that a CPU test checks against the FP64 oracle).  Nothing in
 * the product links or loads it.
 *
 * The same algorithm as the device path, for the terms the shear and large
 * workloads use: a fused pull-stream / bounce-back (periodic x and y, walls
 * on z moving with their wall velocity, LB_kernel_old_08102011.cu:1155-1224)
 * / moments / BGK with the Guo force (LB_kernel.cu:1108-1359) in deviation
 * form; cosine 4-point IB spreading and interpolation (LB_kernel.cu:
 * 1419-1642); Morse between the elements of a platelet over integer cell
 * lists with the cutoff decided as the device decides it; fibrin links (WLC
 * or linear, fiber.cpp:647-733); Bell bonds to wall and fibrin ligands on
 * the shared counter RNG (sem_kernel.cu:1750-2083); the detach test; the
 * marker update x += V dt (sem_kernel.cu:2087-2148).  Other terms (solid
 * flags, velocity faces, mesh springs, area/volume, bending, contact, the
 * kick, BBK) are refused at creation.
 */
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

#include "../include/pfg.h"
#include "../platelet_fiber_gpu_amd/csrc/pfg_rng.h"

namespace {

const int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0};
const int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1};
const int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1};
const int BB[19] = {0, 2, 1, 4, 3, 6, 5, 12, 11, 14, 13, 8, 7, 10, 9, 18, 17, 16, 15};
inline float wgt(int d) { return d == 0 ? 1.f / 3.f : (d < 7 ? 1.f / 18.f : 1.f / 36.f); }

inline float phi(float r) {
  const float a = std::fabs(r);
  return a < 2.f ? 0.25f * (1.f + std::cos(0.5f * (float)M_PI * r)) : 0.f;
}

inline int wrap(int i, int n) { return ((i % n) + n) % n; }

inline float min_image(float d, int n, bool periodic) {
  if (!periodic) return d;
  if (d > 0.5f * n) d -= (float)n;
  else if (d < -0.5f * n) d += (float)n;
  return d;
}

}  // namespace

struct cpu_port {
  pfg_params p{};
  int nx = 0, ny = 0, nz = 0;
  int64_t N = 0;
  bool per[3] = {true, true, false};
  std::vector<float> f[2];
  int cur = 0;
  std::vector<float> F[3], U[3];
  float wall_term[2][19] = {};
  /* markers */
  int n = 0;
  std::vector<float> x, y, z, vx, vy, vz, fx, fy, fz;
  std::vector<int> kind, body;
  std::vector<int> plt_start;
  std::vector<int> tri;
  std::vector<int> l_start, l_nbr; /* CSR marker -> link partners */
  std::vector<float> l_len;
  /* cell lists */
  int nb[3] = {1, 1, 1};
  std::vector<int> bin_start, bin_items, mbin;
  /* bonds */
  int n_rec = 0, n_wlig = 0, n_lig = 0;
  bool fib_lig = false;
  std::vector<int> rec_tri;
  std::vector<float> rec_r1, rec_r2, lig_x, lig_y, lig_z;
  std::vector<int> rec_bond, lig_owner, lig_claim, rec_cand;
  std::vector<float> rec_f, rec_w;
  std::vector<int> lg_start, lg_items; /* wall ligands binned */
  int64_t step = 0, total_bonds = 0;
  bool detached = false;
};

namespace {

void bin_of(const cpu_port &s, float px, float py, float pz, int b[3]) {
  const float pos[3] = {px, py, pz};
  const int dims[3] = {s.nx, s.ny, s.nz};
  for (int a = 0; a < 3; ++a) {
    int c = (int)std::floor(pos[a]);
    if (s.per[a]) c = wrap(c, dims[a]);
    else c = std::min(std::max(c, 0), dims[a] - 1);
    b[a] = c * s.nb[a] / dims[a];
  }
}

inline int bin_index(const cpu_port &s, const int b[3]) {
  return b[0] + s.nb[0] * (b[1] + s.nb[1] * b[2]);
}

/* Calls fn(bin) for the (up to 27) distinct bins around b. */
template <class Fn>
void for_nbr_bins(const cpu_port &s, const int b[3], Fn fn) {
  int lst[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, cnt[3] = {0, 0, 0};
  for (int a = 0; a < 3; ++a) {
    cnt[a] = 0;
    for (int o = -1; o <= 1; ++o) {
      int v = b[a] + o;
      if (s.per[a]) {
        v = wrap(v, s.nb[a]);
      } else if (v < 0 || v >= s.nb[a]) {
        continue;
      }
      bool dup = false;
      for (int k = 0; k < cnt[a]; ++k) dup = dup || lst[a][k] == v;
      if (!dup) lst[a][cnt[a]++] = v;
    }
  }
  for (int k = 0; k < cnt[2]; ++k)
    for (int j = 0; j < cnt[1]; ++j)
      for (int i = 0; i < cnt[0]; ++i) {
        const int q[3] = {lst[0][i], lst[1][j], lst[2][k]};
        fn(bin_index(s, q));
      }
}

void cell_lists(cpu_port &s) {
  const int nbins = s.nb[0] * s.nb[1] * s.nb[2];
  s.bin_start.assign((size_t)nbins + 1, 0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    int b[3];
    bin_of(s, s.x[i], s.y[i], s.z[i], b);
    s.mbin[i] = bin_index(s, b);
  }
  for (int i = 0; i < s.n; ++i) s.bin_start[(size_t)s.mbin[i] + 1]++;
  for (int b = 0; b < nbins; ++b) s.bin_start[(size_t)b + 1] += s.bin_start[(size_t)b];
  std::vector<int> fill(s.bin_start.begin(), s.bin_start.end() - 1);
  for (int i = 0; i < s.n; ++i) s.bin_items[(size_t)fill[(size_t)s.mbin[i]]++] = i;
}

void receptor_pos(const cpu_port &s, int r, float rp[3], int nodes[3]) {
  const int t = s.rec_tri[r];
  for (int k = 0; k < 3; ++k) nodes[k] = s.tri[3 * t + k];
  const float r1 = s.rec_r1[r], r2 = s.rec_r2[r];
  const float c0 = 1.f - r1, c1 = r1 * (1.f - r2), c2 = r1 * r2;
  rp[0] = c0 * s.x[nodes[0]] + c1 * s.x[nodes[1]] + c2 * s.x[nodes[2]];
  rp[1] = c0 * s.y[nodes[0]] + c1 * s.y[nodes[1]] + c2 * s.y[nodes[2]];
  rp[2] = c0 * s.z[nodes[0]] + c1 * s.z[nodes[1]] + c2 * s.z[nodes[2]];
}

void lig_vec(const cpu_port &s, int l, const float rp[3], float d[3]) {
  float L[3];
  if (l < s.n_wlig) {
    L[0] = s.lig_x[l];
    L[1] = s.lig_y[l];
    L[2] = s.lig_z[l];
  } else {
    const int m = l - s.n_wlig;
    L[0] = s.x[m];
    L[1] = s.y[m];
    L[2] = s.z[m];
  }
  const int dims[3] = {s.nx, s.ny, s.nz};
  for (int a = 0; a < 3; ++a) d[a] = min_image(L[a] - rp[a], dims[a], s.per[a]);
}

void bonds(cpu_port &s) {
  const pfg_params &p = s.p;
  const bool ref = p.bond_law == PFG_BOND_REFERENCE;
  const float reach2 = p.bond_reach * p.bond_reach;
  /* decisions from the start-of-step state */
  for (int r = 0; r < s.n_rec; ++r) {
    float rp[3];
    int nodes[3];
    receptor_pos(s, r, rp, nodes);
    s.rec_cand[(size_t)r] = -1;
    const int b = s.rec_bond[(size_t)r];
    if (b >= 0) {
      float d[3];
      lig_vec(s, b, rp, d);
      const float dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      const float ten = p.bond_k * (dist - p.bond_l0);
      const float fo = ref ? std::fabs(ten) : std::max(ten, 0.f);
      const float poff = 1.f - std::exp(-p.k_off0 * std::exp(p.bell_xb_kT * fo));
      if (pfg_uniform(p.seed, s.step, (uint32_t)r, PFG_RNG_BOND_OFF) < poff) s.rec_cand[r] = -2;
      continue;
    }
    int best = INT_MAX;
    float best_d2 = INFINITY;
    int b3[3];
    bin_of(s, rp[0], rp[1], rp[2], b3);
    auto cand = [&](int l, float d2) {
      if (d2 > reach2 || s.lig_owner[(size_t)l] >= 0) return;
      if (d2 < best_d2 || (d2 == best_d2 && l < best)) {
        best_d2 = d2;
        best = l;
      }
    };
    for_nbr_bins(s, b3, [&](int bin) {
      for (int q = s.lg_start[(size_t)bin]; q < s.lg_start[(size_t)bin + 1]; ++q) {
        const int l = s.lg_items[(size_t)q];
        float d[3];
        lig_vec(s, l, rp, d);
        cand(l, d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      }
      if (s.fib_lig)
        for (int q = s.bin_start[(size_t)bin]; q < s.bin_start[(size_t)bin + 1]; ++q) {
          const int m = s.bin_items[(size_t)q];
          if (s.kind[(size_t)m] == PFG_MARKER_ELEMENT) continue;
          float d[3];
          lig_vec(s, s.n_wlig + m, rp, d);
          cand(s.n_wlig + m, d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        }
    });
    if (best == INT_MAX) continue;
    float kon = p.k_on;
    if (ref) {
      const float w = std::fabs(p.bond_l0 - std::sqrt(best_d2)), fo = p.bond_k * w;
      kon = p.k_on * std::exp(p.bell_xb_kT * fo - 0.5f * fo * w / p.bell_kT);
    }
    if (pfg_uniform(p.seed, s.step, (uint32_t)r, PFG_RNG_BOND_ON) < 1.f - std::exp(-kon)) {
      s.rec_cand[(size_t)r] = best;
      s.lig_claim[(size_t)best] = std::min(s.lig_claim[(size_t)best], r);
    }
  }
  /* resolution: ruptures, the lowest claiming receptor, forces */
  s.total_bonds = 0;
  for (int r = 0; r < s.n_rec; ++r) {
    int b = s.rec_bond[(size_t)r];
    const int c = s.rec_cand[(size_t)r];
    if (c == -2) {
      s.lig_owner[(size_t)b] = -1;
      b = -1;
    } else if (c >= 0 && s.lig_claim[(size_t)c] == r) {
      s.lig_owner[(size_t)c] = r;
      b = c;
    }
    s.rec_bond[(size_t)r] = b;
    float F[3] = {0, 0, 0}, w[3] = {0, 0, 0};
    if (b >= 0) {
      float rp[3], d[3];
      int nodes[3];
      receptor_pos(s, r, rp, nodes);
      lig_vec(s, b, rp, d);
      const float dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (dist > 0.f)
        for (int a = 0; a < 3; ++a) F[a] = p.bond_k * (dist - p.bond_l0) / dist * d[a];
      float inv[3], sum = 0.f;
      for (int k = 0; k < 3; ++k) {
        const float ex = rp[0] - s.x[nodes[k]], ey = rp[1] - s.y[nodes[k]],
                    ez = rp[2] - s.z[nodes[k]];
        const float dr = std::sqrt(ex * ex + ey * ey + ez * ez);
        inv[k] = dr > 0.f ? 1.f / dr : 1e30f;
        sum += inv[k];
      }
      for (int k = 0; k < 3; ++k) w[k] = inv[k] / sum;
      s.total_bonds++;
    }
    for (int a = 0; a < 3; ++a) {
      s.rec_f[3 * (size_t)r + a] = F[a];
      s.rec_w[3 * (size_t)r + a] = w[a];
    }
  }
  std::fill(s.lig_claim.begin(), s.lig_claim.end(), INT_MAX);
}

void forces(cpu_port &s) {
  const pfg_params &p = s.p;
  if (s.n_rec > 0 && p.detach_rate > 0.f && s.total_bonds == 0 && !s.detached &&
      pfg_uniform(p.seed, s.step, 0u, PFG_RNG_DETACH) < p.detach_rate)
    s.detached = true;
  const float cut2 = p.morse_cutoff * p.morse_cutoff, r1c = p.morse_cutoff - p.morse_taper;
  const float sgn = p.fiber_ref_sign ? -1.f : 1.f;
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < s.n; ++i) {
    float F[3] = {0, 0, 0};
    const float xi = s.x[(size_t)i], yi = s.y[(size_t)i], zi = s.z[(size_t)i];
    if (s.kind[(size_t)i] == PFG_MARKER_ELEMENT) {
      const int bi = s.body[(size_t)i];
      if (p.morse_u0 != 0.f) {
        int b3[3];
        bin_of(s, xi, yi, zi, b3);
        for_nbr_bins(s, b3, [&](int bin) {
          for (int q = s.bin_start[(size_t)bin]; q < s.bin_start[(size_t)bin + 1]; ++q) {
            const int j = s.bin_items[(size_t)q];
            if (j == i || s.kind[(size_t)j] != PFG_MARKER_ELEMENT || s.body[(size_t)j] != bi)
              continue;
            const float dx = s.x[(size_t)j] - xi, dy = s.y[(size_t)j] - yi,
                        dz = s.z[(size_t)j] - zi;
            const float r2 = std::fmaf(dz, dz, std::fmaf(dy, dy, dx * dx));
            if (r2 >= cut2 || r2 == 0.f) continue;
            const float r = std::sqrt(r2);
            const float e = std::exp(-p.morse_rho * (r - p.morse_req));
            float c = 2.f * p.morse_rho * p.morse_u0 * (e - e * e);
            if (p.morse_taper > 0.f && r > r1c) {
              const float t = (r - r1c) / p.morse_taper;
              c = c * (1.f - t * t * (3.f - 2.f * t)) +
                  p.morse_u0 * (e * e - 2.f * e) * 6.f * t * (t - 1.f) / p.morse_taper;
            }
            c /= r;
            F[0] += c * dx;
            F[1] += c * dy;
            F[2] += c * dz;
          }
        });
      }
    } else {
      for (int q = s.l_start[(size_t)i]; q < s.l_start[(size_t)i + 1]; ++q) {
        const int j = s.l_nbr[(size_t)q];
        const float l0 = s.l_len[(size_t)q];
        const float dx = s.x[(size_t)j] - xi, dy = s.y[(size_t)j] - yi, dz = s.z[(size_t)j] - zi;
        const float l = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (l == 0.f || l0 == 0.f) continue;
        float ten;
        if (p.fiber_law == PFG_FIBER_LINEAR) {
          ten = p.fiber_k * (l - l0) / l0;
        } else {
          const float xs = std::min((l - l0) / l0, p.fiber_xmax), tl = 1.f - xs;
          ten = (p.fiber_k / l0) * (0.25f / (tl * tl) - 0.25f + xs);
        }
        const float c = sgn * ten / l;
        F[0] += c * dx;
        F[1] += c * dy;
        F[2] += c * dz;
      }
      if (s.fib_lig) {
        const int owner = s.lig_owner[(size_t)(s.n_wlig + i)];
        if (owner >= 0)
          for (int a = 0; a < 3; ++a) F[a] -= s.rec_f[3 * (size_t)owner + a];
      }
    }
    s.fx[(size_t)i] = F[0];
    s.fy[(size_t)i] = F[1];
    s.fz[(size_t)i] = F[2];
  }
  /* bond forces onto the triangles' vertices */
  for (int r = 0; r < s.n_rec; ++r) {
    if (s.rec_bond[(size_t)r] < 0) continue;
    const int t = s.rec_tri[(size_t)r];
    for (int k = 0; k < 3; ++k) {
      const int v = s.tri[3 * (size_t)t + k];
      s.fx[(size_t)v] += s.rec_w[3 * (size_t)r + k] * s.rec_f[3 * (size_t)r];
      s.fy[(size_t)v] += s.rec_w[3 * (size_t)r + k] * s.rec_f[3 * (size_t)r + 1];
      s.fz[(size_t)v] += s.rec_w[3 * (size_t)r + k] * s.rec_f[3 * (size_t)r + 2];
    }
  }
}

inline bool coupled(const cpu_port &s, int i) {
  const int k = s.kind[(size_t)i];
  return k == PFG_MARKER_ELEMENT || (k == PFG_MARKER_FIBER && s.p.fiber_move == PFG_FIBER_IB);
}

/* stencil node k along an axis: index (or -1 outside a wall) and weight */
inline int node(const cpu_port &s, int a, float X, int k, float *w) {
  const int dims[3] = {s.nx, s.ny, s.nz};
  if (s.per[a]) X = X - dims[a] * std::floor(X / dims[a]);
  const float fl = std::floor(X);
  *w = phi(X - (fl - 1.f + k));
  int i = (int)fl - 1 + k;
  if (s.per[a]) return wrap(i, dims[a]);
  return (i < 0 || i >= dims[a]) ? -1 : i;
}

void spread(cpu_port &s) {
  const float sc = s.p.ib_force_scale;
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < s.n; ++i) {
    if (!coupled(s, i)) continue;
    int ix[4], iy[4], iz[4];
    float wx[4], wy[4], wz[4];
    for (int k = 0; k < 4; ++k) {
      ix[k] = node(s, 0, s.x[(size_t)i], k, &wx[k]);
      iy[k] = node(s, 1, s.y[(size_t)i], k, &wy[k]);
      iz[k] = node(s, 2, s.z[(size_t)i], k, &wz[k]);
    }
    for (int c = 0; c < 4; ++c)
      for (int b = 0; b < 4; ++b)
        for (int a = 0; a < 4; ++a) {
          if (ix[a] < 0 || iy[b] < 0 || iz[c] < 0) continue;
          const float w = wx[a] * wy[b] * wz[c] * sc;
          const int64_t cell = ix[a] + (int64_t)s.nx * (iy[b] + (int64_t)s.ny * iz[c]);
#pragma omp atomic
          s.F[0][(size_t)cell] += w * s.fx[(size_t)i];
#pragma omp atomic
          s.F[1][(size_t)cell] += w * s.fy[(size_t)i];
#pragma omp atomic
          s.F[2][(size_t)cell] += w * s.fz[(size_t)i];
        }
  }
}

void lb(cpu_port &s) {
  const pfg_params &p = s.p;
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const int64_t N = s.N;
  const float omega = 1.f / p.tau, guo = 1.f - 0.5f * omega;
  const float *fin = s.f[s.cur].data();
  float *fout = s.f[1 - s.cur].data();
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y) {
      for (int x = 0; x < nx; ++x) {
        const int64_t c = x + (int64_t)nx * (y + (int64_t)ny * z);
        float dev[19];
        for (int d = 0; d < 19; ++d) {
          const int sz = z - CZ[d];
          if (!s.per[2] && (sz < 0 || sz >= nz)) {
            dev[d] = fin[(int64_t)BB[d] * N + c] + s.wall_term[sz < 0 ? 0 : 1][d];
            continue;
          }
          const int sx = wrap(x - CX[d], nx), sy = wrap(y - CY[d], ny);
          const int szw = s.per[2] ? wrap(sz, nz) : sz;
          dev[d] = fin[(int64_t)d * N + sx + (int64_t)nx * (sy + (int64_t)ny * szw)];
        }
        float drho = 0.f, j[3] = {0, 0, 0};
        for (int d = 0; d < 19; ++d) {
          drho += dev[d];
          j[0] += CX[d] * dev[d];
          j[1] += CY[d] * dev[d];
          j[2] += CZ[d] * dev[d];
        }
        const float rho = 1.f + drho;
        const float Fv[3] = {p.body_force[0] + s.F[0][(size_t)c], p.body_force[1] + s.F[1][(size_t)c],
                             p.body_force[2] + s.F[2][(size_t)c]};
        s.F[0][(size_t)c] = s.F[1][(size_t)c] = s.F[2][(size_t)c] = 0.f;
        float u[3];
        for (int a = 0; a < 3; ++a) u[a] = (j[a] + 0.5f * Fv[a]) / rho;
        s.U[0][(size_t)c] = u[0];
        s.U[1][(size_t)c] = u[1];
        s.U[2][(size_t)c] = u[2];
        const float usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        for (int d = 0; d < 19; ++d) {
          const float cu = CX[d] * u[0] + CY[d] * u[1] + CZ[d] * u[2];
          const float cF = CX[d] * Fv[0] + CY[d] * Fv[1] + CZ[d] * Fv[2];
          const float uF = u[0] * Fv[0] + u[1] * Fv[1] + u[2] * Fv[2];
          const float w = wgt(d);
          const float feq = w * (drho + rho * (3.f * cu + 4.5f * cu * cu - 1.5f * usq));
          const float src = w * guo * (3.f * (cF - uF) + 9.f * cu * cF);
          fout[(int64_t)d * N + c] = dev[d] - omega * (dev[d] - feq) + src;
        }
      }
    }
  s.cur = 1 - s.cur;
}

void interp_update(cpu_port &s) {
  const float dt = s.p.marker_dt;
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < s.n; ++i) {
    const int k = s.kind[(size_t)i];
    if (k == PFG_MARKER_FIXED) continue;
    float v[3] = {0, 0, 0};
    if (coupled(s, i)) {
      int ix[4], iy[4], iz[4];
      float wx[4], wy[4], wz[4];
      for (int q = 0; q < 4; ++q) {
        ix[q] = node(s, 0, s.x[(size_t)i], q, &wx[q]);
        iy[q] = node(s, 1, s.y[(size_t)i], q, &wy[q]);
        iz[q] = node(s, 2, s.z[(size_t)i], q, &wz[q]);
      }
      for (int c = 0; c < 4; ++c)
        for (int b = 0; b < 4; ++b)
          for (int a = 0; a < 4; ++a) {
            if (ix[a] < 0 || iy[b] < 0 || iz[c] < 0) continue;
            const float w = wx[a] * wy[b] * wz[c];
            const int64_t cell = ix[a] + (int64_t)s.nx * (iy[b] + (int64_t)s.ny * iz[c]);
            for (int ax = 0; ax < 3; ++ax) v[ax] += w * s.U[ax][(size_t)cell];
          }
    }
    s.vx[(size_t)i] = v[0];
    s.vy[(size_t)i] = v[1];
    s.vz[(size_t)i] = v[2];
    s.x[(size_t)i] += v[0] * dt;
    s.y[(size_t)i] += v[1] * dt;
    s.z[(size_t)i] += v[2] * dt;
  }
}

}  // namespace

extern "C" {

/* A port state of the model on the lattice of p, started from
 * f = f_eq(1, u(z)) (one velocity value per z-plane; NULL means 0), on
 * `threads` OpenMP threads (0: the default).  NULL when p or the model uses
 * a term the port does not carry. */
cpu_port *cport_create(const pfg_params *p, const pfg_model *m, const float *ux_z,
                       const float *uy_z, const float *uz_z, int threads) {
  if (!p || !m) return nullptr;
  if (p->use_flags || p->wall[2] == PFG_WALL_VELOCITY || p->k_spring != 0.f || p->ks != 0.f ||
      p->kv != 0.f || p->kbend != 0.f || p->contact_a != 0.f || p->wall_rep_a != 0.f ||
      p->fiber_move == PFG_FIBER_KICK || p->sem_integrator != PFG_SEM_ADVECT ||
      p->wall[0] != PFG_WALL_PERIODIC || p->wall[1] != PFG_WALL_PERIODIC)
    return nullptr;
  if (threads > 0) omp_set_num_threads(threads);
  cpu_port *s = new cpu_port();
  s->p = *p;
  s->nx = p->nx;
  s->ny = p->ny;
  s->nz = p->nz;
  s->N = (int64_t)s->nx * s->ny * s->nz;
  s->per[2] = p->wall[2] == PFG_WALL_PERIODIC;
  for (int b = 0; b < 2; ++b) s->f[b].assign((size_t)19 * s->N, 0.f);
  for (int a = 0; a < 3; ++a) {
    s->F[a].assign((size_t)s->N, 0.f);
    s->U[a].assign((size_t)s->N, 0.f);
  }
  for (int side = 0; side < 2; ++side)
    for (int d = 0; d < 19; ++d) {
      const float *u = side ? p->wall_u_hi[2] : p->wall_u_lo[2];
      s->wall_term[side][d] = 6.f * wgt(d) * (CX[d] * u[0] + CY[d] * u[1] + CZ[d] * u[2]);
    }
  const int64_t plane = (int64_t)s->nx * s->ny;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < s->N; ++c) {
    const int z = (int)(c / plane);
    const float u[3] = {ux_z ? ux_z[z] : 0.f, uy_z ? uy_z[z] : 0.f, uz_z ? uz_z[z] : 0.f};
    const float usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    for (int d = 0; d < 19; ++d) {
      const float cu = CX[d] * u[0] + CY[d] * u[1] + CZ[d] * u[2];
      s->f[0][(size_t)(d * s->N + c)] = wgt(d) * (3.f * cu + 4.5f * cu * cu - 1.5f * usq);
    }
  }
  const int n = m->n_markers;
  s->n = n;
  s->x.assign(m->x, m->x + n);
  s->y.assign(m->y, m->y + n);
  s->z.assign(m->z, m->z + n);
  s->vx.assign((size_t)n, 0.f);
  s->vy = s->vx;
  s->vz = s->vx;
  s->fx = s->vx;
  s->fy = s->vx;
  s->fz = s->vx;
  s->kind.assign(m->kind, m->kind + n);
  s->body.assign(m->body, m->body + n);
  s->tri.assign(m->tri, m->tri + 3 * (size_t)m->n_tri);
  std::vector<std::vector<std::pair<int, float>>> nbr((size_t)n);
  for (int l = 0; l < m->n_links; ++l) {
    const int a = m->link[2 * l], b = m->link[2 * l + 1];
    nbr[(size_t)a].push_back({b, m->link_l0[l]});
    nbr[(size_t)b].push_back({a, m->link_l0[l]});
  }
  s->l_start.assign((size_t)n + 1, 0);
  for (int i = 0; i < n; ++i) {
    s->l_start[(size_t)i + 1] = s->l_start[(size_t)i] + (int)nbr[(size_t)i].size();
    for (auto &e : nbr[(size_t)i]) {
      s->l_nbr.push_back(e.first);
      s->l_len.push_back(e.second);
    }
  }
  const int dims[3] = {s->nx, s->ny, s->nz};
  for (int a = 0; a < 3; ++a) s->nb[a] = std::max(1, dims[a] / std::max(1, p->bin_cells));
  s->mbin.assign((size_t)n, 0);
  s->bin_items.assign((size_t)n, 0);
  s->n_rec = m->n_receptors;
  s->rec_tri.assign(m->rec_tri, m->rec_tri + s->n_rec);
  s->rec_r1.assign(m->rec_r1, m->rec_r1 + s->n_rec);
  s->rec_r2.assign(m->rec_r2, m->rec_r2 + s->n_rec);
  s->n_wlig = m->n_wall_ligands;
  s->fib_lig = m->fibrin_ligands != 0;
  s->n_lig = s->n_wlig + (s->fib_lig ? n : 0);
  for (int l = 0; l < s->n_wlig; ++l) {
    s->lig_x.push_back(m->lig_x[l]);
    s->lig_y.push_back(m->lig_y[l]);
    s->lig_z.push_back(m->lig_z ? m->lig_z[l] : p->wall_z);
  }
  s->rec_bond.assign((size_t)s->n_rec, -1);
  s->rec_cand.assign((size_t)s->n_rec, -1);
  s->rec_f.assign(3 * (size_t)s->n_rec, 0.f);
  s->rec_w.assign(3 * (size_t)s->n_rec, 0.f);
  s->lig_owner.assign((size_t)s->n_lig, -1);
  s->lig_claim.assign((size_t)s->n_lig, INT_MAX);
  /* static wall ligands binned once */
  const int nbins = s->nb[0] * s->nb[1] * s->nb[2];
  s->lg_start.assign((size_t)nbins + 1, 0);
  std::vector<int> lb((size_t)s->n_wlig);
  for (int l = 0; l < s->n_wlig; ++l) {
    int b[3];
    bin_of(*s, s->lig_x[(size_t)l], s->lig_y[(size_t)l], s->lig_z[(size_t)l], b);
    lb[(size_t)l] = bin_index(*s, b);
    s->lg_start[(size_t)lb[(size_t)l] + 1]++;
  }
  for (int b = 0; b < nbins; ++b) s->lg_start[(size_t)b + 1] += s->lg_start[(size_t)b];
  s->lg_items.assign((size_t)s->n_wlig, 0);
  std::vector<int> fill(s->lg_start.begin(), s->lg_start.end() - 1);
  for (int l = 0; l < s->n_wlig; ++l) s->lg_items[(size_t)fill[(size_t)lb[(size_t)l]]++] = l;
  return s;
}

void cport_destroy(cpu_port *s) { delete s; }

/* nsteps coupled steps (fluid only when the model has no markers). */
void cport_step(cpu_port *s, int nsteps) {
  for (int t = 0; t < nsteps; ++t) {
    if (s->n > 0) {
      cell_lists(*s);
      bonds(*s);
      forces(*s);
      spread(*s);
    }
    lb(*s);
    if (s->n > 0) interp_update(*s);
    s->step++;
  }
}

/* Marker positions (n each) and the velocity of the last LB update (N each). */
void cport_markers(const cpu_port *s, float *x, float *y, float *z) {
  std::copy(s->x.begin(), s->x.end(), x);
  std::copy(s->y.begin(), s->y.end(), y);
  std::copy(s->z.begin(), s->z.end(), z);
}

void cport_velocity(const cpu_port *s, float *ux, float *uy, float *uz) {
  std::copy(s->U[0].begin(), s->U[0].end(), ux);
  std::copy(s->U[1].begin(), s->U[1].end(), uy);
  std::copy(s->U[2].begin(), s->U[2].end(), uz);
}

int64_t cport_bonds(const cpu_port *s) { return s->total_bonds; }

}  // extern "C"
