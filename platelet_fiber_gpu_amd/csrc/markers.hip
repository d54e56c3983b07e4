/*
 * markers.hip -- marker-side stages of the coupled step.
 *
 *  CELLS     integer cell lists: bin per marker, counting sort with a
 *            wavefront prefix scan, ascending marker order inside each bin.
 *  MEMBRANE  per-platelet triangle areas/normals and FP64 area/volume sums
 *            (sem_calculate_A_kernel, sem_kernel.cu:724-872) plus the lowest
 *            element for the wall repulsion (sem_kernel.cu:409-430).
 *  BONDS     Bell-model receptor-ligand rupture/formation with the counter RNG
 *            (sem_platelet_wall_kernel, sem_kernel.cu:1750-2083).
 *  FORCES    one gather per marker: Morse over the cell list, mesh-edge springs
 *            (sem_Force_kernel, sem_kernel.cu:320-513), area/volume constraints
 *            (sem_ForceAreaVol_kernel, sem_kernel.cu:1640-1749), wall repulsion,
 *            receptor bond forces, fibrin link forces (fiber.cpp:647-733,
 *            network.cpp:303-437).
 *  UPDATE    x += V dt (sem_update_position_kernel, sem_kernel.cu:2087-2148) or
 *            the damped kick of network.cpp:839-881 for fibrin nodes.
 *
 * Every per-marker force is a gather in a fixed order (cell-list order, CSR
 * order), so forces are deterministic run to run and need no atomics.
 */
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "pfg_internal.h"
#include "pfg_rng.h"

namespace pfg {

struct MArgs {
  Markers m;
  int nx, ny, nz;
  int periodic[3];
  int nb[3];
  /* parameters */
  float morse_u0, morse_rho, morse_req, morse_cut2, morse_cut, morse_taper;
  float k_spring, ks, kv;
  float wall_a, wall_kappa, wall_hc, wall_z;
  float k_on, k_off0, bond_k, bond_l0, bond_reach, bell_xb_kT, detach_rate;
  uint64_t seed;
  int fiber_law, fiber_move, fiber_ref_sign;
  float fiber_k, fiber_xmax, fiber_gamma;
  float marker_dt;
  int bin_cells;
};

constexpr float BOND_EPS_P = 1e-6f;   /* |u - p| below this: near-threshold  */
constexpr float BOND_EPS_R = 1e-4f;   /* |r - reach| below this: near-threshold */

/* ---------------- small helpers ---------------- */

__device__ inline int cell_of(float v, int n, int periodic) {
  int c = (int)floorf(v);
  if (periodic) {
    c %= n;
    if (c < 0) c += n;
  } else {
    c = c < 0 ? 0 : (c >= n ? n - 1 : c);
  }
  return c;
}

__device__ inline int bin_of(const MArgs &a, float x, float y, float z) {
  const int cx = cell_of(x, a.nx, a.periodic[0]);
  const int cy = cell_of(y, a.ny, a.periodic[1]);
  const int cz = cell_of(z, a.nz, a.periodic[2]);
  const int bx = cx * a.nb[0] / a.nx;
  const int by = cy * a.nb[1] / a.ny;
  const int bz = cz * a.nb[2] / a.nz;
  return bx + a.nb[0] * (by + a.nb[1] * bz);
}

/* Minimum-image difference on periodic axes. */
__device__ inline float min_image(float d, int n, int periodic) {
  if (!periodic) return d;
  const float fn = (float)n;
  if (d > 0.5f * fn) d -= fn;
  else if (d < -0.5f * fn) d += fn;
  return d;
}

/* Neighbour bins along one axis: up to three distinct indices. */
__device__ inline int nbr_bins(int b, int nb, int periodic, int out[3]) {
  if (nb <= 3 && periodic) {
    for (int k = 0; k < nb; ++k) out[k] = k;
    return nb;
  }
  int c = 0;
  for (int o = -1; o <= 1; ++o) {
    int v = b + o;
    if (periodic) {
      if (v < 0) v += nb;
      if (v >= nb) v -= nb;
    } else if (v < 0 || v >= nb) {
      continue;
    }
    out[c++] = v;
  }
  return c;
}

__device__ inline int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

/* Exclusive scan of 256 threads x 4 items; returns the block total. */
__device__ inline int block_excl_scan4(int v[4], int *sh_wave) {
  const int s0 = v[0], s1 = s0 + v[1], s2 = s1 + v[2], s3 = s2 + v[3];
  const int incl = wave_incl_scan(s3);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 63) sh_wave[wave] = incl;
  __syncthreads();
  int wave_off = 0, total = 0;
  for (int k = 0; k < 4; ++k) {
    if (k < wave) wave_off += sh_wave[k];
    total += sh_wave[k];
  }
  const int base = wave_off + incl - s3;
  v[0] = base;
  v[1] = base + s0;
  v[2] = base + s1;
  v[3] = base + s2;
  __syncthreads();
  return total;
}

/* ---------------- CELLS ---------------- */

__global__ void bins_kernel(MArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    /* start of the step: reset the bond counter read by the detach test */
    a.m.counters[0] = 0;
  }
  if (i >= a.m.n) return;
  const int b = bin_of(a, a.m.x[i], a.m.y[i], a.m.z[i]);
  a.m.marker_bin[i] = b;
  a.m.marker_slot[i] = atomicAdd(&a.m.bin_count[b], 1);
}

__global__ __launch_bounds__(256) void scan_reduce_kernel(const int32_t *count,
                                                          int64_t n,
                                                          int32_t *partials) {
  __shared__ int sh[4];
  const int64_t base = (int64_t)blockIdx.x * 1024 + threadIdx.x * 4;
  int v[4];
  for (int k = 0; k < 4; ++k) v[k] = base + k < n ? count[base + k] : 0;
  const int total = block_excl_scan4(v, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

/* Exclusive scan of the block totals in one block. */
__global__ __launch_bounds__(256) void scan_top_kernel(int32_t *partials,
                                                       int64_t nblocks,
                                                       int32_t *bin_start,
                                                       int64_t nbins) {
  __shared__ int sh[4];
  int carry = 0;
  for (int64_t off = 0; off < nblocks; off += 1024) {
    int v[4];
    const int64_t base = off + threadIdx.x * 4;
    for (int k = 0; k < 4; ++k) v[k] = base + k < nblocks ? partials[base + k] : 0;
    const int total = block_excl_scan4(v, sh);
    for (int k = 0; k < 4; ++k)
      if (base + k < nblocks) partials[base + k] = v[k] + carry;
    carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) bin_start[nbins] = carry;
}

__global__ __launch_bounds__(256) void scan_apply_kernel(int32_t *count, int64_t n,
                                                         const int32_t *partials,
                                                         int32_t *bin_start) {
  __shared__ int sh[4];
  const int64_t base = (int64_t)blockIdx.x * 1024 + threadIdx.x * 4;
  int v[4];
  for (int k = 0; k < 4; ++k) v[k] = base + k < n ? count[base + k] : 0;
  block_excl_scan4(v, sh);
  const int off = partials[blockIdx.x];
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) {
      bin_start[base + k] = v[k] + off;
      count[base + k] = 0; /* ready for the next step */
    }
  }
}

__global__ void scatter_kernel(MArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.m.n) return;
  const int b = a.m.marker_bin[i];
  a.m.bin_items[a.m.bin_start[b] + a.m.marker_slot[i]] = i;
}

/* Atomic slots arrive in any order: sort each bin by marker index. */
__global__ void sort_bins_kernel(MArgs a) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.m.nbins) return;
  const int s = a.m.bin_start[b], e = a.m.bin_start[b + 1];
  if (e - s < 2) return;
  int32_t *it = a.m.bin_items;
  for (int i = s + 1; i < e; ++i) {
    const int v = it[i];
    int j = i - 1;
    while (j >= s && it[j] > v) {
      it[j + 1] = it[j];
      --j;
    }
    it[j + 1] = v;
  }
}

/* ---------------- MEMBRANE ---------------- */

__global__ __launch_bounds__(256) void membrane_kernel(MArgs a) {
  const int p = blockIdx.x;
  const int t0 = a.m.plt_tri_start[p], t1 = a.m.plt_tri_start[p + 1];
  double S = 0.0, V = 0.0;
  for (int t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
    const int n0 = a.m.tri[3 * t], n1 = a.m.tri[3 * t + 1], n2 = a.m.tri[3 * t + 2];
    const double x0 = a.m.x[n0], y0 = a.m.y[n0], z0 = a.m.z[n0];
    const double x1 = a.m.x[n1], y1 = a.m.y[n1], z1 = a.m.z[n1];
    const double x2 = a.m.x[n2], y2 = a.m.y[n2], z2 = a.m.z[n2];
    const double ax = x1 - x0, ay = y1 - y0, az = z1 - z0;
    const double bx = x2 - x0, by = y2 - y0, bz = z2 - z0;
    double nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
    const double nn = sqrt(nx * nx + ny * ny + nz * nz);
    /* n_elem = -(e12 x e13)/|e12 x e13| (sem_kernel.cu:799-804) */
    nx /= -nn;
    ny /= -nn;
    nz /= -nn;
    const double St = 0.5 * nn;
    const double cx = (x0 + x1 + x2) / 3.0, cy = (y0 + y1 + y2) / 3.0,
                 cz = (z0 + z1 + z2) / 3.0;
    a.m.tri_S[t] = St;
    a.m.tri_n[3 * t] = (float)nx;
    a.m.tri_n[3 * t + 1] = (float)ny;
    a.m.tri_n[3 * t + 2] = (float)nz;
    S += St;
    V += (cx * nx + cy * ny + cz * nz) * St / 3.0;
  }
  __shared__ double shS[256], shV[256];
  __shared__ float shz[256];
  __shared__ int shi[256];
  shS[threadIdx.x] = S;
  shV[threadIdx.x] = V;
  /* lowest element: minimum z, lowest index on ties */
  const int e0 = a.m.plt_elem_start[p], e1 = a.m.plt_elem_start[p + 1];
  float zmin = INFINITY;
  int imin = INT_MAX;
  for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const float z = a.m.z[e];
    if (z < zmin || (z == zmin && e < imin)) {
      zmin = z;
      imin = e;
    }
  }
  shz[threadIdx.x] = zmin;
  shi[threadIdx.x] = imin;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      shS[threadIdx.x] += shS[threadIdx.x + s];
      shV[threadIdx.x] += shV[threadIdx.x + s];
      const float z2 = shz[threadIdx.x + s];
      const int i2 = shi[threadIdx.x + s];
      if (z2 < shz[threadIdx.x] || (z2 == shz[threadIdx.x] && i2 < shi[threadIdx.x])) {
        shz[threadIdx.x] = z2;
        shi[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.m.plt_S[p] = shS[0];
    a.m.plt_V[p] = shV[0];
    a.m.plt_lowest[p] = shi[0] == INT_MAX ? -1 : shi[0];
  }
}

/* ---------------- BONDS ---------------- */

__device__ inline void receptor_pos(const MArgs &a, int r, float out[3],
                                    int nodes[3]) {
  const int t = a.m.rec_tri[r];
  const float r1 = a.m.rec_r1[r], r2 = a.m.rec_r2[r];
  nodes[0] = a.m.tri[3 * t];
  nodes[1] = a.m.tri[3 * t + 1];
  nodes[2] = a.m.tri[3 * t + 2];
  /* (1-r1) x0 + r1 (1-r2) x1 + r1 r2 x2 (sem_kernel.cu:1811, 1840-1841) */
  const float c0 = 1.0f - r1, c1 = r1 * (1.0f - r2), c2 = r1 * r2;
  out[0] = c0 * a.m.x[nodes[0]] + c1 * a.m.x[nodes[1]] + c2 * a.m.x[nodes[2]];
  out[1] = c0 * a.m.y[nodes[0]] + c1 * a.m.y[nodes[1]] + c2 * a.m.y[nodes[2]];
  out[2] = c0 * a.m.z[nodes[0]] + c1 * a.m.z[nodes[1]] + c2 * a.m.z[nodes[2]];
}

/* Vector from the receptor at rp to ligand l. */
__device__ inline void ligand_vec(const MArgs &a, int l, const float rp[3],
                                  float d[3]) {
  if (l < a.m.n_wall_lig) {
    d[0] = min_image(a.m.lig_x[l] - rp[0], a.nx, a.periodic[0]);
    d[1] = min_image(a.m.lig_y[l] - rp[1], a.ny, a.periodic[1]);
    d[2] = min_image(a.m.lig_z[l] - rp[2], a.nz, a.periodic[2]);
  } else {
    const int mk = l - a.m.n_wall_lig;
    d[0] = min_image(a.m.x[mk] - rp[0], a.nx, a.periodic[0]);
    d[1] = min_image(a.m.y[mk] - rp[1], a.ny, a.periodic[1]);
    d[2] = min_image(a.m.z[mk] - rp[2], a.nz, a.periodic[2]);
  }
}

__device__ inline void bump(int64_t *c, int k) {
  atomicAdd((unsigned long long *)(c + k), 1ull);
}

/* Decide ruptures of held bonds and formation candidates of free receptors
 * from the state at the start of the step. */
__global__ void bond_decide_kernel(MArgs a) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.m.n_rec) return;
  const int64_t step = *a.m.d_step;
  float rp[3];
  int nodes[3];
  receptor_pos(a, r, rp, nodes);
  const int b = a.m.rec_bond[r];
  a.m.rec_cand[r] = -1;
  a.m.rec_release[r] = 0;
  if (b >= 0) {
    float d[3];
    ligand_vec(a, b, rp, d);
    const float dist = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const float tension = a.bond_k * (dist - a.bond_l0);
    /* Bell: k_off = k_off0 exp(x_b f / kBT) for a pulling bond */
    const float koff = a.k_off0 * expf(a.bell_xb_kT * fmaxf(tension, 0.0f));
    const float p_off = 1.0f - expf(-koff);
    const float u = pfg_uniform(a.seed, step, (uint32_t)r, PFG_RNG_BOND_OFF);
    if (fabsf(u - p_off) < BOND_EPS_P) bump(a.m.counters, 2);
    a.m.rec_release[r] = u < p_off ? 1 : 0;
    return;
  }
  /* nearest free ligand within reach (ties: lower ligand id) */
  const float reach2 = a.bond_reach * a.bond_reach;
  int best = -1;
  float best_d2 = INFINITY;
  bool near_reach = false;
  const int cx = cell_of(rp[0], a.nx, a.periodic[0]) * a.nb[0] / a.nx;
  const int cy = cell_of(rp[1], a.ny, a.periodic[1]) * a.nb[1] / a.ny;
  const int cz = cell_of(rp[2], a.nz, a.periodic[2]) * a.nb[2] / a.nz;
  int bxs[3], bys[3], bzs[3];
  const int nbx = nbr_bins(cx, a.nb[0], a.periodic[0], bxs);
  const int nby = nbr_bins(cy, a.nb[1], a.periodic[1], bys);
  const int nbz = nbr_bins(cz, a.nb[2], a.periodic[2], bzs);
  if (a.m.n_wall_lig > 0) {
    for (int jz = 0; jz < nbz; ++jz)
      for (int jy = 0; jy < nby; ++jy)
        for (int jx = 0; jx < nbx; ++jx) {
          const int bin = bxs[jx] + a.nb[0] * (bys[jy] + a.nb[1] * bzs[jz]);
          for (int k = a.m.lg_start[bin]; k < a.m.lg_start[bin + 1]; ++k) {
            const int l = a.m.lg_items[k];
            float d[3];
            ligand_vec(a, l, rp, d);
            const float d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (fabsf(sqrtf(d2) - a.bond_reach) < BOND_EPS_R) near_reach = true;
            if (d2 > reach2 || a.m.lig_owner[l] >= 0) continue;
            if (d2 < best_d2 || (d2 == best_d2 && l < best)) {
              best_d2 = d2;
              best = l;
            }
          }
        }
  }
  if (a.m.fibrin_lig) {
    for (int jz = 0; jz < nbz; ++jz)
      for (int jy = 0; jy < nby; ++jy)
        for (int jx = 0; jx < nbx; ++jx) {
          const int bin = bxs[jx] + a.nb[0] * (bys[jy] + a.nb[1] * bzs[jz]);
          for (int k = a.m.bin_start[bin]; k < a.m.bin_start[bin + 1]; ++k) {
            const int mk = a.m.bin_items[k];
            if (a.m.kind[mk] == PFG_MARKER_ELEMENT) continue;
            const int l = a.m.n_wall_lig + mk;
            float d[3];
            ligand_vec(a, l, rp, d);
            const float d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (fabsf(sqrtf(d2) - a.bond_reach) < BOND_EPS_R) near_reach = true;
            if (d2 > reach2 || a.m.lig_owner[l] >= 0) continue;
            if (d2 < best_d2 || (d2 == best_d2 && l < best)) {
              best_d2 = d2;
              best = l;
            }
          }
        }
  }
  if (near_reach) bump(a.m.counters, 2);
  if (best < 0) return;
  const float p_on = 1.0f - expf(-a.k_on);
  const float u = pfg_uniform(a.seed, step, (uint32_t)r, PFG_RNG_BOND_ON);
  if (fabsf(u - p_on) < BOND_EPS_P) bump(a.m.counters, 2);
  if (u < p_on) {
    a.m.rec_cand[r] = best;
    atomicMin(&a.m.lig_claim[best], r);
  }
}

/* Apply ruptures and the lowest-index claim per ligand, then the bond force
 * k (|d| - l0) d/|d| on the receptor and its inverse-distance split over the
 * three vertices (sem_kernel.cu:2039-2065). */
__global__ void bond_resolve_kernel(MArgs a) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.m.n_rec) return;
  int b = a.m.rec_bond[r];
  if (a.m.rec_release[r]) {
    a.m.lig_owner[b] = -1;
    b = -1;
    bump(a.m.counters, 4);
  } else if (a.m.rec_cand[r] >= 0) {
    const int l = a.m.rec_cand[r];
    if (a.m.lig_claim[l] == r) {
      a.m.lig_claim[l] = INT_MAX;
      a.m.lig_owner[l] = r;
      b = l;
      bump(a.m.counters, 4);
    }
  }
  a.m.rec_bond[r] = b;
  float fx = 0.f, fy = 0.f, fz = 0.f, w0 = 0.f, w1 = 0.f, w2 = 0.f;
  if (b >= 0) {
    float rp[3];
    int nodes[3];
    receptor_pos(a, r, rp, nodes);
    float d[3];
    ligand_vec(a, b, rp, d);
    const float dist = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (dist > 0.f) {
      const float c = a.bond_k * (dist - a.bond_l0) / dist;
      fx = c * d[0];
      fy = c * d[1];
      fz = c * d[2];
    }
    float inv[3];
    for (int k = 0; k < 3; ++k) {
      const float ex = rp[0] - a.m.x[nodes[k]];
      const float ey = rp[1] - a.m.y[nodes[k]];
      const float ez = rp[2] - a.m.z[nodes[k]];
      const float dr = sqrtf(ex * ex + ey * ey + ez * ez);
      inv[k] = dr > 0.f ? 1.0f / dr : 1e30f;
    }
    const float sum = inv[0] + inv[1] + inv[2];
    w0 = inv[0] / sum;
    w1 = inv[1] / sum;
    w2 = inv[2] / sum;
    atomicAdd((unsigned long long *)&a.m.counters[0], 1ull);
  }
  a.m.rec_force[3 * r] = fx;
  a.m.rec_force[3 * r + 1] = fy;
  a.m.rec_force[3 * r + 2] = fz;
  a.m.rec_w[3 * r] = w0;
  a.m.rec_w[3 * r + 1] = w1;
  a.m.rec_w[3 * r + 2] = w2;
}

/* ---------------- FORCES ---------------- */

__global__ __launch_bounds__(256) void forces_kernel(MArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && a.m.n_rec > 0 && a.detach_rate > 0.f) {
    /* detach test (semLB.cu:451-455): no bond held after this step's update */
    const int64_t step = *a.m.d_step;
    if (a.m.counters[0] == 0 && a.m.counters[1] == 0) {
      const float u = pfg_uniform(a.seed, step, 0u, PFG_RNG_DETACH);
      if (u < a.detach_rate) {
        a.m.counters[1] = 1;
        *a.m.detach_step = step;
      }
    }
  }
  if (i >= a.m.n) return;
  const int kind = a.m.kind[i];
  const float xi = a.m.x[i], yi = a.m.y[i], zi = a.m.z[i];
  float fx = 0.f, fy = 0.f, fz = 0.f;
  if (kind == PFG_MARKER_ELEMENT) {
    const int body = a.m.body[i];
    /* Morse over the cell list, same platelet only */
    if (a.morse_u0 != 0.f) {
      const int b = a.m.marker_bin[i];
      const int bx = b % a.nb[0], by = (b / a.nb[0]) % a.nb[1], bz = b / (a.nb[0] * a.nb[1]);
      int bxs[3], bys[3], bzs[3];
      const int nbx = nbr_bins(bx, a.nb[0], a.periodic[0], bxs);
      const int nby = nbr_bins(by, a.nb[1], a.periodic[1], bys);
      const int nbz = nbr_bins(bz, a.nb[2], a.periodic[2], bzs);
      const float two_rho_u0 = 2.0f * a.morse_rho * a.morse_u0;
      for (int jz = 0; jz < nbz; ++jz)
        for (int jy = 0; jy < nby; ++jy)
          for (int jx = 0; jx < nbx; ++jx) {
            const int bin = bxs[jx] + a.nb[0] * (bys[jy] + a.nb[1] * bzs[jz]);
            for (int k = a.m.bin_start[bin]; k < a.m.bin_start[bin + 1]; ++k) {
              const int j = a.m.bin_items[k];
              if (j == i || a.m.body[j] != body || a.m.kind[j] != PFG_MARKER_ELEMENT) continue;
              const float dx = a.m.x[j] - xi, dy = a.m.y[j] - yi, dz = a.m.z[j] - zi;
              const float r2 = dx * dx + dy * dy + dz * dz;
              if (r2 >= a.morse_cut2 || r2 == 0.f) continue;
              const float r = sqrtf(r2);
              const float e = expf(-a.morse_rho * (r - a.morse_req));
              /* -d(U S)/dr along (x_j - x_i)/r, U = U0 (e^2 - 2e) */
              float c = two_rho_u0 * (e - e * e);
              const float r1 = a.morse_cut - a.morse_taper;
              if (a.morse_taper > 0.f && r > r1) {
                const float t = (r - r1) / a.morse_taper;
                const float S = 1.0f - t * t * (3.0f - 2.0f * t);
                const float dS = 6.0f * t * (t - 1.0f) / a.morse_taper;
                c = c * S + a.morse_u0 * (e * e - 2.0f * e) * dS;
              }
              c /= r;
              fx += c * dx;
              fy += c * dy;
              fz += c * dz;
            }
          }
    }
    /* mesh-edge Hookean springs: k (r - L)/r (x_j - x_i) (sem_kernel.cu:374-384) */
    if (a.k_spring != 0.f) {
      for (int k = a.m.m2e_start[i]; k < a.m.m2e_start[i + 1]; ++k) {
        const int j = a.m.m2e_nbr[k];
        const float dx = a.m.x[j] - xi, dy = a.m.y[j] - yi, dz = a.m.z[j] - zi;
        const float r = sqrtf(dx * dx + dy * dy + dz * dz);
        if (r == 0.f) continue;
        const float c = a.k_spring * (r - a.m.m2e_l0[k]) / r;
        fx += c * dx;
        fy += c * dy;
        fz += c * dz;
      }
    }
    /* area and volume constraints over the incident triangles */
    if (a.ks != 0.f || a.kv != 0.f) {
      const double Sall = a.m.plt_S[body], S0all = a.m.plt_S0[body];
      const double V = a.m.plt_V[body], V0 = a.m.plt_V0[body];
      const float g_area = (float)((Sall - S0all) / S0all);
      const float cvol = -a.kv * (float)((V - V0) / V0) / 6.0f;
      for (int k = a.m.m2t_start[i]; k < a.m.m2t_start[i + 1]; ++k) {
        const int t = a.m.m2t[k];
        const int n[3] = {a.m.tri[3 * t], a.m.tri[3 * t + 1], a.m.tri[3 * t + 2]};
        const int s = n[0] == i ? 0 : (n[1] == i ? 1 : 2);
        const int s1 = n[(s + 1) % 3], s2 = n[(s + 2) % 3];
        /* L_s = x_{s+1} - x_{s+2} (sem_kernel.cu:1688-1690) */
        const float Lx = a.m.x[s1] - a.m.x[s2], Ly = a.m.y[s1] - a.m.y[s2],
                    Lz = a.m.z[s1] - a.m.z[s2];
        const float nx = a.m.tri_n[3 * t], ny = a.m.tri_n[3 * t + 1], nz = a.m.tri_n[3 * t + 2];
        /* dS/dx_s = -(L_s x n)/2 with the inward n_elem (sem_kernel.cu:1719-1721) */
        const float gx = -0.5f * (Ly * nz - Lz * ny);
        const float gy = -0.5f * (Lz * nx - Lx * nz);
        const float gz = -0.5f * (Lx * ny - Ly * nx);
        const double St = a.m.tri_S[t], S0t = a.m.tri_S0[t];
        const float ca = -a.ks * ((float)((St - S0t) / S0t) + g_area);
        fx += ca * gx;
        fy += ca * gy;
        fz += ca * gz;
        /* dV/dx_s = x_{s+1} x x_{s+2} (the reference's expansion at
         * sem_kernel.cu:1723-1731 reduces to this), taken about x_i: the sum over
         * a closed one-ring is unchanged by the shift and stays exact in FP32 */
        const float ax = a.m.x[s1] - xi, ay = a.m.y[s1] - yi, az = a.m.z[s1] - zi;
        const float bx = a.m.x[s2] - xi, by = a.m.y[s2] - yi, bz = a.m.z[s2] - zi;
        fx += cvol * (ay * bz - az * by);
        fy += cvol * (az * bx - ax * bz);
        fz += cvol * (ax * by - ay * bx);
      }
    }
    /* wall repulsion on the lowest element (sem_kernel.cu:409-430) */
    if (a.wall_a != 0.f && a.m.plt_lowest[body] == i) {
      float h = zi - a.wall_z;
      if (h <= a.wall_hc) {
        h = fmaxf(h, 0.01f * a.wall_hc);
        const float e = expf(-a.wall_kappa * h);
        fz += a.wall_a * a.wall_kappa * e / (1.0f - e);
      }
    }
    /* receptor bond forces, inverse-distance shares */
    for (int k = a.m.m2r_start[i]; k < a.m.m2r_start[i + 1]; ++k) {
      const int e = a.m.m2r[k];
      const int r = e / 3, slot = e % 3;
      const float w = a.m.rec_w[3 * r + slot];
      fx += w * a.m.rec_force[3 * r];
      fy += w * a.m.rec_force[3 * r + 1];
      fz += w * a.m.rec_force[3 * r + 2];
    }
  } else {
    /* fibrin links in node-gather form (fiber.cpp:647-706) */
    const float sgn = a.fiber_ref_sign ? -1.0f : 1.0f;
    for (int k = a.m.m2l_start[i]; k < a.m.m2l_start[i + 1]; ++k) {
      const int j = a.m.m2l_nbr[k];
      const float l0 = a.m.m2l_l0[k];
      const float dx = a.m.x[j] - xi, dy = a.m.y[j] - yi, dz = a.m.z[j] - zi;
      const float l = sqrtf(dx * dx + dy * dy + dz * dz);
      if (l == 0.f || l0 == 0.f) continue;
      float tension;
      if (a.fiber_law == PFG_FIBER_LINEAR) {
        /* Y A (l - l0)/l0 (network.cpp:364-373) */
        tension = a.fiber_k * (l - l0) / l0;
      } else {
        /* WLC (fiber.cpp:721-730, network.cpp:386-405): CC = kB T / Lp with
         * Lp = l0, tension CC (1/(4(1-x)^2) - 1/4 + x), x = (l - l0)/l0 */
        const float x = fminf((l - l0) / l0, a.fiber_xmax);
        const float tl = 1.0f - x;
        tension = (a.fiber_k / l0) * (0.25f / (tl * tl) - 0.25f + x);
      }
      const float c = sgn * tension / l; /* restoring: toward j when stretched */
      fx += c * dx;
      fy += c * dy;
      fz += c * dz;
    }
    /* reaction of a receptor bond held on this node */
    if (a.m.fibrin_lig) {
      const int owner = a.m.lig_owner[a.m.n_wall_lig + i];
      if (owner >= 0) {
        fx -= a.m.rec_force[3 * owner];
        fy -= a.m.rec_force[3 * owner + 1];
        fz -= a.m.rec_force[3 * owner + 2];
      }
    }
  }
  a.m.fx[i] = fx;
  a.m.fy[i] = fy;
  a.m.fz[i] = fz;
}

/* ---------------- UPDATE ---------------- */

__global__ void update_kernel(MArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.m.n) return;
  const int kind = a.m.kind[i];
  if (kind == PFG_MARKER_FIXED) return;
  if (kind == PFG_MARKER_FIBER && a.fiber_move == PFG_FIBER_KICK) {
    /* network.cpp:839-881: v <- e^{-g dt} v + F (1 - e^{-g dt})/g */
    const float dt = a.marker_dt;
    const float vdt = expf(-a.fiber_gamma * dt);
    const float fdt = a.fiber_gamma > 0.f ? (1.0f - vdt) / a.fiber_gamma : dt;
    const float vx = vdt * a.m.vx[i] + a.m.fx[i] * fdt;
    const float vy = vdt * a.m.vy[i] + a.m.fy[i] * fdt;
    const float vz = vdt * a.m.vz[i] + a.m.fz[i] * fdt;
    a.m.vx[i] = vx;
    a.m.vy[i] = vy;
    a.m.vz[i] = vz;
    a.m.x[i] += vx * dt;
    a.m.y[i] += vy * dt;
    a.m.z[i] += vz * dt;
    return;
  }
  /* explicit Euler with the interpolated velocity (sem_kernel.cu:2112-2138);
   * positions stay unwrapped, the wrap is applied where the lattice is indexed */
  a.m.x[i] += a.m.vx[i] * a.marker_dt;
  a.m.y[i] += a.m.vy[i] * a.marker_dt;
  a.m.z[i] += a.m.vz[i] * a.marker_dt;
}

/* ---------------- launchers ---------------- */

static void make_margs(const pfg_handle *h, MArgs *a) {
  const pfg_params &p = h->p;
  a->m = h->m;
  a->nx = p.nx;
  a->ny = p.ny;
  a->nz = p.nz;
  for (int k = 0; k < 3; ++k) {
    a->periodic[k] = p.wall[k] == PFG_WALL_PERIODIC;
    a->nb[k] = h->m.nb[k];
  }
  a->morse_u0 = p.morse_u0;
  a->morse_rho = p.morse_rho;
  a->morse_req = p.morse_req;
  a->morse_cut2 = p.morse_cutoff * p.morse_cutoff;
  a->morse_cut = p.morse_cutoff;
  a->morse_taper = p.morse_taper;
  a->k_spring = p.k_spring;
  a->ks = p.ks;
  a->kv = p.kv;
  a->wall_a = p.wall_rep_a;
  a->wall_kappa = p.wall_rep_kappa;
  a->wall_hc = p.wall_rep_hc;
  a->wall_z = p.wall_z;
  a->k_on = p.k_on;
  a->k_off0 = p.k_off0;
  a->bond_k = p.bond_k;
  a->bond_l0 = p.bond_l0;
  a->bond_reach = p.bond_reach;
  a->bell_xb_kT = p.bell_xb_kT;
  a->detach_rate = p.detach_rate;
  a->seed = p.seed;
  a->fiber_law = p.fiber_law;
  a->fiber_move = p.fiber_move;
  a->fiber_ref_sign = p.fiber_ref_sign;
  a->fiber_k = p.fiber_k;
  a->fiber_xmax = p.fiber_xmax;
  a->fiber_gamma = p.fiber_gamma;
  a->marker_dt = p.marker_dt;
  a->bin_cells = p.bin_cells;
}

static inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

hipError_t launch_cells(pfg_handle *h) {
  const Markers &m = h->m;
  if (m.n == 0) return hipSuccess;
  MArgs a;
  make_margs(h, &a);
  hipLaunchKernelGGL(bins_kernel, grid_for(m.n), dim3(256), 0, h->stream, a);
  hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)m.scan_blocks), dim3(256), 0,
                     h->stream, m.bin_count, m.nbins, m.scan_partials);
  hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(256), 0, h->stream,
                     m.scan_partials, m.scan_blocks, m.bin_start, m.nbins);
  hipLaunchKernelGGL(scan_apply_kernel, dim3((unsigned)m.scan_blocks), dim3(256), 0,
                     h->stream, m.bin_count, m.nbins, m.scan_partials, m.bin_start);
  hipLaunchKernelGGL(scatter_kernel, grid_for(m.n), dim3(256), 0, h->stream, a);
  hipLaunchKernelGGL(sort_bins_kernel, grid_for(m.nbins), dim3(256), 0, h->stream, a);
  return hipGetLastError();
}

hipError_t launch_membrane(pfg_handle *h) {
  if (h->m.n_plt == 0) return hipSuccess;
  MArgs a;
  make_margs(h, &a);
  hipLaunchKernelGGL(membrane_kernel, dim3(h->m.n_plt), dim3(256), 0, h->stream, a);
  return hipGetLastError();
}

hipError_t launch_bonds(pfg_handle *h) {
  if (h->m.n_rec == 0) return hipSuccess;
  MArgs a;
  make_margs(h, &a);
  hipLaunchKernelGGL(bond_decide_kernel, grid_for(h->m.n_rec), dim3(256), 0,
                     h->stream, a);
  hipLaunchKernelGGL(bond_resolve_kernel, grid_for(h->m.n_rec), dim3(256), 0,
                     h->stream, a);
  return hipGetLastError();
}

hipError_t launch_forces(pfg_handle *h) {
  if (h->m.n == 0) return hipSuccess;
  MArgs a;
  make_margs(h, &a);
  hipLaunchKernelGGL(forces_kernel, grid_for(h->m.n), dim3(256), 0, h->stream, a);
  return hipGetLastError();
}

hipError_t launch_update(pfg_handle *h) {
  if (h->m.n == 0) return hipSuccess;
  MArgs a;
  make_margs(h, &a);
  hipLaunchKernelGGL(update_kernel, grid_for(h->m.n), dim3(256), 0, h->stream, a);
  return hipGetLastError();
}

}  // namespace pfg
