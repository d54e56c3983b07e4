/*
 * markers.hip -- marker-side kernels of the coupled step and the immersed-
 * boundary coupling.
 *
 *  CELLS     integer cell lists: bin per marker, counting sort with a
 *            wavefront prefix scan, ascending marker order inside each bin.
 *  MEMBRANE  per-triangle area/normal/volume terms and per-platelet FP64 sums
 *            (sem_calculate_A_kernel, sem_kernel.cu:724-872) plus the lowest
 *            element for the wall repulsion (sem_kernel.cu:409-430).
 *  BONDS     Bell-model receptor-ligand rupture/formation with the counter RNG
 *            (sem_platelet_wall_kernel, sem_kernel.cu:1750-2083).
 *  FORCES    one wavefront per marker gathers Morse pairs over the 27
 *            neighbouring bins, mesh-edge springs (sem_Force_kernel,
 *            sem_kernel.cu:320-513), area/volume constraints
 *            (sem_ForceAreaVol_kernel, sem_kernel.cu:1640-1749), wall repulsion,
 *            receptor bond forces and fibrin links (fiber.cpp:647-733,
 *            network.cpp:303-437), then reduces across the wavefront.
 *  SPREAD    IB force spreading (fluid3d_force_distribute_kernel,
 *            LB_kernel.cu:1419-1520) into the lattice force field with FP32
 *            atomics: per (marker, stencil node), or staged through an LDS
 *            box per chunk of markers when many markers share the lattice;
 *            tile flags.
 *  INTERP/UPDATE  IB interpolation (fluid3d_velocity_distribute_kernel,
 *            LB_kernel.cu:1524-1642) as a wavefront gather over the 4x4x4
 *            stencil, then x += V dt (sem_update_position_kernel,
 *            sem_kernel.cu:2087-2148) or the damped kick of network.cpp:839-881.
 *  CLEAR     zero the force/velocity field in the tiles flagged this step.
 *
 * The stages of one step are fused into few launches (a launch costs several
 * microseconds, more than most of these kernels' work at one platelet): each
 * fused kernel takes a role mask and splits its grid into role ranges.
 *
 * The 4-point cosine kernel is phi(r) = (1 + cos(pi r / 2)) / 4 for |r| < 2
 * (LB_kernel.cu:1477-1478), applied as a product over the three axes.
 *
 * Every force is a gather in a fixed order and every cross-lane reduction is a
 * fixed butterfly, so forces and velocities are run-to-run deterministic; only
 * the spread force field sums FP32 atomics in arrival order.
 */
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "pfg_internal.h"
#include "pfg_rng.h"

namespace pfg {

struct MArgs {
  Markers m;
  /* lattice side of the IB coupling */
  MLat lat;               /* the lattice of a single-lattice handle           */
  const MLat *lats;       /* batches: [K], selected by the marker's lattice   */
  const int64_t *tile0s;  /* batches: first global tile per lattice, [K+1]    */
  int K;
  int64_t ntiles;         /* tiles over all lattices                          */
  int64_t plane;          /* stride between the components of the force field */
  /* parameters */
  float morse_u0, morse_rho, morse_req, morse_cut2, morse_cut, morse_taper;
  float morse_eps2;       /* |r^2 - cutoff^2| below this: near the cutoff      */
  float k_spring, ks, kv;
  float wall_a, wall_kappa, wall_hc, wall_z;
  float k_on, k_off0, bond_k, bond_l0, bond_reach, bell_xb_kT, detach_rate;
  uint64_t seed;
  int fiber_law, fiber_move, fiber_ref_sign;
  float fiber_k, fiber_xmax, fiber_gamma;
  float marker_dt;
  float force_scale;
  float kbend;
  float contact_a, contact_kappa, contact_hc;
  float kick_temp;
  int sem_integrator;
  float sem_mass, sem_gamma, sem_temp, sem_w;
  int bond_law;
  float bell_kT;
  /* role split of the fused launches */
  int roles;
  int halt; /* 1: after the detach test fires, leave bonds, force field and
               markers unchanged (the step loop ends at detach) */
  int mark_rows; /* 1: stamp the x-block rows this step's IB stencils touch */
  int resolve_last; /* 1: the gather launch's last workgroup resolves the bonds */
  int count_spans;  /* 1: the counting launch adds the scan's span totals */
  int bin_elements; /* 0: the cell lists hold the markers other than elements
                       (every element's Morse partners come from its
                       platelet's own lists and nothing else reads them) */
  int spread_scratch; /* 1: the chunks of bodies write scratch boxes (form 4) */
  const int32_t *force_list; /* force_count >= 0: the force pass takes these */
  int force_count;           /* markers only; -1: every marker                */
  const int32_t *move_list;  /* move_count >= 0: move_kernel takes these       */
  int move_count;            /* markers only (and leaves the step counter)    */
  int blocks0, blocks1, blocks2, blocks3; /* blocks of role ranges 0..3 */
};

enum : int {
  ROLE_A = 1, /* first role of a fused kernel  */
  ROLE_B = 2, /* second role                   */
  ROLE_C = 4, /* third role                    */
  ROLE_D = 8  /* fourth role                   */
};

constexpr float BOND_EPS_P = 1e-6f; /* |u - p| below this: near-threshold      */
constexpr float BOND_EPS_R = 1e-4f; /* |r - reach| below this: near-threshold  */

/* ---------------- small helpers ---------------- */

/* The detach test has fired and the run halts there (written by an earlier
 * launch, so every lane reads the same value). */
__device__ inline bool halted(const MArgs &a) { return a.halt && a.m.counters[1] != 0; }

__device__ void spread_direct(const MArgs &a, int i, int lane, float Fx, float Fy, float Fz);
/* body spreading's tile claims: workgroup `block` of the claim blocks the
 * force launch appends (blocks1 bodies, then blocks2 of 256 other markers) */
__device__ void claim_role(const MArgs &a, int block);

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

/* The lattice of marker i. */
__device__ inline MLat lat_of(const MArgs &a, int i) {
  return a.lats ? a.lats[a.m.minst[i]] : a.lat;
}

__device__ inline void bin3_of(const MLat &L, float x, float y, float z, int b[3]) {
  b[0] = cell_of(x, L.nx, L.periodic[0]) * L.nb[0] / L.nx;
  b[1] = cell_of(y, L.ny, L.periodic[1]) * L.nb[1] / L.ny;
  b[2] = cell_of(z, L.nz, L.periodic[2]) * L.nb[2] / L.nz;
}

/* The lattice with the coarse cell-list grid in place of the full one (the
 * lists of the markers other than elements). */
__device__ inline MLat coarse_grid(const MLat &L) {
  MLat c = L;
#pragma unroll
  for (int k = 0; k < 3; ++k) c.nb[k] = L.nbc[k];
  c.bin0 = L.binc0;
  return c;
}

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

/* Global index of the k-th of the (up to 27) bins around bin (b3) of lattice
 * L; -1 when k is past the end. */
__device__ inline int nth_nbr_bin(const MLat &L, const int b3[3], int k) {
  int bx[3], by[3], bz[3];
  const int nx = nbr_bins(b3[0], L.nb[0], L.periodic[0], bx);
  const int ny = nbr_bins(b3[1], L.nb[1], L.periodic[1], by);
  const int nz = nbr_bins(b3[2], L.nb[2], L.periodic[2], bz);
  if (k >= nx * ny * nz) return -1;
  const int jx = k % nx, jy = (k / nx) % ny, jz = k / (nx * ny);
  return L.bin0 + bx[jx] + L.nb[0] * (by[jy] + L.nb[1] * bz[jz]);
}

__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
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

__device__ inline void bump(int64_t *c, int k) {
  atomicAdd((unsigned long long *)(c + k), 1ull);
}

__device__ inline bool ib_coupled(int kind, int fiber_move) {
  return kind == PFG_MARKER_ELEMENT || (kind == PFG_MARKER_FIBER && fiber_move == PFG_FIBER_IB);
}

__device__ inline float wrap_coord(float v, int n, int periodic) {
  if (!periodic) return v;
  const float fn = (float)n;
  float w = v - fn * floorf(v / fn);
  if (w >= fn) w -= fn;
  if (w < 0.f) w += fn;
  return w;
}

/* Lattice index along an axis, or -1 when outside a wall-bounded axis.
 * Stencil and box nodes lie within one period of the lattice, where one
 * add or subtract wraps them; the remainder handles any other index. */
__device__ inline int axis_index(int i, int n, int periodic) {
  if (periodic) {
    if (i < 0) i += n;
    else if (i >= n) i -= n;
    if ((unsigned)i >= (unsigned)n) {
      i %= n;
      if (i < 0) i += n;
    }
    return i;
  }
  return (i < 0 || i >= n) ? -1 : i;
}

/* k / d for 0 <= k < 2048 and 1 <= d <= 512 (a box of at most 2048 nodes is
 * at least 4 nodes along every axis), by the FP32 reciprocal inv = 1/d:
 * (k + 1/2) / d lies at least 1/(2d) >= 2^-10 from an integer, and the
 * product (below 2049) carries an error below 2^-12, so its floor is the
 * quotient. */
static_assert(IB_BOX_CELLS <= 1024, "div_small covers boxes of at most 2048 nodes");
__device__ inline int div_small(int k, float inv) { return (int)(((float)k + 0.5f) * inv); }

/* The 4-point cosine kernel phi(r) = (1 + cos(pi r / 2)) / 4 at the four
 * nodes floor(W) - 1 + k (k = 0..3) of coordinate W: with d = W - floor(W),
 * r = d + 1 - k and cos(pi r / 2) is -sin, cos, sin, -cos of pi d / 2, so
 * one sine-cosine pair gives all four weights. */
__device__ inline void stencil_weights(float W, float w[4]) {
  const float d = W - floorf(W);
  float sn, cs;
  sincospif(0.5f * d, &sn, &cs);
  w[0] = 0.25f * (1.0f - sn);
  w[1] = 0.25f * (1.0f + cs);
  w[2] = 0.25f * (1.0f + sn);
  w[3] = 0.25f * (1.0f - cs);
}

/* Weight k of stencil_weights, for a lane that needs one of them. */
__device__ inline float stencil_weight(float W, int k) {
  const float d = W - floorf(W);
  float sn, cs;
  sincospif(0.5f * d, &sn, &cs);
  const float v = (k & 1) ? cs : sn;
  return 0.25f * ((k == 0 || k == 3) ? 1.0f - v : 1.0f + v);
}

/* Stencil node k (0..3) along one axis for wrapped coordinate X: lattice index
 * (or -1) and weight phi(X - i). */
__device__ inline int stencil_node(float X, int k, int n, int periodic, float *w) {
  *w = stencil_weight(X, k);
  return axis_index((int)floorf(X) - 1 + k, n, periodic);
}

/* ---------------- K1: tile clear + bin counting ---------------- */

/* Lattice of global tile g of a batch (tile0s ascending). */
__device__ inline int lat_of_tile(const MArgs &a, int64_t g) {
  int lo = 0, hi = a.K - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a.tile0s[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

/* Zero the force of the tiles that this step's spreading touched and take
 * their stamps back (a spread that no LB update will consume): block `block`
 * scans 256 tile stamps and zeroes the stamped tiles together. */
__device__ void clear_tiles(const MArgs &a, int block) {
  __shared__ int64_t list[256];
  __shared__ int count;
  if (threadIdx.x == 0) count = 0;
  __syncthreads();
  int32_t *stamps = a.lats ? a.lats[0].tile_stamp : a.lat.tile_stamp; /* all lattices' tiles */
  const int32_t step = (int32_t)*a.m.d_step;
  const int64_t my = (int64_t)block * 256 + threadIdx.x;
  const bool mine = my < a.ntiles && stamps[my] == step;
  if (mine) list[atomicAdd(&count, 1)] = my;
  __syncthreads();
  /* thread t takes cell t % TILE_CELLS of tile k = t / TILE_CELLS (mod the
   * tiles per pass) */
  constexpr int PER = 256 / TILE_CELLS;
  static_assert(PER * TILE_CELLS == 256, "whole tiles per pass");
  const int t = threadIdx.x % TILE_CELLS;
  const int lx = t & (TILE_X - 1);
  const int ly = (t >> TILE_SHIFT_X) & (TILE_Y - 1);
  const int lz = t >> (TILE_SHIFT_X + TILE_SHIFT_Y);
  for (int k = (int)threadIdx.x / TILE_CELLS; k < count; k += PER) {
    const MLat L = a.lats ? a.lats[lat_of_tile(a, list[k])] : a.lat;
    const int64_t tile = list[k] - L.tile0;
    const int tx = (int)(tile % L.tnx);
    const int ty = (int)((tile / L.tnx) % L.tny);
    const int tz = (int)(tile / ((int64_t)L.tnx * L.tny));
    const int x = tx * TILE_X + lx, y = ty * TILE_Y + ly, z = tz * TILE_Z + lz;
    if (x < L.nx && y < L.ny && z < L.nz) {
      const int64_t c = x + (int64_t)L.nx * (y + (int64_t)L.ny * z);
      L.aux[c] = 0.f;
      L.aux[a.plane + c] = 0.f;
      L.aux[2 * a.plane + c] = 0.f;
    }
  }
  if (mine) stamps[my] = -1;
}

__device__ inline void receptor_pos(const MArgs &a, int r, float out[3], int nodes[3]) {
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

/* Tag of a marker in the bin-sorted packed positions: its platelet for an
 * element, -1 for a fibrin node (Morse acts between elements of one body). */
__device__ inline float morse_tag(const MArgs &a, int j) {
  return __int_as_float(a.m.kind[j] == PFG_MARKER_ELEMENT ? a.m.body[j] : -1);
}

/* roles: A = clear this step's force tiles (blocks0 blocks, the CLEAR stage
 * only), B = count bins and pack each marker's position with its Morse tag
 * (blocks1), C = receptor positions (blocks2) */
__global__ __launch_bounds__(256) void begin_kernel(MArgs a) {
  int blk = blockIdx.x;
  if (a.roles & ROLE_A) {
    if (blk < a.blocks0) {
      clear_tiles(a, blk);
      return;
    }
    blk -= a.blocks0;
  }
  if (a.roles & ROLE_C) {
    if (!(a.roles & ROLE_B) || blk >= a.blocks1) {
      const int r = (blk - ((a.roles & ROLE_B) ? a.blocks1 : 0)) * 256 + threadIdx.x;
      if (r < a.m.n_rec) {
        float rp[3];
        int nodes[3];
        receptor_pos(a, r, rp, nodes);
        const int lat = a.lats ? a.m.minst[nodes[0]] : 0;
        a.m.rec_pos[r] = make_float4(rp[0], rp[1], rp[2], __int_as_float(lat));
      }
      return;
    }
  }
  if (!(a.roles & ROLE_B)) return;
  const int i = blk * 256 + threadIdx.x;
  if (i == 0) {
    a.m.counters[0] = 0;          /* bonds held, counted again this step      */
    a.m.counters[6] = *a.m.d_step; /* this step's index, for the later launches
                                     (the last one advances d_step)            */
    a.m.counters[7] = 0;          /* platelet pairs in contact, found again    */
  }
  if (i >= a.m.n) return;
  const MLat L = lat_of(a, i);
  const int kind_i = a.m.kind[i];
  const float4 p = make_float4(a.m.x[i], a.m.y[i], a.m.z[i],
                               __int_as_float(kind_i == PFG_MARKER_ELEMENT ? a.m.body[i] : -1));
  a.m.pos4[i] = p;
  const bool binned = a.bin_elements || kind_i != PFG_MARKER_ELEMENT;
  const MLat G = a.bin_elements ? L : coarse_grid(L); /* the lists' grid */
  int b3[3];
  bin3_of(G, p.x, p.y, p.z, b3);
  const int b = G.bin0 + b3[0] + G.nb[0] * (b3[1] + G.nb[1] * b3[2]);
  a.m.marker_bin[i] = binned ? b : -1;
  if (binned) a.m.marker_slot[i] = atomicAdd(&a.m.bin_count[b], 1);
  if (a.count_spans) {
    /* the scan's per-span totals, one atomic per span a wavefront touches
     * (its markers are consecutive, so mostly one); the scan then needs no
     * reduction pass, and the place kernel clears them for the next step.
     * With many markers the waves meet on few spans and these atomics
     * serialise: the reduction launch is cheaper then. */
    const int span = b / SCAN_SPAN, lane = threadIdx.x & 63;
    uint64_t todo = __ballot(binned);
    while (todo) {
      const int leader = __ffsll((unsigned long long)todo) - 1;
      const int ls = __shfl(span, leader, 64);
      const uint64_t same = __ballot(span == ls) & todo;
      if (lane == leader) atomicAdd(&a.m.scan_partials[ls], (int)__popcll(same));
      todo &= ~same;
    }
  }
  if (a.mark_rows && ib_coupled(a.m.kind[i], a.fiber_move) && !halted(a)) {
    /* stamp the rows of this step's 4x4x4 stencil (the nodes spreading and
     * interpolation use), so that the LB update of every other row may run
     * before the spreading */
    const int step = (int)*a.m.d_step;
    const int ox = (int)floorf(wrap_coord(a.m.x[i], L.nx, L.periodic[0])) - 1;
    const int oy = (int)floorf(wrap_coord(a.m.y[i], L.ny, L.periodic[1])) - 1;
    const int oz = (int)floorf(wrap_coord(a.m.z[i], L.nz, L.periodic[2])) - 1;
    for (int kz = 0; kz < 4; ++kz) {
      const int z = axis_index(oz + kz, L.nz, L.periodic[2]);
      if (z < 0) continue;
      for (int ky = 0; ky < 4; ++ky) {
        const int y = axis_index(oy + ky, L.ny, L.periodic[1]);
        if (y < 0) continue;
        int last = -1;
        for (int kx = 0; kx < 4; ++kx) {
          const int x = axis_index(ox + kx, L.nx, L.periodic[0]);
          if (x < 0 || (x >> 6) == last) continue;
          last = x >> 6;
          L.row_stamp[last + (int64_t)L.nxb * (y + (int64_t)L.ny * z)] = step;
        }
      }
    }
  }
}

/* ---------------- K2..K4: prefix scan of the bin counts ---------------- */

/* Two launches: block sums of SCAN_SPAN bins each, then per block the sum of
 * the preceding block sums (at most a few hundred) and the scan of its own
 * bins.  bin_count is padded with zeros to whole blocks. */
__device__ inline int block_sum(int v, int *sh_wave) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh_wave[threadIdx.x >> 6] = v;
  __syncthreads();
  const int t = sh_wave[0] + sh_wave[1] + sh_wave[2] + sh_wave[3];
  __syncthreads();
  return t;
}

/* Per-span totals of the bin counts (many markers; few markers add them
 * while counting). */
__global__ __launch_bounds__(256) void scan_reduce_kernel(const int32_t *count,
                                                          int32_t *partials) {
  __shared__ int sh[4];
  const int4 *c = reinterpret_cast<const int4 *>(count + (int64_t)blockIdx.x * SCAN_SPAN);
  int s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_SPAN / 1024; ++k) {
    const int4 v = c[k * 256 + threadIdx.x];
    s += v.x + v.y + v.z + v.w;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

/* Each lane scans 32 consecutive bins of the block in registers; one block
 * scan of the lane totals places them. */
__global__ __launch_bounds__(256) void scan_apply_kernel(int32_t *count, int64_t nbins,
                                                         const int32_t *partials,
                                                         int32_t *bin_start) {
  static_assert(SCAN_SPAN == 256 * 32, "32 bins per lane");
  __shared__ int sh[4];
  int off = 0;
  for (int k = threadIdx.x; k < (int)blockIdx.x; k += 256) off += partials[k];
  off = block_sum(off, sh);
  const int64_t base = (int64_t)blockIdx.x * SCAN_SPAN + threadIdx.x * 32;
  int4 *c = reinterpret_cast<int4 *>(count + base);
  int v[32];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int4 q = c[k];
    v[4 * k] = q.x;
    v[4 * k + 1] = q.y;
    v[4 * k + 2] = q.z;
    v[4 * k + 3] = q.w;
  }
  int run = 0;
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    const int t = v[k];
    v[k] = run;
    run += t;
  }
  const int incl = wave_incl_scan(run);
  if ((threadIdx.x & 63) == 63) sh[threadIdx.x >> 6] = incl;
  __syncthreads();
  int lane_off = off + incl - run;
  for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) lane_off += sh[w];
  if (base + 32 <= nbins) {
    int4 *o = reinterpret_cast<int4 *>(bin_start + base);
#pragma unroll
    for (int k = 0; k < 8; ++k)
      o[k] = make_int4(v[4 * k] + lane_off, v[4 * k + 1] + lane_off, v[4 * k + 2] + lane_off,
                       v[4 * k + 3] + lane_off);
  } else {
#pragma unroll
    for (int k = 0; k < 32; ++k)
      if (base + k < nbins) bin_start[base + k] = v[k] + lane_off;
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) c[k] = make_int4(0, 0, 0, 0); /* ready for the next step */
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 255) bin_start[nbins] = lane_off + run;
}

/* ---------------- K5: scatter + triangle geometry ---------------- */

__device__ void triangle_geometry(const MArgs &a, int t) {
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
  const double cx = (x0 + x1 + x2) / 3.0, cy = (y0 + y1 + y2) / 3.0, cz = (z0 + z1 + z2) / 3.0;
  a.m.tri_S[t] = St;
  a.m.tri_V[t] = (cx * nx + cy * ny + cz * nz) * St / 3.0; /* sem_kernel.cu:806-808 */
  /* per vertex slot s: dS/dx_s = -(L_s x n)/2 with L_s = x_{s+1} - x_{s+2}
   * and the inward n_elem (sem_kernel.cu:1688-1690, 1719-1721), and the
   * local area strain; the volume gradient at every vertex of the triangle
   * is (x1 - x0) x (x2 - x0) (the reference's expansion at
   * sem_kernel.cu:1723-1731 reduces to it) */
  /* the gradients below feed the area and volume forces only */
  if (a.ks == 0.f && a.kv == 0.f) return;
  const double lf = (St - a.m.tri_S0[t]) / a.m.tri_S0[t];
  const double L[3][3] = {{x1 - x2, y1 - y2, z1 - z2},
                          {x2 - x0, y2 - y0, z2 - z0},
                          {x0 - x1, y0 - y1, z0 - z1}};
  for (int s = 0; s < 3; ++s) {
    const double gx = -0.5 * (L[s][1] * nz - L[s][2] * ny);
    const double gy = -0.5 * (L[s][2] * nx - L[s][0] * nz);
    const double gz = -0.5 * (L[s][0] * ny - L[s][1] * nx);
    a.m.tri_g[3 * t + s] = make_float4((float)gx, (float)gy, (float)gz, (float)lf);
  }
  a.m.tri_a2[t] = make_float4((float)(ay * bz - az * by), (float)(az * bx - ax * bz),
                              (float)(ax * by - ay * bx), 0.f);
}

/* ---------------- membrane bending ---------------- */

/* Quadratic 6-node triangle: N_k = a xi^2 + b eta^2 + c xi eta + d xi + e eta + f
 * with rows {a, b, c, d, e, f} of the standard quadratic Lagrange triangle
 * (nodes at the vertices, then the midpoints of v0v1, v0v2, v1v2), and the
 * natural coordinates of the 6 nodes (sem_kernel.cu:94-107). */
__constant__ double QN[6][6] = {{2, 2, 4, -3, -3, 1}, {2, 0, 0, -1, 0, 0},
                                {0, 2, 0, 0, -1, 0},  {-4, 0, -4, 4, 0, 0},
                                {0, -4, -4, 0, 4, 0}, {0, 0, 4, 0, 0, 0}};
__constant__ double QXE[6][2] = {{0, 0}, {1, 0}, {0, 1}, {0.5, 0}, {0, 0.5}, {0.5, 0.5}};

/* Forward elimination with partial pivoting and unit pivots on a 3 x 4
 * augmented matrix (gaussian_elimination, sem_kernel.cu:249-284). */
__device__ inline void gauss3x4(double g[12]) {
  int i = 0;
  for (int j = 0; j < 3 && i < 3; ++j) {
    int maxi = i;
    for (int k = i + 1; k < 3; ++k)
      if (fabs(g[k * 4 + j]) > fabs(g[maxi * 4 + j])) maxi = k;
    if (g[maxi * 4 + j] == 0.0) continue;
    double initial = 0.0;
    for (int k = j; k < 4; ++k) {
      const double b = g[i * 4 + k];
      g[i * 4 + k] = g[maxi * 4 + k];
      g[maxi * 4 + k] = b;
      if (k == j) initial = g[i * 4 + j];
      g[i * 4 + k] /= initial;
    }
    for (int k = i + 1; k < 3; ++k) {
      const double f = g[k * 4 + j];
      for (int l = j; l < 4; ++l) g[k * 4 + l] -= f * g[i * 4 + l];
    }
    ++i;
  }
}

__device__ inline void pos_d(const MArgs &a, int i, double x[3]) {
  const float4 p = a.m.pos4[i];
  x[0] = p.x;
  x[1] = p.y;
  x[2] = p.z;
}

/* Unit normal -(e12 x e13)/|e12 x e13| of triangle t (sem_kernel.cu:793-803). */
__device__ inline void tri_unit_normal(const MArgs &a, int t, double n[3]) {
  double p0[3], p1[3], p2[3];
  pos_d(a, a.m.tri[3 * t], p0);
  pos_d(a, a.m.tri[3 * t + 1], p1);
  pos_d(a, a.m.tri[3 * t + 2], p2);
  const double e12[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
  const double e13[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
  n[0] = e12[1] * e13[2] - e12[2] * e13[1];
  n[1] = e12[2] * e13[0] - e12[0] * e13[2];
  n[2] = e12[0] * e13[1] - e12[1] * e13[0];
  const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  n[0] /= -nn;
  n[1] /= -nn;
  n[2] /= -nn;
}

/* Node normal k (k < n: vertex k; else the midpoint of triangle edge k - n):
 * mean of the unit normals of its triangles, normalised
 * (sem_calculate_A_kernel and sem_surface_tension_kernel,
 * sem_kernel.cu:810-820, 874-900). */
__device__ void bend_normal(const MArgs &a, int k) {
  double s[3] = {0.0, 0.0, 0.0}, nt[3];
  int cnt = 0;
  double *out;
  if (k < a.m.n) {
    for (int q = a.m.m2t_start[k]; q < a.m.m2t_start[k + 1]; ++q, ++cnt) {
      tri_unit_normal(a, a.m.m2t[q] / 3, nt);
      s[0] += nt[0];
      s[1] += nt[1];
      s[2] += nt[2];
    }
    out = a.m.vn + 3 * (int64_t)k;
  } else {
    const int e = k - a.m.n;
    for (int j = 0; j < 2; ++j) {
      const int q = a.m.edge_tri[2 * e + j];
      if (q < 0) continue;
      tri_unit_normal(a, q / 3, nt);
      s[0] += nt[0];
      s[1] += nt[1];
      s[2] += nt[2];
      ++cnt;
    }
    out = a.m.en + 3 * (int64_t)e;
  }
  if (cnt == 0) return;
  for (int d = 0; d < 3; ++d) s[d] /= cnt;
  const double nn = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  for (int d = 0; d < 3; ++d) out[d] = s[d] / nn;
}

/* Mean curvature of vertex v: at its node of every incident 6-node triangle
 * (midpoints at the vertex averages) solve K x_xi = n_xi, K x_eta = n_eta,
 * K n = 0 row by row (solve_linear_eqns); km = trace of the mean K / 2
 * (sem_calculate_Kapa_kernel and sem_calculate_m_kernel,
 * sem_kernel.cu:950-1100; the symmetrisation there keeps the trace). */
__device__ void bend_km(const MArgs &a, int v) {
  const int q0 = a.m.m2t_start[v], cnt = a.m.m2t_start[v + 1] - q0;
  if (cnt == 0) return;
  double tr = 0.0;
  for (int q = q0; q < q0 + cnt; ++q) {
    const int t = a.m.m2t[q] / 3, s = a.m.m2t[q] % 3;
    double xc[6][3], nc[6][3];
    for (int k = 0; k < 3; ++k) {
      const int vk = a.m.tri[3 * t + k];
      pos_d(a, vk, xc[k]);
      for (int d = 0; d < 3; ++d) nc[k][d] = a.m.vn[3 * (int64_t)vk + d];
      const int e = a.m.tri_edge[3 * t + k];
      for (int d = 0; d < 3; ++d) nc[3 + k][d] = a.m.en[3 * (int64_t)e + d];
    }
    for (int d = 0; d < 3; ++d) {
      xc[3][d] = 0.5 * (xc[0][d] + xc[1][d]);
      xc[4][d] = 0.5 * (xc[0][d] + xc[2][d]);
      xc[5][d] = 0.5 * (xc[1][d] + xc[2][d]);
    }
    const double xi = QXE[s][0], et = QXE[s][1];
    double A1[3] = {0, 0, 0}, A2[3] = {0, 0, 0}, A3[3] = {0, 0, 0}, A4[3] = {0, 0, 0};
    for (int k = 0; k < 6; ++k) {
      const double dxi = 2.0 * QN[k][0] * xi + QN[k][2] * et + QN[k][3];
      const double det = 2.0 * QN[k][1] * et + QN[k][2] * xi + QN[k][4];
      for (int d = 0; d < 3; ++d) {
        A1[d] += xc[k][d] * dxi;
        A2[d] += nc[k][d] * dxi;
        A3[d] += xc[k][d] * det;
        A4[d] += nc[k][d] * det;
      }
    }
    for (int r = 0; r < 3; ++r) { /* row r of K; only K[r][r] enters km */
      double g[12];
      for (int c = 0; c < 3; ++c) {
        g[c] = A1[c];
        g[4 + c] = A3[c];
        g[8 + c] = nc[s][c];
      }
      g[3] = A2[r];
      g[7] = A4[r];
      g[11] = 0.0;
      gauss3x4(g);
      const double k2 = g[11];
      const double k1 = g[7] - k2 * g[6];
      const double k0 = g[3] - k2 * g[2] - k1 * g[1];
      tr += r == 0 ? k0 : (r == 1 ? k1 : k2);
    }
  }
  a.m.km[v] = 0.5 * tr / cnt;
}

__device__ inline double dist_d(const double a[3], const double b[3]) {
  return sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
              (a[2] - b[2]) * (a[2] - b[2]));
}

/* Ring entry q of vertex i (neighbour j, previous x1, next x2): the cotangent
 * term 0.5 (cot a + cot b)(km_j - km_i) and the area of vertex i's region in
 * triangle (i, j, x2) (sem_Laplace_Kapa_kernel, sem_kernel.cu:1533-1628). */
__device__ void ring_term(const MArgs &a, int i, int r0, int c, int q, double *sum,
                          double *area) {
  const int j = a.m.ring[r0 + q];
  const int v1 = a.m.ring[r0 + (q == 0 ? c - 1 : q - 1)];
  const int v2 = a.m.ring[r0 + (q == c - 1 ? 0 : q + 1)];
  double xi[3], xj[3], x1[3], x2[3];
  pos_d(a, i, xi);
  pos_d(a, j, xj);
  pos_d(a, v1, x1);
  pos_d(a, v2, x2);
  double ri1[3], r1j[3], rij[3], ri2[3], r2j[3], cij[3], ci2[3], c2j[3], ce[3];
  for (int k = 0; k < 3; ++k) {
    ri1[k] = xi[k] - x1[k];
    r1j[k] = x1[k] - xj[k];
    rij[k] = xi[k] - xj[k];
    ri2[k] = xi[k] - x2[k];
    r2j[k] = x2[k] - xj[k];
    cij[k] = 0.5 * (xi[k] + xj[k]);
    ci2[k] = 0.5 * (xi[k] + x2[k]);
    c2j[k] = 0.5 * (x2[k] + xj[k]);
  }
  auto dot = [](const double u[3], const double w[3]) {
    return u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
  };
  const double ai = dot(rij, ri2), a2 = -dot(ri2, r2j), aj = dot(rij, r2j),
               a1 = -dot(r1j, ri1);
  if (ai < 0) {
    for (int k = 0; k < 3; ++k) ce[k] = c2j[k];
  } else if (a2 < 0) {
    for (int k = 0; k < 3; ++k) ce[k] = cij[k];
  } else if (aj < 0) {
    for (int k = 0; k < 3; ++k) ce[k] = ci2[k];
  } else { /* circumcentre: two bisector planes and the triangle's plane */
    double g[12] = {0};
    for (int k = 0; k < 3; ++k) {
      g[k] = rij[k];
      g[4 + k] = ri2[k];
      g[3] += cij[k] * rij[k];
      g[7] += ci2[k] * ri2[k];
    }
    g[8] = (xj[1] - xi[1]) * (x2[2] - xi[2]) - (xj[2] - xi[2]) * (x2[1] - xi[1]);
    g[9] = (xj[2] - xi[2]) * (x2[0] - xi[0]) - (xj[0] - xi[0]) * (x2[2] - xi[2]);
    g[10] = (xj[0] - xi[0]) * (x2[1] - xi[1]) - (xj[1] - xi[1]) * (x2[0] - xi[0]);
    g[11] = g[8] * xi[0] + g[9] * xi[1] + g[10] * xi[2];
    gauss3x4(g);
    ce[2] = g[11];
    ce[1] = g[7] - ce[2] * g[6];
    ce[0] = g[3] - ce[2] * g[2] - ce[1] * g[1];
  }
  double L1 = dist_d(xi, cij), L2 = dist_d(ce, cij);
  const double L3 = dist_d(xi, ce);
  double P = 0.5 * (L1 + L2 + L3);
  double A = sqrt(P * (P - L1) * (P - L2) * (P - L3));
  L1 = dist_d(xi, ci2);
  L2 = dist_d(ce, ci2);
  P = 0.5 * (L1 + L2 + L3);
  A += sqrt(P * (P - L1) * (P - L2) * (P - L3));
  const double alpha = 1.0 / tan(acos(a1 / (dist_d(xi, x1) * dist_d(x1, xj))));
  const double beta = 1.0 / tan(acos(a2 / (dist_d(xi, x2) * dist_d(x2, xj))));
  *sum = 0.5 * (alpha + beta) * (a.m.km[j] - a.m.km[i]);
  *area = A;
}

template <int G>
__device__ inline double group_sum_d(double v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}

/* roles: A = scatter markers into bins (blocks0), B = triangle terms (blocks1),
 * C = bending node normals, vertices then edge midpoints (blocks2) */
__global__ __launch_bounds__(256) void place_kernel(MArgs a) {
  int blk = blockIdx.x;
  if (a.roles & ROLE_A) {
    if (blk < a.blocks0) {
      const int i = blk * 256 + threadIdx.x;
      if (i < a.m.n) {
        const int b = a.m.marker_bin[i];
        if (b >= 0) a.m.bin_scatter[a.m.bin_start[b] + a.m.marker_slot[i]] = i;
      }
      /* the scan has read the span totals: clear them for the next count */
      for (int64_t k = i; k < a.m.scan_blocks; k += (int64_t)a.blocks0 * 256)
        a.m.scan_partials[k] = 0;
      return;
    }
    blk -= a.blocks0;
  }
  if (a.roles & ROLE_B) {
    const int t = blk * 256 + threadIdx.x;
    if (t < a.m.n_tri) triangle_geometry(a, t);
  }
}

/* Membrane bending, launched only when it is on (separate kernels keep the
 * FP64 solves out of the register budget of the force pass):
 * bend_normal_kernel after the triangle pass, bend_km_kernel after it,
 * bend_force_kernel after the force pass (it adds to the marker forces). */
__global__ __launch_bounds__(256) void bend_normal_kernel(MArgs a) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < a.m.n + a.m.n_tedges) bend_normal(a, k);
}

__global__ __launch_bounds__(256) void bend_km_kernel(MArgs a) {
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v < a.m.n) bend_km(a, v);
}

/* Four lanes per vertex over its ring (sem_Laplace_Kapa_kernel,
 * sem_kernel.cu:1499-1638); lane 0 adds kbend (sum / A) n_i. */
__global__ __launch_bounds__(256) void bend_force_kernel(MArgs a) {
  constexpr int G = 4;
  const int i = (blockIdx.x * 256 + threadIdx.x) / G, lane = threadIdx.x & (G - 1);
  if (i >= a.m.n) return;
  const int r0 = a.m.ring_start[i], c = a.m.ring_start[i + 1] - r0;
  if (c == 0) return; /* the whole group leaves together */
  double sum = 0.0, area = 0.0;
  for (int q = lane; q < c; q += G) {
    double sq, aq;
    ring_term(a, i, r0, c, q, &sq, &aq);
    sum += sq;
    area += aq;
  }
  sum = group_sum_d<G>(sum);
  area = group_sum_d<G>(area);
  if (lane == 0) {
    const double coef = (double)a.kbend / area * sum;
    a.m.fx[i] += (float)(coef * a.m.vn[3 * (int64_t)i]);
    a.m.fy[i] += (float)(coef * a.m.vn[3 * (int64_t)i + 1]);
    a.m.fz[i] += (float)(coef * a.m.vn[3 * (int64_t)i + 2]);
  }
}

/* ---------------- K6: sort bins + platelet sums + bond decisions ---------------- */

/* Platelet p's area and volume (sums of the triangle terms) and its lowest
 * element (ties to the lower index); one block, the loads of four strides
 * issued together. */
__device__ void platelet_sums(const MArgs &a, int p) {
  const int t0 = a.m.plt_tri_start[p], t1 = a.m.plt_tri_start[p + 1];
  const int e0 = a.m.plt_elem_start[p], e1 = a.m.plt_elem_start[p + 1];
  double S[4] = {0.0, 0.0, 0.0, 0.0}, V[4] = {0.0, 0.0, 0.0, 0.0};
  int t = t0 + threadIdx.x;
  for (; t + 768 < t1; t += 1024) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      S[k] += a.m.tri_S[t + 256 * k];
      V[k] += a.m.tri_V[t + 256 * k];
    }
  }
  for (; t < t1; t += 256) {
    S[0] += a.m.tri_S[t];
    V[0] += a.m.tri_V[t];
  }
  double Ssum = (S[0] + S[1]) + (S[2] + S[3]), Vsum = (V[0] + V[1]) + (V[2] + V[3]);
  float zmin = INFINITY;
  int imin = INT_MAX;
  for (int e = e0 + threadIdx.x; e < e1; e += 256) {
    const float z = a.m.z[e];
    if (z < zmin || (z == zmin && e < imin)) {
      zmin = z;
      imin = e;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    Ssum += __shfl_xor(Ssum, o, 64);
    Vsum += __shfl_xor(Vsum, o, 64);
    const float z2 = __shfl_xor(zmin, o, 64);
    const int i2 = __shfl_xor(imin, o, 64);
    if (z2 < zmin || (z2 == zmin && i2 < imin)) {
      zmin = z2;
      imin = i2;
    }
  }
  __shared__ double shS[4], shV[4];
  __shared__ float shz[4];
  __shared__ int shi[4];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    shS[w] = Ssum;
    shV[w] = Vsum;
    shz[w] = zmin;
    shi[w] = imin;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k) {
      if (shz[k] < shz[0] || (shz[k] == shz[0] && shi[k] < shi[0])) {
        shz[0] = shz[k];
        shi[0] = shi[k];
      }
    }
    a.m.plt_S[p] = (shS[0] + shS[1]) + (shS[2] + shS[3]);
    a.m.plt_V[p] = (shV[0] + shV[1]) + (shV[2] + shV[3]);
    a.m.plt_lowest[p] = shi[0] == INT_MAX ? -1 : shi[0];
  }
}

/* The neighbouring bins of a point as ranges of a bin-ordered array with
 * per-bin starts: bins are numbered x fastest, so the (up to 3) x-neighbours
 * of one (y, z) pair are one range unless a periodic wrap splits them in
 * two; at most 2 x 3 x 3 ranges. */
constexpr int NBR_RANGES = 18;

/* Neighbour bins of bin b along an axis of nb bins: the j-th of *count
 * (the same set as nbr_bins, without an indexed array). */
__device__ inline int nbr_bin_j(int b, int nb, int periodic, int j, int *count) {
  if (periodic && nb <= 3) {
    *count = nb;
    return j;
  }
  if (periodic) {
    *count = 3;
    const int v = b - 1 + j;
    return v < 0 ? v + nb : (v >= nb ? v - nb : v);
  }
  const int lo = max(b - 1, 0), hi = min(b + 1, nb - 1);
  *count = hi - lo + 1;
  return lo + j;
}

/* Lane r < the returned count gets range r as [*s, *e) of the array whose
 * bins start at starts[]. */
__device__ inline int nbr_range(const MLat &L, const int32_t *starts, const int b3[3], int r,
                                int *s, int *e) {
  /* x-runs of consecutive bins: two when a periodic wrap splits the three */
  const int bx = b3[0], nbx = L.nb[0];
  int runs = 1, f0, l0, f1 = 0, l1 = 0;
  if (L.periodic[0] && nbx <= 3) {
    f0 = 0;
    l0 = nbx - 1;
  } else if (L.periodic[0] && (bx == 0 || bx == nbx - 1)) {
    runs = 2;
    f0 = bx == 0 ? nbx - 1 : nbx - 2;
    l0 = nbx - 1;
    l1 = bx == 0 ? 1 : 0;
  } else {
    f0 = max(bx - 1, 0);
    l0 = min(bx + 1, nbx - 1);
  }
  int cy, cz;
  nbr_bin_j(b3[1], L.nb[1], L.periodic[1], 0, &cy);
  nbr_bin_j(b3[2], L.nb[2], L.periodic[2], 0, &cz);
  const int n = runs * cy * cz;
  if (r < n) {
    const int run = r % runs, jy = (r / runs) % cy, jz = r / (runs * cy);
    const int y = nbr_bin_j(b3[1], L.nb[1], L.periodic[1], jy, &cy);
    const int z = nbr_bin_j(b3[2], L.nb[2], L.periodic[2], jz, &cz);
    const int row = L.bin0 + L.nb[0] * (y + L.nb[1] * z);
    *s = starts[row + (run == 0 ? f0 : f1)];
    *e = starts[row + (run == 0 ? l0 : l1) + 1];
  }
  return n;
}

/* Lanes r < nr hold range r as [s, e); the ranges become one list of
 * candidates c = 0..total-1 over the wavefront, at array index
 * c + off[r] for the r with ends[r-1] <= c < ends[r] (tables in this
 * wavefront's LDS).  All 64 lanes call it. */
__device__ inline int flatten_ranges(int nr, int s, int e, int *off, int *ends) {
  const int lane = threadIdx.x & 63;
  const int len = lane < nr ? e - s : 0;
  const int end = wave_incl_scan(len);
  const int total = __shfl(end, 63, 64);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (lane < nr) {
    off[lane] = s - (end - len);
    ends[lane] = end;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return total;
}

__device__ inline void ligand_vec(const MArgs &a, const MLat &L, int l, const float rp[3],
                                  float d[3]) {
  if (l < a.m.n_wall_lig) {
    d[0] = min_image(a.m.lig_x[l] - rp[0], L.nx, L.periodic[0]);
    d[1] = min_image(a.m.lig_y[l] - rp[1], L.ny, L.periodic[1]);
    d[2] = min_image(a.m.lig_z[l] - rp[2], L.nz, L.periodic[2]);
  } else {
    const int mk = l - a.m.n_wall_lig;
    d[0] = min_image(a.m.x[mk] - rp[0], L.nx, L.periodic[0]);
    d[1] = min_image(a.m.y[mk] - rp[1], L.ny, L.periodic[1]);
    d[2] = min_image(a.m.z[mk] - rp[2], L.nz, L.periodic[2]);
  }
}

/* Candidate ligand l at offset (dx, dy, dz) from the receptor: the nearest
 * free one within reach wins, ties to the lower id. */
__device__ inline void ligand_candidate(const MArgs &a, int l, float dx, float dy, float dz,
                                        float reach2, float &best_d2, int &best,
                                        bool &near_reach) {
  const float d2 = dx * dx + dy * dy + dz * dz;
  if (fabsf(sqrtf(d2) - a.bond_reach) < BOND_EPS_R) near_reach = true;
  if (d2 > reach2 || a.m.lig_owner[l] >= 0) return;
  if (d2 < best_d2 || (d2 == best_d2 && l < best)) {
    best_d2 = d2;
    best = l;
  }
}

/* v of lane (i XOR s), read with a DPP lane move instead of a trip through
 * the LDS crossbar.  Groups of 4 lanes: the quad permutation is the XOR
 * itself.  Groups of 16 (one DPP row), with s taken in the order 8, 4, 2, 1
 * of a butterfly sum: rotating the row by s gives lane (i - s) mod 16, which
 * is lane i XOR s for s = 8 and, for the smaller s, a lane holding the same
 * partial sum as lane i XOR s (after the steps above s the partial sums
 * repeat with period 2s, each computed as a commutative pair), so the sums
 * are bit for bit the XOR butterfly's. */
template <int G, int S>
__device__ inline float dpp_partner(float v) {
  static_assert((G == 4 && (S == 2 || S == 1)) || (G == 16 && S >= 1 && S <= 8), "DPP group");
  constexpr int ctrl = G == 4 ? (S == 2 ? 0x4E /* quad_perm 2,3,0,1 */ : 0xB1 /* 1,0,3,2 */)
                              : 0x120 + S /* row_ror:S */;
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctrl, 0xF, 0xF, false));
}

/* Lanes per receptor in the bond decisions, receptors per workgroup of 256. */
constexpr int REC_LANES = 16, REC_PER_BLOCK = 256 / REC_LANES;

/* flatten_ranges for a group of REC_LANES lanes: lane l holds ranges l and
 * l + 16 of nbr_range (at most NBR_RANGES = 18); the group's offsets and
 * running ends go to its off/ends; returns the group's candidate count. */
__device__ inline int group_flatten_ranges(int nr, int s0, int e0, int s1, int e1, int *off,
                                           int *ends) {
  static_assert(REC_LANES == 16 && NBR_RANGES <= 2 * REC_LANES, "two ranges per lane");
  const int gl = threadIdx.x & (REC_LANES - 1);
  const int len0 = gl < nr ? e0 - s0 : 0;
  int end0 = len0;
#pragma unroll
  for (int d = 1; d < REC_LANES; d <<= 1) {
    const int v = __shfl_up(end0, d, REC_LANES);
    if (gl >= d) end0 += v;
  }
  const int tot0 = __shfl(end0, REC_LANES - 1, REC_LANES);
  const int len1 = gl + REC_LANES < nr ? e1 - s1 : 0;
  const int l16 = __shfl(len1, 0, REC_LANES);
  const int end1 = tot0 + (gl == 0 ? len1 : l16 + len1); /* ranges 16 and 17 */
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (gl < nr) {
    off[gl] = s0 - (end0 - len0);
    ends[gl] = end0;
  }
  if (gl + REC_LANES < nr) {
    off[gl + REC_LANES] = s1 - (end1 - len1);
    ends[gl + REC_LANES] = end1;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return nr > REC_LANES ? __shfl(end1, nr - REC_LANES - 1, REC_LANES) : tot0;
}

/* The neighbour ranges of bin b3 for a group of REC_LANES lanes (lane l
 * takes ranges l and l + 16), flattened into off/ends. */
__device__ inline int group_ranges(const MLat &L, const int32_t *starts, const int b3[3],
                                   int *off, int *ends) {
  const int gl = threadIdx.x & (REC_LANES - 1);
  int s0 = 0, e0 = 0, s1 = 0, e1 = 0;
  const int nr = nbr_range(L, starts, b3, gl, &s0, &e0);
  if (nr > REC_LANES) nbr_range(L, starts, b3, gl + REC_LANES, &s1, &e1);
  return group_flatten_ranges(nr, s0, e0, s1, e1, off, ends);
}

/* (d2, id) of lane (i XOR S) within a row of 16 lanes, by row rotation (the
 * arg-min butterfly's pairs repeat with period 2S after the steps above S,
 * as in group_sum). */
template <int S>
__device__ inline void argmin_partner16(float &d2, int &id) {
  const float od = dpp_partner<16, S>(d2);
  const int ol = __builtin_amdgcn_mov_dpp(id, 0x120 + S, 0xF, 0xF, false);
  if (od < d2 || (od == d2 && ol < id)) {
    d2 = od;
    id = ol;
  }
}

/* Ruptures of held bonds and formation candidates of free receptors, from the
 * state at the start of the step; a group of REC_LANES lanes per receptor,
 * the lanes over the ligands of the neighbouring bins (wall ligands from
 * their bin-ordered packed positions, fibrin ligands from the scatter order
 * of the markers).  The decision depends on no lane order: the arg-min is
 * over (distance^2, ligand id). */
__device__ void bond_decide(const MArgs &a, int r, int *off, int *ends) {
  const int lane = threadIdx.x & (REC_LANES - 1);
  const int64_t step = *a.m.d_step;
  const float4 rq = a.m.rec_pos[r];
  const float rp[3] = {rq.x, rq.y, rq.z};
  const MLat L = a.lats ? a.lats[__float_as_int(rq.w)] : a.lat;
  const int b = a.m.rec_bond[r];
  if (b >= 0) {
    if (lane == 0) {
      a.m.rec_cand[r] = -1;
      float d[3];
      ligand_vec(a, L, b, rp, d);
      const float dist = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      const float tension = a.bond_k * (dist - a.bond_l0);
      /* Bell: k_off = k_off0 exp(x_b f / kBT) for a pulling bond; the
       * reference's law takes the force magnitude k |l0 - r|
       * (sem_kernel.cu:1852-1863) */
      const float f = a.bond_law == PFG_BOND_REFERENCE ? fabsf(tension) : fmaxf(tension, 0.0f);
      const float koff = a.k_off0 * expf(a.bell_xb_kT * f);
      const float p_off = 1.0f - expf(-koff);
      const float u = pfg_uniform(L.seed, step, (uint32_t)(r - L.rec0), PFG_RNG_BOND_OFF);
      if (fabsf(u - p_off) < BOND_EPS_P) bump(a.m.counters, 2);
      a.m.rec_release[r] = u < p_off ? 1 : 0;
    }
    return;
  }
  const float reach2 = a.bond_reach * a.bond_reach;
  int best = INT_MAX;
  float best_d2 = INFINITY;
  bool near_reach = false;
  int b3[3];
  bin3_of(L, rp[0], rp[1], rp[2], b3);
  if (a.m.n_wall_lig > 0) {
    const int total = group_ranges(L, a.m.lg_start, b3, off, ends);
    int q = 0;
    for (int c = lane; c < total; c += REC_LANES) {
      while (c >= ends[q]) ++q;
      const float4 g = a.m.lg_pos[c + off[q]];
      ligand_candidate(a, __float_as_int(g.w), min_image(g.x - rp[0], L.nx, L.periodic[0]),
                       min_image(g.y - rp[1], L.ny, L.periodic[1]),
                       min_image(g.z - rp[2], L.nz, L.periodic[2]), reach2, best_d2, best,
                       near_reach);
    }
  }
  if (a.m.fibrin_lig) {
    /* markers other than elements are ligands (index n_wall_lig + marker);
     * the scatter order is final before this launch and order-free here;
     * the lists' grid is the coarse one when they leave the elements out */
    const MLat G = a.bin_elements ? L : coarse_grid(L);
    int g3[3];
    bin3_of(G, rp[0], rp[1], rp[2], g3);
    const int total = group_ranges(G, a.m.bin_start, g3, off, ends);
    int q = 0;
    for (int c = lane; c < total; c += REC_LANES) {
      while (c >= ends[q]) ++q;
      const int mk = a.m.bin_scatter[c + off[q]];
      const float4 g = a.m.pos4[mk];
      if (__float_as_int(g.w) != -1) continue; /* an element */
      ligand_candidate(a, a.m.n_wall_lig + mk, min_image(g.x - rp[0], L.nx, L.periodic[0]),
                       min_image(g.y - rp[1], L.ny, L.periodic[1]),
                       min_image(g.z - rp[2], L.nz, L.periodic[2]), reach2, best_d2, best,
                       near_reach);
    }
  }
  /* the group's arg-min on (distance^2, ligand id) */
  argmin_partner16<8>(best_d2, best);
  argmin_partner16<4>(best_d2, best);
  argmin_partner16<2>(best_d2, best);
  argmin_partner16<1>(best_d2, best);
  const int sh = threadIdx.x & 63 & ~(REC_LANES - 1);
  const bool any_near = ((__ballot(near_reach) >> sh) & ((1ull << REC_LANES) - 1)) != 0;
  if (lane != 0) return;
  a.m.rec_release[r] = 0;
  a.m.rec_cand[r] = -1;
  if (any_near) bump(a.m.counters, 2);
  if (best == INT_MAX) return;
  float kon = a.k_on;
  if (a.bond_law == PFG_BOND_REFERENCE) {
    /* k_on e^{f (x_b - w/2) / kT}, w = |l0 - r|, f = k w (sem_kernel.cu:1852-1858) */
    const float w = fabsf(a.bond_l0 - sqrtf(best_d2)), f = a.bond_k * w;
    kon = a.k_on * expf(a.bell_xb_kT * f - 0.5f * f * w / a.bell_kT);
  }
  const float p_on = 1.0f - expf(-kon);
  const float u = pfg_uniform(L.seed, step, (uint32_t)(r - L.rec0), PFG_RNG_BOND_ON);
  if (fabsf(u - p_on) < BOND_EPS_P) bump(a.m.counters, 2);
  if (u < p_on) {
    a.m.rec_cand[r] = best;
    atomicMin(&a.m.lig_claim[best], r);
  }
}

/* roles: A = sort bins (blocks0, one lane per marker: each marker takes its
 * rank by index among the scatter order of its bin, and writes its index and
 * packed position there), B = platelet sums (blocks1 = one block per platelet),
 * C = bond decisions (blocks2, four receptors per block; they read the
 * scatter order, which no role of this launch writes) */
__device__ void gather_body(const MArgs &a) {
  int blk = blockIdx.x;
  if (a.roles & ROLE_A) {
    if (blk < a.blocks0) {
      const int i = blk * 256 + threadIdx.x;
      if (i < a.m.n) {
        /* rank of i among its bin's markers by index (an element left out
         * of the lists has none) */
        const int b = a.m.marker_bin[i];
        if (b < 0) return;
        const int s = a.m.bin_start[b], e = a.m.bin_start[b + 1];
        int k = s;
        for (int q = s; q < e; ++q) k += a.m.bin_scatter[q] < i;
        a.m.bin_items[k] = i;
        a.m.sorted_pos[k] = a.m.pos4[i];
      }
      return;
    }
    blk -= a.blocks0;
  }
  if (a.roles & ROLE_B) {
    if (blk < a.blocks1) {
      platelet_sums(a, blk);
      return;
    }
    blk -= a.blocks1;
  }
  if (a.roles & ROLE_C) {
    __shared__ int sh_off[REC_PER_BLOCK][NBR_RANGES], sh_end[REC_PER_BLOCK][NBR_RANGES];
    const int w = threadIdx.x / REC_LANES;
    const int r = blk * REC_PER_BLOCK + w;
    if (r < a.m.n_rec && !halted(a)) bond_decide(a, r, sh_off[w], sh_end[w]);
  }
}

/* ---------------- K7: bond resolution ---------------- */

/* Apply ruptures and the lowest-index claim per ligand, then the bond force
 * k (|d| - l0) d/|d| on the receptor and its inverse-distance split over the
 * three vertices (sem_kernel.cu:2039-2065). */
__device__ void resolve_one(const MArgs &a, int r) {
  if (halted(a)) return;
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
    ligand_vec(a, lat_of(a, nodes[0]), b, rp, d);
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

__global__ void bond_resolve_kernel(MArgs a) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < a.m.n_rec) resolve_one(a, r);
}

/* Cell-list sort, platelet sums and bond decisions; with resolve_last the
 * workgroup that finishes last also resolves the bonds (few receptors: one
 * launch fewer).  Each workgroup publishes its writes with a device-scope
 * fence before it takes its ticket; the last one fences again before it
 * reads them, and clears the ticket for the next launch. */
__global__ __launch_bounds__(256) void gather_kernel(MArgs a) {
  gather_body(a);
  if (!a.resolve_last) return;
  __shared__ int last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    last = atomicAdd(a.m.ticket, 1) == (int)gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  for (int r = threadIdx.x; r < a.m.n_rec; r += 256) resolve_one(a, r);
  if (threadIdx.x == 0) *a.m.ticket = 0;
}

/* ---------------- K8: forces ---------------- */

/* Morse force on element (xi, yi, zi) from candidate p of the bin-sorted
 * packed positions: only elements of the same body (tag) within the cutoff;
 * the pair with itself is the r = 0 one and is skipped. */
__device__ inline void morse_pair(const MArgs &a, float4 p, float tag, float xi, float yi,
                                  float zi, float &fx, float &fy, float &fz) {
  if (__float_as_int(p.w) != __float_as_int(tag)) return;
  const float dx = p.x - xi, dy = p.y - yi, dz = p.z - zi;
  /* fixed evaluation order, so that the cutoff decision is reproducible
   * from the same FP32 positions elsewhere */
  const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
  if (fabsf(r2 - a.morse_cut2) < a.morse_eps2) bump(a.m.counters, 8);
  if (r2 >= a.morse_cut2 || r2 == 0.f) return;
  const float r = sqrtf(r2);
  const float e = expf(-a.morse_rho * (r - a.morse_req));
  /* -d(U S)/dr along (x_j - x_i)/r, U = U0 (e^2 - 2e) */
  float c = 2.0f * a.morse_rho * a.morse_u0 * (e - e * e);
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

/* Area and volume constraint force on the vertex at slot e % 3 of triangle
 * e / 3 (sem_ForceAreaVol_kernel, sem_kernel.cu:1640-1749), from the
 * per-triangle terms of this step. */
__device__ inline void area_volume_term(const MArgs &a, int e, float g_area, float cvol,
                                        float &fx, float &fy, float &fz) {
  const float4 g = a.m.tri_g[e];
  const float4 v = a.m.tri_a2[e / 3];
  const float ca = -a.ks * (g.w + g_area);
  fx += ca * g.x + cvol * v.x;
  fy += ca * g.y + cvol * v.y;
  fz += ca * g.z + cvol * v.z;
}

__device__ inline float link_tension(const MArgs &a, float l, float l0) {
  if (a.fiber_law == PFG_FIBER_LINEAR) {
    /* Y A (l - l0)/l0 (network.cpp:364-373) */
    return a.fiber_k * (l - l0) / l0;
  }
  /* WLC (fiber.cpp:721-730, network.cpp:386-405): CC = kB T / Lp with Lp = l0,
   * tension CC (1/(4(1-x)^2) - 1/4 + x), x = (l - l0)/l0 */
  const float x = fminf((l - l0) / l0, a.fiber_xmax);
  const float tl = 1.0f - x;
  return (a.fiber_k / l0) * (0.25f / (tl * tl) - 0.25f + x);
}

/* Calls f(s, e) for each neighbour range of bin b3 in the array whose bins
 * start at starts[] (the ranges of nbr_range, without an indexed array). */
template <typename F>
__device__ inline void for_nbr_ranges(const MLat &L, const int32_t *starts, const int b3[3],
                                      F f) {
  const int bx = b3[0], nbx = L.nb[0];
  int runs = 1, f0, l0, f1 = 0, l1 = 0;
  if (L.periodic[0] && nbx <= 3) {
    f0 = 0;
    l0 = nbx - 1;
  } else if (L.periodic[0] && (bx == 0 || bx == nbx - 1)) {
    runs = 2;
    f0 = bx == 0 ? nbx - 1 : nbx - 2;
    l0 = nbx - 1;
    l1 = bx == 0 ? 1 : 0;
  } else {
    f0 = max(bx - 1, 0);
    l0 = min(bx + 1, nbx - 1);
  }
  int cy, cz;
  nbr_bin_j(b3[1], L.nb[1], L.periodic[1], 0, &cy);
  nbr_bin_j(b3[2], L.nb[2], L.periodic[2], 0, &cz);
  for (int jz = 0; jz < cz; ++jz) {
    const int z = nbr_bin_j(b3[2], L.nb[2], L.periodic[2], jz, &cz);
    for (int jy = 0; jy < cy; ++jy) {
      const int y = nbr_bin_j(b3[1], L.nb[1], L.periodic[1], jy, &cy);
      const int row = L.bin0 + L.nb[0] * (y + L.nb[1] * z);
      f(starts[row + f0], starts[row + l0 + 1]);
      if (runs == 2) f(starts[row + f1], starts[row + l1 + 1]);
    }
  }
}

/* Sum over an aligned group of G lanes (XOR butterfly, largest stride
 * first); every lane of the group gets the sum. */
template <int G>
__device__ inline float group_sum(float v) {
  if constexpr (G == 4) {
    v += dpp_partner<4, 2>(v);
    v += dpp_partner<4, 1>(v);
  } else if constexpr (G == 16) {
    v += dpp_partner<16, 8>(v);
    v += dpp_partner<16, 4>(v);
    v += dpp_partner<16, 2>(v);
    v += dpp_partner<16, 1>(v);
  } else {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  }
  return v;
}

/* The element's force terms other than Morse, lane `lane` of G taking every
 * G-th entry of its CSR lists: area and volume constraints, mesh-edge
 * springs, receptor bond shares, and (lane 0) the wall repulsion on the
 * platelet's lowest element. */
template <int G>
__device__ inline void element_terms(const MArgs &a, int i, int lane, float xi, float yi,
                                     float zi, int body, float &fx, float &fy, float &fz) {
  /* the lists of terms whose coefficients are zero are not read */
  const bool area_vol = a.ks != 0.f || a.kv != 0.f, springs = a.k_spring != 0.f;
  const int t0 = area_vol ? a.m.m2t_start[i] : 0, nt = area_vol ? a.m.m2t_start[i + 1] - t0 : 0;
  const int e0 = springs ? a.m.m2e_start[i] : 0, ne = springs ? a.m.m2e_start[i + 1] - e0 : 0;
  const int r0 = a.m.n_rec > 0 ? a.m.m2r_start[i] : 0;
  const int nr = a.m.n_rec > 0 ? a.m.m2r_start[i + 1] - r0 : 0;
  float g_area = 0.f, cvol = 0.f;
  if (nt > 0) {
    const double Sall = a.m.plt_S[body], S0all = a.m.plt_S0[body];
    const double V = a.m.plt_V[body], V0 = a.m.plt_V0[body];
    g_area = (float)((Sall - S0all) / S0all);
    cvol = -a.kv * (float)((V - V0) / V0) / 6.0f;
  }
  for (int k = lane; k < nt + ne + nr; k += G) {
    if (k < nt) {
      area_volume_term(a, a.m.m2t[t0 + k], g_area, cvol, fx, fy, fz);
    } else if (k < nt + ne) {
      /* mesh-edge Hookean spring k (r - L)/r (x_j - x_i) (sem_kernel.cu:374-384) */
      const int q = e0 + k - nt;
      const float4 pj = a.m.pos4[a.m.m2e_nbr[q]];
      const float dx = pj.x - xi, dy = pj.y - yi, dz = pj.z - zi;
      const float r = sqrtf(dx * dx + dy * dy + dz * dz);
      if (r > 0.f) {
        const float c = a.k_spring * (r - a.m.m2e_l0[q]) / r;
        fx += c * dx;
        fy += c * dy;
        fz += c * dz;
      }
    } else {
      /* receptor bond force, inverse-distance share of this vertex */
      const int e = a.m.m2r[r0 + k - nt - ne];
      const int r = e / 3, slot = e % 3;
      const float w = a.m.rec_w[3 * r + slot];
      fx += w * a.m.rec_force[3 * r];
      fy += w * a.m.rec_force[3 * r + 1];
      fz += w * a.m.rec_force[3 * r + 2];
    }
  }
  /* wall repulsion on the lowest element (sem_kernel.cu:409-430) */
  if (lane == 0 && a.wall_a != 0.f && a.m.plt_lowest[body] == i) {
    float h = zi - a.wall_z;
    if (h <= a.wall_hc) {
      h = fmaxf(h, 0.01f * a.wall_hc);
      const float e = expf(-a.wall_kappa * h);
      fz += a.wall_a * a.wall_kappa * e / (1.0f - e);
    }
  }
}

/* A group of G lanes per marker (G = 64: one wavefront, markers in index
 * order, the Morse candidates of all neighbour ranges flattened over the
 * lanes; G < 64: markers in bin-sorted order, so that the groups of a
 * wavefront mostly share their neighbour bins and read the same packed
 * positions, each group striding through the ranges).  The lanes of a group
 * then stride over the marker's CSR lists (triangles, edges, receptor
 * shares, fibrin links) and a butterfly reduction sums them. */
template <int G, bool SPREAD = false>
__global__ __launch_bounds__(256) void forces_kernel(MArgs a) {
  static_assert(!SPREAD || G == 64, "spreading takes a lane per stencil node");
  if ((int)blockIdx.x >= a.blocks0 && a.blocks1 + a.blocks2 > 0) { /* whole workgroups */
    claim_role(a, (int)blockIdx.x - a.blocks0);
    return;
  }
  const int lane = threadIdx.x & (G - 1);
  if (blockIdx.x == 0 && threadIdx.x == 0 && a.m.n_rec > 0 && a.detach_rate > 0.f) {
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
  __shared__ int sh_off[4][NBR_RANGES], sh_end[4][NBR_RANGES];
  const int k = (blockIdx.x * 256 + threadIdx.x) / G;
  /* every marker (in index or bin order), or those of a list */
  if (k >= (a.force_count < 0 ? a.m.n : a.force_count)) return;
  const int i = a.force_count >= 0 ? a.force_list[k] : (G == 64 ? k : a.m.bin_items[k]);
  const int kind = a.m.kind[i];
  const float4 pi = a.m.pos4[i];
  const float xi = pi.x, yi = pi.y, zi = pi.z;
  float fx = 0.f, fy = 0.f, fz = 0.f;
  if (kind == PFG_MARKER_ELEMENT) {
    const int body = a.m.body[i];
    if (a.morse_u0 != 0.f) {
      const MLat L = lat_of(a, i);
      const int b = a.m.marker_bin[i] - L.bin0;
      const int b3[3] = {b % L.nb[0], (b / L.nb[0]) % L.nb[1], b / (L.nb[0] * L.nb[1])};
      const float tag = pi.w;
      if constexpr (G == 64) {
        /* lane r writes range r's offset and end into this wave's LDS table */
        int s = 0, e = 0;
        const int nr = nbr_range(L, a.m.bin_start, b3, lane, &s, &e);
        int *off = sh_off[threadIdx.x >> 6], *ends = sh_end[threadIdx.x >> 6];
        const int total = flatten_ranges(nr, s, e, off, ends);
        int r = 0;
        for (int c = lane; c < total; c += 64) {
          while (c >= ends[r]) ++r;
          morse_pair(a, a.m.sorted_pos[c + off[r]], tag, xi, yi, zi, fx, fy, fz);
        }
      } else {
        for_nbr_ranges(L, a.m.bin_start, b3, [&](int s, int e) {
          for (int c = s + lane; c < e; c += G)
            morse_pair(a, a.m.sorted_pos[c], tag, xi, yi, zi, fx, fy, fz);
        });
      }
    }
    element_terms<G>(a, i, lane, xi, yi, zi, body, fx, fy, fz);
  } else {
    /* fibrin links in node-gather form (fiber.cpp:647-706) */
    const float sgn = a.fiber_ref_sign ? -1.0f : 1.0f;
    const int l0i = a.m.m2l_start[i], nl = a.m.m2l_start[i + 1] - l0i;
    for (int k = lane; k < nl; k += G) {
      const float4 pj = a.m.pos4[a.m.m2l_nbr[l0i + k]];
      const float l0 = a.m.m2l_l0[l0i + k];
      const float dx = pj.x - xi, dy = pj.y - yi, dz = pj.z - zi;
      const float l = sqrtf(dx * dx + dy * dy + dz * dz);
      if (l == 0.f || l0 == 0.f) continue;
      const float c = sgn * link_tension(a, l, l0) / l; /* restoring when sgn = 1 */
      fx += c * dx;
      fy += c * dy;
      fz += c * dz;
    }
    /* reaction of a receptor bond held on this node */
    if (lane == 0 && a.m.fibrin_lig) {
      const int owner = a.m.lig_owner[a.m.n_wall_lig + i];
      if (owner >= 0) {
        fx -= a.m.rec_force[3 * owner];
        fy -= a.m.rec_force[3 * owner + 1];
        fz -= a.m.rec_force[3 * owner + 2];
      }
    }
  }
  fx = group_sum<G>(fx);
  fy = group_sum<G>(fy);
  fz = group_sum<G>(fz);
  if (lane == 0) {
    a.m.fx[i] = fx;
    a.m.fy[i] = fy;
    a.m.fz[i] = fz;
  }
  /* the wavefront holds the marker's whole force: spread it at once, a
   * lane per stencil node (the direct form, without another launch) */
  if constexpr (SPREAD) spread_direct(a, i, lane, fx, fy, fz);
}

/* Force pass by body: a workgroup of BODY_THREADS lanes per body (a
 * platelet of at most FB_MAX elements, whose Morse partners are its own
 * elements), a lane per element, with a neighbour list per element kept in
 * device memory between steps.
 * - The list of element i holds every element j of the body with
 *   |x_j - x_i| < R = cutoff + skin at the positions of its last build
 *   (fb_ref).  It still holds every pair now within the cutoff while no
 *   element has moved more than skin / 2 away from the body's mean
 *   displacement since (the distance of a pair changes by at most the
 *   difference of its two displacements); otherwise the body rebuilds it
 *   (so does a body whose positions were uploaded or whose list is new:
 *   fb_ref is NaN then).
 * - A build sorts the positions into a cell list of the body's own box in
 *   LDS, cells no smaller than R (the cell size doubles until the cells fit
 *   FB_BINS), each cell's elements in index order, and lists each element's
 *   partners in cell order (deterministic).  A list longer than FB_NBR_MAX
 *   leaves the body without lists: it sums from the cell list this step and
 *   builds again the next.
 * - Each element then sums the Morse pairs of its list from the positions
 *   in LDS (the pair test and FP32 distance of the per-marker form, so the
 *   same cutoff decisions), adds its other terms and writes its force. */
__global__ __launch_bounds__(BODY_THREADS) void forces_body_build_kernel(MArgs a) {
  constexpr int NT = BODY_THREADS, NW = NT / 64, PER = (FB_MAX + NT - 1) / NT;
  __shared__ float4 sp[FB_MAX];          /* positions, index order           */
  __shared__ int16_t sidx[FB_MAX];       /* elements in cell order           */
  __shared__ int cend[FB_BINS + 1];      /* counts, then ends of the cells   */
  __shared__ float shf[NW][6];
  __shared__ int sflag[2];               /* [0] rebuild, [1] a list overflowed */
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int g = blockIdx.x;
  const int i0 = a.m.fbody_range[2 * g], cnt = a.m.fbody_range[2 * g + 1] - i0;
  const int64_t n = a.m.n;
  const float cut = a.morse_cut, skin = FB_SKIN * a.morse_cut, R = cut + skin;
  const bool morse = a.morse_u0 != 0.f && cut > 0.f;
  float4 pr[PER], rr[PER];
  float dsp[3] = {0.f, 0.f, 0.f};
  bool bad = false; /* a reference position that is not a number */
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int j = t + NT * k;
    pr[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    rr[k] = pr[k];
    if (j < cnt) {
      pr[k] = a.m.pos4[i0 + j];
      sp[j] = pr[k];
      if (morse) {
        rr[k] = a.m.fb_ref[i0 + j];
        dsp[0] += pr[k].x - rr[k].x;
        dsp[1] += pr[k].y - rr[k].y;
        dsp[2] += pr[k].z - rr[k].z;
        bad = bad || !(rr[k].x == rr[k].x);
      }
    }
  }
  for (int b = t; b <= FB_BINS; b += NT) cend[b] = 0;
  if (t == 0) {
    sflag[0] = 0;
    sflag[1] = 0;
  }
  __syncthreads();
  if (cnt <= 0) return;
  /* the sum of the displacements (fixed order: lanes, then wavefronts) */
  bool rebuild = false;
  if (morse) {
#pragma unroll
    for (int k = 0; k < 3; ++k) dsp[k] = wave_sum(dsp[k]);
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < 3; ++k) shf[wave][k] = dsp[k];
    }
    if (__any(bad) && lane == 0) sflag[0] = 1;
    __syncthreads();
    float mean[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float v = 0.f;
      for (int w = 0; w < NW; ++w) v += shf[w][k];
      mean[k] = v / (float)cnt;
    }
    /* the largest displacement away from the mean, against skin / 2 */
    bool far = false;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int j = t + NT * k;
      if (j >= cnt) continue;
      const float ex = pr[k].x - rr[k].x - mean[0], ey = pr[k].y - rr[k].y - mean[1],
                  ez = pr[k].z - rr[k].z - mean[2];
      /* skin / 2 less 2 %, a margin over FP32 rounding in these distances */
      far = far || !(ex * ex + ey * ey + ez * ez <= 0.2401f * skin * skin);
    }
    if (__any(far) && lane == 0) sflag[0] = 1;
    __syncthreads();
    rebuild = sflag[0] != 0;
  }
  /* the platelet's box, for a build only (the decision is the workgroup's) */
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (rebuild) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      if (t + NT * k >= cnt) continue;
      lo[0] = fminf(lo[0], pr[k].x);
      lo[1] = fminf(lo[1], pr[k].y);
      lo[2] = fminf(lo[2], pr[k].z);
      hi[0] = fmaxf(hi[0], pr[k].x);
      hi[1] = fmaxf(hi[1], pr[k].y);
      hi[2] = fmaxf(hi[2], pr[k].z);
    }
#pragma unroll
    for (int s2 = 32; s2 > 0; s2 >>= 1) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        lo[k] = fminf(lo[k], __shfl_xor(lo[k], s2, 64));
        hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], s2, 64));
      }
    }
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        shf[wave][k] = lo[k];
        shf[wave][3 + k] = hi[k];
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[k] = shf[0][k];
      hi[k] = shf[0][3 + k];
      for (int w = 1; w < NW; ++w) {
        lo[k] = fminf(lo[k], shf[w][k]);
        hi[k] = fmaxf(hi[k], shf[w][3 + k]);
      }
    }
  }
  /* a cell list: for a build, or to sum from it when a list overflows */
  float cs = 0.f, inv = 0.f;
  int d[3] = {1, 1, 1};
  if (rebuild) {
    cs = fmaxf(R, 1e-3f);
    for (;;) {
      int64_t nb = 1;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        d[k] = (int)fminf((hi[k] - lo[k]) / cs, 1e6f) + 1;
        nb *= d[k];
      }
      if (nb <= FB_BINS) break;
      cs *= 2.0f;
    }
    inv = 1.0f / cs;
    const int nbins = d[0] * d[1] * d[2];
    int cell[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int bx = min(max((int)((pr[k].x - lo[0]) * inv), 0), d[0] - 1);
      const int by = min(max((int)((pr[k].y - lo[1]) * inv), 0), d[1] - 1);
      const int bz = min(max((int)((pr[k].z - lo[2]) * inv), 0), d[2] - 1);
      cell[k] = bx + d[0] * (by + d[1] * bz);
      if (t + NT * k < cnt) atomicAdd(&cend[cell[k] + 1], 1);
    }
    __syncthreads();
    if (t < 64) { /* inclusive scan: cend[b + 1] = end of cell b = start of b + 1 */
      int run = 0;
      for (int b0 = 1; b0 <= nbins; b0 += 64) {
        const int b = b0 + lane;
        const int v = b <= nbins ? cend[b] : 0;
        const int incl = wave_incl_scan(v);
        if (b <= nbins) cend[b] = run + incl;
        run += __shfl(incl, 63, 64);
      }
    }
    __syncthreads();
    /* place: cell b's fill pointer runs from its start to its end, after
     * which cend[b] ends cell b and cell b starts at cend[b - 1] */
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int j = t + NT * k;
      if (j < cnt) sidx[atomicAdd(&cend[cell[k]], 1)] = (int16_t)j;
    }
    __syncthreads();
    for (int b = t; b < nbins; b += NT) { /* each cell in index order */
      const int s0 = b == 0 ? 0 : cend[b - 1], s1 = cend[b];
      for (int q = s0 + 1; q < s1; ++q) {
        const int16_t key = sidx[q];
        int r = q - 1;
        while (r >= s0 && sidx[r] > key) {
          sidx[r + 1] = sidx[r];
          --r;
        }
        sidx[r + 1] = key;
      }
    }
    __syncthreads();
    /* the lists, and the new reference positions */
    const float R2 = R * R;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int j = t + NT * k;
      if (j >= cnt) continue;
      const float4 p = pr[k];
      const int bx = min(max((int)((p.x - lo[0]) * inv), 0), d[0] - 1);
      const int by = min(max((int)((p.y - lo[1]) * inv), 0), d[1] - 1);
      const int bz = min(max((int)((p.z - lo[2]) * inv), 0), d[2] - 1);
      int len = 0;
      for (int z = max(bz - 1, 0); z <= min(bz + 1, d[2] - 1); ++z)
        for (int y = max(by - 1, 0); y <= min(by + 1, d[1] - 1); ++y) {
          const int row = d[0] * (y + d[1] * z);
          const int c0 = row + max(bx - 1, 0), c1 = row + min(bx + 1, d[0] - 1);
          const int s0 = c0 == 0 ? 0 : cend[c0 - 1], s1 = cend[c1];
          for (int c = s0; c < s1; ++c) {
            const int q = sidx[c];
            if (q == j) continue;
            const float4 o = sp[q];
            const float dx = o.x - p.x, dy = o.y - p.y, dz = o.z - p.z;
            if (fmaf(dz, dz, fmaf(dy, dy, dx * dx)) >= R2) continue;
            if (len < FB_NBR_MAX) a.m.fb_nbr[(int64_t)len * n + i0 + j] = (int16_t)q;
            ++len;
          }
        }
      a.m.fb_nlen[i0 + j] = (int16_t)min(len, FB_NBR_MAX);
      if (len > FB_NBR_MAX) sflag[1] = 1;
    }
    __syncthreads();
    const bool overflow = sflag[1] != 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) { /* an overflowed body builds again next step */
      const int j = t + NT * k;
      if (j < cnt)
        a.m.fb_ref[i0 + j] = overflow ? make_float4(NAN, NAN, NAN, NAN) : pr[k];
    }
  }
  /* a body whose lists overflowed sums its forces here, from the cell list,
   * and the list pass leaves it alone this step */
  if (!(rebuild && sflag[1] != 0)) return;
  if (t == 0) a.m.fb_done[g] = (int32_t)*a.m.d_step;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int j = t + NT * k;
    if (j >= cnt) continue;
    const int i = i0 + j;
    const float4 pi = pr[k];
    const float xi = pi.x, yi = pi.y, zi = pi.z;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    if (morse) {
      {
        const int bx = min(max((int)((xi - lo[0]) * inv), 0), d[0] - 1);
        const int by = min(max((int)((yi - lo[1]) * inv), 0), d[1] - 1);
        const int bz = min(max((int)((zi - lo[2]) * inv), 0), d[2] - 1);
        for (int z = max(bz - 1, 0); z <= min(bz + 1, d[2] - 1); ++z)
          for (int y = max(by - 1, 0); y <= min(by + 1, d[1] - 1); ++y) {
            const int row = d[0] * (y + d[1] * z);
            const int c0 = row + max(bx - 1, 0), c1 = row + min(bx + 1, d[0] - 1);
            const int s0 = c0 == 0 ? 0 : cend[c0 - 1], s1 = cend[c1];
            for (int c = s0; c < s1; ++c) morse_pair(a, sp[sidx[c]], pi.w, xi, yi, zi, fx, fy, fz);
          }
      }
    }
    element_terms<1>(a, i, 0, xi, yi, zi, a.m.body[i], fx, fy, fz);
    a.m.fx[i] = fx;
    a.m.fy[i] = fy;
    a.m.fz[i] = fz;
  }
}

/* The list pass of the force pass by body: a workgroup of 256 lanes per
 * body, after forces_body_build_kernel has rebuilt the lists it had to.
 * The body's positions go to LDS; each element sums the Morse pairs of
 * its list (eight list entries read at a time, the pairs summed in list
 * order), adds its other terms and writes its force.  A body whose lists
 * overflowed this step was summed by the build kernel. */
__global__ __launch_bounds__(256) void forces_body_kernel(MArgs a) {
  __shared__ float4 sp[FB_MAX];
  const int t = threadIdx.x, g = blockIdx.x;
  if (a.m.fb_done[g] == (int32_t)*a.m.d_step) return;
  const int i0 = a.m.fbody_range[2 * g], cnt = a.m.fbody_range[2 * g + 1] - i0;
  const int64_t n = a.m.n;
  const bool morse = a.morse_u0 != 0.f && a.morse_cut > 0.f;
  for (int j = t; j < cnt; j += 256) sp[j] = a.m.pos4[i0 + j];
  __syncthreads();
  for (int j = t; j < cnt; j += 256) {
    const int i = i0 + j;
    const float4 pi = sp[j];
    const float xi = pi.x, yi = pi.y, zi = pi.z;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    if (morse) {
      const int len = a.m.fb_nlen[i];
      for (int q0 = 0; q0 < len; q0 += 8) {
        int16_t nb[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
          nb[u] = q0 + u < len ? a.m.fb_nbr[(int64_t)(q0 + u) * n + i] : (int16_t)-1;
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (nb[u] >= 0) morse_pair(a, sp[nb[u]], pi.w, xi, yi, zi, fx, fy, fz);
      }
    }
    element_terms<1>(a, i, 0, xi, yi, zi, a.m.body[i], fx, fy, fz);
    a.m.fx[i] = fx;
    a.m.fy[i] = fy;
    a.m.fz[i] = fz;
  }
}

/* ---------------- platelet-platelet contact ---------------- */

/* The non-bonded term of sem_Force_kernel (sem_kernel.cu:436-509): for every
 * two platelets the closest pair of elements (lowest element of the lower
 * platelet, then lowest partner, on ties) repels with
 * a kappa e^{-kappa h} / (1 - e^{-kappa h}), h = max(distance, hc / 100),
 * here only below the cutoff hc and found on the cell lists.  First launch:
 * each element, for each other platelet with an element within hc,
 * proposes the distance to its nearest element there; the key (distance
 * bits, element of the lower platelet) goes to the pair's slot by atomicMin
 * (distances >= 0 order as their bits), and the first proposal of a pair
 * appends the pair to the list.  An element reduces its candidates over up
 * to four other platelets in registers; a candidate of a fifth is proposed
 * on its own (any candidate pair is a valid proposal, so the result is the
 * same) and counted in counters[9]. */
constexpr uint64_t CONTACT_EMPTY = ~(uint64_t)0;

__device__ inline void contact_propose(const MArgs &a, int p, int q, int i, int j, float d2) {
  const int lo = min(p, q), hi = max(p, q);
  const uint32_t elem_lo = (uint32_t)(p < q ? i : j);
  const uint64_t key = ((uint64_t)__float_as_uint(sqrtf(d2)) << 32) | elem_lo;
  const int64_t pair = (int64_t)lo * a.m.n_plt + hi;
  const uint64_t old = atomicMin((unsigned long long *)&a.m.contact_slot[pair],
                                 (unsigned long long)key);
  if (old == CONTACT_EMPTY) {
    const int idx = (int)atomicAdd((unsigned long long *)&a.m.counters[7], 1ull);
    a.m.contact_list[idx] = (int32_t)pair;
  }
}

__global__ __launch_bounds__(256) void contact_pairs_kernel(MArgs a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.m.n || a.m.kind[i] != PFG_MARKER_ELEMENT) return;
  const int p = a.m.body[i];
  const float4 pi = a.m.pos4[i];
  const MLat L = lat_of(a, i);
  const int b = a.m.marker_bin[i] - L.bin0;
  const int b3[3] = {b % L.nb[0], (b / L.nb[0]) % L.nb[1], b / (L.nb[0] * L.nb[1])};
  const float hc2 = a.contact_hc * a.contact_hc;
  int qb[4] = {-1, -1, -1, -1}, qj[4] = {0, 0, 0, 0};
  float qd[4] = {0.f, 0.f, 0.f, 0.f};
  for_nbr_ranges(L, a.m.bin_start, b3, [&](int s, int e) {
    for (int c = s; c < e; ++c) {
      const float4 q = a.m.sorted_pos[c];
      const int tag = __float_as_int(q.w);
      if (tag < 0 || tag == p) continue;
      const float dx = q.x - pi.x, dy = q.y - pi.y, dz = q.z - pi.z;
      const float d2 = dx * dx + dy * dy + dz * dz;
      if (d2 >= hc2) continue;
      const int j = a.m.bin_items[c];
      bool held = false;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (qb[k] == tag) {
          if (d2 < qd[k] || (d2 == qd[k] && j < qj[k])) {
            qd[k] = d2;
            qj[k] = j;
          }
          held = true;
          break;
        }
        if (qb[k] < 0) {
          qb[k] = tag;
          qd[k] = d2;
          qj[k] = j;
          held = true;
          break;
        }
      }
      if (!held) { /* a fifth platelet: propose this pair on its own */
        contact_propose(a, p, tag, i, j, d2);
        bump(a.m.counters, 9);
      }
    }
  });
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (qb[k] >= 0) contact_propose(a, p, qb[k], i, qj[k], qd[k]);
}

/* Second launch, after the force pass: each listed pair takes its winning
 * element e of the lower platelet and e's nearest element j of the higher
 * one (lowest index on ties), adds the repulsion to both and empties the
 * slot.  A fixed grid strides over the list. */
__global__ __launch_bounds__(256) void contact_force_kernel(MArgs a) {
  const int64_t cnt = a.m.counters[7];
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < cnt;
       k += (int64_t)gridDim.x * 256) {
    const int pair = a.m.contact_list[k];
    const int hi = pair % a.m.n_plt;
    const uint64_t key = a.m.contact_slot[pair];
    a.m.contact_slot[pair] = CONTACT_EMPTY;
    const int e = (int)(uint32_t)(key & 0xffffffffu);
    const float4 pe = a.m.pos4[e];
    const MLat L = lat_of(a, e);
    const int b = a.m.marker_bin[e] - L.bin0;
    const int b3[3] = {b % L.nb[0], (b / L.nb[0]) % L.nb[1], b / (L.nb[0] * L.nb[1])};
    float best = INFINITY;
    int j = -1;
    for_nbr_ranges(L, a.m.bin_start, b3, [&](int s, int t) {
      for (int c = s; c < t; ++c) {
        const float4 q = a.m.sorted_pos[c];
        if (__float_as_int(q.w) != hi) continue;
        const float dx = q.x - pe.x, dy = q.y - pe.y, dz = q.z - pe.z;
        const float d2 = dx * dx + dy * dy + dz * dz;
        const int jj = a.m.bin_items[c];
        if (d2 < best || (d2 == best && jj < j)) {
          best = d2;
          j = jj;
        }
      }
    });
    if (j < 0) continue;
    const float4 pj = a.m.pos4[j];
    const float dx = pe.x - pj.x, dy = pe.y - pj.y, dz = pe.z - pj.z;
    const float d = sqrtf(dx * dx + dy * dy + dz * dz);
    if (!(d > 0.f)) continue;
    const float h = fmaxf(d, 0.01f * a.contact_hc);
    const float ex = expf(-a.contact_kappa * h);
    const float F = a.contact_a * a.contact_kappa * ex / (1.0f - ex) / d;
    atomicAdd(&a.m.fx[e], F * dx);
    atomicAdd(&a.m.fy[e], F * dy);
    atomicAdd(&a.m.fz[e], F * dz);
    atomicAdd(&a.m.fx[j], -F * dx);
    atomicAdd(&a.m.fy[j], -F * dy);
    atomicAdd(&a.m.fz[j], -F * dz);
  }
}

/* ---------------- K9/K10: spreading, interpolation, update ---------------- */

/* Spreading has two forms.  Direct: one wavefront per marker, a lane per
 * stencil node, FP32 atomics into the lattice force field.  Staged: a
 * workgroup per chunk (a run of consecutive markers of one platelet) stages
 * the chunk's stencils in LDS, sums them node by node over the box of nodes
 * they cover, and adds the box's nonzero nodes to the global field; a
 * platelet's stencils overlap, so this cuts the device-wide atomics several
 * times over when many markers share the lattice.  A chunk whose box is
 * large (markers scattered through the lattice) is spread directly. */

/* Stamp the tiles of a marker's stencil with this step (the LB update reads
 * the force of stamped tiles only): lane (kx, ky, kz) holds node (x, y, z)
 * and stores the stamp, without reading it, when its node starts a run of
 * the stencil's nodes inside one tile along every axis (the stencil's first
 * node, or a node at a tile boundary, which a wrap or a wall-bounded axis's
 * first node also is), so at most eight lanes store and every tile the
 * stencil meets is stamped. */
__device__ inline void stamp_stencil_tile(const MLat &L, int kx, int ky, int kz, int x, int y,
                                          int z, int32_t step) {
  if ((kx == 0 || (x & (TILE_X - 1)) == 0) && (ky == 0 || (y & (TILE_Y - 1)) == 0) &&
      (kz == 0 || (z & (TILE_Z - 1)) == 0))
    L.tile_stamp[(x >> TILE_SHIFT_X) + L.tnx * ((y >> TILE_SHIFT_Y) + L.tny * (z >> TILE_SHIFT_Z))] =
        step;
}

/* Direct spreading of marker i: lane = stencil node (consecutive lanes take
 * consecutive x, so one atomic instruction covers 16 runs of 4 nodes). */
__device__ void spread_direct(const MArgs &a, int i, int lane, float Fx, float Fy, float Fz) {
  if (!ib_coupled(a.m.kind[i], a.fiber_move)) return;
  const MLat L = lat_of(a, i);
  const float X = wrap_coord(a.m.x[i], L.nx, L.periodic[0]);
  const float Y = wrap_coord(a.m.y[i], L.ny, L.periodic[1]);
  const float Z = wrap_coord(a.m.z[i], L.nz, L.periodic[2]);
  float wx, wy, wz;
  const int ix = stencil_node(X, lane & 3, L.nx, L.periodic[0], &wx);
  const int iy = stencil_node(Y, (lane >> 2) & 3, L.ny, L.periodic[1], &wy);
  const int iz = stencil_node(Z, lane >> 4, L.nz, L.periodic[2], &wz);
  if (ix < 0 || iy < 0 || iz < 0) return;
  stamp_stencil_tile(L, lane & 3, (lane >> 2) & 3, lane >> 4, ix, iy, iz, (int32_t)*a.m.d_step);
  if (Fx == 0.f && Fy == 0.f && Fz == 0.f) return; /* adds nothing; the tiles are stamped */
  const float w = wx * wy * wz * a.force_scale;
  const int64_t c = ix + (int64_t)L.nx * (iy + (int64_t)L.ny * iz);
  atomicAdd(L.aux + c, w * Fx);
  atomicAdd(L.aux + a.plane + c, w * Fy);
  atomicAdd(L.aux + 2 * a.plane + c, w * Fz);
}

__device__ inline void spread_direct(const MArgs &a, int i, int lane) {
  spread_direct(a, i, lane, a.m.fx[i], a.m.fy[i], a.m.fz[i]);
}

/* One wavefront per marker. */
__global__ __launch_bounds__(256) void spread_kernel(MArgs a) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= a.m.n || halted(a)) return;
  spread_direct(a, i, threadIdx.x & 63);
}

/* Wrapped coordinate (gives the kernel weights exactly as the direct form)
 * and the first stencil node in the periodic image nearest to the chunk's
 * reference point `ref` (the wrapped position of its first marker), so that
 * one platelet's stencils stay contiguous across a periodic boundary. */
__device__ inline void stencil_origin(const MLat &L, const float v[3], const float ref[3],
                                      float W[3], int o[3]) {
  const int n[3] = {L.nx, L.ny, L.nz};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    W[k] = wrap_coord(v[k], n[k], L.periodic[k]);
    const int img = L.periodic[k] ? -(int)rintf((W[k] - ref[k]) / (float)n[k]) * n[k] : 0;
    o[k] = (int)floorf(W[k]) - 1 + img;
  }
}


/* Two wavefronts per chunk (at most IB_CHUNK markers of one body).  The
 * workgroup first stages the chunk in LDS: per marker its first stencil
 * node, its 3 x 4 kernel weights and its force.  When the chunk's stencils
 * span at most IB_BOX_CELLS nodes, each wavefront then adds its half of the
 * markers, one after the other, into its own FP32 box of those nodes in
 * LDS, lane l taking node (l & 3, (l >> 2) & 3, l >> 4) of the marker's
 * 4 x 4 x 4 stencil: the 64 nodes of one marker are distinct, and a
 * wavefront's LDS accesses are carried out in the order it issues them, so
 * the read-add-write of each node needs no atomic and the sums follow
 * marker order whatever the timing.  The two boxes are added (first half,
 * then second) and go to the lattice field with one device-wide atomic per
 * nonzero node and component.  A box of up to twice IB_BOX_CELLS nodes (a
 * chunk that rotated or deformed since load) is summed by wavefront 0
 * alone, marker after marker, in the two boxes' memory taken as one.  A
 * scattered chunk (fibrin nodes far apart) adds each marker's 64 nodes with
 * device-wide atomics, as the direct form does; so do the markers listed
 * for direct spreading, IB_CHUNK per workgroup before the chunks', staged
 * the same way (each with its own lattice).  Contributions are the direct form's bit for bit; only the
 * order in which they are added differs. */
__global__ __launch_bounds__(128) void spread_chunk_kernel(MArgs a) {
  static_assert(IB_CHUNK <= 128, "a thread per marker in the staging pass");
  __shared__ int4 so[IB_CHUNK];      /* first stencil node; x = INT_MIN: not coupled */
  __shared__ float sw[IB_CHUNK][12]; /* weights along x (0..3), y (4..7), z (8..11) */
  __shared__ float sf[IB_CHUNK][3];
  /* a box per wavefront (wavefront w's at w * IB_BOX_CELLS), or one wide
   * box; the three components in arrays of their own, so that a marker's
   * three read-add-writes are independent */
  __shared__ float bxs[2 * IB_BOX_CELLS], bys[2 * IB_BOX_CELLS], bzs[2 * IB_BOX_CELLS];
  __shared__ int sh[2][6];
  __shared__ int sidx[IB_CHUNK];
  if (halted(a)) return;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  /* the workgroups of the direct list come first, so that their atomics
   * run beside the chunks' summation instead of after it (spreading them
   * evenly among the chunks' workgroups measured no faster: the atomics and
   * the chunks' summation do not overlap either way) */
  const int ndirect = (a.m.n_direct + IB_CHUNK - 1) / IB_CHUNK;
  const bool chunk = (int)blockIdx.x >= ndirect;
  const int cid = (int)blockIdx.x - ndirect; /* the chunk */
  int c0, cnt;
  if (chunk) {
    c0 = a.m.chunk_start[cid];
    cnt = a.m.chunk_start[cid + 1] - c0;
  } else { /* IB_CHUNK markers of the direct list (of any lattices) */
    const int k0 = IB_CHUNK * (int)blockIdx.x;
    c0 = a.m.chunk_start[a.m.n_chunks] + k0;
    cnt = min(IB_CHUNK, a.m.n_direct - k0);
  }
  if (cnt <= 0) return;
  const int *items = a.m.chunk_items + c0;
  /* body spreading: a chunk of a body writes its box to its scratch box;
   * a chunk's stencils are placed in the periodic image nearest to its
   * reference marker (for a chunk of a body the body's first marker, the
   * frame of the body's box) */
  const int cbody = chunk && a.spread_scratch ? a.m.chunk_body[cid] : -1;
  const int m0 = chunk ? a.m.chunk_ref[cid] : items[0];
  const MLat L = lat_of(a, m0); /* a chunk never spans two lattices */
  float ref[3];
  ref[0] = wrap_coord(a.m.x[m0], L.nx, L.periodic[0]);
  ref[1] = wrap_coord(a.m.y[m0], L.ny, L.periodic[1]);
  ref[2] = wrap_coord(a.m.z[m0], L.nz, L.periodic[2]);
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  if (t < cnt) {
    const int i = items[t];
    sidx[t] = i;
    const bool has = ib_coupled(a.m.kind[i], a.fiber_move);
    int o[3] = {0, 0, 0};
    if (has) {
      float W[3];
      const float v[3] = {a.m.x[i], a.m.y[i], a.m.z[i]};
      if (chunk) {
        stencil_origin(L, v, ref, W, o);
      } else { /* its own lattice and image */
        const MLat Li = lat_of(a, i);
        float own[3];
        own[0] = wrap_coord(v[0], Li.nx, Li.periodic[0]);
        own[1] = wrap_coord(v[1], Li.ny, Li.periodic[1]);
        own[2] = wrap_coord(v[2], Li.nz, Li.periodic[2]);
        stencil_origin(Li, v, own, W, o);
      }
      stencil_weights(W[0], &sw[t][0]);
      stencil_weights(W[1], &sw[t][4]);
      stencil_weights(W[2], &sw[t][8]);
      sf[t][0] = a.m.fx[i];
      sf[t][1] = a.m.fy[i];
      sf[t][2] = a.m.fz[i];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        lo[k] = o[k];
        hi[k] = o[k] + 3;
      }
    } else { /* summed into the box as zeros, without a branch */
#pragma unroll
      for (int k = 0; k < 12; ++k) sw[t][k] = 0.f;
      sf[t][0] = sf[t][1] = sf[t][2] = 0.f;
    }
    so[t] = make_int4(has ? o[0] : INT_MIN, o[1], o[2], 0);
  }
#pragma unroll
  for (int s2 = 32; s2 > 0; s2 >>= 1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[k] = min(lo[k], __shfl_xor(lo[k], s2, 64));
      hi[k] = max(hi[k], __shfl_xor(hi[k], s2, 64));
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      sh[wave][k] = lo[k];
      sh[wave][3 + k] = hi[k];
    }
  }
  __syncthreads(); /* the staged chunk and both halves' bounds */
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lo[k] = min(sh[0][k], sh[1][k]);
    hi[k] = max(sh[0][3 + k], sh[1][3 + k]);
  }
  const int dim0 = hi[0] - lo[0] + 1, dim1 = hi[1] - lo[1] + 1, dim2 = hi[2] - lo[2] + 1;
  const int64_t vol64 = (int64_t)dim0 * dim1 * dim2;
  const int32_t step = (int32_t)*a.m.d_step;
  if (cbody >= 0 && t == 0) {
    /* the scratch box written below, or none (extent 0): no coupled
     * marker, or a box too large, spread with atomics (the body's nodes then
     * all go with atomics) */
    const bool boxed = lo[0] <= hi[0] && vol64 <= 2 * IB_BOX_CELLS;
    a.m.chunk_info[2 * cid] = make_int4(lo[0], lo[1], lo[2], 0);
    a.m.chunk_info[2 * cid + 1] =
        boxed ? make_int4(dim0, dim1, dim2, (int)vol64) : make_int4(0, 0, 0, 0);
    if (lo[0] <= hi[0] && !boxed) a.m.body_fallback[cbody] = step;
  }
  if (lo[0] > hi[0]) return; /* no IB-coupled marker in the chunk */
  const int kx = lane & 3, ky = (lane >> 2) & 3, kz = lane >> 4;
  /* a box of IB_BOX_CELLS to 2 IB_BOX_CELLS nodes (a chunk that rotated or
   * deformed since load) is summed by wavefront 0 alone in the two
   * wavefronts' boxes taken as one */
  const bool wide = vol64 > IB_BOX_CELLS;
  if (!chunk || vol64 > 2 * IB_BOX_CELLS) { /* scattered markers: device-wide atomics per node */
    for (int j = wave; j < cnt; j += 2) {
      if (so[j].x == INT_MIN) continue;
      const MLat Lj = chunk ? L : lat_of(a, sidx[j]);
      const int ix = axis_index(so[j].x + kx, Lj.nx, Lj.periodic[0]);
      const int iy = axis_index(so[j].y + ky, Lj.ny, Lj.periodic[1]);
      const int iz = axis_index(so[j].z + kz, Lj.nz, Lj.periodic[2]);
      if (ix < 0 || iy < 0 || iz < 0) continue;
      stamp_stencil_tile(Lj, kx, ky, kz, ix, iy, iz, step);
      /* a marker without force adds nothing (its tiles are stamped still:
       * the interpolation reads the velocity there) */
      if (sf[j][0] == 0.f && sf[j][1] == 0.f && sf[j][2] == 0.f) continue;
      const float w = sw[j][kx] * sw[j][4 + ky] * sw[j][8 + kz] * a.force_scale;
      const int64_t c = ix + (int64_t)Lj.nx * (iy + (int64_t)Lj.ny * iz);
      atomicAdd(Lj.aux + c, w * sf[j][0]);
      atomicAdd(Lj.aux + a.plane + c, w * sf[j][1]);
      atomicAdd(Lj.aux + 2 * a.plane + c, w * sf[j][2]);
    }
    return;
  }
  const int vol = (int)vol64;
  for (int k = t; k < vol; k += 128) {
    bxs[k] = 0.f;
    bys[k] = 0.f;
    bzs[k] = 0.f;
    if (!wide) {
      bxs[IB_BOX_CELLS + k] = 0.f;
      bys[IB_BOX_CELLS + k] = 0.f;
      bzs[IB_BOX_CELLS + k] = 0.f;
    }
  }
  /* each marker's first box node (0 for a marker not coupled: its weights
   * and force are zero), and each lane's offset of its stencil node from it */
  if (t < cnt) {
    const int4 oj = so[t];
    so[t].w = oj.x == INT_MIN ? 0 : (oj.x - lo[0]) + dim0 * ((oj.y - lo[1]) + dim1 * (oj.z - lo[2]));
  }
  const int off = kx + dim0 * (ky + dim1 * kz);
  __syncthreads();
  const int boff = wide ? 0 : wave * IB_BOX_CELLS;
  const int j0 = wide ? 0 : 64 * wave, j1 = wide ? (wave == 0 ? cnt : 0) : min(cnt, 64 * (wave + 1));
#pragma unroll 4
  for (int j = j0; j < j1; ++j) {
    const float w = sw[j][kx] * sw[j][4 + ky] * sw[j][8 + kz] * a.force_scale;
    const int c = boff + so[j].w + off;
    bxs[c] += w * sf[j][0];
    bys[c] += w * sf[j][1];
    bzs[c] += w * sf[j][2];
  }
  __syncthreads();
  if (cbody >= 0) { /* the chunk's sums to its scratch box (x, y, z planes of vol), zeros included */
    float *out = a.m.chunk_sum + (int64_t)cid * (6 * IB_BOX_CELLS);
    for (int k = t; k < vol; k += 128) {
      out[k] = wide ? bxs[k] : bxs[k] + bxs[IB_BOX_CELLS + k];
      out[vol + k] = wide ? bys[k] : bys[k] + bys[IB_BOX_CELLS + k];
      out[2 * vol + k] = wide ? bzs[k] : bzs[k] + bzs[IB_BOX_CELLS + k];
    }
    return;
  }
  const float inv0 = 1.0f / (float)dim0, inv1 = 1.0f / (float)dim1; /* vol <= 2 IB_BOX_CELLS */
  /* Every tile the box meets is stamped (a superset of the stencils' tiles:
   * interpolation reads the velocity of all their nodes).  Along each axis
   * the box nodes inside one tile form runs that start at the box's first
   * node or at a tile boundary (a wrap lands on node 0, and a wall-bounded
   * axis resumes at node 0), so the corner nodes of those runs reach every
   * tile while only a few lanes read and write the stamps. */
  for (int k = t; k < vol; k += 128) {
    const int r = div_small(k, inv0), bx = k - r * dim0;
    const int bz = div_small(r, inv1), by = r - bz * dim1;
    const int gx = axis_index(lo[0] + bx, L.nx, L.periodic[0]);
    const int gy = axis_index(lo[1] + by, L.ny, L.periodic[1]);
    const int gz = axis_index(lo[2] + bz, L.nz, L.periodic[2]);
    if (gx < 0 || gy < 0 || gz < 0) continue;
    if ((bx == 0 || (gx & (TILE_X - 1)) == 0) && (by == 0 || (gy & (TILE_Y - 1)) == 0) &&
        (bz == 0 || (gz & (TILE_Z - 1)) == 0)) /* few per chunk: stored without a read */
      L.tile_stamp[(gx >> TILE_SHIFT_X) + L.tnx * ((gy >> TILE_SHIFT_Y) + L.tny * (gz >> TILE_SHIFT_Z))] = step;
    const int64_t c = gx + (int64_t)L.nx * (gy + (int64_t)L.ny * gz);
    const float fx = wide ? bxs[k] : bxs[k] + bxs[IB_BOX_CELLS + k];
    const float fy = wide ? bys[k] : bys[k] + bys[IB_BOX_CELLS + k];
    const float fz = wide ? bzs[k] : bzs[k] + bzs[IB_BOX_CELLS + k];
    if (fx != 0.f) atomicAdd(L.aux + c, fx);
    if (fy != 0.f) atomicAdd(L.aux + a.plane + c, fy);
    if (fz != 0.f) atomicAdd(L.aux + 2 * a.plane + c, fz);
  }
}

/* ---------------- body spreading (form 4) ---------------- */

/* Per tile, the step (high word) and the body that claimed the tile alone
 * (low word), or OWNER_SHARED: another body, a marker outside the bodies,
 * or a body that spreads with atomics meets the tile this step. */
constexpr uint32_t OWNER_SHARED = 0xffffffffu;

__device__ inline uint64_t owner_word(uint32_t step, uint32_t who) {
  return ((uint64_t)step << 32) | who;
}

/* Distinct tiles (at most three: a periodic axis whose length is no
 * multiple of the tile's extent wraps from a partial tile) of stencil nodes
 * o..o+3 along one axis; nodes outside a wall-bounded axis are dropped. */
__device__ inline int stencil_tiles3(int o, int n, int periodic, int shift, int t[3]) {
  int cnt = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int i = axis_index(o + k, n, periodic);
    if (i < 0) continue;
    const int tt = i >> shift;
    bool seen = false;
    for (int q = 0; q < cnt; ++q) seen = seen || t[q] == tt;
    if (!seen && cnt < 3) t[cnt++] = tt;
  }
  return cnt;
}

/* Mark every tile of the stencil with first node o shared this step. */
__device__ inline void claim_stencil_shared(const MLat &L, const int o[3], uint32_t step) {
  int tx[3], ty[3], tz[3];
  const int nx = stencil_tiles3(o[0], L.nx, L.periodic[0], TILE_SHIFT_X, tx);
  const int ny = stencil_tiles3(o[1], L.ny, L.periodic[1], TILE_SHIFT_Y, ty);
  const int nz = stencil_tiles3(o[2], L.nz, L.periodic[2], TILE_SHIFT_Z, tz);
  const uint64_t shared = owner_word(step, OWNER_SHARED);
  for (int c = 0; c < nz; ++c)
    for (int b = 0; b < ny; ++b)
      for (int q = 0; q < nx; ++q)
        atomicExch((unsigned long long *)(L.tile_owner + tx[q] + (int64_t)L.tnx * (ty[b] + (int64_t)L.tny * tz[c])),
                   (unsigned long long)shared);
}

/* Claim tile gt for body g: it stays the body's only while no other
 * contributor claims it this step. */
__device__ inline void claim_tile(uint64_t *owner, uint32_t step, uint32_t g) {
  const uint64_t mine = owner_word(step, g), shared = owner_word(step, OWNER_SHARED);
  unsigned long long *p = (unsigned long long *)owner;
  unsigned long long cur = atomicCAS(p, (unsigned long long)mine, (unsigned long long)mine);
  while (cur != (unsigned long long)mine) {
    const unsigned long long want = (cur >> 32) == step ? (unsigned long long)shared
                                                       : (unsigned long long)mine;
    if (cur == want) break; /* already shared */
    const unsigned long long prev = atomicCAS(p, cur, want);
    if (prev == cur) break;
    cur = prev;
  }
}

/* Stencil box of body g (markers i0 .. i0+cnt-1): the first stencil nodes
 * of its IB-coupled markers in the periodic image nearest to its first
 * marker, reduced over the workgroup of 256 lanes; lo > hi: none.  The
 * claims and the spreading both compute it from the same positions. */
__device__ void body_box(const MArgs &a, int i0, int cnt, const MLat &L, const float ref[3],
                         int lo[3], int hi[3]) {
  __shared__ int shb[4][6];
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lo[k] = INT_MAX;
    hi[k] = INT_MIN;
  }
#pragma unroll 4
  for (int j = t; j < cnt; j += 256) {
    const int i = i0 + j;
    if (!ib_coupled(a.m.kind[i], a.fiber_move)) continue;
    const float v[3] = {a.m.x[i], a.m.y[i], a.m.z[i]};
    float W[3];
    int o[3];
    stencil_origin(L, v, ref, W, o);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[k] = min(lo[k], o[k]);
      hi[k] = max(hi[k], o[k] + 3);
    }
  }
#pragma unroll
  for (int s2 = 32; s2 > 0; s2 >>= 1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[k] = min(lo[k], __shfl_xor(lo[k], s2, 64));
      hi[k] = max(hi[k], __shfl_xor(hi[k], s2, 64));
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      shb[wave][k] = lo[k];
      shb[wave][3 + k] = hi[k];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lo[k] = min(min(shb[0][k], shb[1][k]), min(shb[2][k], shb[3][k]));
    hi[k] = max(max(shb[0][3 + k], shb[1][3 + k]), max(shb[2][3 + k], shb[3][3 + k]));
  }
  __syncthreads(); /* shb may be written again */
}

/* The tiles of a body box, as runs of box nodes along each axis: run[ax][b]
 * is the run of box node b, tile[ax][r] the lattice tile coordinate of run
 * r (-1: nodes outside a wall-bounded axis).  Consecutive nodes in one tile
 * form a run.  Lanes 0, 64 and 128 build one axis each. */
struct BoxRuns {
  uint8_t run[3][BODY_BOX_DIM];
  int16_t tile[3][BODY_BOX_DIM];
  int16_t first[3][BODY_BOX_DIM + 1]; /* first box node of each run, then dim */
  int nrun[3];
};

__device__ void box_runs(const MLat &L, const int lo[3], const int dim[3], BoxRuns &R) {
  const int t = threadIdx.x;
  if ((t & 63) == 0 && (t >> 6) < 3) {
    const int ax = t >> 6;
    const int n = ax == 0 ? L.nx : ax == 1 ? L.ny : L.nz;
    const int shift = ax == 0 ? TILE_SHIFT_X : ax == 1 ? TILE_SHIFT_Y : TILE_SHIFT_Z;
    int r = -1, prev = INT_MIN;
    for (int b = 0; b < dim[ax]; ++b) {
      const int g = axis_index(lo[ax] + b, n, L.periodic[ax]);
      const int tt = g < 0 ? -1 : g >> shift;
      if (b == 0 || tt != prev) {
        ++r;
        R.tile[ax][r] = (int16_t)tt;
        R.first[ax][r] = (int16_t)b;
        prev = tt;
      }
      R.run[ax][b] = (uint8_t)r;
    }
    R.nrun[ax] = r + 1;
    R.first[ax][r + 1] = (int16_t)dim[ax];
  }
  __syncthreads();
}

/* Whether a body's box can take tile ownership: at most BODY_BOX_DIM nodes
 * along each axis, none longer than a periodic axis (each box node is then
 * one lattice node). */
__device__ inline bool body_box_fits(const MLat &L, const int dim[3]) {
  const int n[3] = {L.nx, L.ny, L.nz};
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 3; ++k)
    ok = ok && dim[k] <= BODY_BOX_DIM && !(L.periodic[k] && dim[k] > n[k]);
  return ok;
}

__device__ void claim_role(const MArgs &a, int block) {
  __shared__ BoxRuns R;
  const uint32_t step = (uint32_t)*a.m.d_step;
  const int t = threadIdx.x;
  if (block >= a.blocks1) { /* the markers outside the bodies: their tiles are shared */
    const int k = 256 * (block - a.blocks1) + t;
    const int nitems = a.m.rchunk_start[a.m.n_rchunks] + a.m.n_rdirect;
    if (k >= nitems) return;
    const int i = a.m.rchunk_items[k];
    if (!ib_coupled(a.m.kind[i], a.fiber_move)) return;
    const MLat L = lat_of(a, i);
    const float v[3] = {a.m.x[i], a.m.y[i], a.m.z[i]};
    float own[3], W[3];
    int o[3];
    own[0] = wrap_coord(v[0], L.nx, L.periodic[0]);
    own[1] = wrap_coord(v[1], L.ny, L.periodic[1]);
    own[2] = wrap_coord(v[2], L.nz, L.periodic[2]);
    stencil_origin(L, v, own, W, o);
    claim_stencil_shared(L, o, step);
    return;
  }
  const int g = block;
  const int i0 = a.m.body_range[4 * g], cnt = a.m.body_range[4 * g + 1] - i0;
  const MLat L = lat_of(a, i0);
  float ref[3];
  ref[0] = wrap_coord(a.m.x[i0], L.nx, L.periodic[0]);
  ref[1] = wrap_coord(a.m.y[i0], L.ny, L.periodic[1]);
  ref[2] = wrap_coord(a.m.z[i0], L.nz, L.periodic[2]);
  int lo[3], hi[3];
  body_box(a, i0, cnt, L, ref, lo, hi);
  if (lo[0] > hi[0]) return;
  const int dim[3] = {hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
  bool fits = body_box_fits(L, dim);
  if (fits) box_runs(L, lo, dim, R);
  const int nt = fits ? R.nrun[0] * R.nrun[1] * R.nrun[2] : 0;
  fits = fits && nt <= BODY_BOX_TILES;
  if (!fits) { /* spread with atomics: every tile of its stencils is shared */
    for (int j = t; j < cnt; j += 256) {
      const int i = i0 + j;
      if (!ib_coupled(a.m.kind[i], a.fiber_move)) continue;
      const float v[3] = {a.m.x[i], a.m.y[i], a.m.z[i]};
      float W[3];
      int o[3];
      stencil_origin(L, v, ref, W, o);
      claim_stencil_shared(L, o, step);
    }
    return;
  }
  for (int lt = t; lt < nt; lt += 256) {
    const int rx = lt % R.nrun[0], r2 = lt / R.nrun[0];
    const int ry = r2 % R.nrun[1], rz = r2 / R.nrun[1];
    const int tx = R.tile[0][rx], ty = R.tile[1][ry], tz = R.tile[2][rz];
    if (tx < 0 || ty < 0 || tz < 0) continue;
    claim_tile(L.tile_owner + tx + (int64_t)L.tnx * (ty + (int64_t)L.tny * tz), step, (uint32_t)g);
  }
}

/* Second pass of the body spreading: a workgroup of 256 lanes per body.
 * The first pass (spread_chunk_kernel with spread_scratch) left each chunk
 * of the body's sums in a scratch box (its extents in chunk_info); the
 * body's box is their union (the box the force launch claimed tiles for).
 * Each tile of it is stamped with the step; each node sums the chunks'
 * boxes that hold it, in chunk order, and a nonzero sum goes to the
 * lattice field with a plain store when the force launch left its tile
 * claimed by this body alone (no other stencil meets it, so the node is
 * this body's sum and was zero), else with a device-wide atomic.  A body
 * spread once this step, or one of whose chunks spread with atomics, adds
 * with atomics only.  Contributions are the direct form's bit for bit;
 * only the order of the sums differs, and it is fixed. */
__global__ __launch_bounds__(256) void spread_body_kernel(MArgs a) {
  __shared__ BoxRuns R;
  __shared__ uint8_t owned[BODY_BOX_TILES];
  /* per tile of the box, the chunks (of the first 32) whose boxes meet it */
  __shared__ uint32_t tmask[BODY_BOX_TILES];
  __shared__ int4 cinfo[BODY_MAX_CHUNKS][2];
  constexpr int NT = 256, NW = NT / 64;
  __shared__ int shb[NW][6];
  __shared__ int may_store;
  if (halted(a)) return;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int g = blockIdx.x;
  const int i0 = a.m.body_range[4 * g];
  const int c0 = a.m.body_range[4 * g + 2], nc = a.m.body_range[4 * g + 3] - c0;
  const MLat L = lat_of(a, i0);
  const uint32_t step = (uint32_t)*a.m.d_step;
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  for (int q = t; q < nc; q += NT) {
    const int4 l = a.m.chunk_info[2 * (c0 + q)], d = a.m.chunk_info[2 * (c0 + q) + 1];
    cinfo[q][0] = l;
    cinfo[q][1] = d;
    if (d.w > 0) {
      lo[0] = min(lo[0], l.x);
      lo[1] = min(lo[1], l.y);
      lo[2] = min(lo[2], l.z);
      hi[0] = max(hi[0], l.x + d.x - 1);
      hi[1] = max(hi[1], l.y + d.y - 1);
      hi[2] = max(hi[2], l.z + d.z - 1);
    }
  }
#pragma unroll
  for (int s2 = 32; s2 > 0; s2 >>= 1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[k] = min(lo[k], __shfl_xor(lo[k], s2, 64));
      hi[k] = max(hi[k], __shfl_xor(hi[k], s2, 64));
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      shb[wave][k] = lo[k];
      shb[wave][3 + k] = hi[k];
    }
  }
  if (t == 0) { /* the first spreading of this body in this step stores */
    may_store = a.m.body_spread[g] != (int32_t)step && a.m.body_fallback[g] != (int32_t)step;
    a.m.body_spread[g] = (int32_t)step;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lo[k] = shb[0][k];
    hi[k] = shb[0][3 + k];
    for (int w = 1; w < NW; ++w) {
      lo[k] = min(lo[k], shb[w][k]);
      hi[k] = max(hi[k], shb[w][3 + k]);
    }
  }
  if (lo[0] > hi[0]) return; /* no chunk wrote a box */
  const int dim[3] = {hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
  bool fits = body_box_fits(L, dim);
  if (fits) box_runs(L, lo, dim, R);
  const int nt = fits ? R.nrun[0] * R.nrun[1] * R.nrun[2] : 0;
  fits = fits && nt <= BODY_BOX_TILES;
  if (!fits) { /* chunk by chunk, every nonzero node with atomics */
    for (int q = 0; q < nc; ++q) {
      const int4 l = cinfo[q][0], d = cinfo[q][1];
      const float *src = a.m.chunk_sum + (int64_t)(c0 + q) * (6 * IB_BOX_CELLS);
      for (int k = t; k < d.w; k += NT) {
        const int r = k / d.x, bx = k - r * d.x;
        const int bz = r / d.y, by = r - bz * d.y;
        const int gx = axis_index(l.x + bx, L.nx, L.periodic[0]);
        const int gy = axis_index(l.y + by, L.ny, L.periodic[1]);
        const int gz = axis_index(l.z + bz, L.nz, L.periodic[2]);
        if (gx < 0 || gy < 0 || gz < 0) continue;
        if ((bx == 0 || (gx & (TILE_X - 1)) == 0) && (by == 0 || (gy & (TILE_Y - 1)) == 0) &&
            (bz == 0 || (gz & (TILE_Z - 1)) == 0))
          L.tile_stamp[(gx >> TILE_SHIFT_X) + L.tnx * ((gy >> TILE_SHIFT_Y) + L.tny * (gz >> TILE_SHIFT_Z))] =
              (int32_t)step;
        const int64_t c = gx + (int64_t)L.nx * (gy + (int64_t)L.ny * gz);
        const float fx = src[k], fy = src[d.w + k], fz = src[2 * d.w + k];
        if (fx != 0.f) atomicAdd(L.aux + c, fx);
        if (fy != 0.f) atomicAdd(L.aux + a.plane + c, fy);
        if (fz != 0.f) atomicAdd(L.aux + 2 * a.plane + c, fz);
      }
    }
    return;
  }
  const uint64_t mine = owner_word(step, (uint32_t)g);
  /* stamps and ownership of the box's tiles: a tile is stamped when one of
   * the chunks' boxes (which hold all stencils) meets it, so the corners of
   * the platelet's box that no stencil reaches stay out of the LB update's
   * IB work */
  for (int lt = t; lt < nt; lt += NT) {
    const int rx = lt % R.nrun[0], r2 = lt / R.nrun[0];
    const int ry = r2 % R.nrun[1], rz = r2 / R.nrun[1];
    const int tx = R.tile[0][rx], ty = R.tile[1][ry], tz = R.tile[2][rz];
    bool own = false;
    uint32_t mask = 0;
    if (tx >= 0 && ty >= 0 && tz >= 0) {
      /* the tile's box nodes, in the frame of the chunks' boxes */
      const int x0 = lo[0] + R.first[0][rx], x1 = lo[0] + R.first[0][rx + 1] - 1;
      const int y0 = lo[1] + R.first[1][ry], y1 = lo[1] + R.first[1][ry + 1] - 1;
      const int z0 = lo[2] + R.first[2][rz], z1 = lo[2] + R.first[2][rz + 1] - 1;
      bool met = false;
      for (int q = 0; q < nc; ++q) {
        const int4 l = cinfo[q][0], d = cinfo[q][1];
        const bool m = d.w > 0 && x0 < l.x + d.x && l.x <= x1 && y0 < l.y + d.y &&
                       l.y <= y1 && z0 < l.z + d.z && l.z <= z1;
        met = met || m;
        if (m && q < 32) mask |= 1u << q;
      }
      if (met) {
        const int64_t gt = tx + (int64_t)L.tnx * (ty + (int64_t)L.tny * tz);
        L.tile_stamp[gt] = (int32_t)step;
        own = may_store && L.tile_owner[gt] == mine;
      }
    }
    owned[lt] = own ? 1 : 0;
    tmask[lt] = mask;
  }
  __syncthreads();
  const int vol = dim[0] * dim[1] * dim[2];
  /* node k = bx + dim0 (by + dim1 bz), stepped by NT without divisions */
  const int sx = NT % dim[0], sq = NT / dim[0], sy = sq % dim[1], sz = sq / dim[1];
  int nbx = t % dim[0], nby = (t / dim[0]) % dim[1], nbz = t / dim[0] / dim[1];
  for (int k = t; k < vol; k += NT) {
    const int bx = nbx, by = nby, bz = nbz;
    nbx += sx;
    nby += sy;
    nbz += sz;
    if (nbx >= dim[0]) {
      nbx -= dim[0];
      ++nby;
    }
    if (nby >= dim[1]) {
      nby -= dim[1];
      ++nbz;
    }
    const int ux = lo[0] + bx, uy = lo[1] + by, uz = lo[2] + bz;
    const int lt = R.run[0][bx] + R.nrun[0] * (R.run[1][by] + R.nrun[1] * R.run[2][bz]);
    float fx = 0.f, fy = 0.f, fz = 0.f;
    /* the chunks whose boxes meet the node's tile (all chunks when there are
     * more than 32), four or eight at a time: their loads are issued
     * together, then summed in chunk order; a chunk whose box does not hold
     * the node adds nothing (+0, which leaves a sum begun at +0 unchanged) */
    for (uint32_t mask = nc <= 32 ? tmask[lt] : 0u; mask != 0u;) {
      float vx[4], vy[4], vz[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        vx[u] = vy[u] = vz[u] = 0.f;
        if (mask == 0u) continue;
        const int q = __builtin_ctz(mask);
        mask &= mask - 1u;
        const int4 d = cinfo[q][1];
        const int4 l = cinfo[q][0];
        const int px = ux - l.x, py = uy - l.y, pz = uz - l.z;
        if ((unsigned)px >= (unsigned)d.x || (unsigned)py >= (unsigned)d.y ||
            (unsigned)pz >= (unsigned)d.z)
          continue;
        const float *src = a.m.chunk_sum + (int64_t)(c0 + q) * (6 * IB_BOX_CELLS);
        const int p = px + d.x * (py + d.y * pz);
        vx[u] = src[p];
        vy[u] = src[d.w + p];
        vz[u] = src[2 * d.w + p];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        fx += vx[u];
        fy += vy[u];
        fz += vz[u];
      }
    }
    for (int q0 = 0; q0 < (nc > 32 ? nc : 0); q0 += 8) {
      float vx[8], vy[8], vz[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        vx[u] = vy[u] = vz[u] = 0.f;
        const int q = q0 + u;
        if (q >= nc) continue;
        const int4 d = cinfo[q][1];
        const int4 l = cinfo[q][0];
        const int px = ux - l.x, py = uy - l.y, pz = uz - l.z;
        if (d.w == 0 || (unsigned)px >= (unsigned)d.x || (unsigned)py >= (unsigned)d.y ||
            (unsigned)pz >= (unsigned)d.z)
          continue;
        const float *src = a.m.chunk_sum + (int64_t)(c0 + q) * (6 * IB_BOX_CELLS);
        const int p = px + d.x * (py + d.y * pz);
        vx[u] = src[p];
        vy[u] = src[d.w + p];
        vz[u] = src[2 * d.w + p];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        fx += vx[u];
        fy += vy[u];
        fz += vz[u];
      }
    }
    if (fx == 0.f && fy == 0.f && fz == 0.f) continue;
    const int gx = axis_index(ux, L.nx, L.periodic[0]);
    const int gy = axis_index(uy, L.ny, L.periodic[1]);
    const int gz = axis_index(uz, L.nz, L.periodic[2]);
    if (gx < 0 || gy < 0 || gz < 0) continue;
    const int64_t c = gx + (int64_t)L.nx * (gy + (int64_t)L.ny * gz);
    if (owned[lt]) {
      if (fx != 0.f) L.aux[c] = fx;
      if (fy != 0.f) L.aux[a.plane + c] = fy;
      if (fz != 0.f) L.aux[2 * a.plane + c] = fz;
    } else {
      if (fx != 0.f) atomicAdd(L.aux + c, fx);
      if (fy != 0.f) atomicAdd(L.aux + a.plane + c, fy);
      if (fz != 0.f) atomicAdd(L.aux + 2 * a.plane + c, fz);
    }
  }
}

/* Tile-gather spreading.  A first pass, over the markers in cell-list
 * order, writes each IB-coupled marker's record (first stencil node, the
 * 3 x 4 kernel weights and the force) into its cell-list slot and stamps the
 * 4x4x4 tiles its 4x4x4 stencil meets with the step, listing every tile
 * once (the lane that changes a stamp lists the tile).  Then a wavefront
 * per listed tile finds the markers whose stencils can reach it on the cell
 * lists (the bins of the cells within two of the tile), stages 64 of their
 * records in LDS at a time and adds them, one marker after the other, to
 * its copy of the tile in LDS, lane l taking stencil node (l & 3,
 * (l >> 2) & 3, l >> 4) (a marker's 64 nodes are distinct, and a
 * wavefront's LDS adds complete in issue order), then writes the tile with
 * plain stores.  No device-wide atomics, and the sums follow the cell-list
 * order, so the field is the same from run to run.
 * Each contribution is the direct form's w_x w_y w_z s F bit for bit.
 * Needs every periodic axis at least twice the tile's extent along it (a
 * node then lies in a tile in one periodic image only). */

/* Distinct tiles (at most two) of the stencil nodes o..o+3 along one axis;
 * nodes outside a wall-bounded axis are dropped. */
__device__ inline int stencil_tiles(int o, int n, int periodic, int shift, int t[2]) {
  int cnt = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int i = axis_index(o + k, n, periodic);
    if (i < 0) continue;
    const int tt = i >> shift;
    if (cnt == 0 || (tt != t[cnt - 1] && (cnt < 2 || tt != t[0]))) {
      if (cnt < 2) t[cnt++] = tt;
    }
  }
  return cnt;
}

__global__ __launch_bounds__(256) void spread_prep_kernel(MArgs a) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int tx[2] = {0, 0}, ty[2] = {0, 0}, tz[2] = {0, 0};
  int ntx = 0, nty = 0, ntz = 0;
  int64_t tile0 = 0;
  int tnx = 1, tny = 1;
  if (s < a.m.n && !halted(a)) {
    const int i = a.m.bin_items[s];
    SpreadRec r;
    r.o = make_int4(0, 0, 0, 0);
    if (ib_coupled(a.m.kind[i], a.fiber_move)) {
      const MLat L = lat_of(a, i);
      const float W[3] = {wrap_coord(a.m.x[i], L.nx, L.periodic[0]),
                          wrap_coord(a.m.y[i], L.ny, L.periodic[1]),
                          wrap_coord(a.m.z[i], L.nz, L.periodic[2])};
      const int ox = (int)floorf(W[0]) - 1, oy = (int)floorf(W[1]) - 1,
                oz = (int)floorf(W[2]) - 1;
      r.o = make_int4(ox, oy, oz, 1);
      float w4[4];
      stencil_weights(W[0], w4);
      r.wx = make_float4(w4[0], w4[1], w4[2], w4[3]);
      stencil_weights(W[1], w4);
      r.wy = make_float4(w4[0], w4[1], w4[2], w4[3]);
      stencil_weights(W[2], w4);
      r.wz = make_float4(w4[0], w4[1], w4[2], w4[3]);
      r.f = make_float4(a.m.fx[i], a.m.fy[i], a.m.fz[i], 0.f);
      ntx = stencil_tiles(ox, L.nx, L.periodic[0], TILE_SHIFT_X, tx);
      nty = stencil_tiles(oy, L.ny, L.periodic[1], TILE_SHIFT_Y, ty);
      ntz = stencil_tiles(oz, L.nz, L.periodic[2], TILE_SHIFT_Z, tz);
      tile0 = L.tile0;
      tnx = L.tnx;
      tny = L.tny;
    }
    a.m.spread_rec[s] = r;
  }
  /* the wavefront's distinct tiles, one per lane (neighbouring slots hold
   * neighbouring markers, so a wavefront meets few), then each stamped and
   * listed by its lane */
  const int32_t step = (int32_t)*a.m.d_step;
  int32_t *stamps = a.lats ? a.lats[0].tile_stamp : a.lat.tile_stamp; /* all lattices' tiles */
  int64_t mine = -1;
  int nd = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int q = c & 1, b = (c >> 1) & 1, z = c >> 2;
    int64_t t = (q < ntx && b < nty && z < ntz)
                    ? tile0 + tx[q] + (int64_t)tnx * (ty[b] + (int64_t)tny * tz[z])
                    : -1;
    for (;;) {
      const uint64_t todo = __ballot(t >= 0);
      if (todo == 0) break;
      const int leader = __ffsll((unsigned long long)todo) - 1;
      const int64_t g = __shfl(t, leader, 64);
      if (t == g) t = -1;
      if (__ballot(mine == g) == 0) { /* not collected yet */
        if (lane == nd) mine = g;
        if (++nd == 64) { /* full: stamp these before collecting more */
          if (atomicExch(&stamps[mine], step) != step)
            a.m.tile_list[atomicAdd(&a.m.tile_count[0], 1)] = (int32_t)mine;
          mine = -1;
          nd = 0;
        }
      }
    }
  }
  if (mine >= 0 && atomicExch(&stamps[mine], step) != step)
    a.m.tile_list[atomicAdd(&a.m.tile_count[0], 1)] = (int32_t)mine;
}

/* Cell ranges [c0, c1] (at most two, wrapped into [0, n) on a periodic axis,
 * clamped on a wall-bounded one, where markers outside sit in the edge
 * cells' bins) -> bin ranges. */
__device__ inline int bin_ranges(int c0, int c1, int n, int periodic, int nb, int lo[2],
                                 int hi[2]) {
  int r = 0;
  if (!periodic) {
    c0 = max(c0, 0);
    c1 = min(c1, n - 1);
    if (c0 > c1) return 0;
    lo[0] = c0 * nb / n;
    hi[0] = c1 * nb / n;
    return 1;
  }
  if (c1 - c0 + 1 >= n) {
    lo[0] = 0;
    hi[0] = nb - 1;
    return 1;
  }
  int a0 = c0 % n;
  if (a0 < 0) a0 += n;
  const int a1 = a0 + (c1 - c0);
  if (a1 < n) {
    lo[r] = a0 * nb / n;
    hi[r++] = a1 * nb / n;
  } else {
    lo[r] = a0 * nb / n;
    hi[r++] = nb - 1;
    lo[r] = 0;
    hi[r++] = (a1 - n) * nb / n;
  }
  return r;
}

/* Offset of a marker's first stencil node from the tile origin t0 along one
 * axis, in the periodic image nearest to the tile. */
__device__ inline int tile_offset(int o, int t0, int n, int periodic) {
  int r = o - t0;
  if (periodic) {
    if (r > n / 2) r -= n;
    else if (r < -(n / 2)) r += n;
  }
  return r;
}

__device__ inline float pick4(const float4 &v, int k) {
  return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}

constexpr int SPREAD_TILE_SEGS = 128; /* bin-row segments per tile: <= 7 x 7 x 2 */
/* a wavefront's copy of the tile in LDS: node (x, y, z) at x + 36 y + 144 z,
 * so the 32 nodes one half-wavefront adds (4 x 4 x 2 of a stencil) fall in
 * 32 distinct banks */
constexpr int ACC_SY = 36, ACC_SZ = 144, ACC_N = 560;
static_assert(TILE_X <= 16 && TILE_Y <= 4 && TILE_Z <= 4 &&
                  (TILE_X - 1) + ACC_SY * (TILE_Y - 1) + ACC_SZ * (TILE_Z - 1) < ACC_N,
              "the padded tile fits its copy");

/* One wavefront per listed tile (a grid of wavefronts taking the list in
 * turn). */
__global__ __launch_bounds__(256) void spread_tiles_kernel(MArgs a) {
  __shared__ float sacc[4][3 * ACC_N];
  __shared__ SpreadRec sstage[4][64];
  __shared__ int sfirst[4][SPREAD_TILE_SEGS + 1]; /* candidates before segment s */
  __shared__ int sslot[4][SPREAD_TILE_SEGS];      /* first cell-list slot of segment s */
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  float *acc = sacc[wave];
  SpreadRec *stage = sstage[wave];
  int *seg_first = sfirst[wave], *seg_slot = sslot[wave];
  const int kx = lane & 3, ky = (lane >> 2) & 3, kz = lane >> 4;
  const int count = a.m.tile_count[0];
  const int nwaves = (int)gridDim.x * 4;
  const int32_t *bstart = a.m.bin_start;
  for (int k = (int)blockIdx.x * 4 + wave; k < count; k += nwaves) {
    const int64_t g = a.m.tile_list[k];
    const MLat L = a.lats ? a.lats[lat_of_tile(a, g)] : a.lat;
    const int64_t tile = g - L.tile0;
    const int x0 = (int)(tile % L.tnx) * TILE_X;
    const int y0 = (int)((tile / L.tnx) % L.tny) * TILE_Y;
    const int z0 = (int)(tile / ((int64_t)L.tnx * L.tny)) * TILE_Z;
    for (int q = lane; q < 3 * ACC_N; q += 64) acc[q] = 0.f;
    /* the markers whose cell c has a stencil node c-1..c+2 in the tile:
     * c in [t0 - 2, t0 + extent] along each axis */
    int xl[2], xh[2], yl[2], yh[2], zl[2], zh[2];
    const int ex = min(TILE_X, L.nx - x0), ey = min(TILE_Y, L.ny - y0),
              ez = min(TILE_Z, L.nz - z0);
    const int nxr = bin_ranges(x0 - 2, x0 + ex, L.nx, L.periodic[0], L.nb[0], xl, xh);
    const int nyr = bin_ranges(y0 - 2, y0 + ey, L.ny, L.periodic[1], L.nb[1], yl, yh);
    const int nzr = bin_ranges(z0 - 2, z0 + ez, L.nz, L.periodic[2], L.nb[2], zl, zh);
    /* the bin-row segments (x-bin range of one (y, z) bin row), two per
     * lane, and their candidate counts scanned into seg_first */
    const int ly0 = yh[0] - yl[0] + 1, lz0 = zh[0] - zl[0] + 1;
    const int nyb = ly0 + (nyr > 1 ? yh[1] - yl[1] + 1 : 0);
    const int nzb = lz0 + (nzr > 1 ? zh[1] - zl[1] + 1 : 0);
    const int ns = min(nxr * nyb * nzb, SPREAD_TILE_SEGS);
    int carry = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int sg = lane + 64 * h;
      int cnt = 0, s0 = 0;
      if (sg < ns) {
        const int rx = sg % nxr, iy = (sg / nxr) % nyb, iz = sg / (nxr * nyb);
        const int by = iy < ly0 ? yl[0] + iy : yl[1] + (iy - ly0);
        const int bz = iz < lz0 ? zl[0] + iz : zl[1] + (iz - lz0);
        const int row = L.bin0 + L.nb[0] * (by + L.nb[1] * bz);
        s0 = bstart[row + xl[rx]];
        cnt = bstart[row + xh[rx] + 1] - s0;
      }
      int v = cnt; /* inclusive scan across the wavefront */
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
      }
      if (sg < ns) {
        seg_first[sg + 1] = carry + v;
        seg_slot[sg] = s0;
      }
      carry += __shfl(v, 63, 64);
    }
    if (lane == 0) seg_first[0] = 0;
    __builtin_amdgcn_wave_barrier();
    const int total = carry;
    for (int base = 0; base < total; base += 64) {
      const int c = base + lane;
      if (c < total) {
        int lo = 0, hi = ns - 1; /* the segment holding candidate c */
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (seg_first[mid] <= c) lo = mid;
          else hi = mid - 1;
        }
        stage[lane] = a.m.spread_rec[seg_slot[lo] + (c - seg_first[lo])];
      }
      __builtin_amdgcn_wave_barrier(); /* the records, in issue order */
      const int n = min(64, total - base);
      for (int j = 0; j < n; ++j) {
        const SpreadRec r = stage[j];
        if (!r.o.w) continue; /* not IB-coupled */
        const int px = tile_offset(r.o.x, x0, L.nx, L.periodic[0]) + kx;
        const int py = tile_offset(r.o.y, y0, L.ny, L.periodic[1]) + ky;
        const int pz = tile_offset(r.o.z, z0, L.nz, L.periodic[2]) + kz;
        if (px < 0 || px >= ex || py < 0 || py >= ey || pz < 0 || pz >= ez) continue;
        const float w = pick4(r.wx, kx) * pick4(r.wy, ky) * pick4(r.wz, kz) * a.force_scale;
        const int node = px + ACC_SY * py + ACC_SZ * pz;
        /* LDS adds of one wavefront complete in issue order: the sums follow
         * the candidates' (cell-list) order */
        atomicAdd(&acc[node], w * r.f.x);
        atomicAdd(&acc[ACC_N + node], w * r.f.y);
        atomicAdd(&acc[2 * ACC_N + node], w * r.f.z);
      }
      __builtin_amdgcn_wave_barrier(); /* read before the next group overwrites */
    }
    for (int q = lane; q < TILE_CELLS; q += 64) {
      const int lx = q & (TILE_X - 1), ly = (q >> TILE_SHIFT_X) & (TILE_Y - 1),
                lz = q >> (TILE_SHIFT_X + TILE_SHIFT_Y);
      if (lx < ex && ly < ey && lz < ez) {
        const int64_t cell = (x0 + lx) + (int64_t)L.nx * ((y0 + ly) + (int64_t)L.ny * (z0 + lz));
        const int node = lx + ACC_SY * ly + ACC_SZ * lz;
        L.aux[cell] = acc[node];
        L.aux[a.plane + cell] = acc[ACC_N + node];
        L.aux[2 * a.plane + cell] = acc[2 * ACC_N + node];
      }
    }
    __builtin_amdgcn_wave_barrier(); /* the copy is read before it is cleared again */
  }
  /* the last wavefront to finish empties the list for the next spread */
  if (lane == 0) {
    __threadfence();
    if (atomicAdd(&a.m.tile_count[1], 1) == nwaves - 1) {
      a.m.tile_count[0] = 0;
      a.m.tile_count[1] = 0;
      __threadfence();
    }
  }
}

/* Interpolated velocity of marker i by a group of 16 lanes: lane j takes
 * node (x, y) = (j & 3, j >> 2) of the four z-layers of the stencil, so one
 * load instruction of the wavefront covers 4 markers x 4 rows of 4
 * consecutive nodes; summed over the group (nodes outside a wall-bounded
 * axis contribute nothing). */
__device__ void interp_marker(const MArgs &a, int i, int j, float *v) {
  const MLat L = lat_of(a, i);
  const float X = wrap_coord(a.m.x[i], L.nx, L.periodic[0]);
  const float Y = wrap_coord(a.m.y[i], L.ny, L.periodic[1]);
  const float Z = wrap_coord(a.m.z[i], L.nz, L.periodic[2]);
  float wx, wy, ux = 0.f, uy = 0.f, uz = 0.f;
  const int ix = stencil_node(X, j & 3, L.nx, L.periodic[0], &wx);
  const int iy = stencil_node(Y, j >> 2, L.ny, L.periodic[1], &wy);
  if (ix >= 0 && iy >= 0) {
    const int64_t col = ix + (int64_t)L.nx * iy;
    const int64_t layer = (int64_t)L.nx * L.ny;
    const float wxy = wx * wy;
    float wz4[4];
    stencil_weights(Z, wz4);
    const int oz = (int)floorf(Z) - 1;
#pragma unroll
    for (int kz = 0; kz < 4; ++kz) {
      const int iz = axis_index(oz + kz, L.nz, L.periodic[2]);
      if (iz < 0) continue;
      const float w = wxy * wz4[kz];
      const int64_t c = col + layer * iz;
      const float4 v = L.vel[c];
      ux += w * v.x;
      uy += w * v.y;
      uz += w * v.z;
    }
  }
  ux = group_sum<16>(ux);
  uy = group_sum<16>(uy);
  uz = group_sum<16>(uz);
  v[0] = ux;
  v[1] = uy;
  v[2] = uz;
}

/* Langevin leapfrog of a platelet element (the BBK half kicks of verlet_1 /
 * verlet_3, sem_kernel.cu.orig:251-266, 547-560): two exact
 * Ornstein-Uhlenbeck half steps with this step's force, then the drift
 * with the weighted fluid velocity u.  The velocity is kept at half steps. */
__device__ void bbk_element(const MArgs &a, int i, float ux, float uy, float uz) {
  const float dt = a.marker_dt, g = a.sem_gamma, m = a.sem_mass;
  const float vdt = expf(-0.5f * g * dt);
  const float fdt = g > 0.f ? (1.0f - vdt) / g : 0.5f * dt;
  const float fm = fdt / m;
  float v[3] = {a.m.vx[i], a.m.vy[i], a.m.vz[i]};
  const float F[3] = {a.m.fx[i], a.m.fy[i], a.m.fz[i]};
  double za[3] = {0.0, 0.0, 0.0}, zb[3] = {0.0, 0.0, 0.0};
  float sig = 0.f;
  if (a.sem_temp > 0.f && g > 0.f) {
    sig = (float)sqrt((double)a.sem_temp / m * (1.0 - exp(-(double)g * dt)));
    const int64_t step = a.m.counters[6];
    const MLat L = lat_of(a, i); /* the lattice's own key */
    pfg_normal3(L.seed, step, (uint32_t)(i - L.marker0), PFG_RNG_BBK_A, za);
    pfg_normal3(L.seed, step, (uint32_t)(i - L.marker0), PFG_RNG_BBK_B, zb);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    v[k] = fmaf(vdt, v[k], fm * F[k]) + sig * (float)za[k];
    v[k] = fmaf(vdt, v[k], fm * F[k]) + sig * (float)zb[k];
  }
  const float w = a.sem_w, wv = 1.0f - w;
  a.m.vx[i] = v[0];
  a.m.vy[i] = v[1];
  a.m.vz[i] = v[2];
  a.m.x[i] += dt * (wv * v[0] + w * ux);
  a.m.y[i] += dt * (wv * v[1] + w * uy);
  a.m.z[i] += dt * (wv * v[2] + w * uz);
}

__device__ inline bool bbk_kind(const MArgs &a, int kind) {
  return kind == PFG_MARKER_ELEMENT && a.sem_integrator == PFG_SEM_BBK;
}

__device__ void update_marker(const MArgs &a, int i, float vx, float vy, float vz) {
  const int kind = a.m.kind[i];
  if (kind == PFG_MARKER_FIXED) return;
  if (bbk_kind(a, kind)) {
    bbk_element(a, i, vx, vy, vz);
    return;
  }
  const float dt = a.marker_dt;
  if (kind == PFG_MARKER_FIBER && a.fiber_move == PFG_FIBER_KICK) {
    /* v <- e^{-g dt} v + F (1 - e^{-g dt})/g */
    const float vdt = expf(-a.fiber_gamma * dt);
    const float fdt = a.fiber_gamma > 0.f ? (1.0f - vdt) / a.fiber_gamma : dt;
    vx = vdt * a.m.vx[i] + a.m.fx[i] * fdt;
    vy = vdt * a.m.vy[i] + a.m.fy[i] * fdt;
    vz = vdt * a.m.vz[i] + a.m.fz[i] * fdt;
    if (a.kick_temp > 0.f) {
      /* thermal noise of the exact Ornstein-Uhlenbeck step, unit mass (the
       * BBK form of sem_kernel.cu.orig:251-266), drawn for this step (the
       * step index the first launch of the step recorded) */
      const double g = a.fiber_gamma, sig = sqrt((double)a.kick_temp *
                                                 (1.0 - exp(-2.0 * g * (double)dt)));
      double z[3];
      const MLat L = lat_of(a, i); /* the lattice's own key */
      pfg_normal3(L.seed, a.m.counters[6], (uint32_t)(i - L.marker0), PFG_RNG_KICK, z);
      vx += (float)(sig * z[0]);
      vy += (float)(sig * z[1]);
      vz += (float)(sig * z[2]);
    }
    a.m.vx[i] = vx;
    a.m.vy[i] = vy;
    a.m.vz[i] = vz;
  }
  a.m.x[i] += vx * dt;
  a.m.y[i] += vy * dt;
  a.m.z[i] += vz * dt;
}

/* After the interpolation of marker i by its group of 16 lanes (v, ROLE_A):
 * store the velocity, then (ROLE_B) move the marker; lane 0 of the group
 * writes. */
__device__ inline void finish_marker(const MArgs &a, int i, int lane, int kind, float v[3]) {
  const bool bbk = bbk_kind(a, kind);
  if (a.roles & ROLE_A) {
    if (lane == 0 && bbk) { /* the element keeps its own velocity */
      a.m.uib[i] = make_float4(v[0], v[1], v[2], 0.f);
    } else if (lane == 0 && (ib_coupled(kind, a.fiber_move) || kind == PFG_MARKER_FIXED)) {
      a.m.vx[i] = v[0];
      a.m.vy[i] = v[1];
      a.m.vz[i] = v[2];
    }
  } else if (lane == 0) {
    if (bbk) {
      const float4 u = a.m.uib[i];
      v[0] = u.x;
      v[1] = u.y;
      v[2] = u.z;
    } else {
      v[0] = a.m.vx[i];
      v[1] = a.m.vy[i];
      v[2] = a.m.vz[i];
    }
  }
  if ((a.roles & ROLE_B) && lane == 0) update_marker(a, i, v[0], v[1], v[2]);
}

/* roles: A = interpolate V at the markers, B = move the markers and advance
 * the step counter.  Four markers per wavefront, 16 lanes each. */
__global__ __launch_bounds__(256) void move_kernel(MArgs a) {
  const int k = (blockIdx.x * 256 + threadIdx.x) >> 4;
  const int lane = threadIdx.x & 15;
  if (halted(a)) return;
  /* every marker, advancing the step counter, or the markers of a list
   * (the chunk launch of the same step advances it) */
  if ((a.roles & ROLE_B) && a.move_count < 0 && blockIdx.x == 0 && threadIdx.x == 0)
    *a.m.d_step += 1;
  if (k >= (a.move_count < 0 ? a.m.n : a.move_count)) return;
  const int i = a.move_count < 0 ? k : a.move_list[k];
  const int kind = a.m.kind[i];
  float v[3] = {0.f, 0.f, 0.f};
  if ((a.roles & ROLE_A) && ib_coupled(kind, a.fiber_move)) interp_marker(a, i, lane, v);
  finish_marker(a, i, lane, kind, v);
}

/* Interpolation of marker i by its group of 16 lanes from its staged first
 * stencil node o and kernel weights w (x 0..3, y 4..7, z 8..11), reading
 * the field: the nodes and weights of interp_marker, summed in its order. */
__device__ inline void interp_staged(const MLat &L, int4 o, const float *w, int j, float v[3]) {
  const int jx = j & 3, jy = j >> 2;
  const int ix = axis_index(o.x + jx, L.nx, L.periodic[0]);
  const int iy = axis_index(o.y + jy, L.ny, L.periodic[1]);
  float ux = 0.f, uy = 0.f, uz = 0.f;
  if (ix >= 0 && iy >= 0) {
    const int64_t col = ix + (int64_t)L.nx * iy;
    const int64_t layer = (int64_t)L.nx * L.ny;
    const float wxy = w[jx] * w[4 + jy];
#pragma unroll
    for (int kz = 0; kz < 4; ++kz) {
      const int iz = axis_index(o.z + kz, L.nz, L.periodic[2]);
      if (iz < 0) continue;
      const float ww = wxy * w[8 + kz];
      const float4 u = L.vel[col + layer * iz];
      ux += ww * u.x;
      uy += ww * u.y;
      uz += ww * u.z;
    }
  }
  ux = group_sum<16>(ux);
  uy = group_sum<16>(uy);
  uz = group_sum<16>(uz);
  v[0] = ux;
  v[1] = uy;
  v[2] = uz;
}

/* Chunk-staged interpolation (ROLE_A, always) and move (ROLE_B): a
 * workgroup of MC_THREADS lanes per IB chunk, then one per IB_CHUNK markers
 * listed for direct spreading (the first IB_CHUNK lanes stage the markers;
 * every lane loads the velocity box and interpolates, a marker per group of
 * 16 lanes).  Every global read of a marker (index, kind,
 * position) is made once, by a lane per marker, into LDS together with the
 * marker's first stencil node and kernel weights (the staged spreading's),
 * so that the loop over the markers reads only LDS and writes.  When a
 * chunk's stencils span at most IB_BOX_CELLS nodes, the velocity of that
 * box of nodes is read once into LDS, row by row (nodes outside a
 * wall-bounded axis read as zero), and each group of 16 lanes interpolates
 * one marker from it: lane j takes column (j & 3, j >> 2) over the four
 * z-layers and the group sums as interp_marker sums, so the velocities
 * equal the per-marker form's bit for bit while each node is read from the
 * field once per chunk.  Scattered chunks and the listed markers read their
 * stencils from the field with the staged weights.  Without the Langevin
 * leapfrog and the fibrin kick, the move x += v dt uses the staged
 * position; otherwise update_marker reads what it needs. */
constexpr int MC_THREADS = 256;
static_assert(IB_CHUNK <= MC_THREADS && IB_CHUNK <= 128, "the markers fill the first two wavefronts");
template <bool FAST>
__global__ __launch_bounds__(MC_THREADS, 4) void move_chunk_kernel(MArgs a) {
  __shared__ int4 so[IB_CHUNK];      /* first stencil node (x = INT_MIN: not coupled);
                                        .w: its box node */
  __shared__ float sw[IB_CHUNK][12]; /* weights along x (0..3), y (4..7), z (8..11) */
  __shared__ float4 sp[IB_CHUNK];    /* position; .w: marker index (int bits) */
  __shared__ int skind[IB_CHUNK];
  __shared__ float4 box[IB_BOX_CELLS];
  __shared__ int sh[MC_THREADS / 64][6];
  if (halted(a)) return;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int grp = t >> 4, j16 = t & 15;
  if ((a.roles & ROLE_B) && blockIdx.x == 0 && t == 0) *a.m.d_step += 1;
  const bool chunk = (int)blockIdx.x < a.m.n_chunks;
  int c0, cnt;
  if (chunk) {
    c0 = a.m.chunk_start[blockIdx.x];
    cnt = a.m.chunk_start[blockIdx.x + 1] - c0;
  } else { /* IB_CHUNK markers of the direct list (of any lattices) */
    const int k0 = IB_CHUNK * ((int)blockIdx.x - a.m.n_chunks);
    c0 = a.m.chunk_start[a.m.n_chunks] + k0;
    cnt = min(IB_CHUNK, a.m.n_direct - k0);
  }
  if (cnt <= 0) return;
  const int *items = a.m.chunk_items + c0;
  const int m0 = chunk ? a.m.chunk_ref[blockIdx.x] : items[0];
  const MLat L0 = lat_of(a, m0); /* a chunk never spans two lattices */
  float ref[3];
  ref[0] = wrap_coord(a.m.x[m0], L0.nx, L0.periodic[0]);
  ref[1] = wrap_coord(a.m.y[m0], L0.ny, L0.periodic[1]);
  ref[2] = wrap_coord(a.m.z[m0], L0.nz, L0.periodic[2]);
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  if (t < cnt) {
    const int i = items[t];
    const int kind = a.m.kind[i];
    const float v[3] = {a.m.x[i], a.m.y[i], a.m.z[i]};
    sp[t] = make_float4(v[0], v[1], v[2], __int_as_float(i));
    skind[t] = kind;
    const bool has = ib_coupled(kind, a.fiber_move);
    int o[3] = {0, 0, 0};
    if (has) {
      const MLat L = chunk ? L0 : lat_of(a, i);
      float W[3];
      if (chunk) {
        stencil_origin(L, v, ref, W, o);
      } else { /* its own image */
        float own[3];
        own[0] = wrap_coord(v[0], L.nx, L.periodic[0]);
        own[1] = wrap_coord(v[1], L.ny, L.periodic[1]);
        own[2] = wrap_coord(v[2], L.nz, L.periodic[2]);
        stencil_origin(L, v, own, W, o);
      }
      stencil_weights(W[0], &sw[t][0]);
      stencil_weights(W[1], &sw[t][4]);
      stencil_weights(W[2], &sw[t][8]);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        lo[k] = o[k];
        hi[k] = o[k] + 3;
      }
    }
    so[t] = make_int4(has ? o[0] : INT_MIN, o[1], o[2], -1);
  }
#pragma unroll
  for (int s2 = 32; s2 > 0; s2 >>= 1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[k] = min(lo[k], __shfl_xor(lo[k], s2, 64));
      hi[k] = max(hi[k], __shfl_xor(hi[k], s2, 64));
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      sh[wave][k] = lo[k];
      sh[wave][3 + k] = hi[k];
    }
  }
  __syncthreads(); /* the staged markers and both halves' bounds */
#pragma unroll
  for (int k = 0; k < 3; ++k) { /* the markers are in the first two wavefronts */
    lo[k] = min(sh[0][k], sh[1][k]);
    hi[k] = max(sh[0][3 + k], sh[1][3 + k]);
  }
  const int64_t dim0 = (int64_t)hi[0] - lo[0] + 1, dim1 = (int64_t)hi[1] - lo[1] + 1,
                dim2 = (int64_t)hi[2] - lo[2] + 1;
  /* a chunk with IB-coupled markers (lo <= hi) whose box fits LDS */
  const bool boxed = chunk && lo[0] <= hi[0] && dim0 * dim1 * dim2 <= IB_BOX_CELLS;
  if (boxed) {
    const int d0 = (int)dim0, d1 = (int)dim1, vol = (int)(dim0 * dim1 * dim2);
    const float inv0 = 1.0f / (float)d0, inv1 = 1.0f / (float)d1; /* vol <= IB_BOX_CELLS */
    for (int k = t; k < vol; k += MC_THREADS) {
      const int r = div_small(k, inv0), bx = k - r * d0;
      const int bz = div_small(r, inv1), by = r - bz * d1;
      const int gx = axis_index(lo[0] + bx, L0.nx, L0.periodic[0]);
      const int gy = axis_index(lo[1] + by, L0.ny, L0.periodic[1]);
      const int gz = axis_index(lo[2] + bz, L0.nz, L0.periodic[2]);
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (gx >= 0 && gy >= 0 && gz >= 0) v = L0.vel[gx + (int64_t)L0.nx * (gy + (int64_t)L0.ny * gz)];
      box[k] = v;
    }
    if (t < cnt && so[t].x != INT_MIN) {
      const int4 oj = so[t];
      so[t].w = (oj.x - lo[0]) + d0 * ((oj.y - lo[1]) + d1 * (oj.z - lo[2]));
    }
    __syncthreads(); /* the box and the box nodes */
  }
  /* FAST: launched without the Langevin leapfrog and the fibrin kick */
  const int jx = j16 & 3, jy = j16 >> 2;
  const int col = boxed ? jx + (int)dim0 * jy : 0, layer = boxed ? (int)(dim0 * dim1) : 0;
  for (int j = grp; j < cnt; j += MC_THREADS / 16) {
    const float4 p = sp[j];
    const int i = __float_as_int(p.w);
    const int kind = skind[j];
    const bool coupled = ib_coupled(kind, a.fiber_move);
    float v[3] = {0.f, 0.f, 0.f};
    if (coupled) {
      if (boxed) {
        const int base = so[j].w + col;
        const float wxy = sw[j][jx] * sw[j][4 + jy];
        float ux = 0.f, uy = 0.f, uz = 0.f;
#pragma unroll
        for (int kz = 0; kz < 4; ++kz) {
          const float w = wxy * sw[j][8 + kz];
          const float4 u = box[base + layer * kz];
          ux += w * u.x;
          uy += w * u.y;
          uz += w * u.z;
        }
        ux = group_sum<16>(ux);
        uy = group_sum<16>(uy);
        uz = group_sum<16>(uz);
        v[0] = ux;
        v[1] = uy;
        v[2] = uz;
      } else {
        interp_staged(chunk ? L0 : lat_of(a, i), so[j], sw[j], j16, v);
      }
    }
    if constexpr (!FAST) {
      finish_marker(a, i, j16, kind, v);
    } else if (j16 == 0) { /* finish_marker's stores, update_marker's Euler move */
      if (coupled || kind == PFG_MARKER_FIXED) {
        a.m.vx[i] = v[0];
        a.m.vy[i] = v[1];
        a.m.vz[i] = v[2];
      }
      if ((a.roles & ROLE_B) && kind != PFG_MARKER_FIXED) {
        float x = p.x, y = p.y, z = p.z;
        const float dt = a.marker_dt;
        x += v[0] * dt;
        y += v[1] * dt;
        z += v[2] * dt;
        a.m.x[i] = x;
        a.m.y[i] = y;
        a.m.z[i] = z;
      }
    }
  }
}

/* ---------------- launchers ---------------- */

void fill_mlat(const pfg_handle *h, int k, MLat *L) {
  const Inst &I = h->inst[(size_t)k];
  L->aux = h->aux + I.cell0;
  L->vel = h->vel + I.cell0;
  L->tile_stamp = h->tile_stamp + I.tile0;
  L->tile_owner = h->m.tile_owner ? h->m.tile_owner + I.tile0 : nullptr;
  L->tnx = I.tnx;
  L->tny = I.tny;
  L->tnz = I.tnz;
  L->nx = I.p.nx;
  L->ny = I.p.ny;
  L->nz = I.p.nz;
  for (int ax = 0; ax < 3; ++ax) {
    L->periodic[ax] = I.p.wall[ax] == PFG_WALL_PERIODIC;
    L->nb[ax] = I.nb[ax];
    L->nbc[ax] = I.nbc[ax];
  }
  L->bin0 = (int)I.bin0;
  L->binc0 = (int)I.binc0;
  L->tile0 = I.tile0;
  L->nxb = I.nxb;
  L->row_stamp = h->row_stamp ? h->row_stamp + I.fblk0 : nullptr;
  L->seed = I.p.seed;
  L->rec0 = I.rec0;
  L->marker0 = I.marker0;
}

static void make_margs(const pfg_handle *h, MArgs *a) {
  const pfg_params &p = h->p;
  a->m = h->m;
  fill_mlat(h, 0, &a->lat);
  a->lats = h->K > 1 ? h->d_lats : nullptr;
  a->tile0s = h->K > 1 ? h->d_tile0 : nullptr;
  a->K = h->K;
  a->ntiles = h->ntiles;
  a->morse_u0 = p.morse_u0;
  a->morse_rho = p.morse_rho;
  a->morse_req = p.morse_req;
  a->morse_cut2 = p.morse_cutoff * p.morse_cutoff;
  a->morse_cut = p.morse_cutoff;
  a->morse_taper = p.morse_taper;
  a->morse_eps2 = 2.0f * p.morse_cutoff * MORSE_EPS_R;
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
  a->force_scale = p.ib_force_scale;
  a->plane = h->plane;
  a->kbend = p.kbend;
  a->contact_a = p.contact_a;
  a->contact_kappa = p.contact_kappa;
  a->contact_hc = p.contact_hc;
  a->kick_temp = p.kick_temp;
  a->sem_integrator = p.sem_integrator;
  a->sem_mass = p.sem_mass;
  a->sem_gamma = p.sem_gamma;
  a->sem_temp = p.sem_temp;
  a->sem_w = p.sem_fluid_weight;
  a->bond_law = p.bond_law;
  a->bell_kT = p.bell_kT;
  a->roles = 0;
  a->halt = h->halt_on_detach ? 1 : 0;
  a->mark_rows = h->overlap && h->row_stamp != nullptr ? 1 : 0;
  a->resolve_last = 0;
  a->count_spans = 0;
  a->spread_scratch = 0;
  a->bin_elements = bin_elements(h) ? 1 : 0;
  a->force_list = nullptr;
  a->force_count = -1;
  a->move_list = nullptr;
  a->move_count = -1;
  a->blocks0 = a->blocks1 = a->blocks2 = a->blocks3 = 0;
}

static inline int blocks_for(int64_t n, int per_block) {
  return (int)((n + per_block - 1) / per_block);
}

/* Few markers: the counting launch adds the span totals itself. */
static bool count_spans(const pfg_handle *h) { return h->m.n < SPAN_COUNT_MAX; }

hipError_t launch_begin(pfg_handle *h, bool clear, bool count) {
  MArgs a;
  make_margs(h, &a);
  const bool recs = count && h->m.n_rec > 0;
  count = count && h->m.n > 0;
  a.count_spans = count && count_spans(h);
  a.roles = (clear ? ROLE_A : 0) | (count ? ROLE_B : 0) | (recs ? ROLE_C : 0);
  a.blocks0 = clear ? blocks_for(h->ntiles, 256) : 0;
  a.blocks1 = count ? blocks_for(h->m.n, 256) : 0;
  a.blocks2 = recs ? blocks_for(h->m.n_rec, 256) : 0;
  const int grid = a.blocks0 + a.blocks1 + a.blocks2;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(begin_kernel, dim3(grid), dim3(256), 0, h->stream, a);
  return hipGetLastError();
}

hipError_t launch_scan(pfg_handle *h) {
  const Markers &m = h->m;
  if (m.n == 0) return hipSuccess;
  const bool full = bin_elements(h); /* else the coarse grid */
  const int64_t blocks = full ? m.scan_blocks : m.scan_blocks_c;
  if (!count_spans(h))
    hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0, h->stream,
                       m.bin_count, m.scan_partials);
  hipLaunchKernelGGL(scan_apply_kernel, dim3((unsigned)blocks), dim3(256), 0, h->stream,
                     m.bin_count, full ? m.nbins : m.nbinsc, m.scan_partials, m.bin_start);
  return hipGetLastError();
}

hipError_t launch_place(pfg_handle *h, bool scatter, bool geometry) {
  MArgs a;
  make_margs(h, &a);
  scatter = scatter && h->m.n > 0;
  geometry = geometry && h->m.n_tri > 0;
  a.roles = (scatter ? ROLE_A : 0) | (geometry ? ROLE_B : 0);
  a.blocks0 = scatter ? blocks_for(h->m.n, 256) : 0;
  a.blocks1 = geometry ? blocks_for(h->m.n_tri, 256) : 0;
  const int grid = a.blocks0 + a.blocks1;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(place_kernel, dim3(grid), dim3(256), 0, h->stream, a);
  if (geometry && h->m.bend) { /* node normals, then the mean curvature */
    hipLaunchKernelGGL(bend_normal_kernel,
                       dim3(blocks_for((int64_t)h->m.n + h->m.n_tedges, 256)), dim3(256), 0,
                       h->stream, a);
    hipLaunchKernelGGL(bend_km_kernel, dim3(blocks_for(h->m.n, 256)), dim3(256), 0,
                       h->stream, a);
  }
  return hipGetLastError();
}

hipError_t launch_gather(pfg_handle *h, bool sort, bool sums, bool decide, bool *resolved) {
  MArgs a;
  make_margs(h, &a);
  sort = sort && h->m.n > 0;
  sums = sums && h->m.n_plt > 0;
  decide = decide && h->m.n_rec > 0;
  a.roles = (sort ? ROLE_A : 0) | (sums ? ROLE_B : 0) | (decide ? ROLE_C : 0);
  /* the bonds of a few receptors are resolved by the last workgroup */
  a.resolve_last = resolved && decide && h->m.n_rec <= RESOLVE_IN_GATHER_MAX;
  if (resolved) *resolved = a.resolve_last != 0;
  a.blocks0 = sort ? blocks_for(h->m.n, 256) : 0;
  a.blocks1 = sums ? h->m.n_plt : 0;
  a.blocks2 = decide ? blocks_for(h->m.n_rec, REC_PER_BLOCK) : 0;
  const int grid = a.blocks0 + a.blocks1 + a.blocks2;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(gather_kernel, dim3(grid), dim3(256), 0, h->stream, a);
  return hipGetLastError();
}

hipError_t launch_resolve(pfg_handle *h) {
  if (h->m.n_rec == 0) return hipSuccess;
  MArgs a;
  make_margs(h, &a);
  hipLaunchKernelGGL(bond_resolve_kernel, dim3(blocks_for(h->m.n_rec, 256)), dim3(256), 0,
                     h->stream, a);
  return hipGetLastError();
}

/* spreading form: 1 staged chunks, 2 direct, 3 tile gather, 4 by body */
static int spread_form(const pfg_handle *h) {
  if (h->ib_staging == 3 && !h->m.tile_gather_ok) return 1;
  if (h->ib_staging == 4 && h->m.n_bodies == 0) return 1;
  if (h->ib_staging) return h->ib_staging;
  return h->m.n >= IB_STAGED_MIN ? (h->m.n_bodies > 0 ? 4 : 1) : 2;
}

static bool staged_spreading(const pfg_handle *h) { return spread_form(h) != 2; }

static int forces_lanes(const pfg_handle *h) {
  /* few markers: a wavefront each, for parallelism; many: 4 lanes each in
   * bin order, for fewer instructions per marker */
  return h->forces_lanes ? h->forces_lanes : (h->m.n >= FORCES_SORTED_MIN ? 4 : 64);
}

/* the force pass by body: automatic with many markers, when the model has
 * such bodies */
static bool forces_by_body(const pfg_handle *h) {
  if (h->m.n_fbodies == 0 || h->forces_body == 2) return false;
  return h->forces_body == 1 || h->m.n >= FORCES_SORTED_MIN;
}

static bool contact_on(const pfg_handle *h) {
  return h->m.contact_slot && h->p.contact_a != 0.f && h->m.n_plt >= 2;
}

/* Whether the cell lists hold the elements: not when the step's force pass
 * takes every element's Morse partners from its platelet's own lists (the
 * pass by body, which the step runs unless the force pass spreads too) and
 * no contact search or tile gather reads the lists (then only the fibrin
 * ligand search does). */
bool bin_elements(const pfg_handle *h) {
  const bool body_pass = forces_by_body(h) && !forces_can_spread(h);
  return !(body_pass && h->m.fb_all_elements && !contact_on(h) && spread_form(h) != 3);
}

/* The force pass can spread each marker's force itself when it has a
 * wavefront per marker, spreading is direct, no later pass adds to the
 * forces (bending, contact) and no detach halt can stop the step midway. */
bool forces_can_spread(const pfg_handle *h) {
  return h->m.n > 0 && forces_lanes(h) == 64 && !staged_spreading(h) && !h->m.bend &&
         !contact_on(h) && !h->halt_on_detach;
}

hipError_t launch_forces(pfg_handle *h, bool spread) {
  if (h->m.n == 0) return hipSuccess;
  MArgs a;
  make_margs(h, &a);
  const int G = forces_lanes(h);
  const bool by_body = !spread && forces_by_body(h);
  if (by_body) { /* the per-marker pass over the markers outside those bodies */
    a.force_list = h->m.frest_items;
    a.force_count = h->m.n_frest;
  }
  /* at least one workgroup: the first one makes the detach test */
  unsigned grid = (unsigned)std::max(
      1, blocks_for((int64_t)(by_body ? h->m.n_frest : h->m.n) * G, 256));
  if (!spread && spread_form(h) == 4) {
    /* the body spreading's tile claims, in workgroups after the markers' */
    a.blocks0 = (int)grid;
    a.blocks1 = h->m.n_bodies;
    a.blocks2 = blocks_for(h->m.n_ritems, 256);
    grid += (unsigned)(a.blocks1 + a.blocks2);
  }
  if (spread) {
    if (!forces_can_spread(h)) return hipErrorInvalidValue;
    hipLaunchKernelGGL((forces_kernel<64, true>), dim3(grid), dim3(256), 0, h->stream, a);
    return hipGetLastError();
  }
  switch (G) {
    case 1: hipLaunchKernelGGL(forces_kernel<1>, dim3(grid), dim3(256), 0, h->stream, a); break;
    case 4: hipLaunchKernelGGL(forces_kernel<4>, dim3(grid), dim3(256), 0, h->stream, a); break;
    case 2: hipLaunchKernelGGL(forces_kernel<2>, dim3(grid), dim3(256), 0, h->stream, a); break;
    case 8: hipLaunchKernelGGL(forces_kernel<8>, dim3(grid), dim3(256), 0, h->stream, a); break;
    case 16: hipLaunchKernelGGL(forces_kernel<16>, dim3(grid), dim3(256), 0, h->stream, a); break;
    case 32: hipLaunchKernelGGL(forces_kernel<32>, dim3(grid), dim3(256), 0, h->stream, a); break;
    default: hipLaunchKernelGGL(forces_kernel<64>, dim3(grid), dim3(256), 0, h->stream, a); break;
  }
  if (by_body) { /* lists rebuilt where needed, then the list pass */
    hipLaunchKernelGGL(forces_body_build_kernel, dim3(h->m.n_fbodies), dim3(BODY_THREADS), 0,
                       h->stream, a);
    hipLaunchKernelGGL(forces_body_kernel, dim3(h->m.n_fbodies), dim3(256), 0, h->stream, a);
  }
  if (h->m.bend)
    hipLaunchKernelGGL(bend_force_kernel, dim3(blocks_for((int64_t)h->m.n * 4, 256)),
                       dim3(256), 0, h->stream, a);
  if (contact_on(h)) {
    hipLaunchKernelGGL(contact_pairs_kernel, dim3(blocks_for(h->m.n, 256)), dim3(256), 0,
                       h->stream, a);
    hipLaunchKernelGGL(contact_force_kernel, dim3(64), dim3(256), 0, h->stream, a);
  }
  return hipGetLastError();
}

hipError_t launch_spread(pfg_handle *h) {
  if (h->m.n == 0) return hipSuccess;
  MArgs a;
  make_margs(h, &a);
  /* staged through LDS when many markers share the lattice */
  const int form = spread_form(h);
  if (form == 2) {
    hipLaunchKernelGGL(spread_kernel, dim3(blocks_for(h->m.n, 4)), dim3(256), 0, h->stream, a);
    return hipGetLastError();
  }
  if (form == 3) {
    /* marker records, stamps and the tile list, then a workgroup per listed
     * tile (a grid of SPREAD_TILE_GRID workgroups taking the list in turn) */
    hipLaunchKernelGGL(spread_prep_kernel, dim3(blocks_for(h->m.n, 256)), dim3(256), 0,
                       h->stream, a);
    const int64_t grid = std::min<int64_t>(SPREAD_TILE_GRID, std::max<int64_t>(h->ntiles, 1));
    hipLaunchKernelGGL(spread_tiles_kernel, dim3((unsigned)grid), dim3(256), 0, h->stream, a);
    return hipGetLastError();
  }
  if (form == 4) {
    /* the chunk launch (chunks of bodies to their scratch boxes, every other
     * marker as in form 1), then a workgroup per body adds its boxes */
    a.spread_scratch = 1;
    hipLaunchKernelGGL(spread_chunk_kernel,
                       dim3(h->m.n_chunks + (h->m.n_direct + IB_CHUNK - 1) / IB_CHUNK),
                       dim3(128), 0, h->stream, a);
    hipLaunchKernelGGL(spread_body_kernel, dim3(h->m.n_bodies), dim3(256), 0, h->stream, a);
    return hipGetLastError();
  }
  /* one workgroup per IB_CHUNK markers listed for direct spreading, then
   * one per chunk, in a single launch */
  hipLaunchKernelGGL(spread_chunk_kernel,
                     dim3(h->m.n_chunks + (h->m.n_direct + IB_CHUNK - 1) / IB_CHUNK),
                     dim3(128), 0, h->stream, a);
  return hipGetLastError();
}

hipError_t launch_move(pfg_handle *h, bool interp, bool update) {
  if (h->m.n == 0) return hipSuccess;
  MArgs a;
  make_margs(h, &a);
  a.roles = (interp ? ROLE_A : 0) | (update ? ROLE_B : 0);
  if (interp && staged_spreading(h) && h->m.n_chunks > 0) {
    /* interpolation staged by chunk, as the spreading is (this launch
     * advances the step counter); the markers listed for direct spreading,
     * scattered, take a group of lanes each in a launch of their own, where
     * the many groups hide their loads' latency */
    const dim3 grid(h->m.n_chunks);
    if (h->p.sem_integrator != PFG_SEM_BBK && h->p.fiber_move != PFG_FIBER_KICK)
      hipLaunchKernelGGL(move_chunk_kernel<true>, grid, dim3(MC_THREADS), 0, h->stream, a);
    else
      hipLaunchKernelGGL(move_chunk_kernel<false>, grid, dim3(MC_THREADS), 0, h->stream, a);
    if (h->m.n_direct > 0) {
      a.move_list = h->m.chunk_items + h->m.n_chunk_items;
      a.move_count = h->m.n_direct;
      hipLaunchKernelGGL(move_kernel, dim3(blocks_for(h->m.n_direct, 16)), dim3(256), 0,
                         h->stream, a);
    }
    return hipGetLastError();
  }
  hipLaunchKernelGGL(move_kernel, dim3(blocks_for(h->m.n, 16)), dim3(256), 0, h->stream, a);
  return hipGetLastError();
}

}  // namespace pfg
