/*
 * lb.hip -- fused D3Q19 pull-stream + bounce-back + moments + BGK/Guo collide.
 *
 * One pass per step replaces the reference's four full-grid passes
 * (fluid3d_collision_kernel LB_kernel.cu:1203-1304, fluid3d_stream_kernel
 * LB_kernel.cu:1312-1359, fluid3d_velocity_density_kernel LB_kernel.cu:1108-1199
 * and the wall kernels LB_kernel.cu:395-780): each cell pulls its 19
 * post-collision populations from the upstream neighbours of the previous step
 * (bounce-back where the upstream cell is a wall or a solid cell), takes its
 * moments with the current force, relaxes and writes its 19 new populations.
 * Each population is read once and written once: 152 B per cell update in
 * FP32 storage, 76 B in 16-bit storage.
 *
 * Storage holds deviations f_i - w_i (FP32), (f_i - w_i) * 2^s (FP16) or
 * f_i - w_i (BF16).  Arithmetic is FP32 in deviation form throughout
 * (rho - 1 = sum of deviations), so FP32 rounding acts on the deviations and
 * not on the O(1) weights; FP16 values stay in the scaled space and only the
 * four moments are unscaled.
 *
 * Layout (blocked SoA): the cells of a lattice row are grouped in x-blocks of
 * 64 (one wavefront).  Block b = (x/64) + nxb*(y + ny*z) stores its 19
 * directions as 19 consecutive chunks of 64 values:
 *     element(d, x, y, z) = (b * 19 + d) * 64 + x % 64.
 * A wavefront updates one x-block: its 19 stores fill one contiguous
 * 19*64-value segment, and each of its loads comes from one chunk of an
 * upstream block (one of nine rows), so DRAM pages are reused across the
 * directions instead of streaming 38 separate planes.  The chunk base is
 * wavefront-uniform; the lane offset is one of three values per lane.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <vector>

#include "pfg_internal.h"

namespace pfg {

/* ---------------- storage formats ---------------- */

/* load: the stored value as float (FP16 values stay scaled by 2^s);
 * put: store an already scaled value. */
template <int S>
struct Store;

template <>
struct Store<PFG_STORE_F32> {
  using T = float;
  static constexpr bool scaled = false;
  __device__ static inline float load(const T *p) { return *p; }
  __device__ static inline void put(T *p, float v) { __builtin_nontemporal_store(v, p); }
};

template <>
struct Store<PFG_STORE_F16> {
  using T = uint16_t;
  static constexpr bool scaled = true;
  __device__ static inline float load(const T *p) {
    return (float)__builtin_bit_cast(_Float16, *p);
  }
  __device__ static inline void put(T *p, float v) {
    const _Float16 hv = (_Float16)v; /* round to nearest even */
    __builtin_nontemporal_store(__builtin_bit_cast(uint16_t, hv), p);
  }
};

template <>
struct Store<PFG_STORE_BF16> {
  using T = uint16_t;
  static constexpr bool scaled = false;
  __device__ static inline float load(const T *p) {
    return __builtin_bit_cast(float, ((uint32_t)*p) << 16);
  }
  __device__ static inline void put(T *p, float v) {
    uint32_t u = __builtin_bit_cast(uint32_t, v);
    u += 0x7FFFu + ((u >> 16) & 1u); /* round to nearest even */
    __builtin_nontemporal_store((uint16_t)(u >> 16), p);
  }
};

__device__ __host__ inline int64_t felem(int nxb, int ny, int d, int x, int y, int z) {
  const int64_t b = (int64_t)(x >> 6) + (int64_t)nxb * (y + (int64_t)ny * z);
  return (b * Q + d) * XBLOCK + (x & (XBLOCK - 1));
}

/* ---------------- the fused kernel ---------------- */

/* STEP: every row, with the IB force in flagged tiles.  FIELDS: rho and u
 * only.  BULK / REGION: the rows whose stamp is not / is the current step
 * (rows that this step's IB stencils touch), without / with the IB force;
 * one z-row per wavefront. */
enum { LB_MODE_STEP = 0, LB_MODE_FIELDS = 1, LB_MODE_BULK = 2, LB_MODE_REGION = 3 };
/* Flag bit of MODE: the lattice has velocity faces on z (compiled only into
 * the variant that needs it, which keeps the scalar registers and the
 * occupancy of the default kernels). */
enum { LB_ZVEL = 4 };

/* Velocity face z = 0 (s = +1) or z = nz-1 (s = -1) on the pulled
 * deviations (fluid3d_in_out_flow_boundary_kernel,
 * LB_kernel_old_08102011.cu:586-1090): face velocity = wall velocity x the
 * duct profile 4 q (L - q) / L^2, q = i + 1/2, over each bounded transverse
 * axis; a cell next to a transverse wall takes feq(1, u); the others
 * Zou/He on the populations entering from outside,
 *   rho = (1 + sum_{cz=0} dev + 2 sum_{cz=-s} dev) / (1 - s uz),
 *   dev_{s z} = dev_{-s z} + s rho uz / 3,
 *   dev_{s z, t} = dev_{opp} + rho (s uz + c_t u_t) / 6 - c_t N_t,
 *   N_t = 1/2 sum_{cz=0} c_t dev - rho u_t / 3
 * (the weights cancel between opposite directions, so deviations take the
 * place of populations except in rho).  dev in storage units. */
template <class ST>
__device__ __forceinline__ void velocity_face(const LBArgs &a, int x, int y, int s,
                                              float dev[Q]) {
  const float *U = s > 0 ? a.zvel_u[0] : a.zvel_u[1];
  float shape = 1.0f;
  if (!a.periodic[0]) {
    const float q = (float)x + 0.5f;
    shape *= 4.0f * q * ((float)a.nx - q) / ((float)a.nx * (float)a.nx);
  }
  if (!a.periodic[1]) {
    const float q = (float)y + 0.5f;
    shape *= 4.0f * q * ((float)a.ny - q) / ((float)a.ny * (float)a.ny);
  }
  const float ux = U[0] * shape, uy = U[1] * shape, uz = U[2] * shape;
  const float sc = ST::scaled ? a.scale : 1.0f, isc = ST::scaled ? a.inv_scale : 1.0f;
  if ((!a.periodic[0] && (x == 0 || x == a.nx - 1)) ||
      (!a.periodic[1] && (y == 0 || y == a.ny - 1))) {
    const float usq = ux * ux + uy * uy + uz * uz;
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      const float cu = 3.0f * (CX[d] * ux + CY[d] * uy + CZ[d] * uz);
      dev[d] = sc * weight(d) * (cu + 0.5f * cu * cu - 1.5f * usq);
    }
    return;
  }
  float S0 = 0.f, Sm = 0.f, Nx = 0.f, Ny = 0.f;
#pragma unroll
  for (int d = 0; d < Q; ++d) {
    if (CZ[d] == 0) {
      S0 += dev[d];
      Nx += CX[d] * dev[d];
      Ny += CY[d] * dev[d];
    } else if (CZ[d] == -s) {
      Sm += dev[d];
    }
  }
  const float rho = (1.0f + (S0 + 2.0f * Sm) * isc) / (1.0f - (float)s * uz);
  Nx = 0.5f * Nx * isc - rho * ux / 3.0f;
  Ny = 0.5f * Ny * isc - rho * uy / 3.0f;
#pragma unroll
  for (int d = 0; d < Q; ++d) {
    if (CZ[d] != s) continue;
    float t;
    if (CX[d] != 0) t = rho * ((float)s * uz + CX[d] * ux) / 6.0f - CX[d] * Nx;
    else if (CY[d] != 0) t = rho * ((float)s * uz + CY[d] * uy) / 6.0f - CY[d] * Ny;
    else t = (float)s * rho * uz / 3.0f;
    dev[d] = dev[BB[d]] + sc * t;
  }
}

/* One wavefront per (x-block, y, CPT consecutive z); workgroup `blk` of the
 * lattice described by `a` (in memory that the whole workgroup reads), the
 * wavefront's place in it `wave` (default: its place in the launch's
 * workgroup). */
template <int S, int CPT, int MODE>
__device__ __forceinline__ void lb_update(const LBArgs &a, uint32_t blk, float *orho,
                                          float *oux, float *ouy, float *ouz, int wave = -1) {
  constexpr int M = MODE & 3;
  constexpr bool ZV = (MODE & LB_ZVEL) != 0;
  using ST = Store<S>;
  using T = typename ST::T;
  constexpr int R = Q * XBLOCK; /* values per x-block */
  const T *__restrict__ fin = (const T *)a.fin;
  T *__restrict__ fout = (T *)a.fout;
  if (M != LB_MODE_FIELDS && a.halt != nullptr && *a.halt != 0) return;
  const int lane = threadIdx.x & (XBLOCK - 1);
  /* wavefront-uniform task decomposition in 32-bit arithmetic (x-blocks per
   * lattice stay below 2^31) */
  const uint32_t task =
      blk * 4u + __builtin_amdgcn_readfirstlane(wave < 0 ? (int)(threadIdx.x >> 6) : wave);
  const uint32_t rows = (uint32_t)a.nxb * (uint32_t)a.ny;
  const uint32_t zc = task / rows;
  if (zc * CPT >= (uint32_t)a.nz) return;
  /* IB force tiles and rows are stamped with the step that touches them */
  const int32_t step32 = M == LB_MODE_FIELDS ? 0 : (int32_t)*a.step;
  if (M == LB_MODE_BULK || M == LB_MODE_REGION) {
    static_assert(M < LB_MODE_BULK || CPT == 1, "row modes update one z-row per wave");
    const bool touched = a.row_stamp[task] == step32; /* task = x-block row */
    if (touched != (M == LB_MODE_REGION)) return;
  }
  const uint32_t rem = task - zc * rows;
  const int y = (int)(rem / (uint32_t)a.nxb);
  const int xb = (int)(rem - (uint32_t)y * a.nxb);
  const int x = xb * XBLOCK + lane;
  if (x >= a.nx) return;

  /* lane offsets from the upstream chunk base (placed one block early, so all
   * offsets are >= 0): centre, upstream x-1, upstream x+1.  Lanes at the ends
   * of the row take their x-neighbour in the edge pass below. */
  const bool xlo = x == 0, xhi = x == a.nx - 1;
  const uint32_t o_c = R + lane;
  const uint32_t o_m = xlo ? o_c : (lane > 0 ? R + lane - 1 : XBLOCK - 1);
  const uint32_t o_p = xhi ? o_c : (lane < XBLOCK - 1 ? R + lane + 1 : 2 * R);

  /* upstream rows, wavefront-uniform */
  const int y0 = y == 0, y1 = y == a.ny - 1;
  const int ym = y0 ? (a.periodic[1] ? a.ny - 1 : y) : y - 1;
  const int yp = y1 ? (a.periodic[1] ? 0 : y) : y + 1;
  const bool oym = y0 && !a.periodic[1], oyp = y1 && !a.periodic[1];

  const float isc = ST::scaled ? a.inv_scale : 1.0f;

#pragma unroll
  for (int c = 0; c < CPT; ++c) {
    const int z = zc * CPT + c;
    if (z >= a.nz) return;
    const int64_t idx = x + (int64_t)a.nx * (y + (int64_t)a.ny * z); /* cell index */
    uint32_t mask = 0;
    if (a.bbmask != nullptr) {
      mask = a.bbmask[idx];
      if (mask >> 31) continue; /* solid cell */
    }
    const int z0 = z == 0, z1 = z == a.nz - 1;
    const int zm = z0 ? (a.periodic[2] ? a.nz - 1 : z) : z - 1;
    const int zp = z1 ? (a.periodic[2] ? 0 : z) : z + 1;
    const bool ozm = z0 && !a.periodic[2], ozp = z1 && !a.periodic[2];
    const uint64_t bown = (uint32_t)xb + (uint32_t)a.nxb * (uint32_t)(y + a.ny * z);

    /* the IB force of the tiles this step's stencils meet, read (and
     * consumed: the next step's spreading adds into zeros) before the
     * populations, so that its latency overlaps theirs */
    bool flagged = false;
    float ax = 0.f, ay = 0.f, az = 0.f;
    if (M == LB_MODE_STEP || M == LB_MODE_REGION) {
      const int tile = (x >> TILE_SHIFT_X) + a.tnx * ((y >> TILE_SHIFT_Y) + a.tny * (z >> TILE_SHIFT_Z));
      flagged = a.tile_stamp[tile] == step32;
      if (flagged) {
        ax = a.aux[idx];
        ay = a.aux[a.plane + idx];
        az = a.aux[2 * a.plane + idx];
        a.aux[idx] = 0.f;
        a.aux[a.plane + idx] = 0.f;
        a.aux[2 * a.plane + idx] = 0.f;
      }
    }

    float dev[Q];
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      const int sy = CY[d] > 0 ? ym : (CY[d] < 0 ? yp : y);
      const int sz = CZ[d] > 0 ? zm : (CZ[d] < 0 ? zp : z);
      const uint64_t bsrc = (uint32_t)xb + (uint32_t)a.nxb * (uint32_t)(sy + a.ny * sz);
      const T *base = fin + ((int64_t)(bsrc * Q + d) * XBLOCK - R);
      const uint32_t o = CX[d] > 0 ? o_m : (CX[d] < 0 ? o_p : o_c);
      dev[d] = ST::load(base + o);
    }
    /* edge pass: rows next to a wall (wavefront-uniform), row ends, solid
     * neighbours.  Upstream cell outside a wall or solid: halfway bounce-back,
     * the population leaving this cell returns plus the moving-wall momentum
     * 6 w (c . u_wall) (z wall over y wall over x wall); otherwise the
     * periodic x-neighbour. */
    const bool edge = xlo || xhi || oym || oyp || ozm || ozp || mask != 0;
    if (edge) {
#pragma unroll
      for (int d = 1; d < Q; ++d) {
        const bool oz = CZ[d] > 0 ? ozm : (CZ[d] < 0 ? ozp : false);
        const bool oy = CY[d] > 0 ? oym : (CY[d] < 0 ? oyp : false);
        const bool xe = CX[d] > 0 ? xlo : (CX[d] < 0 ? xhi : false);
        const bool ox = xe && !a.periodic[0];
        if (oz || oy || ox || ((mask >> d) & 1u)) {
          const int sxs = CX[d] > 0 ? 0 : 1, sys = CY[d] > 0 ? 0 : 1, szs = CZ[d] > 0 ? 0 : 1;
          const float term = oz ? a.wall_term[2][szs][d]
                                : (oy ? a.wall_term[1][sys][d] : (ox ? a.wall_term[0][sxs][d] : 0.f));
          dev[d] = ST::load(fin + (bown * Q + BB[d]) * XBLOCK + lane) + term;
        } else if (xe) { /* periodic wrap along x */
          const int sx = CX[d] > 0 ? a.nx - 1 : 0;
          const int sy = CY[d] > 0 ? ym : (CY[d] < 0 ? yp : y);
          const int sz = CZ[d] > 0 ? zm : (CZ[d] < 0 ? zp : z);
          dev[d] = ST::load(fin + felem(a.nxb, a.ny, d, sx, sy, sz));
        }
      }
    }
    if (ZV && (z0 || z1)) velocity_face<ST>(a, x, y, z0 ? 1 : -1, dev);

    /* moments: rho - 1 and momentum from deviations (sum_i w_i c_i = 0),
     * over the nine pairs (d, opposite of d): the pair sum feeds the density,
     * the pair difference the momentum.  -0.0f starts are exact no-ops. */
    float drho = dev[0], jx = -0.0f, jy = -0.0f, jz = -0.0f;
#pragma unroll
    for (int d = 1; d < Q; ++d) {
      if (d < BB[d]) {
        const float s = dev[d] + dev[BB[d]], df = dev[d] - dev[BB[d]];
        drho += s;
        if (CX[d] > 0) jx += df; else if (CX[d] < 0) jx -= df;
        if (CY[d] > 0) jy += df; else if (CY[d] < 0) jy -= df;
        if (CZ[d] > 0) jz += df; else if (CZ[d] < 0) jz -= df;
      }
    }
    if (ST::scaled) {
      drho *= isc;
      jx *= isc;
      jy *= isc;
      jz *= isc;
    }
    const float rho = 1.0f + drho;
    float Fx = a.g[0], Fy = a.g[1], Fz = a.g[2];
    if (flagged) {
      Fx += ax;
      Fy += ay;
      Fz += az;
    }
    const float inv_rho = __builtin_amdgcn_rcpf(rho); /* hardware reciprocal, 1 ulp */
    const float ux = (jx + 0.5f * Fx) * inv_rho;
    const float uy = (jy + 0.5f * Fy) * inv_rho;
    const float uz = (jz + 0.5f * Fz) * inv_rho;

    if (M == LB_MODE_FIELDS) {
      if (orho) orho[idx] = rho;
      if (oux) oux[idx] = ux;
      if (ouy) ouy[idx] = uy;
      if (ouz) ouz[idx] = uz;
      continue;
    }
    if (flagged) { /* velocity for the IB interpolation of this step */
      a.vel[idx] = make_float4(ux, uy, uz, 0.f);
    }

    /* BGK with Guo forcing, in deviation form and storage units:
     *   out_d = (1-omega) dev_d + omega (feq_d - w_d) + guo w_d (3 (c_d-u).F + 9 (c_d.u)(c_d.F))
     * split into the part even in c_d,
     *   E = w_d [K0 + cu (4.5 omega rho cu + 9 guo cF)],
     *   K0 = omega (drho - 1.5 rho u.u) - 3 guo u.F,
     * and the odd part O = w_d (3 omega rho cu + 3 guo cF); the pair
     * (d, opposite) takes E + O and E - O. */
    const float usq = ux * ux + uy * uy + uz * uz;
    const float uF = ux * Fx + uy * Fy + uz * Fz;
    const float K0 = fmaf(-a.guo3, uF, a.omega * fmaf(-1.5f * rho, usq, drho));
    const float Ab = a.om45 * rho, Cp = a.om3 * rho;
    T *own = fout + bown * R + lane;
    ST::put(own, fmaf(a.keep, dev[0], a.ws[0] * K0));
#pragma unroll
    for (int d = 1; d < Q; ++d) {
      if (d < BB[d]) {
        const int k = d < 7 ? 1 : 2;
        float cu = -0.0f, cF = -0.0f;
        if (CX[d] > 0) { cu += ux; cF += Fx; } else if (CX[d] < 0) { cu -= ux; cF -= Fx; }
        if (CY[d] > 0) { cu += uy; cF += Fy; } else if (CY[d] < 0) { cu -= uy; cF -= Fy; }
        if (CZ[d] > 0) { cu += uz; cF += Fz; } else if (CZ[d] < 0) { cu -= uz; cF -= Fz; }
        const float t = fmaf(a.gB[k], cF, (a.ws[k] * Ab) * cu);
        const float E = fmaf(cu, t, a.ws[k] * K0);
        const float O = fmaf(a.gD[k], cF, (a.ws[k] * Cp) * cu);
        ST::put(own + d * XBLOCK, fmaf(a.keep, dev[d], E + O));
        ST::put(own + BB[d] * XBLOCK, fmaf(a.keep, dev[BB[d]], E - O));
      }
    }
  }
}

template <int S, int CPT, int MODE>
__global__ __launch_bounds__(256) void lb_fused_kernel(LBArgs a, float *orho, float *oux,
                                                       float *ouy, float *ouz) {
  /* Workgroups are dealt round-robin to the 8 XCDs.  With xcd_rows each
   * XCD takes a contiguous range of rows instead, so that neighbouring rows
   * (which read each other's populations) share its L2; the autotuner picks
   * the order (faster on large lattices, slower on small ones). */
  uint32_t blk = blockIdx.x;
  if (a.xcd_rows) {
    const uint32_t q = gridDim.x / 8u;
    if (blk < 8u * q) blk = (blk % 8u) * q + blk / 8u;
  }
  lb_update<S, CPT, MODE>(a, blk, orho, oux, ouy, ouz);
}

/* A batch of lattices in one launch: workgroup -> lattice through a table,
 * each lattice's arguments read from device memory.  xcd_rows: each XCD
 * takes a contiguous range of the table (as in the single-lattice kernel). */
template <int S, int CPT, int MODE>
__global__ __launch_bounds__(256) void lb_batch_kernel(const LBArgs *args,
                                                       const int32_t *block_inst,
                                                       int xcd_rows) {
  uint32_t blk = blockIdx.x;
  if (xcd_rows) {
    const uint32_t q = gridDim.x / 8u;
    if (blk < 8u * q) blk = (blk % 8u) * q + blk / 8u;
  }
  const LBArgs &a = args[block_inst[blk]];
  lb_update<S, CPT, MODE>(a, (uint32_t)(blk - a.block0), nullptr, nullptr, nullptr, nullptr);
}

/* A batch of lattices in one persistent launch: a grid of resident
 * workgroups whose wavefronts take the table's x-block rows (of any
 * lattice; row r is wavefront r % 4 of workgroup task r / 4) from a device
 * work queue until it is empty.  The rows are cut into LB_QUEUES contiguous
 * ranges, one per XCD (the hardware deals workgroup w to XCD w % 8), each
 * with its own counter, so neighbouring rows stay in one L2 and the counters
 * are not one hot spot; a grab takes LB_QUEUE_GRAB consecutive rows, and a
 * wavefront whose range is used up moves on to the next range.  Wavefronts
 * run on their own (no barrier).  queue[LB_QUEUES] counts the wavefronts
 * that are done; the last one resets every counter for the next launch (all
 * others have made their last grab by then), so a captured step needs no
 * clearing node. */
template <int S, int CPT, int MODE>
__global__ __launch_bounds__(256) void lb_batch_queue_kernel(const LBArgs *args,
                                                             const int32_t *block_inst,
                                                             uint32_t ntasks, uint32_t *queue) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t nrows = ntasks * 4u;
  const uint32_t r0 = blockIdx.x % LB_QUEUES;
  for (uint32_t k = 0; k < LB_QUEUES; ++k) {
    const uint32_t r = (r0 + k) % LB_QUEUES;
    const uint32_t lo = (uint32_t)((uint64_t)nrows * r / LB_QUEUES);
    const uint32_t hi = (uint32_t)((uint64_t)nrows * (r + 1) / LB_QUEUES);
    for (;;) {
      uint32_t got = 0;
      if (lane == 0) got = atomicAdd(&queue[r], (uint32_t)LB_QUEUE_GRAB);
      got = __builtin_amdgcn_readfirstlane(__shfl(got, 0, 64));
      if (got >= hi - lo) break;
      const uint32_t t0 = lo + got;
      const uint32_t t1 = t0 + LB_QUEUE_GRAB < hi ? t0 + LB_QUEUE_GRAB : hi;
      for (uint32_t row = t0; row < t1; ++row) {
        const uint32_t blk = row >> 2;
        const LBArgs &a = args[block_inst[blk]];
        lb_update<S, CPT, MODE>(a, (uint32_t)(blk - a.block0), nullptr, nullptr, nullptr,
                                nullptr, (int)(row & 3u));
      }
    }
  }
  if (lane == 0) {
    __threadfence();
    if (atomicAdd(&queue[LB_QUEUES], 1u) == gridDim.x * 4u - 1u) {
      for (int r = 0; r <= LB_QUEUES; ++r) queue[r] = 0u;
      __threadfence();
    }
  }
}

/* ---------------- fields through a fixed device array ---------------- */

/* Staging of pfg_download_fields_noalloc: the fields of a window of whole
 * z-planes, [field][cell of the window].  One copy per device, held until the
 * process ends; calls on one device share it. */
static __device__ __align__(256) float scratch[FIELDS_SCRATCH_FLOATS];

/* The fields pass over the x-block rows [t0, t1) (the rows of z-planes
 * [z0, z1)), one wavefront per row; o* point at cell 0 of the window's first
 * plane minus z0 planes, so that each cell lands at its offset in the window. */
template <int S>
__global__ __launch_bounds__(256) void lb_fields_window_kernel(LBArgs a, uint32_t blk0,
                                                               uint32_t t0, uint32_t t1,
                                                               float *orho, float *oux,
                                                               float *ouy, float *ouz) {
  const uint32_t blk = blk0 + blockIdx.x;
  const uint32_t task = blk * 4u + (threadIdx.x >> 6);
  if (task < t0 || task >= t1) return;
  lb_update<S, 1, LB_MODE_FIELDS>(a, blk, orho, oux, ouy, ouz);
}

__global__ __launch_bounds__(256) void fill_zero_kernel(float *p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

/* Device address of `scratch` per device, looked up once with the device
 * current (the array's name in host code is a host-side stand-in). */
static std::mutex g_scratch_mu;
static float *g_scratch_ptr[FIELDS_SCRATCH_MAX_DEVICES];

hipError_t fields_scratch_setup(int device, float **out) {
  if (device < 0 || device >= FIELDS_SCRATCH_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != device) hipSetDevice(cur);
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = (float *)ptr;
  }
  if (out) *out = g_scratch_ptr[device];
  return hipSuccess;
}

template <int S>
static hipError_t fields_window_s(pfg_handle *h, const LBArgs &a, int z0, int z1,
                                  float *const win[4]) {
  const int64_t plane = (int64_t)a.nx * a.ny;
  const uint32_t rows = (uint32_t)a.nxb * (uint32_t)a.ny;
  const uint32_t t0 = (uint32_t)z0 * rows, t1 = (uint32_t)z1 * rows;
  const uint32_t blk0 = t0 / 4u, blk1 = (t1 + 3u) / 4u;
  float *o[4];
  for (int c = 0; c < 4; ++c) o[c] = win[c] ? win[c] - (int64_t)z0 * plane : nullptr;
  hipLaunchKernelGGL(lb_fields_window_kernel<S>, dim3(blk1 - blk0), dim3(256), 0, h->stream, a,
                     blk0, t0, t1, o[0], o[1], o[2], o[3]);
  return hipGetLastError();
}

hipError_t download_fields_noalloc(pfg_handle *h, float *const host[4]) {
  float *base = nullptr;
  hipError_t e = fields_scratch_setup(h->device, &base);
  if (e != hipSuccess) return e;
  LBArgs a;
  make_lb_args(h, 0, h->cur, &a);
  int nf = 0;
  for (int c = 0; c < 4; ++c) nf += host[c] != nullptr;
  if (nf == 0) return hipSuccess;
  const int64_t plane = (int64_t)a.nx * a.ny;
  const int64_t zmax = FIELDS_SCRATCH_FLOATS / (plane * nf); /* planes per window */
  if (zmax < 1) return hipErrorInvalidValue;
  for (int z0 = 0; z0 < a.nz && e == hipSuccess; z0 += (int)zmax) {
    const int z1 = (int)std::min<int64_t>(a.nz, z0 + zmax);
    const int64_t n = plane * (z1 - z0);
    float *win[4] = {nullptr, nullptr, nullptr, nullptr};
    int slot = 0;
    for (int c = 0; c < 4; ++c)
      if (host[c]) win[c] = base + n * slot++;
    /* solid cells are not written by the pass: they read 0 */
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((n * nf + 255) / 256)), dim3(256), 0,
                       h->stream, base, n * nf);
    e = hipGetLastError();
    if (e != hipSuccess) break;
    switch (h->p.store) {
      case PFG_STORE_F16: e = fields_window_s<PFG_STORE_F16>(h, a, z0, z1, win); break;
      case PFG_STORE_BF16: e = fields_window_s<PFG_STORE_BF16>(h, a, z0, z1, win); break;
      default: e = fields_window_s<PFG_STORE_F32>(h, a, z0, z1, win);
    }
    for (int c = 0; c < 4 && e == hipSuccess; ++c)
      if (host[c])
        e = hipMemcpyAsync(host[c] + (int64_t)z0 * plane, win[c], sizeof(float) * n,
                           hipMemcpyDeviceToHost, h->stream);
    /* the next window reuses the array */
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  }
  return e;
}

/* ---------------- auxiliary kernels ---------------- */

/* Per-cell bounce mask from the solid flags (built once per flag upload). */
__global__ void bbmask_kernel(LBArgs a, const uint8_t *solid, uint32_t *mask) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nxy = (int64_t)a.nx * a.ny;
  if (idx >= nxy * a.nz) return;
  if (solid[idx]) {
    mask[idx] = 1u << 31;
    return;
  }
  const int x = (int)(idx % a.nx), y = (int)((idx / a.nx) % a.ny), z = (int)(idx / nxy);
  uint32_t m = 0;
#pragma unroll
  for (int d = 1; d < Q; ++d) {
    int s[3] = {x - CX[d], y - CY[d], z - CZ[d]};
    const int n[3] = {a.nx, a.ny, a.nz};
    bool outside = false;
    for (int k = 0; k < 3; ++k) {
      if (s[k] < 0) {
        if (a.periodic[k]) s[k] += n[k]; else outside = true;
      } else if (s[k] >= n[k]) {
        if (a.periodic[k]) s[k] -= n[k]; else outside = true;
      }
    }
    if (!outside && solid[s[0] + (int64_t)a.nx * s[1] + nxy * s[2]]) m |= 1u << d;
  }
  mask[idx] = m;
}

__device__ inline void cell_xyz(const LBArgs &a, int64_t idx, int &x, int &y, int &z) {
  x = (int)(idx % a.nx);
  y = (int)((idx / a.nx) % a.ny);
  z = (int)(idx / ((int64_t)a.nx * a.ny));
}

/* per_z: the inputs hold one value per z-plane (a velocity that varies
 * along z only, like the reference's linear shear start) instead of one per
 * cell. */
template <int S>
__global__ void init_eq_kernel(LBArgs a, void *dst0, void *dst1, const float *rho_in,
                               const float *ux_in, const float *uy_in, const float *uz_in,
                               int64_t n, int per_z) {
  using ST = Store<S>;
  using T = typename ST::T;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  int x, y, z;
  cell_xyz(a, idx, x, y, z);
  const int64_t src = per_z ? (int64_t)z : idx;
  const float rho = rho_in ? rho_in[src] : 1.0f;
  const float ux = ux_in ? ux_in[src] : 0.f;
  const float uy = uy_in ? uy_in[src] : 0.f;
  const float uz = uz_in ? uz_in[src] : 0.f;
  const float usq15 = 1.5f * (ux * ux + uy * uy + uz * uz);
  const float drho = rho - 1.0f;
  const float sc = ST::scaled ? a.scale : 1.0f;
#pragma unroll
  for (int d = 0; d < Q; ++d) {
    const float w = weight(d);
    const float cu = CX[d] * ux + CY[d] * uy + CZ[d] * uz;
    const float dev = w * drho + w * rho * (3.0f * cu + 4.5f * cu * cu - usq15);
    const int64_t e = felem(a.nxb, a.ny, d, x, y, z);
    ST::put((T *)dst0 + e, dev * sc);
    ST::put((T *)dst1 + e, dev * sc);
  }
}

/* direction d: to_float -> buf[i] = w_d + dev; else stored <- buf[i] - w_d */
template <int S>
__global__ void convert_plane_kernel(LBArgs a, void *stored, float *buf, int d, int to_float,
                                     int64_t n) {
  using ST = Store<S>;
  using T = typename ST::T;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  int x, y, z;
  cell_xyz(a, idx, x, y, z);
  T *p = (T *)stored + felem(a.nxb, a.ny, d, x, y, z);
  const float w = weight(d);
  if (to_float) {
    buf[idx] = w + ST::load(p) * (ST::scaled ? a.inv_scale : 1.0f);
  } else {
    ST::put(p, (buf[idx] - w) * (ST::scaled ? a.scale : 1.0f));
  }
}

/* FP64 totals of mass and momentum over fluid cells, one partial per block. */
template <int S>
__global__ __launch_bounds__(256) void conservation_kernel(LBArgs a, const void *stored,
                                                          double *partials, int64_t n) {
  using ST = Store<S>;
  using T = typename ST::T;
  const T *p = (const T *)stored;
  const double isc = ST::scaled ? (double)a.inv_scale : 1.0;
  double m = 0, jx = 0, jy = 0, jz = 0;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    if (a.solid != nullptr && a.solid[idx]) continue;
    int x, y, z;
    cell_xyz(a, idx, x, y, z);
    const int64_t e0 = felem(a.nxb, a.ny, 0, x, y, z);
    double cm = 0, cx = 0, cy = 0, cz = 0;
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      const double v = (double)ST::load(p + e0 + d * XBLOCK);
      cm += v;
      cx += CX[d] * v;
      cy += CY[d] * v;
      cz += CZ[d] * v;
    }
    m += 1.0 + cm * isc;
    jx += cx * isc;
    jy += cy * isc;
    jz += cz * isc;
  }
  __shared__ double sh[4][256];
  const int t = threadIdx.x;
  sh[0][t] = m;
  sh[1][t] = jx;
  sh[2][t] = jy;
  sh[3][t] = jz;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      for (int k = 0; k < 4; ++k) sh[k][t] += sh[k][t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    for (int k = 0; k < 4; ++k) partials[4 * blockIdx.x + k] = sh[k][0];
  }
}

/* ---------------- host launchers ---------------- */

/* Arguments of lattice k of the handle, starting from copy `cur` of f. */
void make_lb_args(const pfg_handle *h, int k, int cur, LBArgs *a) {
  const Inst &L = h->inst[(size_t)k];
  const pfg_params &p = L.p;
  const size_t foff = (size_t)L.fblk0 * Q * XBLOCK * h->esize;
  a->fin = (const char *)h->f[cur] + foff;
  a->fout = (char *)h->f[1 - cur] + foff;
  a->aux = h->aux + L.cell0;
  a->plane = h->plane;
  a->vel = h->vel + L.cell0;
  a->tile_stamp = h->tile_stamp + L.tile0;
  a->solid = h->solid ? h->solid + L.cell0 : nullptr;
  a->bbmask = h->bbmask ? h->bbmask + L.cell0 : nullptr;
  a->halt = h->halt_on_detach ? h->m.counters + 1 : nullptr;
  a->row_stamp = h->row_stamp ? h->row_stamp + L.fblk0 : nullptr;
  a->step = h->m.d_step;
  a->nx = p.nx;
  a->ny = p.ny;
  a->nz = p.nz;
  a->nxb = L.nxb;
  a->block0 = L.lb_block0;
  a->tnx = L.tnx;
  a->tny = L.tny;
  for (int ax = 0; ax < 3; ++ax) a->periodic[ax] = p.wall[ax] == PFG_WALL_PERIODIC;
  a->omega = 1.0f / p.tau;
  a->guo = 1.0f - 0.5f / p.tau;
  for (int ax = 0; ax < 3; ++ax) a->g[ax] = p.body_force[ax];
  if (h->p.store == PFG_STORE_F16) {
    a->scale = ldexpf(1.0f, h->p.f16_scale_log2);
    a->inv_scale = ldexpf(1.0f, -h->p.f16_scale_log2);
  } else {
    a->scale = 1.0f;
    a->inv_scale = 1.0f;
  }
  a->keep = 1.0f - a->omega;
  a->om15 = 1.5f * a->omega;
  a->om45 = 4.5f * a->omega;
  a->om3 = 3.0f * a->omega;
  a->guo3 = 3.0f * a->guo;
  const int rep[3] = {0, 1, 7}; /* rest, axis, diagonal */
  for (int c = 0; c < 3; ++c) {
    a->ws[c] = weight(rep[c]) * a->scale;
    a->gB[c] = 9.0f * a->guo * a->ws[c];
    a->gD[c] = 3.0f * a->guo * a->ws[c];
  }
  /* halfway bounce-back with a moving wall (Ladd): the population entering
   * from the wall is f_bb(d) + 6 w_d rho_w (c_d . u_w), rho_w = 1; in the
   * storage units of the format */
  for (int ax = 0; ax < 3; ++ax)
    for (int s = 0; s < 2; ++s)
      for (int d = 0; d < Q; ++d) {
        const float *u = s == 0 ? p.wall_u_lo[ax] : p.wall_u_hi[ax];
        const float cu = CX[d] * u[0] + CY[d] * u[1] + CZ[d] * u[2];
        a->wall_term[ax][s][d] = 6.0f * weight(d) * cu * a->scale;
      }
  a->zvel = p.wall[2] == PFG_WALL_VELOCITY ? 1 : 0;
  a->xcd_rows = h->lb_xcd_rows;
  for (int k = 0; k < 3; ++k) {
    a->zvel_u[0][k] = p.wall_u_lo[2][k];
    a->zvel_u[1][k] = p.wall_u_hi[2][k];
  }
}

template <int S, int CPT, int MODE>
static void launch_fused(const LBArgs &a, dim3 grid, float *rho, float *ux, float *uy,
                         float *uz, hipStream_t st) {
  if (a.zvel)
    hipLaunchKernelGGL((lb_fused_kernel<S, CPT, MODE | LB_ZVEL>), grid, dim3(256), 0, st, a,
                       rho, ux, uy, uz);
  else
    hipLaunchKernelGGL((lb_fused_kernel<S, CPT, MODE>), grid, dim3(256), 0, st, a, rho, ux,
                       uy, uz);
}

template <int S, int CPT>
static hipError_t launch_lb_t(pfg_handle *h, const LBArgs &a, int mode, float *rho,
                              float *ux, float *uy, float *uz, hipStream_t st) {
  const int64_t waves = (int64_t)a.nxb * a.ny * ((a.nz + CPT - 1) / CPT);
  const dim3 grid((unsigned)((waves + 3) / 4));
  if (mode == LB_MODE_STEP)
    launch_fused<S, CPT, LB_MODE_STEP>(a, grid, nullptr, nullptr, nullptr, nullptr, st);
  else if (mode == LB_MODE_FIELDS)
    launch_fused<S, CPT, LB_MODE_FIELDS>(a, grid, rho, ux, uy, uz, st);
  else if (CPT != 1)
    return hipErrorInvalidValue;
  else if (mode == LB_MODE_BULK)
    launch_fused<S, 1, LB_MODE_BULK>(a, grid, nullptr, nullptr, nullptr, nullptr, st);
  else
    launch_fused<S, 1, LB_MODE_REGION>(a, grid, nullptr, nullptr, nullptr, nullptr, st);
  return hipGetLastError();
}

template <int S>
static hipError_t launch_lb_s(pfg_handle *h, const LBArgs &a, int mode, float *rho, float *ux,
                              float *uy, float *uz, hipStream_t st) {
  const int cpt = mode >= LB_MODE_BULK ? 1 : h->lb_cpt;
  switch (cpt) {
    case 2: return launch_lb_t<S, 2>(h, a, mode, rho, ux, uy, uz, st);
    case 4: return launch_lb_t<S, 4>(h, a, mode, rho, ux, uy, uz, st);
    default: return launch_lb_t<S, 1>(h, a, mode, rho, ux, uy, uz, st);
  }
}

static hipError_t dispatch_lb(pfg_handle *h, int k, int mode, float *rho, float *ux, float *uy,
                              float *uz, hipStream_t st) {
  LBArgs a;
  make_lb_args(h, k, h->cur, &a);
  switch (h->p.store) {
    case PFG_STORE_F16: return launch_lb_s<PFG_STORE_F16>(h, a, mode, rho, ux, uy, uz, st);
    case PFG_STORE_BF16: return launch_lb_s<PFG_STORE_BF16>(h, a, mode, rho, ux, uy, uz, st);
    default: return launch_lb_s<PFG_STORE_F32>(h, a, mode, rho, ux, uy, uz, st);
  }
}

/* Resident workgroups of the persistent batched launch: what the CUs hold at
 * once (occupancy x CUs), at most the tasks there are. */
template <class KER>
static unsigned queue_grid(pfg_handle *h, KER kernel) {
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess ||
      per_cu <= 0)
    per_cu = 4;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device) !=
          hipSuccess ||
      cus <= 0)
    cus = 256;
  const int64_t g = std::min<int64_t>((int64_t)per_cu * cus, h->lb_blocks);
  return (unsigned)std::max<int64_t>(g, 1);
}

template <int S, int CPT, int MODE>
static void launch_batch(pfg_handle *h, hipStream_t st) {
  const dim3 grid((unsigned)h->lb_blocks), block(256);
  const LBArgs *args = h->d_lbargs[h->cur];
  if (h->lb_persistent && MODE == LB_MODE_STEP && h->lb_queue) {
    if (h->p.wall[2] == PFG_WALL_VELOCITY) {
      auto k = lb_batch_queue_kernel<S, CPT, MODE | LB_ZVEL>;
      hipLaunchKernelGGL(k, dim3(queue_grid(h, k)), block, 0, st, args, h->lb_block_inst,
                         (uint32_t)h->lb_blocks, h->lb_queue);
    } else {
      auto k = lb_batch_queue_kernel<S, CPT, MODE>;
      hipLaunchKernelGGL(k, dim3(queue_grid(h, k)), block, 0, st, args, h->lb_block_inst,
                         (uint32_t)h->lb_blocks, h->lb_queue);
    }
    return;
  }
  const int xcd = h->lb_xcd_rows;
  if (h->p.wall[2] == PFG_WALL_VELOCITY) /* the lattices of a batch share the walls */
    hipLaunchKernelGGL((lb_batch_kernel<S, CPT, MODE | LB_ZVEL>), grid, block, 0, st, args,
                       h->lb_block_inst, xcd);
  else
    hipLaunchKernelGGL((lb_batch_kernel<S, CPT, MODE>), grid, block, 0, st, args,
                       h->lb_block_inst, xcd);
}

/* The workgroup table follows h->lb_cpt; the row-split passes take one
 * z-row per wavefront. */
template <int S>
static hipError_t launch_lb_batch_s(pfg_handle *h, int mode, hipStream_t st) {
  if (mode != LB_MODE_STEP && h->lb_cpt != 1) return hipErrorInvalidValue;
  switch (mode) {
    case LB_MODE_BULK: launch_batch<S, 1, LB_MODE_BULK>(h, st); break;
    case LB_MODE_REGION: launch_batch<S, 1, LB_MODE_REGION>(h, st); break;
    default:
      switch (h->lb_cpt) {
        case 2: launch_batch<S, 2, LB_MODE_STEP>(h, st); break;
        case 4: launch_batch<S, 4, LB_MODE_STEP>(h, st); break;
        default: launch_batch<S, 1, LB_MODE_STEP>(h, st);
      }
  }
  return hipGetLastError();
}

/* All lattices of a batch in one launch (one z-cell per lane). */
static hipError_t launch_lb_batch(pfg_handle *h, int mode, hipStream_t st) {
  switch (h->p.store) {
    case PFG_STORE_F16: return launch_lb_batch_s<PFG_STORE_F16>(h, mode, st);
    case PFG_STORE_BF16: return launch_lb_batch_s<PFG_STORE_BF16>(h, mode, st);
    default: return launch_lb_batch_s<PFG_STORE_F32>(h, mode, st);
  }
}

hipError_t launch_lb(pfg_handle *h) {
  hipError_t e = h->K > 1 ? launch_lb_batch(h, LB_MODE_STEP, h->stream)
                          : dispatch_lb(h, 0, LB_MODE_STEP, nullptr, nullptr, nullptr, nullptr,
                                        h->stream);
  if (e == hipSuccess) h->cur = 1 - h->cur;
  return e;
}

hipError_t launch_lb_part(pfg_handle *h, int part, hipStream_t st) {
  if (!h->row_stamp) return hipErrorInvalidValue;
  const int mode = part == 1 ? LB_MODE_BULK : LB_MODE_REGION;
  hipError_t e = h->K > 1 ? launch_lb_batch(h, mode, st)
                          : dispatch_lb(h, 0, mode, nullptr, nullptr, nullptr, nullptr, st);
  if (e == hipSuccess && part == 2) h->cur = 1 - h->cur;
  return e;
}

hipError_t launch_lb_fields(pfg_handle *h, int k, float *rho, float *ux, float *uy, float *uz) {
  return dispatch_lb(h, k, LB_MODE_FIELDS, rho, ux, uy, uz, h->stream);
}

hipError_t launch_bbmask(pfg_handle *h, int k) {
  if (!h->solid || !h->bbmask) return hipSuccess;
  LBArgs a;
  make_lb_args(h, k, h->cur, &a);
  const Inst &L = h->inst[(size_t)k];
  hipLaunchKernelGGL(bbmask_kernel, dim3((unsigned)((L.N + 255) / 256)), dim3(256), 0,
                     h->stream, a, a.solid, h->bbmask + L.cell0);
  return hipGetLastError();
}

hipError_t launch_init_eq(pfg_handle *h, int k, const float *rho, const float *ux,
                          const float *uy, const float *uz, bool per_z) {
  LBArgs a;
  make_lb_args(h, k, 0, &a);
  const int64_t n = h->inst[(size_t)k].N;
  void *d0 = (void *)a.fin, *d1 = a.fout; /* both copies of lattice k */
  dim3 block(256), grid((unsigned)((n + 255) / 256));
  switch (h->p.store) {
    case PFG_STORE_F16:
      hipLaunchKernelGGL(init_eq_kernel<PFG_STORE_F16>, grid, block, 0, h->stream, a, d0, d1,
                         rho, ux, uy, uz, n, per_z ? 1 : 0);
      break;
    case PFG_STORE_BF16:
      hipLaunchKernelGGL(init_eq_kernel<PFG_STORE_BF16>, grid, block, 0, h->stream, a, d0, d1,
                         rho, ux, uy, uz, n, per_z ? 1 : 0);
      break;
    default:
      hipLaunchKernelGGL(init_eq_kernel<PFG_STORE_F32>, grid, block, 0, h->stream, a, d0, d1,
                         rho, ux, uy, uz, n, per_z ? 1 : 0);
  }
  return hipGetLastError();
}

/* [19][N] full f of a single-lattice handle to or from the current copy. */
hipError_t launch_convert_f(pfg_handle *h, float *f_out, const float *f_in) {
  LBArgs a;
  make_lb_args(h, 0, h->cur, &a);
  const int64_t n = h->N;
  float *buf = nullptr;
  hipError_t e = hipMalloc(&buf, sizeof(float) * (size_t)n);
  if (e != hipSuccess) return e;
  dim3 block(256), grid((unsigned)((n + 255) / 256));
  void *stored = h->f[h->cur];
  for (int d = 0; d < Q && e == hipSuccess; ++d) {
    const int to_float = f_out != nullptr;
    if (!to_float) {
      e = hipMemcpyAsync(buf, f_in + (int64_t)d * n, sizeof(float) * n, hipMemcpyHostToDevice,
                         h->stream);
      if (e != hipSuccess) break;
    }
    switch (h->p.store) {
      case PFG_STORE_F16:
        hipLaunchKernelGGL(convert_plane_kernel<PFG_STORE_F16>, grid, block, 0, h->stream, a,
                           stored, buf, d, to_float, n);
        break;
      case PFG_STORE_BF16:
        hipLaunchKernelGGL(convert_plane_kernel<PFG_STORE_BF16>, grid, block, 0, h->stream, a,
                           stored, buf, d, to_float, n);
        break;
      default:
        hipLaunchKernelGGL(convert_plane_kernel<PFG_STORE_F32>, grid, block, 0, h->stream, a,
                           stored, buf, d, to_float, n);
    }
    e = hipGetLastError();
    if (e != hipSuccess) break;
    if (to_float) {
      e = hipMemcpyAsync(f_out + (int64_t)d * n, buf, sizeof(float) * n, hipMemcpyDeviceToHost,
                         h->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  }
  hipFree(buf);
  return e;
}

hipError_t launch_conservation(pfg_handle *h, int k, double out[4]) {
  LBArgs a;
  make_lb_args(h, k, h->cur, &a);
  const int64_t n = h->inst[(size_t)k].N;
  const int blocks = (int)std::min<int64_t>(h->red_blocks, (n + 255) / 256);
  switch (h->p.store) {
    case PFG_STORE_F16:
      hipLaunchKernelGGL(conservation_kernel<PFG_STORE_F16>, dim3(blocks), dim3(256), 0,
                         h->stream, a, a.fin, h->red, n);
      break;
    case PFG_STORE_BF16:
      hipLaunchKernelGGL(conservation_kernel<PFG_STORE_BF16>, dim3(blocks), dim3(256), 0,
                         h->stream, a, a.fin, h->red, n);
      break;
    default:
      hipLaunchKernelGGL(conservation_kernel<PFG_STORE_F32>, dim3(blocks), dim3(256), 0,
                         h->stream, a, a.fin, h->red, n);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  std::vector<double> part((size_t)blocks * 4);
  e = hipMemcpyAsync(part.data(), h->red, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                     h->stream);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return e;
  for (int c = 0; c < 4; ++c) out[c] = 0.0;
  for (int b = 0; b < blocks; ++b)
    for (int c = 0; c < 4; ++c) out[c] += part[4 * (size_t)b + c];
  return hipSuccess;
}

}  // namespace pfg
