/*
 * ib.hip -- immersed-boundary coupling between markers and the lattice.
 *
 * Spreading (reference fluid3d_force_distribute_kernel, LB_kernel.cu:1419-1520)
 * and interpolation (fluid3d_velocity_distribute_kernel, LB_kernel.cu:1524-1642)
 * use the 4-point cosine kernel phi(r) = (1 + cos(pi r / 2)) / 4, |r| < 2, as a
 * product over the three axes (LB_kernel.cu:1477-1478).  The reference loops every
 * lattice node over every marker (O(N*M)); here each marker visits only its
 * 4x4x4 stencil.
 *
 * Force field: aux[3][plane] holds the spread force between SPREAD and LB, and
 * the LB kernel overwrites it with the velocity u = (sum c f + F/2)/rho in the
 * tiles it finds flagged; INTERP reads that velocity.  Spreading marks every
 * 16x4x4 tile its stencils touch and appends newly marked tiles to a list, so
 * LB reads force only where some exists and CLEAR resets only those tiles.
 *
 * Spreading accumulates with hardware FP32 atomics: the value of each lattice
 * node is exact up to the order of the FP32 additions.  Interpolation is a
 * gather and deterministic.
 */
#include <hip/hip_runtime.h>

#include "pfg_internal.h"

namespace pfg {

struct IBArgs {
  const float *x, *y, *z;
  const float *fx, *fy, *fz;
  float *vx, *vy, *vz;
  const int32_t *kind;
  int n;
  int fiber_ib;          /* free fibrin nodes are IB-coupled                 */
  float force_scale;
  float *aux;
  int64_t plane;
  uint8_t *tile_flag;
  int32_t *tile_list;    /* [0] = count, then tile ids                        */
  int32_t *tile_done;    /* block counter of the clear kernel                 */
  int64_t *d_step;
  int nx, ny, nz;
  int tnx, tny, tnz;
  int periodic[3];
};

__device__ inline bool ib_coupled(int kind, int fiber_ib) {
  return kind == PFG_MARKER_ELEMENT || (kind == PFG_MARKER_FIBER && fiber_ib);
}

/* Wrap a coordinate into [0, n) on a periodic axis. */
__device__ inline float wrap_coord(float v, int n, int periodic) {
  if (!periodic) return v;
  const float fn = (float)n;
  float w = v - fn * floorf(v / fn);
  if (w >= fn) w -= fn; /* v slightly negative rounds up to n */
  if (w < 0.f) w += fn;
  return w;
}

/* Stencil base cell and the four weights phi(X - i), i = base .. base+3. */
__device__ inline int stencil_1d(float X, float w[4]) {
  const float fl = floorf(X);
  const float r = X - fl; /* in [0, 1) */
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float d = r + 1.0f - (float)k; /* X - (fl - 1 + k) */
    w[k] = 0.25f * (1.0f + cospif(0.5f * d));
  }
  return (int)fl - 1;
}

/* Lattice index along an axis, or -1 when outside a wall-bounded axis. */
__device__ inline int axis_index(int i, int n, int periodic) {
  if (periodic) {
    i %= n;
    return i < 0 ? i + n : i;
  }
  return (i < 0 || i >= n) ? -1 : i;
}

__global__ void spread_kernel(IBArgs a) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= a.n) return;
  if (!ib_coupled(a.kind[m], a.fiber_ib)) return;
  const float X = wrap_coord(a.x[m], a.nx, a.periodic[0]);
  const float Y = wrap_coord(a.y[m], a.ny, a.periodic[1]);
  const float Z = wrap_coord(a.z[m], a.nz, a.periodic[2]);
  float wx[4], wy[4], wz[4];
  const int bx = stencil_1d(X, wx), by = stencil_1d(Y, wy), bz = stencil_1d(Z, wz);
  const float Fx = a.fx[m] * a.force_scale;
  const float Fy = a.fy[m] * a.force_scale;
  const float Fz = a.fz[m] * a.force_scale;
  int ix[4], iy[4], iz[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    ix[k] = axis_index(bx + k, a.nx, a.periodic[0]);
    iy[k] = axis_index(by + k, a.ny, a.periodic[1]);
    iz[k] = axis_index(bz + k, a.nz, a.periodic[2]);
  }
  /* mark tiles: every stencil node's tile (at most 2 per axis, wraps aside) */
#pragma unroll
  for (int kz = 0; kz < 4; ++kz) {
    if (iz[kz] < 0) continue;
    if (kz > 0 && kz < 3 && (iz[kz] >> TILE_SHIFT_Z) == (iz[kz - 1] >> TILE_SHIFT_Z)) continue;
#pragma unroll
    for (int ky = 0; ky < 4; ++ky) {
      if (iy[ky] < 0) continue;
      if (ky > 0 && ky < 3 && (iy[ky] >> TILE_SHIFT_Y) == (iy[ky - 1] >> TILE_SHIFT_Y)) continue;
#pragma unroll
      for (int kx = 0; kx < 4; ++kx) {
        if (ix[kx] < 0) continue;
        if (kx > 0 && kx < 3 && (ix[kx] >> TILE_SHIFT_X) == (ix[kx - 1] >> TILE_SHIFT_X)) continue;
        const int tile = (ix[kx] >> TILE_SHIFT_X) +
                         a.tnx * ((iy[ky] >> TILE_SHIFT_Y) + a.tny * (iz[kz] >> TILE_SHIFT_Z));
        if (a.tile_flag[tile] == 0) {
          /* first marker to reach this tile appends it to the list */
          unsigned int *word = (unsigned int *)(a.tile_flag + (tile & ~3));
          const unsigned int bit = 1u << (8 * (tile & 3));
          const unsigned int old = atomicOr(word, bit);
          if ((old & bit) == 0) {
            const int slot = atomicAdd(&a.tile_list[0], 1);
            a.tile_list[1 + slot] = tile;
          }
        }
      }
    }
  }
#pragma unroll
  for (int kz = 0; kz < 4; ++kz) {
    if (iz[kz] < 0) continue;
#pragma unroll
    for (int ky = 0; ky < 4; ++ky) {
      if (iy[ky] < 0) continue;
      const float wyz = wy[ky] * wz[kz];
      const int64_t row = (int64_t)a.nx * (iy[ky] + (int64_t)a.ny * iz[kz]);
#pragma unroll
      for (int kx = 0; kx < 4; ++kx) {
        if (ix[kx] < 0) continue;
        const float w = wx[kx] * wyz;
        const int64_t c = row + ix[kx];
        atomicAdd(a.aux + c, w * Fx);
        atomicAdd(a.aux + a.plane + c, w * Fy);
        atomicAdd(a.aux + 2 * a.plane + c, w * Fz);
      }
    }
  }
}

__global__ void interp_kernel(IBArgs a) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= a.n) return;
  const int kind = a.kind[m];
  if (!ib_coupled(kind, a.fiber_ib)) {
    if (kind == PFG_MARKER_FIXED) {
      a.vx[m] = 0.f;
      a.vy[m] = 0.f;
      a.vz[m] = 0.f;
    }
    return;
  }
  const float X = wrap_coord(a.x[m], a.nx, a.periodic[0]);
  const float Y = wrap_coord(a.y[m], a.ny, a.periodic[1]);
  const float Z = wrap_coord(a.z[m], a.nz, a.periodic[2]);
  float wx[4], wy[4], wz[4];
  const int bx = stencil_1d(X, wx), by = stencil_1d(Y, wy), bz = stencil_1d(Z, wz);
  int ix[4], iy[4], iz[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    ix[k] = axis_index(bx + k, a.nx, a.periodic[0]);
    iy[k] = axis_index(by + k, a.ny, a.periodic[1]);
    iz[k] = axis_index(bz + k, a.nz, a.periodic[2]);
  }
  float ux = 0.f, uy = 0.f, uz = 0.f;
#pragma unroll
  for (int kz = 0; kz < 4; ++kz) {
    if (iz[kz] < 0) continue;
#pragma unroll
    for (int ky = 0; ky < 4; ++ky) {
      if (iy[ky] < 0) continue;
      const float wyz = wy[ky] * wz[kz];
      const int64_t row = (int64_t)a.nx * (iy[ky] + (int64_t)a.ny * iz[kz]);
#pragma unroll
      for (int kx = 0; kx < 4; ++kx) {
        if (ix[kx] < 0) continue;
        const float w = wx[kx] * wyz;
        const int64_t c = row + ix[kx];
        ux += w * a.aux[c];
        uy += w * a.aux[a.plane + c];
        uz += w * a.aux[2 * a.plane + c];
      }
    }
  }
  a.vx[m] = ux;
  a.vy[m] = uy;
  a.vz[m] = uz;
}

/* Zero the force/velocity field in every listed tile and unmark it.  A fixed
 * grid of blocks walks the list; the last block to finish resets the list for
 * the next step and advances the device step counter. */
__global__ __launch_bounds__(256) void clear_kernel(IBArgs a) {
  __shared__ int count;
  if (threadIdx.x == 0) count = a.tile_list[0];
  __syncthreads();
  const int t = threadIdx.x;
  const int lx = t & (TILE_X - 1);
  const int ly = (t >> TILE_SHIFT_X) & (TILE_Y - 1);
  const int lz = t >> (TILE_SHIFT_X + TILE_SHIFT_Y);
  for (int i = blockIdx.x; i < count; i += gridDim.x) {
    const int tile = a.tile_list[1 + i];
    const int tx = tile % a.tnx;
    const int ty = (tile / a.tnx) % a.tny;
    const int tz = tile / (a.tnx * a.tny);
    const int x = tx * TILE_X + lx, y = ty * TILE_Y + ly, z = tz * TILE_Z + lz;
    if (x < a.nx && y < a.ny && z < a.nz) {
      const int64_t c = x + (int64_t)a.nx * (y + (int64_t)a.ny * z);
      a.aux[c] = 0.f;
      a.aux[a.plane + c] = 0.f;
      a.aux[2 * a.plane + c] = 0.f;
    }
    if (t == 0) a.tile_flag[tile] = 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const int done = atomicAdd(a.tile_done, 1);
    if (done == (int)gridDim.x - 1) {
      a.tile_list[0] = 0;
      *a.tile_done = 0;
      if (a.d_step) *a.d_step += 1;
      __threadfence();
    }
  }
}

static void make_ib_args(pfg_handle *h, IBArgs *a) {
  const Markers &m = h->m;
  a->x = m.x;
  a->y = m.y;
  a->z = m.z;
  a->fx = m.fx;
  a->fy = m.fy;
  a->fz = m.fz;
  a->vx = m.vx;
  a->vy = m.vy;
  a->vz = m.vz;
  a->kind = m.kind;
  a->n = m.n;
  a->fiber_ib = h->p.fiber_move == PFG_FIBER_IB;
  a->force_scale = h->p.ib_force_scale;
  a->aux = h->aux;
  a->plane = h->plane;
  a->tile_flag = h->tile_flag;
  a->tile_list = h->tile_list;
  a->tile_done = h->tile_done;
  a->d_step = m.d_step;
  a->nx = h->p.nx;
  a->ny = h->p.ny;
  a->nz = h->p.nz;
  a->tnx = h->tnx;
  a->tny = h->tny;
  a->tnz = h->tnz;
  for (int k = 0; k < 3; ++k) a->periodic[k] = h->p.wall[k] == PFG_WALL_PERIODIC;
}

hipError_t launch_spread(pfg_handle *h) {
  if (h->m.n == 0) return hipSuccess;
  IBArgs a;
  make_ib_args(h, &a);
  hipLaunchKernelGGL(spread_kernel, dim3((h->m.n + 255) / 256), dim3(256), 0,
                     h->stream, a);
  return hipGetLastError();
}

hipError_t launch_interp(pfg_handle *h) {
  if (h->m.n == 0) return hipSuccess;
  IBArgs a;
  make_ib_args(h, &a);
  hipLaunchKernelGGL(interp_kernel, dim3((h->m.n + 255) / 256), dim3(256), 0,
                     h->stream, a);
  return hipGetLastError();
}

hipError_t launch_clear(pfg_handle *h) {
  IBArgs a;
  make_ib_args(h, &a);
  hipLaunchKernelGGL(clear_kernel, dim3(CLEAR_BLOCKS), dim3(256), 0, h->stream, a);
  return hipGetLastError();
}

}  // namespace pfg
