/*
 * pfg_lattice.h -- D3Q19 velocity set and the cosine immersed-boundary kernel,
 * shared by device and host code.
 *
 * Direction order, weights and bounce-back partners follow LB_kernel.cu:40-47
 * (the order of the kernels the reference builds; fiber.cpp:19-21 uses another
 * order that the fibrin path of the reference alone relies on).
 */
#ifndef PFG_LATTICE_H
#define PFG_LATTICE_H

#include <stdint.h>

#if defined(__HIPCC__)
#define PFG_HDI __host__ __device__ inline
#else
#define PFG_HDI inline
#endif

namespace pfg {

constexpr int Q = 19;

/* e_i of LB_kernel.cu:43-45 */
constexpr int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0};
constexpr int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1};
constexpr int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1};
/* opposite direction, LB_kernel.cu:46 */
constexpr int BB[Q] = {0, 2, 1, 4, 3, 6, 5, 12, 11, 14, 13, 8, 7, 10, 9, 18, 17, 16, 15};

PFG_HDI constexpr float weight(int d) {
  return d == 0 ? 1.0f / 3.0f : (d < 7 ? 1.0f / 18.0f : 1.0f / 36.0f);
}

/* Cells per x-block of the blocked population layout (one wavefront). */
constexpr int XBLOCK = 64;

/* Force-field tile (cells) used to skip untouched IB force/velocity data:
 * the LB update reads the force and writes the velocity only in the tiles
 * this step's stencils meet.  4 x 4 x 4 cells: a 4 x 4 x 4 stencil meets
 * 1-8 of them, and the stamped cells stay close to the cells the stencils
 * cover (15 % of the 1024x512x512 lattice with 16 x 4 x 4 tiles, 10 % with
 * these). */
constexpr int TILE_X = 4, TILE_Y = 4, TILE_Z = 4;
constexpr int TILE_SHIFT_X = 2, TILE_SHIFT_Y = 2, TILE_SHIFT_Z = 2;
constexpr int TILE_CELLS = TILE_X * TILE_Y * TILE_Z;

/* 4-point cosine delta of LB_kernel.cu:1477-1478: phi(r) = (1 + cos(pi r/2))/4
 * for |r| < 2, r in lattice units. */
}  // namespace pfg

#endif
