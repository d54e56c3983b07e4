/*
 * pfg_rng.h -- counter-based random numbers shared by the HIP kernels and the
 * host code.  Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64-bit
 * multiplies and a Weyl-sequence key schedule.  A draw is a pure function of
 * (key = seed, counter = {step, step>>32, item, stream}), so any thread, any
 * launch order and the CPU produce the same values.
 *
 * Replaces the reference's host-side gsl_rng_ranlxs2 refill of randNum every
 * step (semLB.cu:348-365, gpuSEM_LB.c:732-747) and the detach draw
 * (semLB.cu:451-455).
 *
 * Plain C99 + HIP: usable from C (host) and from device code.
 */
#ifndef PFG_RNG_H
#define PFG_RNG_H

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define PFG_HD __host__ __device__ static inline
#else
#define PFG_HD static inline
#endif

#define PFG_PHILOX_M0 0xD2511F53u
#define PFG_PHILOX_M1 0xCD9E8D57u
#define PFG_PHILOX_W0 0x9E3779B9u
#define PFG_PHILOX_W1 0xBB67AE85u

typedef struct pfg_u32x4 { uint32_t v[4]; } pfg_u32x4;

PFG_HD void pfg_philox_round(uint32_t c[4], const uint32_t k[2]) {
  uint64_t p0 = (uint64_t)PFG_PHILOX_M0 * (uint64_t)c[0];
  uint64_t p1 = (uint64_t)PFG_PHILOX_M1 * (uint64_t)c[2];
  uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  uint32_t n0 = hi1 ^ c[1] ^ k[0];
  uint32_t n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0;
  c[1] = lo1;
  c[2] = n2;
  c[3] = lo0;
}

/* Philox4x32-10 of counter ctr under key key. */
PFG_HD pfg_u32x4 pfg_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k[0] += PFG_PHILOX_W0;
      k[1] += PFG_PHILOX_W1;
    }
    pfg_philox_round(c, k);
  }
  pfg_u32x4 out;
  out.v[0] = c[0];
  out.v[1] = c[1];
  out.v[2] = c[2];
  out.v[3] = c[3];
  return out;
}

/* 24-bit uniform in [0, 1): exactly representable in float and double, so the
 * device and the host see the same value. */
PFG_HD float pfg_u01_from_u32(uint32_t r) {
  return (float)(r >> 8) * (1.0f / 16777216.0f);
}

/* Streams drawn per (step, item). */
enum {
  PFG_RNG_BOND_OFF = 0,
  PFG_RNG_BOND_ON = 1,
  PFG_RNG_DETACH = 2,
  PFG_RNG_KICK = 3,
  PFG_RNG_BBK_A = 4, /* first half kick of a step (completes the previous one) */
  PFG_RNG_BBK_B = 5  /* second half kick of a step                           */
};

/* One uniform for (seed, step, item, stream). */
PFG_HD float pfg_uniform(uint64_t seed, int64_t step, uint32_t item,
                         uint32_t stream) {
  uint32_t ctr[4] = {(uint32_t)(uint64_t)step, (uint32_t)((uint64_t)step >> 32),
                     item, stream};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  pfg_u32x4 r = pfg_philox4x32_10(ctr, key);
  return pfg_u01_from_u32(r.v[0]);
}

/* Three standard normals for (seed, step, item, stream): the Box-Muller
 * transform (Box and Muller, Ann. Math. Stat. 29, 1958) of the four 24-bit
 * uniforms of one Philox block, u in (0, 1] for the logarithm; FP64 on both
 * the device and the host. */
PFG_HD void pfg_normal3(uint64_t seed, int64_t step, uint32_t item, uint32_t stream,
                        double out[3]) {
  uint32_t ctr[4] = {(uint32_t)(uint64_t)step, (uint32_t)((uint64_t)step >> 32),
                     item, stream};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  pfg_u32x4 r = pfg_philox4x32_10(ctr, key);
  const double s24 = 1.0 / 16777216.0, two_pi = 6.283185307179586;
  const double u0 = ((double)(r.v[0] >> 8) + 1.0) * s24, u1 = (double)(r.v[1] >> 8) * s24;
  const double u2 = ((double)(r.v[2] >> 8) + 1.0) * s24, u3 = (double)(r.v[3] >> 8) * s24;
  const double m0 = sqrt(-2.0 * log(u0)), m1 = sqrt(-2.0 * log(u2));
  out[0] = m0 * cos(two_pi * u1);
  out[1] = m0 * sin(two_pi * u1);
  out[2] = m1 * cos(two_pi * u3);
}

#endif /* PFG_RNG_H */
