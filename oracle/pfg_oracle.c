/*
 * pfg_oracle.c -- FP64 CPU restatement of the coupled step (test infrastructure).
 *
 * What pins each function (tests under tests/ hold the checks):
 *   D3Q19 tables, orc_init_eq   feq moments (sum feq = rho, sum c feq = rho u,
 *                               sum cc feq = rho uu + rho/3 I) and rest-state
 *                               fixed point.
 *   orc_lb_step                 exact mass conservation with walls; momentum gain
 *                               of exactly N*F per step in a periodic box; Guo
 *                               source moments (0, (1-1/2tau)F, ...); Poiseuille
 *                               between plates and Couette profiles against their
 *                               closed forms; the reference's tensor form of the
 *                               Guo term (LB_kernel.cu:1254-1296) is coded here and
 *                               the device uses the expanded form, so agreement
 *                               pins the algebra.
 *   orc_phi / spread / interp   partition of unity and sum phi^2 = 3/8 exactly;
 *                               spread/interp adjointness; the reference's own
 *                               node-loops-over-markers form (orc_ib_spread_scan)
 *                               against the stencil form.
 *   orc_cell_list               brute-force all-pairs neighbour search.
 *   membrane (area/volume)      closed forms for the unit cube mesh and a right
 *                               triangle; ellipsoid volume against 4/3 pi abc.
 *   forces                      Morse: finite differences of U, zero at r_eq;
 *                               area/volume forces: finite differences of the
 *                               constraint energies; WLC / linear links against
 *                               network::getSpringForce compiled from the
 *                               reference (oracle/_ref) and Newton's third law.
 *   bonds                       Philox4x32-10 known-answer vectors
 *                               (Random123); p = 1 - exp(-k dt) closed form;
 *                               the reference law's rates (sem_kernel.cu:
 *                               1852-1863) in detailed balance with the bond's
 *                               spring energy, k_on/k_off = (Kf0/Kr0) e^{-E/kT},
 *                               and Bell's off rate in the force k|l0 - r| for
 *                               stretched and compressed bonds (rates measured
 *                               from the decisions);
 *                               the bond spring as minus the gradient of its
 *                               energy under rigid translation (finite
 *                               differences), shares proportional to 1/r_k,
 *                               all of it on a vertex the receptor sits on,
 *                               equal shares at an equilateral centroid.
 *   wall repulsion              minus the gradient of U(h) = -a ln(1 - e^{-kh})
 *                               (finite differences), the a/h small-gap limit,
 *                               zero above hc, clamped below hc/100, on the
 *                               lowest element only.
 *   Langevin leapfrog (BBK)     free decay of the velocity against the closed
 *                               form e^{-g t}; without friction the discrete
 *                               harmonic oscillator's exact recursion;
 *                               equipartition <m v^2> = T; with drift weight 1
 *                               the element follows the fluid exactly.
 *   update / kick               update: free markers in a uniform flow (an
 *                               exact fixed point of the lattice) read the
 *                               flow velocity and move by u dt per step,
 *                               anchored ones stay; kick against
 *                               network::solverkick compiled from the
 *                               reference (oracle/_ref).
 *   bending chain               mean curvature of spheres and ellipsoids against
 *                               their closed forms; the ring Laplacian exact on
 *                               linear and quadratic functions over planar
 *                               patches; the hexagonal Voronoi area; the force
 *                               independent of the triangle winding.
 *   contact                     the closest element pair by brute force and the
 *                               closed-form magnitude; zero beyond the cutoff.
 *   kick noise                  Box-Muller moments (mean, variance, kurtosis);
 *                               equipartition of the Ornstein-Uhlenbeck step.
 *   velocity faces              exact face velocities after a step; plane
 *                               Poiseuille profile and pressure gradient.
 *
 * The lattice loops (init, LB step, moments) and the IB interpolation run
 * under OpenMP when it is enabled: every cell and marker is computed on its
 * own, so the results do not depend on the thread count.
 */
#define _DEFAULT_SOURCE /* M_PI */
#include "pfg_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../platelet_fiber_gpu_amd/csrc/pfg_rng.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------- D3Q19 (LB_kernel.cu:40-47) ---------------- */

static const int CXo[19] = {0, 1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0};
static const int CYo[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1};
static const int CZo[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1};
static const int BBo[19] = {0, 2, 1, 4, 3, 6, 5, 12, 11, 14, 13, 8, 7, 10, 9, 18, 17, 16, 15};
static const double Wo[19] = {1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                              1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                              1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                              1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};

double orc_w(int d) { return Wo[d]; }
int orc_c(int d, int axis) { return axis == 0 ? CXo[d] : (axis == 1 ? CYo[d] : CZo[d]); }
int orc_bb(int d) { return BBo[d]; }

void orc_lattice_from_params(const pfg_params *p, const uint8_t *solid, orc_lattice *L) {
  L->nx = p->nx;
  L->ny = p->ny;
  L->nz = p->nz;
  for (int a = 0; a < 3; ++a) {
    L->periodic[a] = p->wall[a] == PFG_WALL_PERIODIC;
    for (int k = 0; k < 3; ++k) {
      L->wall_u_lo[a][k] = p->wall_u_lo[a][k];
      L->wall_u_hi[a][k] = p->wall_u_hi[a][k];
    }
    L->g[a] = p->body_force[a];
  }
  L->tau = p->tau;
  L->solid = p->use_flags ? solid : NULL;
  L->zvel = p->wall[2] == PFG_WALL_VELOCITY;
}

static inline int64_t cidx(const orc_lattice *L, int x, int y, int z) {
  return (int64_t)x + (int64_t)L->nx * ((int64_t)y + (int64_t)L->ny * z);
}

/* feq_d(rho, u) with C = 1 (LB_kernel.cu:1281) */
static inline double feq(int d, double rho, double ux, double uy, double uz) {
  const double cu = 3.0 * (CXo[d] * ux + CYo[d] * uy + CZo[d] * uz);
  return rho * Wo[d] * (1.0 + cu + 0.5 * cu * cu - 1.5 * (ux * ux + uy * uy + uz * uz));
}

void orc_init_eq(const orc_lattice *L, const double *rho, const double *ux,
                 const double *uy, const double *uz, double *f) {
  const int64_t N = (int64_t)L->nx * L->ny * L->nz;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    const double r = rho ? rho[i] : 1.0;
    const double a = ux ? ux[i] : 0.0, b = uy ? uy[i] : 0.0, c = uz ? uz[i] : 0.0;
    for (int d = 0; d < 19; ++d) f[d * N + i] = feq(d, r, a, b, c);
  }
}

/* Pull one population for cell (x,y,z): f_d(x) = f*_d(x - c_d), or the
 * bounced f*_bb(d)(x) + 6 w_d (c_d . u_wall) when x - c_d is a wall
 * (z wall over y wall over x wall at edges) or a solid cell. */
static double pull(const orc_lattice *L, const double *fin, int64_t N, int x, int y,
                   int z, int d) {
  int s[3] = {x - CXo[d], y - CYo[d], z - CZo[d]};
  const int n[3] = {L->nx, L->ny, L->nz};
  int wall = -1, side = 0;
  for (int a = 0; a < 3; ++a) {
    if (s[a] < 0) {
      if (L->periodic[a]) s[a] += n[a];
      else { wall = a; side = 0; }
    } else if (s[a] >= n[a]) {
      if (L->periodic[a]) s[a] -= n[a];
      else { wall = a; side = 1; }
    }
  }
  const int64_t self = cidx(L, x, y, z);
  if (d == 0) return fin[self];
  if (wall >= 0) {
    const double *u = side == 0 ? L->wall_u_lo[wall] : L->wall_u_hi[wall];
    const double cu = CXo[d] * u[0] + CYo[d] * u[1] + CZo[d] * u[2];
    return fin[BBo[d] * N + self] + 6.0 * Wo[d] * cu;
  }
  const int64_t src = cidx(L, s[0], s[1], s[2]);
  if (L->solid && L->solid[src]) return fin[BBo[d] * N + self];
  return fin[d * N + src];
}

/* Velocity face z = 0 (s = +1) or z = nz-1 (s = -1)
 * (fluid3d_in_out_flow_boundary_kernel, LB_kernel_old_08102011.cu:586-1090,
 * in the current direction order and lattice units): the face velocity is
 * the wall velocity times the duct profile 16 s1(L1-s1) s2(L2-s2)/(L1 L2)^2
 * over the bounded transverse axes; a cell next to a transverse wall is set
 * to feq(1, u) (the reference's fix for those cells); otherwise Zou/He:
 *   rho = (sum_{cz=0} f + 2 sum_{cz=-s} f) / (1 - s uz),
 *   f_{s z} = f_{-s z} + s rho uz / 3,
 *   f_{s z, sig t} = f_{-s z, -sig t} + rho (s uz + sig u_t) / 6 - sig N_t,
 *   N_t = 1/2 sum_{cz=0} c_t f - rho u_t / 3. */
static void velocity_face(const orc_lattice *L, int x, int y, int z, double f[19]) {
  const int s = z == 0 ? 1 : -1;
  const double *U = s > 0 ? L->wall_u_lo[2] : L->wall_u_hi[2];
  double shape = 1.0;
  if (!L->periodic[0]) {
    const double q = x + 0.5;
    shape *= 4.0 * q * (L->nx - q) / ((double)L->nx * L->nx);
  }
  if (!L->periodic[1]) {
    const double q = y + 0.5;
    shape *= 4.0 * q * (L->ny - q) / ((double)L->ny * L->ny);
  }
  const double u[3] = {U[0] * shape, U[1] * shape, U[2] * shape};
  if ((!L->periodic[0] && (x == 0 || x == L->nx - 1)) ||
      (!L->periodic[1] && (y == 0 || y == L->ny - 1))) {
    for (int d = 0; d < 19; ++d) f[d] = feq(d, 1.0, u[0], u[1], u[2]);
    return;
  }
  double S0 = 0, Sm = 0, Nx = 0, Ny = 0;
  for (int d = 0; d < 19; ++d) {
    if (CZo[d] == 0) {
      S0 += f[d];
      Nx += CXo[d] * f[d];
      Ny += CYo[d] * f[d];
    } else if (CZo[d] == -s) {
      Sm += f[d];
    }
  }
  const double rho = (S0 + 2.0 * Sm) / (1.0 - s * u[2]);
  Nx = 0.5 * Nx - rho * u[0] / 3.0;
  Ny = 0.5 * Ny - rho * u[1] / 3.0;
  for (int d = 0; d < 19; ++d) {
    if (CZo[d] != s) continue;
    const int o = BBo[d];
    if (CXo[d] != 0)
      f[d] = f[o] + rho * (s * u[2] + CXo[d] * u[0]) / 6.0 - CXo[d] * Nx;
    else if (CYo[d] != 0)
      f[d] = f[o] + rho * (s * u[2] + CYo[d] * u[1]) / 6.0 - CYo[d] * Ny;
    else
      f[d] = f[o] + s * rho * u[2] / 3.0;
  }
}

void orc_lb_step(const orc_lattice *L, const double *fin, double *fout, const double *Fu,
                 const double *Fs, double *u_out) {
  const int64_t N = (int64_t)L->nx * L->ny * L->nz;
  const double tau = L->tau;
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < L->nz; ++z)
    for (int y = 0; y < L->ny; ++y)
      for (int x = 0; x < L->nx; ++x) {
        const int64_t i = cidx(L, x, y, z);
        if (L->solid && L->solid[i]) continue;
        double f[19];
        for (int d = 0; d < 19; ++d) f[d] = pull(L, fin, N, x, y, z, d);
        if (L->zvel && (z == 0 || z == L->nz - 1)) velocity_face(L, x, y, z, f);
        /* moments (LB_kernel.cu:1137-1163), C = 1 */
        double rho = 0, jx = 0, jy = 0, jz = 0;
        for (int d = 0; d < 19; ++d) {
          rho += f[d];
          jx += CXo[d] * f[d];
          jy += CYo[d] * f[d];
          jz += CZo[d] * f[d];
        }
        double Fa[3] = {L->g[0], L->g[1], L->g[2]}, Fb[3] = {L->g[0], L->g[1], L->g[2]};
        if (Fu)
          for (int k = 0; k < 3; ++k) Fa[k] += Fu[k * N + i];
        if (Fs)
          for (int k = 0; k < 3; ++k) Fb[k] += Fs[k * N + i];
        const double u[3] = {(jx + 0.5 * Fa[0]) / rho, (jy + 0.5 * Fa[1]) / rho,
                             (jz + 0.5 * Fa[2]) / rho};
        if (u_out)
          for (int k = 0; k < 3; ++k) u_out[k * N + i] = u[k];
        /* Guo source in the reference's tensor form (LB_kernel.cu:1254-1296):
         * uF = u F + F u, ee = e e - I/3, fdt = w (1 - 1/(2 tau))
         * (3 e.F + 4.5 uF:ee) */
        double uF[3][3];
        for (int l = 0; l < 3; ++l)
          for (int m = 0; m < 3; ++m) uF[l][m] = u[l] * Fb[m] + Fb[l] * u[m];
        for (int d = 0; d < 19; ++d) {
          const double e[3] = {CXo[d], CYo[d], CZo[d]};
          double sum = 0;
          for (int l = 0; l < 3; ++l)
            for (int m = 0; m < 3; ++m) {
              const double ee = e[l] * e[m] - (l == m ? 1.0 / 3.0 : 0.0);
              sum += uF[l][m] * ee;
            }
          const double fdt = Wo[d] * (1.0 - 0.5 / tau) *
                             (3.0 * (e[0] * Fb[0] + e[1] * Fb[1] + e[2] * Fb[2]) + 4.5 * sum);
          const double fe = feq(d, rho, u[0], u[1], u[2]);
          fout[d * N + i] = f[d] - (f[d] - fe) / tau + fdt;
        }
      }
}

/* rho and u of the pulled distributions with body force and optional extra
 * force (velocity_density after streaming, LB_kernel.cu:1108-1199). */
static void pull_moments(const orc_lattice *L, const double *f, const double *Fx3,
                         double *rho_o, double *ux, double *uy, double *uz) {
  const int64_t N = (int64_t)L->nx * L->ny * L->nz;
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < L->nz; ++z)
    for (int y = 0; y < L->ny; ++y)
      for (int x = 0; x < L->nx; ++x) {
        const int64_t i = cidx(L, x, y, z);
        if (L->solid && L->solid[i]) {
          if (rho_o) rho_o[i] = 0;
          if (ux) ux[i] = 0;
          if (uy) uy[i] = 0;
          if (uz) uz[i] = 0;
          continue;
        }
        double rho = 0, j[3] = {0, 0, 0}, fp[19];
        for (int d = 0; d < 19; ++d) fp[d] = pull(L, f, N, x, y, z, d);
        if (L->zvel && (z == 0 || z == L->nz - 1)) velocity_face(L, x, y, z, fp);
        for (int d = 0; d < 19; ++d) {
          const double v = fp[d];
          rho += v;
          j[0] += CXo[d] * v;
          j[1] += CYo[d] * v;
          j[2] += CZo[d] * v;
        }
        double F[3] = {L->g[0], L->g[1], L->g[2]};
        if (Fx3)
          for (int k = 0; k < 3; ++k) F[k] += Fx3[k * N + i];
        if (rho_o) rho_o[i] = rho;
        if (ux) ux[i] = (j[0] + 0.5 * F[0]) / rho;
        if (uy) uy[i] = (j[1] + 0.5 * F[1]) / rho;
        if (uz) uz[i] = (j[2] + 0.5 * F[2]) / rho;
      }
}

void orc_fields(const orc_lattice *L, const double *f, double *rho, double *ux, double *uy,
                double *uz) {
  pull_moments(L, f, NULL, rho, ux, uy, uz);
}

void orc_conservation(const orc_lattice *L, const double *f, double out[4]) {
  const int64_t N = (int64_t)L->nx * L->ny * L->nz;
  out[0] = out[1] = out[2] = out[3] = 0;
  for (int64_t i = 0; i < N; ++i) {
    if (L->solid && L->solid[i]) continue;
    for (int d = 0; d < 19; ++d) {
      const double v = f[d * N + i];
      out[0] += v;
      out[1] += CXo[d] * v;
      out[2] += CYo[d] * v;
      out[3] += CZo[d] * v;
    }
  }
}

/* ---------------- immersed boundary ---------------- */

double orc_phi(double r) {
  const double a = fabs(r);
  if (a >= 2.0) return 0.0;
  return 0.25 * (1.0 + cos(0.5 * M_PI * r));
}

static double wrap_pos(double v, int n, int periodic) {
  if (!periodic) return v;
  double w = fmod(v, (double)n);
  if (w < 0) w += n;
  if (w >= n) w -= n;
  return w;
}

static int axis_idx(int i, int n, int periodic) {
  if (periodic) {
    i %= n;
    return i < 0 ? i + n : i;
  }
  return (i < 0 || i >= n) ? -1 : i;
}

/* Visit the 4x4x4 stencil of a marker: weights and lattice indices. */
typedef struct {
  int ix[4], iy[4], iz[4];
  double wx[4], wy[4], wz[4];
} stencil;

static void make_stencil(const orc_lattice *L, double x, double y, double z, stencil *s) {
  const double P[3] = {wrap_pos(x, L->nx, L->periodic[0]), wrap_pos(y, L->ny, L->periodic[1]),
                       wrap_pos(z, L->nz, L->periodic[2])};
  const int n[3] = {L->nx, L->ny, L->nz};
  int *I[3] = {s->ix, s->iy, s->iz};
  double *W[3] = {s->wx, s->wy, s->wz};
  for (int a = 0; a < 3; ++a) {
    const int b = (int)floor(P[a]) - 1;
    for (int k = 0; k < 4; ++k) {
      I[a][k] = axis_idx(b + k, n[a], L->periodic[a]);
      W[a][k] = orc_phi(P[a] - (double)(b + k));
    }
  }
}

void orc_ib_spread(const orc_lattice *L, int n, const double *x, const double *y,
                   const double *z, const double *fx, const double *fy, const double *fz,
                   const int *coupled, double scale, double *F) {
  const int64_t N = (int64_t)L->nx * L->ny * L->nz;
  for (int m = 0; m < n; ++m) {
    if (coupled && !coupled[m]) continue;
    stencil s;
    make_stencil(L, x[m], y[m], z[m], &s);
    for (int c = 0; c < 4; ++c) {
      if (s.iz[c] < 0) continue;
      for (int b = 0; b < 4; ++b) {
        if (s.iy[b] < 0) continue;
        for (int a = 0; a < 4; ++a) {
          if (s.ix[a] < 0) continue;
          const double w = s.wx[a] * s.wy[b] * s.wz[c];
          const int64_t i = cidx(L, s.ix[a], s.iy[b], s.iz[c]);
          F[i] += w * fx[m] * scale;
          F[N + i] += w * fy[m] * scale;
          F[2 * N + i] += w * fz[m] * scale;
        }
      }
    }
  }
}

/* LB_kernel.cu:1419-1520: each lattice node loops over all markers and takes
 * those within 2 cells per axis; periodic images are taken on every periodic
 * axis (the reference takes them in y only). */
void orc_ib_spread_scan(const orc_lattice *L, int n, const double *x, const double *y,
                        const double *z, const double *fx, const double *fy,
                        const double *fz, const int *coupled, double scale, double *F) {
  const int64_t N = (int64_t)L->nx * L->ny * L->nz;
  for (int k = 0; k < L->nz; ++k)
    for (int j = 0; j < L->ny; ++j)
      for (int i = 0; i < L->nx; ++i) {
        const int64_t c = cidx(L, i, j, k);
        for (int m = 0; m < n; ++m) {
          if (coupled && !coupled[m]) continue;
          const double P[3] = {wrap_pos(x[m], L->nx, L->periodic[0]),
                               wrap_pos(y[m], L->ny, L->periodic[1]),
                               wrap_pos(z[m], L->nz, L->periodic[2])};
          const int node[3] = {i, j, k};
          const int nn[3] = {L->nx, L->ny, L->nz};
          double w = 1.0;
          for (int a = 0; a < 3 && w != 0.0; ++a) {
            double r = P[a] - node[a];
            if (L->periodic[a]) {
              if (r > 0.5 * nn[a]) r -= nn[a];
              if (r < -0.5 * nn[a]) r += nn[a];
            }
            if (fabs(r) >= 2.0) w = 0.0;
            else w *= 0.25 * (1.0 + cos(0.5 * M_PI * r));
          }
          if (w == 0.0) continue;
          F[c] += w * fx[m] * scale;
          F[N + c] += w * fy[m] * scale;
          F[2 * N + c] += w * fz[m] * scale;
        }
      }
}

void orc_ib_interp(const orc_lattice *L, int n, const double *x, const double *y,
                   const double *z, const double *u, const int *coupled, double *vx,
                   double *vy, double *vz) {
  const int64_t N = (int64_t)L->nx * L->ny * L->nz;
#pragma omp parallel for schedule(static)
  for (int m = 0; m < n; ++m) {
    if (coupled && !coupled[m]) continue;
    stencil s;
    make_stencil(L, x[m], y[m], z[m], &s);
    double v[3] = {0, 0, 0};
    for (int c = 0; c < 4; ++c) {
      if (s.iz[c] < 0) continue;
      for (int b = 0; b < 4; ++b) {
        if (s.iy[b] < 0) continue;
        for (int a = 0; a < 4; ++a) {
          if (s.ix[a] < 0) continue;
          const double w = s.wx[a] * s.wy[b] * s.wz[c];
          const int64_t i = cidx(L, s.ix[a], s.iy[b], s.iz[c]);
          v[0] += w * u[i];
          v[1] += w * u[N + i];
          v[2] += w * u[2 * N + i];
        }
      }
    }
    vx[m] = v[0];
    vy[m] = v[1];
    vz[m] = v[2];
  }
}

/* ---------------- cell lists ---------------- */

static int cell_of_f(float v, int n, int periodic) {
  int c = (int)floorf(v);
  if (periodic) {
    c %= n;
    if (c < 0) c += n;
  } else {
    c = c < 0 ? 0 : (c >= n ? n - 1 : c);
  }
  return c;
}

static int bin_of_f(float x, float y, float z, int nx, int ny, int nz, const int per[3],
                    const int nb[3]) {
  const int bx = cell_of_f(x, nx, per[0]) * nb[0] / nx;
  const int by = cell_of_f(y, ny, per[1]) * nb[1] / ny;
  const int bz = cell_of_f(z, nz, per[2]) * nb[2] / nz;
  return bx + nb[0] * (by + nb[1] * bz);
}

void orc_cell_list(int n, const float *x, const float *y, const float *z, int nx, int ny,
                   int nz, const int periodic[3], const int nb[3], int *bin_start,
                   int *items) {
  const int64_t nbins = (int64_t)nb[0] * nb[1] * nb[2];
  memset(bin_start, 0, sizeof(int) * (size_t)(nbins + 1));
  int *bin = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) {
    bin[i] = bin_of_f(x[i], y[i], z[i], nx, ny, nz, periodic, nb);
    bin_start[bin[i] + 1]++;
  }
  for (int64_t b = 0; b < nbins; ++b) bin_start[b + 1] += bin_start[b];
  int *fillp = (int *)malloc(sizeof(int) * (size_t)(nbins > 0 ? nbins : 1));
  memcpy(fillp, bin_start, sizeof(int) * (size_t)nbins);
  for (int i = 0; i < n; ++i) items[fillp[bin[i]]++] = i;
  free(fillp);
  free(bin);
}

/* ---------------- full state ---------------- */

struct orc_state {
  pfg_params p;
  orc_lattice L;
  int64_t N;
  int order;
  uint8_t *solid;
  double *f, *f2, *F, *Fprev, *u;
  int n;
  double *mk[9]; /* x y z vx vy vz fx fy fz */
  int *kind, *body, *coupled;
  int n_plt, n_tri, n_edges, n_rec, n_wlig, fib_lig, n_lig, n_links;
  int *plt_elem_start, *plt_tri_start, *tri;
  double *tri_S0, *plt_S0, *plt_V0;
  double *tri_S, *tri_n, *plt_S, *plt_V;
  int *lowest;
  int *edge;
  double *edge_l0;
  int *rec_tri;
  double *rec_r1, *rec_r2;
  double *lig_x, *lig_y, *lig_z;
  int *link;
  double *link_l0;
  int *rec_bond, *rec_cand, *rec_release, *lig_owner, *lig_claim;
  double *rec_force, *rec_w;
  int nb[3];
  int *bin_start, *items;
  int64_t step, total_bonds, detached, detach_step, near, events;
  int64_t near_cut; /* Morse pair visits within 1e-4 of the cutoff */
  int64_t *plt_near_cut; /* the same per platelet */
  double *uib; /* 3n: fluid velocity at the markers (PFG_SEM_BBK) */
  /* bending chain: 6-node triangles (3 vertices, 3 edge-midpoint nodes
   * numbered n + edge), triangles per node, ordered vertex rings */
  int n_enode, bend_ok;
  int *tri6, *node_cnt, *ring_start, *ring, *ring_ok;
  double *node_n, *kapa, *km, *fb;
};

static void *xcalloc(size_t n, size_t sz) { return calloc(n > 0 ? n : 1, sz); }

/* ---- bending topology (gpuSEM_LB.c:846-936) ----
 * Edge-midpoint nodes: the edge of local vertices (0,1) goes to slot 3,
 * (0,2) to slot 4, (1,2) to slot 5 (the switch on the sum of the local
 * indices, gpuSEM_LB.c:859-871).  node_cnt: triangles touching a node
 * (node_nbrElemNum).  Vertex rings (node_nbrNodes): start at the vertex after
 * i in its first triangle, then repeatedly take, among the triangles of i
 * not yet used, the first one holding the current neighbour and step to the
 * vertex before it there (gpuSEM_LB.c:900-936).  Edge ids in order of the
 * sorted vertex pairs.  ring_ok[i] = 0 when the ring of i does not close (a
 * vertex on the border of an open mesh): it gets no bending force. */
static int edge_key_cmp(const void *a, const void *b) {
  const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

static void bend_topology(orc_state *s) {
  const int n = s->n, T = s->n_tri;
  s->bend_ok = T > 0;
  if (T == 0) return;
  /* edges: sort (key, first appearance) to give each edge an id */
  int64_t *keys = (int64_t *)xcalloc((size_t)3 * T, 2 * sizeof(int64_t));
  static const int sl[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int t = 0; t < T; ++t)
    for (int k = 0; k < 3; ++k) {
      int a = s->tri[3 * t + sl[k][0]], b = s->tri[3 * t + sl[k][1]];
      if (a > b) { const int c = a; a = b; b = c; }
      keys[2 * (3 * t + k)] = (int64_t)a * n + b;
      keys[2 * (3 * t + k) + 1] = 3 * t + k;
    }
  int64_t (*kv)[2] = (int64_t (*)[2])keys;
  qsort(kv, (size_t)3 * T, 2 * sizeof(int64_t), edge_key_cmp);
  int ne = 0;
  int *slot_edge = (int *)xcalloc((size_t)3 * T, sizeof(int));
  for (int i = 0; i < 3 * T;) {
    int j = i;
    while (j < 3 * T && kv[j][0] == kv[i][0]) ++j;
    for (int a = i; a < j; ++a) slot_edge[kv[a][1]] = ne;
    ++ne;
    i = j;
  }
  s->n_enode = ne;
  s->tri6 = (int *)xcalloc((size_t)6 * T, sizeof(int));
  s->node_cnt = (int *)xcalloc((size_t)n + ne, sizeof(int));
  for (int t = 0; t < T; ++t) {
    for (int k = 0; k < 3; ++k) s->tri6[6 * t + k] = s->tri[3 * t + k];
    for (int k = 0; k < 3; ++k) s->tri6[6 * t + 3 + k] = n + slot_edge[3 * t + k];
    for (int k = 0; k < 6; ++k) s->node_cnt[s->tri6[6 * t + k]]++;
  }
  /* triangles of each vertex in ascending order (node_share_Elem) */
  int *vstart = (int *)xcalloc((size_t)n + 1, sizeof(int));
  for (int t = 0; t < T; ++t)
    for (int k = 0; k < 3; ++k) vstart[s->tri[3 * t + k] + 1]++;
  for (int i = 0; i < n; ++i) vstart[i + 1] += vstart[i];
  int *vtri = (int *)xcalloc((size_t)vstart[n], sizeof(int));
  int *fill = (int *)xcalloc((size_t)n, sizeof(int));
  for (int t = 0; t < T; ++t)
    for (int k = 0; k < 3; ++k) {
      const int v = s->tri[3 * t + k];
      vtri[vstart[v] + fill[v]++] = t;
    }
  s->ring_start = vstart;
  s->ring = (int *)xcalloc((size_t)vstart[n], sizeof(int));
  s->ring_ok = (int *)xcalloc((size_t)n, sizeof(int));
  for (int i = 0; i < n; ++i) {
    const int c = vstart[i + 1] - vstart[i];
    if (c == 0) continue;
    int *el = vtri + vstart[i];
    int *rg = s->ring + vstart[i];
    int cur = -1, ok = 1;
    const int e0 = el[0];
    for (int j = 0; j < 3; ++j)
      if (s->tri[3 * e0 + j] == i) cur = s->tri[3 * e0 + (j == 2 ? 0 : j + 1)];
    rg[0] = cur;
    for (int j = 1; j < c; ++j) {
      int found = 0;
      for (int k = j; k < c && !found; ++k) {
        const int e = el[k];
        for (int l = 0; l < 3; ++l) {
          if (s->tri[3 * e + l] == cur) {
            found = 1;
            cur = s->tri[3 * e + (l == 0 ? 2 : l - 1)];
            rg[j] = cur;
            const int tmp = el[j];
            el[j] = el[k];
            el[k] = tmp;
            break;
          }
        }
      }
      if (!found) {
        ok = 0;
        rg[j] = i;
      }
    }
    /* closed: the last neighbour and the first share a triangle of i */
    if (ok && c > 2) {
      int shared = 0;
      for (int k = 0; k < c && !shared; ++k) {
        const int e = el[k];
        int hl = 0, hf = 0;
        for (int l = 0; l < 3; ++l) {
          hl |= s->tri[3 * e + l] == rg[c - 1];
          hf |= s->tri[3 * e + l] == rg[0];
        }
        shared = hl && hf;
      }
      ok = shared;
    }
    s->ring_ok[i] = ok && c > 2;
  }
  s->node_n = (double *)xcalloc((size_t)3 * (n + ne), sizeof(double));
  s->kapa = (double *)xcalloc((size_t)9 * (n + ne), sizeof(double));
  s->km = (double *)xcalloc((size_t)n + ne, sizeof(double));
  s->fb = (double *)xcalloc((size_t)3 * n, sizeof(double));
  free(keys);
  free(slot_edge);
  free(vtri);
  free(fill);
}

orc_state *orc_create(const pfg_params *p, const pfg_model *m, const uint8_t *flags) {
  orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
  s->p = *p;
  s->N = (int64_t)p->nx * p->ny * p->nz;
  if (p->use_flags && flags) {
    s->solid = (uint8_t *)malloc((size_t)s->N);
    memcpy(s->solid, flags, (size_t)s->N);
  }
  orc_lattice_from_params(p, s->solid, &s->L);
  s->f = (double *)xcalloc((size_t)19 * s->N, sizeof(double));
  s->f2 = (double *)xcalloc((size_t)19 * s->N, sizeof(double));
  s->F = (double *)xcalloc((size_t)3 * s->N, sizeof(double));
  s->Fprev = (double *)xcalloc((size_t)3 * s->N, sizeof(double));
  s->u = (double *)xcalloc((size_t)3 * s->N, sizeof(double));
  orc_init_eq(&s->L, NULL, NULL, NULL, NULL, s->f);
  const int n = m ? m->n_markers : 0;
  s->n = n;
  for (int k = 0; k < 9; ++k) s->mk[k] = (double *)xcalloc((size_t)n, sizeof(double));
  s->kind = (int *)xcalloc((size_t)n, sizeof(int));
  s->uib = (double *)xcalloc((size_t)3 * n + 1, sizeof(double));
  s->body = (int *)xcalloc((size_t)n, sizeof(int));
  s->coupled = (int *)xcalloc((size_t)n, sizeof(int));
  for (int i = 0; i < n; ++i) {
    s->mk[0][i] = m->x[i];
    s->mk[1][i] = m->y[i];
    s->mk[2][i] = m->z[i];
    s->kind[i] = m->kind[i];
    s->body[i] = m->body[i];
    s->coupled[i] = m->kind[i] == PFG_MARKER_ELEMENT ||
                    (m->kind[i] == PFG_MARKER_FIBER && p->fiber_move == PFG_FIBER_IB);
  }
  s->n_plt = m ? m->n_platelets : 0;
  s->n_tri = m ? m->n_tri : 0;
  s->n_edges = m ? m->n_edges : 0;
  s->n_rec = m ? m->n_receptors : 0;
  s->n_wlig = m ? m->n_wall_ligands : 0;
  s->fib_lig = m ? (m->fibrin_ligands != 0) : 0;
  s->n_lig = s->n_wlig + (s->fib_lig ? n : 0);
  s->n_links = m ? m->n_links : 0;
  s->plt_elem_start = (int *)xcalloc((size_t)s->n_plt + 1, sizeof(int));
  s->plt_tri_start = (int *)xcalloc((size_t)s->n_plt + 1, sizeof(int));
  for (int q = 0; q <= s->n_plt && s->n_plt > 0; ++q) {
    s->plt_elem_start[q] = m->plt_elem_start[q];
    s->plt_tri_start[q] = m->plt_tri_start[q];
  }
  s->tri = (int *)xcalloc((size_t)3 * s->n_tri, sizeof(int));
  s->tri_S0 = (double *)xcalloc((size_t)s->n_tri, sizeof(double));
  s->tri_S = (double *)xcalloc((size_t)s->n_tri, sizeof(double));
  s->tri_n = (double *)xcalloc((size_t)3 * s->n_tri, sizeof(double));
  for (int t = 0; t < s->n_tri; ++t) {
    for (int k = 0; k < 3; ++k) s->tri[3 * t + k] = m->tri[3 * t + k];
    s->tri_S0[t] = m->tri_S0[t];
  }
  s->plt_S0 = (double *)xcalloc((size_t)s->n_plt, sizeof(double));
  s->plt_near_cut = (int64_t *)xcalloc((size_t)s->n_plt + 1, sizeof(int64_t));
  s->plt_V0 = (double *)xcalloc((size_t)s->n_plt, sizeof(double));
  s->plt_S = (double *)xcalloc((size_t)s->n_plt, sizeof(double));
  s->plt_V = (double *)xcalloc((size_t)s->n_plt, sizeof(double));
  s->lowest = (int *)xcalloc((size_t)s->n_plt, sizeof(int));
  for (int q = 0; q < s->n_plt; ++q) {
    s->plt_S0[q] = m->plt_S0[q];
    s->plt_V0[q] = m->plt_V0[q];
    s->lowest[q] = -1;
  }
  s->edge = (int *)xcalloc((size_t)2 * s->n_edges, sizeof(int));
  s->edge_l0 = (double *)xcalloc((size_t)s->n_edges, sizeof(double));
  for (int e = 0; e < s->n_edges; ++e) {
    s->edge[2 * e] = m->edge[2 * e];
    s->edge[2 * e + 1] = m->edge[2 * e + 1];
    s->edge_l0[e] = m->edge_l0[e];
  }
  s->rec_tri = (int *)xcalloc((size_t)s->n_rec, sizeof(int));
  s->rec_r1 = (double *)xcalloc((size_t)s->n_rec, sizeof(double));
  s->rec_r2 = (double *)xcalloc((size_t)s->n_rec, sizeof(double));
  s->rec_bond = (int *)xcalloc((size_t)s->n_rec, sizeof(int));
  s->rec_cand = (int *)xcalloc((size_t)s->n_rec, sizeof(int));
  s->rec_release = (int *)xcalloc((size_t)s->n_rec, sizeof(int));
  s->rec_force = (double *)xcalloc((size_t)3 * s->n_rec, sizeof(double));
  s->rec_w = (double *)xcalloc((size_t)3 * s->n_rec, sizeof(double));
  for (int r = 0; r < s->n_rec; ++r) {
    s->rec_tri[r] = m->rec_tri[r];
    s->rec_r1[r] = m->rec_r1[r];
    s->rec_r2[r] = m->rec_r2[r];
    s->rec_bond[r] = -1;
  }
  s->lig_x = (double *)xcalloc((size_t)s->n_wlig, sizeof(double));
  s->lig_y = (double *)xcalloc((size_t)s->n_wlig, sizeof(double));
  s->lig_z = (double *)xcalloc((size_t)s->n_wlig, sizeof(double));
  for (int l = 0; l < s->n_wlig; ++l) {
    s->lig_x[l] = m->lig_x[l];
    s->lig_y[l] = m->lig_y[l];
    s->lig_z[l] = m->lig_z ? (double)m->lig_z[l] : (double)p->wall_z;
  }
  s->lig_owner = (int *)xcalloc((size_t)s->n_lig, sizeof(int));
  s->lig_claim = (int *)xcalloc((size_t)s->n_lig, sizeof(int));
  for (int l = 0; l < s->n_lig; ++l) {
    s->lig_owner[l] = -1;
    s->lig_claim[l] = INT_MAX;
  }
  s->link = (int *)xcalloc((size_t)2 * s->n_links, sizeof(int));
  s->link_l0 = (double *)xcalloc((size_t)s->n_links, sizeof(double));
  for (int l = 0; l < s->n_links; ++l) {
    s->link[2 * l] = m->link[2 * l];
    s->link[2 * l + 1] = m->link[2 * l + 1];
    s->link_l0[l] = m->link_l0[l];
  }
  const int nn[3] = {p->nx, p->ny, p->nz};
  for (int k = 0; k < 3; ++k) {
    s->nb[k] = nn[k] / p->bin_cells;
    if (s->nb[k] < 1) s->nb[k] = 1;
  }
  s->bin_start = (int *)xcalloc((size_t)s->nb[0] * s->nb[1] * s->nb[2] + 1, sizeof(int));
  s->items = (int *)xcalloc((size_t)n, sizeof(int));
  s->detach_step = -1;
  bend_topology(s);
  return s;
}

void orc_destroy(orc_state *s) {
  if (!s) return;
  free(s->solid);
  free(s->f);
  free(s->f2);
  free(s->F);
  free(s->Fprev);
  free(s->u);
  for (int k = 0; k < 9; ++k) free(s->mk[k]);
  free(s->kind);
  free(s->uib);
  free(s->body);
  free(s->coupled);
  free(s->plt_elem_start);
  free(s->plt_tri_start);
  free(s->tri);
  free(s->tri_S0);
  free(s->tri_S);
  free(s->tri_n);
  free(s->plt_S0);
  free(s->plt_near_cut);
  free(s->plt_V0);
  free(s->plt_S);
  free(s->plt_V);
  free(s->lowest);
  free(s->edge);
  free(s->edge_l0);
  free(s->rec_tri);
  free(s->rec_r1);
  free(s->rec_r2);
  free(s->rec_bond);
  free(s->rec_cand);
  free(s->rec_release);
  free(s->rec_force);
  free(s->rec_w);
  free(s->lig_x);
  free(s->lig_y);
  free(s->lig_z);
  free(s->lig_owner);
  free(s->lig_claim);
  free(s->link);
  free(s->link_l0);
  free(s->bin_start);
  free(s->items);
  free(s->tri6);
  free(s->node_cnt);
  free(s->ring_start);
  free(s->ring);
  free(s->ring_ok);
  free(s->node_n);
  free(s->kapa);
  free(s->km);
  free(s->fb);
  free(s);
}

void orc_set_order(orc_state *s, int order) { s->order = order; }

int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}
void orc_set_step(orc_state *s, int64_t step) { s->step = step; }

/* ---- membrane: sem_calculate_A_kernel (sem_kernel.cu:724-872) ---- */
static void stage_membrane(orc_state *s) {
  const double *X = s->mk[0], *Y = s->mk[1], *Z = s->mk[2];
  for (int q = 0; q < s->n_plt; ++q) {
    double S = 0, V = 0;
    for (int t = s->plt_tri_start[q]; t < s->plt_tri_start[q + 1]; ++t) {
      const int *nd = s->tri + 3 * t;
      double c[3][3];
      for (int k = 0; k < 3; ++k) {
        c[k][0] = X[nd[k]];
        c[k][1] = Y[nd[k]];
        c[k][2] = Z[nd[k]];
      }
      double e12[3], e13[3], ne[3];
      for (int k = 0; k < 3; ++k) {
        e12[k] = c[1][k] - c[0][k];
        e13[k] = c[2][k] - c[0][k];
      }
      ne[0] = e12[1] * e13[2] - e12[2] * e13[1];
      ne[1] = e12[2] * e13[0] - e12[0] * e13[2];
      ne[2] = e12[0] * e13[1] - e12[1] * e13[0];
      const double nn = sqrt(ne[0] * ne[0] + ne[1] * ne[1] + ne[2] * ne[2]);
      for (int k = 0; k < 3; ++k) ne[k] /= -nn;
      const double St = 0.5 * nn;
      double cen[3];
      for (int k = 0; k < 3; ++k) cen[k] = (c[0][k] + c[1][k] + c[2][k]) / 3.0;
      s->tri_S[t] = St;
      for (int k = 0; k < 3; ++k) s->tri_n[3 * t + k] = ne[k];
      S += St;
      V += (cen[0] * ne[0] + cen[1] * ne[1] + cen[2] * ne[2]) * St / 3.0;
    }
    s->plt_S[q] = S;
    s->plt_V[q] = V;
    int lo = -1;
    double zmin = INFINITY;
    for (int e = s->plt_elem_start[q]; e < s->plt_elem_start[q + 1]; ++e) {
      if ((float)Z[e] < zmin) {
        zmin = (float)Z[e];
        lo = e;
      }
    }
    s->lowest[q] = lo;
  }
}

/* ---- cells ---- */
static void stage_cells(orc_state *s) {
  float *x = (float *)malloc(sizeof(float) * (size_t)(s->n > 0 ? s->n : 1));
  float *y = (float *)malloc(sizeof(float) * (size_t)(s->n > 0 ? s->n : 1));
  float *z = (float *)malloc(sizeof(float) * (size_t)(s->n > 0 ? s->n : 1));
  for (int i = 0; i < s->n; ++i) {
    x[i] = (float)s->mk[0][i];
    y[i] = (float)s->mk[1][i];
    z[i] = (float)s->mk[2][i];
  }
  orc_cell_list(s->n, x, y, z, s->p.nx, s->p.ny, s->p.nz, s->L.periodic, s->nb,
                s->bin_start, s->items);
  free(x);
  free(y);
  free(z);
  s->total_bonds = 0;
}

/* ---- bonds: Bell model (sem_platelet_wall_kernel, sem_kernel.cu:1750-2083) ---- */
static double min_img(double d, int n, int periodic) {
  if (!periodic) return d;
  if (d > 0.5 * n) d -= n;
  else if (d < -0.5 * n) d += n;
  return d;
}

static void receptor_pos(const orc_state *s, int r, double rp[3], int nodes[3]) {
  const int t = s->rec_tri[r];
  const double r1 = s->rec_r1[r], r2 = s->rec_r2[r];
  for (int k = 0; k < 3; ++k) nodes[k] = s->tri[3 * t + k];
  /* sem_kernel.cu:1811, 1840-1841 */
  const double c0 = 1.0 - r1, c1 = r1 * (1.0 - r2), c2 = r1 * r2;
  for (int a = 0; a < 3; ++a)
    rp[a] = c0 * s->mk[a][nodes[0]] + c1 * s->mk[a][nodes[1]] + c2 * s->mk[a][nodes[2]];
}

static void ligand_vec(const orc_state *s, int l, const double rp[3], double d[3]) {
  double L3[3];
  if (l < s->n_wlig) {
    L3[0] = s->lig_x[l];
    L3[1] = s->lig_y[l];
    L3[2] = s->lig_z[l];
  } else {
    const int mk = l - s->n_wlig;
    for (int a = 0; a < 3; ++a) L3[a] = s->mk[a][mk];
  }
  const int nn[3] = {s->p.nx, s->p.ny, s->p.nz};
  for (int a = 0; a < 3; ++a) d[a] = min_img(L3[a] - rp[a], nn[a], s->L.periodic[a]);
}

static void stage_bonds(orc_state *s) {
  const double eps_p = 1e-6, eps_r = 1e-4;
  const pfg_params *p = &s->p;
  /* decide from the start-of-step state */
  for (int r = 0; r < s->n_rec; ++r) {
    double rp[3];
    int nodes[3];
    receptor_pos(s, r, rp, nodes);
    s->rec_cand[r] = -1;
    s->rec_release[r] = 0;
    const int b = s->rec_bond[r];
    if (b >= 0) {
      double d[3];
      ligand_vec(s, b, rp, d);
      const double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      const double tension = (double)p->bond_k * (dist - (double)p->bond_l0);
      /* Bell on the pulling force, or the reference's law on the force
       * magnitude k |l0 - r| (sem_kernel.cu:1852-1863) */
      const double f = p->bond_law == PFG_BOND_REFERENCE ? fabs(tension)
                                                          : (tension > 0 ? tension : 0);
      const double koff = (double)p->k_off0 * exp((double)p->bell_xb_kT * f);
      const double p_off = 1.0 - exp(-koff);
      const double u = pfg_uniform(p->seed, s->step, (uint32_t)r, PFG_RNG_BOND_OFF);
      if (fabs(u - p_off) < eps_p) s->near++;
      s->rec_release[r] = u < p_off;
      continue;
    }
    int best = -1;
    double best_d2 = INFINITY;
    int near_reach = 0;
    const double reach = p->bond_reach, reach2 = (double)p->bond_reach * p->bond_reach;
    for (int l = 0; l < s->n_lig; ++l) {
      if (l >= s->n_wlig && s->kind[l - s->n_wlig] == PFG_MARKER_ELEMENT) continue;
      double d[3];
      ligand_vec(s, l, rp, d);
      const double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (d2 > 4.0 * reach2 + 1.0) continue;
      if (fabs(sqrt(d2) - reach) < eps_r) near_reach = 1;
      if (d2 > reach2 || s->lig_owner[l] >= 0) continue;
      if (d2 < best_d2) {
        best_d2 = d2;
        best = l;
      }
    }
    if (near_reach) s->near++;
    if (best < 0) continue;
    double kon = p->k_on;
    if (p->bond_law == PFG_BOND_REFERENCE) {
      /* Kf = Kf0 e^{f (x_b - w/2) / kT}, w = |l0 - r|, f = k w
       * (sem_kernel.cu:1852-1858) */
      const double w = fabs((double)p->bond_l0 - sqrt(best_d2)), f = (double)p->bond_k * w;
      kon = (double)p->k_on * exp((double)p->bell_xb_kT * f - 0.5 * f * w / (double)p->bell_kT);
    }
    const double p_on = 1.0 - exp(-kon);
    const double u = pfg_uniform(p->seed, s->step, (uint32_t)r, PFG_RNG_BOND_ON);
    if (fabs(u - p_on) < eps_p) s->near++;
    if (u < p_on) {
      s->rec_cand[r] = best;
      if (r < s->lig_claim[best]) s->lig_claim[best] = r;
    }
  }
  /* resolve: ruptures, lowest-index claims, forces */
  for (int r = 0; r < s->n_rec; ++r) {
    int b = s->rec_bond[r];
    if (s->rec_release[r]) {
      s->lig_owner[b] = -1;
      b = -1;
      s->events++;
    } else if (s->rec_cand[r] >= 0) {
      const int l = s->rec_cand[r];
      if (s->lig_claim[l] == r) {
        s->lig_owner[l] = r;
        b = l;
        s->events++;
      }
    }
    s->rec_bond[r] = b;
    double F[3] = {0, 0, 0}, w[3] = {0, 0, 0};
    if (b >= 0) {
      double rp[3];
      int nodes[3];
      receptor_pos(s, r, rp, nodes);
      double d[3];
      ligand_vec(s, b, rp, d);
      const double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (dist > 0) {
        const double c = (double)p->bond_k * (dist - (double)p->bond_l0) / dist;
        for (int a = 0; a < 3; ++a) F[a] = c * d[a];
      }
      /* inverse-distance split (sem_kernel.cu:2051-2065) */
      double inv[3], sum = 0;
      for (int k = 0; k < 3; ++k) {
        double e2 = 0;
        for (int a = 0; a < 3; ++a) {
          const double e = rp[a] - s->mk[a][nodes[k]];
          e2 += e * e;
        }
        inv[k] = e2 > 0 ? 1.0 / sqrt(e2) : 1e30;
        sum += inv[k];
      }
      for (int k = 0; k < 3; ++k) w[k] = inv[k] / sum;
      s->total_bonds++;
    }
    for (int a = 0; a < 3; ++a) {
      s->rec_force[3 * r + a] = F[a];
      s->rec_w[3 * r + a] = w[a];
    }
  }
  for (int l = 0; l < s->n_lig; ++l) s->lig_claim[l] = INT_MAX;
}

/* ---- forces ---- */
static void bend_chain(orc_state *s);

static void stage_forces(orc_state *s) {
  const pfg_params *p = &s->p;
  double *X = s->mk[0], *Y = s->mk[1], *Z = s->mk[2];
  double *FX = s->mk[6], *FY = s->mk[7], *FZ = s->mk[8];
  /* detach test (semLB.cu:451-455) after this step's bond update */
  if (s->n_rec > 0 && p->detach_rate > 0 && s->total_bonds == 0 && !s->detached) {
    const double u = pfg_uniform(p->seed, s->step, 0u, PFG_RNG_DETACH);
    if (u < (double)p->detach_rate) {
      s->detached = 1;
      s->detach_step = s->step;
    }
  }
  for (int i = 0; i < s->n; ++i) FX[i] = FY[i] = FZ[i] = 0;
  /* Morse over all pairs of elements of one platelet within the cutoff.
   * The cutoff decision is the device's: the squared distance of the
   * FP32-rounded positions in FP32, in the same order of operations, against
   * the FP32 squared cutoff, so that both take the same pairs while their
   * positions agree to FP32; pairs within 1e-4 of the cutoff, which a
   * rounding difference can switch, are counted.  The force is FP64. */
  if (p->morse_u0 != 0) {
    const double U0 = p->morse_u0, rho = p->morse_rho, re = p->morse_req;
    const float cut2f = p->morse_cutoff * p->morse_cutoff;
    const float eps2f = 2.0f * p->morse_cutoff * 1e-4f;
    for (int q = 0; q < s->n_plt; ++q)
      for (int i = s->plt_elem_start[q]; i < s->plt_elem_start[q + 1]; ++i)
        for (int j = s->plt_elem_start[q]; j < s->plt_elem_start[q + 1]; ++j) {
          if (j == i) continue;
          const float dxf = (float)X[j] - (float)X[i], dyf = (float)Y[j] - (float)Y[i],
                      dzf = (float)Z[j] - (float)Z[i];
          const float r2f = fmaf(dzf, dzf, fmaf(dyf, dyf, dxf * dxf));
          if (fabsf(r2f - cut2f) < eps2f) {
            s->near_cut++;
            s->plt_near_cut[q]++;
          }
          if (r2f >= cut2f || r2f == 0.f) continue;
          const double dx = X[j] - X[i], dy = Y[j] - Y[i], dz = Z[j] - Z[i];
          const double r = sqrt(dx * dx + dy * dy + dz * dz);
          /* U = U0 (e^{-2 rho (r - re)} - 2 e^{-rho (r - re)}) times the
           * smoothstep S that falls to 0 over [rc - taper, rc];
           * F_i = -d(U S)/dr (x_i - x_j)/r */
          const double em = exp(-rho * (r - re));
          const double U = U0 * (em * em - 2.0 * em);
          double mdUdr = 2.0 * rho * U0 * (em - em * em); /* -dU/dr */
          const double r1 = (double)p->morse_cutoff - (double)p->morse_taper;
          if (p->morse_taper > 0 && r > r1) {
            const double t = (r - r1) / (double)p->morse_taper;
            const double S = 1.0 - t * t * (3.0 - 2.0 * t);
            const double dS = 6.0 * t * (t - 1.0) / (double)p->morse_taper;
            mdUdr = mdUdr * S + U * dS;
          }
          FX[i] += mdUdr * dx / r;
          FY[i] += mdUdr * dy / r;
          FZ[i] += mdUdr * dz / r;
        }
  }
  /* mesh-edge Hookean springs (sem_kernel.cu:356-396) */
  if (p->k_spring != 0) {
    for (int e = 0; e < s->n_edges; ++e) {
      const int a = s->edge[2 * e], b = s->edge[2 * e + 1];
      const double dx = X[b] - X[a], dy = Y[b] - Y[a], dz = Z[b] - Z[a];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r == 0) continue;
      const double c = (double)p->k_spring * (r - s->edge_l0[e]) / r;
      FX[a] += c * dx;
      FY[a] += c * dy;
      FZ[a] += c * dz;
      FX[b] -= c * dx;
      FY[b] -= c * dy;
      FZ[b] -= c * dz;
    }
  }
  /* area / volume constraints, sem_ForceAreaVol_kernel (sem_kernel.cu:1640-1749),
   * in the reference's expanded form */
  if (p->ks != 0 || p->kv != 0) {
    for (int q = 0; q < s->n_plt; ++q) {
      const double Sall = s->plt_S[q], S0all = s->plt_S0[q], V = s->plt_V[q], V0 = s->plt_V0[q];
      for (int t = s->plt_tri_start[q]; t < s->plt_tri_start[q + 1]; ++t) {
        const int *nd = s->tri + 3 * t;
        double c[3][3], Lm[3][3], cen[3], ne[3];
        for (int k = 0; k < 3; ++k) {
          c[k][0] = X[nd[k]];
          c[k][1] = Y[nd[k]];
          c[k][2] = Z[nd[k]];
        }
        for (int a = 0; a < 3; ++a) {
          Lm[0][a] = c[1][a] - c[2][a];
          Lm[1][a] = c[2][a] - c[0][a];
          Lm[2][a] = c[0][a] - c[1][a];
          cen[a] = (c[0][a] + c[1][a] + c[2][a]) / 3.0;
          ne[a] = s->tri_n[3 * t + a];
        }
        const double Se = s->tri_S[t], S0e = s->tri_S0[t];
        for (int i = 0; i < 3; ++i) {
          double dS[3];
          dS[0] = -0.5 * (Lm[i][1] * ne[2] - Lm[i][2] * ne[1]);
          dS[1] = -0.5 * (Lm[i][2] * ne[0] - Lm[i][0] * ne[2]);
          dS[2] = -0.5 * (Lm[i][0] * ne[1] - Lm[i][1] * ne[0]);
          const double ca = -(double)p->ks * ((Se - S0e) / S0e + (Sall - S0all) / S0all);
          FX[nd[i]] += ca * dS[0];
          FY[nd[i]] += ca * dS[1];
          FZ[nd[i]] += ca * dS[2];
        }
        double dV[3][3];
        const double L1xL0_0 = Lm[1][1] * Lm[0][2] - Lm[1][2] * Lm[0][1];
        const double L1xL0_1 = Lm[1][2] * Lm[0][0] - Lm[1][0] * Lm[0][2];
        const double L1xL0_2 = Lm[1][0] * Lm[0][1] - Lm[1][1] * Lm[0][0];
        for (int i = 0; i < 3; ++i) {
          dV[i][0] = -L1xL0_0 / 3.0 - Lm[i][2] * cen[1] + Lm[i][1] * cen[2];
          dV[i][1] = -L1xL0_1 / 3.0 + Lm[i][2] * cen[0] - Lm[i][0] * cen[2];
          dV[i][2] = -L1xL0_2 / 3.0 - Lm[i][1] * cen[0] + Lm[i][0] * cen[1];
          const double cv = -(double)p->kv * (V - V0) / V0 / 6.0;
          FX[nd[i]] += cv * dV[i][0];
          FY[nd[i]] += cv * dV[i][1];
          FZ[nd[i]] += cv * dV[i][2];
        }
      }
    }
  }
  /* wall repulsion on the lowest element (sem_kernel.cu:409-430) */
  if (p->wall_rep_a != 0) {
    for (int q = 0; q < s->n_plt; ++q) {
      const int e = s->lowest[q];
      if (e < 0) continue;
      double h = Z[e] - p->wall_z;
      if (h <= p->wall_rep_hc) {
        if (h < 0.01 * p->wall_rep_hc) h = 0.01 * (double)p->wall_rep_hc;
        const double ex = exp(-(double)p->wall_rep_kappa * h);
        FZ[e] += (double)p->wall_rep_a * p->wall_rep_kappa * ex / (1.0 - ex);
      }
    }
  }
  /* receptor bond forces onto the triangle vertices */
  for (int r = 0; r < s->n_rec; ++r) {
    const int t = s->rec_tri[r];
    for (int k = 0; k < 3; ++k) {
      const int v = s->tri[3 * t + k];
      FX[v] += s->rec_w[3 * r + k] * s->rec_force[3 * r];
      FY[v] += s->rec_w[3 * r + k] * s->rec_force[3 * r + 1];
      FZ[v] += s->rec_w[3 * r + k] * s->rec_force[3 * r + 2];
    }
  }
  /* fibrin links (fiber.cpp:647-733, network.cpp:378-418) */
  const double sgn = p->fiber_ref_sign ? -1.0 : 1.0;
  for (int l = 0; l < s->n_links; ++l) {
    const int a = s->link[2 * l], b = s->link[2 * l + 1];
    const double l0 = s->link_l0[l];
    const double dx = X[b] - X[a], dy = Y[b] - Y[a], dz = Z[b] - Z[a];
    const double len = sqrt(dx * dx + dy * dy + dz * dz);
    if (len == 0 || l0 == 0) continue;
    double tension;
    if (p->fiber_law == PFG_FIBER_LINEAR) {
      /* factor = Y A / rl, force factor (ol - rl) (network.cpp:364-373) */
      tension = (double)p->fiber_k / l0 * (len - l0);
    } else {
      /* wormlike_chain_force (fiber.cpp:710-733): PerL = ini_l, CC = En/PerL,
       * tl = 1 - (cur-ini)/ini, IP1 = 0.25/tl^2 - 0.25 + (cur-ini)/ini */
      double x = (len - l0) / l0;
      if (x > p->fiber_xmax) x = p->fiber_xmax;
      const double tl = 1.0 - x;
      const double CC = (double)p->fiber_k / l0;
      tension = CC * (0.25 / (tl * tl) - 0.25 + x);
    }
    const double c = sgn * tension / len;
    FX[a] += c * dx;
    FY[a] += c * dy;
    FZ[a] += c * dz;
    FX[b] -= c * dx;
    FY[b] -= c * dy;
    FZ[b] -= c * dz;
  }
  /* platelet-platelet contact (sem_kernel.cu:436-509): for every two
   * platelets the closest element pair, the first in (i, j) order on ties,
   * repels with a kappa e^{-kappa h}/(1 - e^{-kappa h}), h = max(d, hc/100),
   * below the cutoff hc */
  if (p->contact_a != 0) {
    const double hc = p->contact_hc, ka = p->contact_kappa, A = p->contact_a;
    for (int q1 = 0; q1 < s->n_plt; ++q1)
      for (int q2 = q1 + 1; q2 < s->n_plt; ++q2) {
        double best = INFINITY;
        int bi = -1, bj = -1;
        for (int i = s->plt_elem_start[q1]; i < s->plt_elem_start[q1 + 1]; ++i)
          for (int j = s->plt_elem_start[q2]; j < s->plt_elem_start[q2 + 1]; ++j) {
            const double dx = X[i] - X[j], dy = Y[i] - Y[j], dz = Z[i] - Z[j];
            const double d = sqrt(dx * dx + dy * dy + dz * dz);
            if (d < best) {
              best = d;
              bi = i;
              bj = j;
            }
          }
        if (bi < 0 || !(best < hc) || !(best > 0)) continue;
        const double h = best > 0.01 * hc ? best : 0.01 * hc;
        const double e = exp(-ka * h);
        const double F = A * ka * e / (1.0 - e) / best;
        const double dx = X[bi] - X[bj], dy = Y[bi] - Y[bj], dz = Z[bi] - Z[bj];
        FX[bi] += F * dx;
        FY[bi] += F * dy;
        FZ[bi] += F * dz;
        FX[bj] -= F * dx;
        FY[bj] -= F * dy;
        FZ[bj] -= F * dz;
      }
  }
  /* membrane bending (sem_Laplace_Kapa_kernel after the area/volume forces) */
  if (p->kbend != 0 && s->bend_ok) {
    bend_chain(s);
    for (int i = 0; i < s->n; ++i) {
      FX[i] += s->fb[3 * i];
      FY[i] += s->fb[3 * i + 1];
      FZ[i] += s->fb[3 * i + 2];
    }
  }
  /* bond reaction on fibrin-node ligands */
  if (s->fib_lig) {
    for (int i = 0; i < s->n; ++i) {
      if (s->kind[i] == PFG_MARKER_ELEMENT) continue;
      const int owner = s->lig_owner[s->n_wlig + i];
      if (owner < 0) continue;
      FX[i] -= s->rec_force[3 * owner];
      FY[i] -= s->rec_force[3 * owner + 1];
      FZ[i] -= s->rec_force[3 * owner + 2];
    }
  }
}

/* ---- membrane bending chain ----
 * calculate_A (sem_kernel.cu:724-820): unit triangle normals
 *   n_t = -(e12 x e13)/|e12 x e13| summed onto the 6 nodes of the triangle;
 * surface_tension (874-900): node normal = normalize(sum / count);
 * calculate_Kapa (950-1055): at each node of the quadratic 6-node triangle
 *   (midpoints at the vertex averages) solve K dx/dxi = dn/dxi,
 *   K dx/deta = dn/deta, K n = 0 row by row and add K^T to the node;
 * calculate_m (1057-1100): K / count, symmetrised, km = trace / 2;
 * Laplace_Kapa (1499-1638): at each vertex, over its ring, the cotangent sum
 *   0.5 (cot a + cot b)(km_j - km_i) and the area of the region bounded by
 *   edge midpoints and circumcentres (an obtuse triangle's long-edge
 *   midpoint instead), both sides by Heron's formula; force
 *   kbend (sum / A) n on the element. */
static const double QN[6][6] = {{2, 2, 4, -3, -3, 1}, {2, 0, 0, -1, 0, 0},
                                {0, 2, 0, 0, -1, 0},  {-4, 0, -4, 4, 0, 0},
                                {0, -4, -4, 0, 4, 0}, {0, 0, 4, 0, 0, 0}};
static const double QXE[6][2] = {{0, 0}, {1, 0}, {0, 1}, {0.5, 0}, {0, 0.5}, {0.5, 0.5}};

/* Forward elimination with partial pivoting and unit pivots on an m x n
 * augmented matrix (gaussian_elimination, sem_kernel.cu:249-284). */
static void gauss_elim(double *g, int m, int n) {
  int i = 0, j = 0;
  while (i < m && j < n - 1) {
    int maxi = i;
    for (int k = i + 1; k < m; ++k)
      if (fabs(g[k * n + j]) > fabs(g[maxi * n + j])) maxi = k;
    if (g[maxi * n + j] != 0) {
      double initial = 0;
      for (int k = j; k < n; ++k) {
        const double b = g[i * n + k];
        g[i * n + k] = g[maxi * n + k];
        g[maxi * n + k] = b;
        if (k == j) initial = g[i * n + j];
        g[i * n + k] /= initial;
      }
      for (int k = i + 1; k < m; ++k) {
        const double f = g[k * n + j];
        for (int l = j; l < n; ++l) g[k * n + l] -= f * g[i * n + l];
      }
      ++i;
    }
    ++j;
  }
}

/* Rows of K from [A1; A3; A5] K[r] = [A2[r]; A4[r]; A6[r]]
 * (solve_linear_eqns, sem_kernel.cu:286-305). */
static void solve_rows(double K[3][3], const double A1[3], const double A2[3],
                       const double A3[3], const double A4[3], const double A5[3],
                       const double A6[3]) {
  for (int r = 0; r < 3; ++r) {
    double g[12];
    for (int c = 0; c < 3; ++c) {
      g[c] = A1[c];
      g[4 + c] = A3[c];
      g[8 + c] = A5[c];
    }
    g[3] = A2[r];
    g[7] = A4[r];
    g[11] = A6[r];
    gauss_elim(g, 3, 4);
    K[r][2] = g[11];
    K[r][1] = g[7] - K[r][2] * g[6];
    K[r][0] = g[3] - K[r][2] * g[2] - K[r][1] * g[1];
  }
}

static double dist3(const double a[3], const double b[3]) {
  return sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
              (a[2] - b[2]) * (a[2] - b[2]));
}

static double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* Ring sums of vertex i for the per-vertex values f (sem_kernel.cu:1533-1628):
 * *sum = sum 0.5 (cot a + cot b)(f_j - f_i), *area = A_i. */
static void ring_laplace(const orc_state *s, int i, const double *f, double *sum,
                         double *area) {
  const double *X = s->mk[0], *Y = s->mk[1], *Z = s->mk[2];
  const int r0 = s->ring_start[i], c = s->ring_start[i + 1] - r0;
  const double xi[3] = {X[i], Y[i], Z[i]};
  double Ai = 0, sm = 0;
  for (int q = 0; q < c; ++q) {
    const int j = s->ring[r0 + q];
    const int v1 = s->ring[r0 + (q == 0 ? c - 1 : q - 1)];
    const int v2 = s->ring[r0 + (q == c - 1 ? 0 : q + 1)];
    const double xj[3] = {X[j], Y[j], Z[j]}, x1[3] = {X[v1], Y[v1], Z[v1]},
                 x2[3] = {X[v2], Y[v2], Z[v2]};
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
    const double ai = dot3(rij, ri2), a2 = -dot3(ri2, r2j), aj = dot3(rij, r2j),
                 a1 = -dot3(r1j, ri1);
    if (ai < 0) {
      for (int k = 0; k < 3; ++k) ce[k] = c2j[k];
    } else if (a2 < 0) {
      for (int k = 0; k < 3; ++k) ce[k] = cij[k];
    } else if (aj < 0) {
      for (int k = 0; k < 3; ++k) ce[k] = ci2[k];
    } else { /* circumcentre of (xi, xj, x2): two bisector planes and the plane */
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
      gauss_elim(g, 3, 4);
      ce[2] = g[11];
      ce[1] = g[7] - ce[2] * g[6];
      ce[0] = g[3] - ce[2] * g[2] - ce[1] * g[1];
    }
    double L1 = dist3(xi, cij), L2 = dist3(ce, cij);
    const double L3 = dist3(xi, ce);
    double P = 0.5 * (L1 + L2 + L3);
    Ai += sqrt(P * (P - L1) * (P - L2) * (P - L3));
    L1 = dist3(xi, ci2);
    L2 = dist3(ce, ci2);
    P = 0.5 * (L1 + L2 + L3);
    Ai += sqrt(P * (P - L1) * (P - L2) * (P - L3));
    const double alpha = 1.0 / tan(acos(a1 / (dist3(xi, x1) * dist3(x1, xj))));
    const double beta = 1.0 / tan(acos(a2 / (dist3(xi, x2) * dist3(x2, xj))));
    sm += 0.5 * (alpha + beta) * (f[j] - f[i]);
  }
  *sum = sm;
  *area = Ai;
}

/* Node normals, mean curvature km at every node and the bending force fb of
 * every element vertex from the current positions. */
static void bend_chain(orc_state *s) {
  const int n = s->n, T = s->n_tri, nodes = n + s->n_enode;
  const double *X = s->mk[0], *Y = s->mk[1], *Z = s->mk[2];
  memset(s->node_n, 0, sizeof(double) * 3 * nodes);
  memset(s->kapa, 0, sizeof(double) * 9 * nodes);
  for (int t = 0; t < T; ++t) {
    const int a = s->tri[3 * t], b = s->tri[3 * t + 1], c = s->tri[3 * t + 2];
    const double e12[3] = {X[b] - X[a], Y[b] - Y[a], Z[b] - Z[a]};
    const double e13[3] = {X[c] - X[a], Y[c] - Y[a], Z[c] - Z[a]};
    double nt[3] = {e12[1] * e13[2] - e12[2] * e13[1], e12[2] * e13[0] - e12[0] * e13[2],
                    e12[0] * e13[1] - e12[1] * e13[0]};
    const double nn = sqrt(dot3(nt, nt));
    for (int k = 0; k < 3; ++k) nt[k] /= -nn;
    for (int k = 0; k < 6; ++k)
      for (int d = 0; d < 3; ++d) s->node_n[3 * s->tri6[6 * t + k] + d] += nt[d];
  }
  for (int v = 0; v < nodes; ++v) {
    if (s->node_cnt[v] == 0) continue;
    double *r = s->node_n + 3 * v;
    for (int d = 0; d < 3; ++d) r[d] /= s->node_cnt[v];
    const double nn = sqrt(dot3(r, r));
    for (int d = 0; d < 3; ++d) r[d] /= nn;
  }
  for (int t = 0; t < T; ++t) {
    double xc[6][3], nc[6][3];
    for (int k = 0; k < 3; ++k) {
      const int v = s->tri[3 * t + k];
      xc[k][0] = X[v];
      xc[k][1] = Y[v];
      xc[k][2] = Z[v];
    }
    for (int d = 0; d < 3; ++d) {
      xc[3][d] = 0.5 * (xc[0][d] + xc[1][d]);
      xc[4][d] = 0.5 * (xc[0][d] + xc[2][d]);
      xc[5][d] = 0.5 * (xc[1][d] + xc[2][d]);
    }
    for (int k = 0; k < 6; ++k)
      for (int d = 0; d < 3; ++d) nc[k][d] = s->node_n[3 * s->tri6[6 * t + k] + d];
    for (int i = 0; i < 6; ++i) {
      const double xi = QXE[i][0], et = QXE[i][1];
      double A1[3] = {0}, A2[3] = {0}, A3[3] = {0}, A4[3] = {0}, A5[3], A6[3] = {0, 0, 0};
      for (int k = 0; k < 6; ++k) {
        const double dxi = 2 * QN[k][0] * xi + QN[k][2] * et + QN[k][3];
        const double det = 2 * QN[k][1] * et + QN[k][2] * xi + QN[k][4];
        for (int d = 0; d < 3; ++d) {
          A1[d] += xc[k][d] * dxi;
          A2[d] += nc[k][d] * dxi;
          A3[d] += xc[k][d] * det;
          A4[d] += nc[k][d] * det;
        }
      }
      for (int d = 0; d < 3; ++d) A5[d] = nc[i][d];
      double K[3][3];
      solve_rows(K, A1, A2, A3, A4, A5, A6);
      double *M = s->kapa + 9 * s->tri6[6 * t + i];
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) M[3 * k + l] += K[l][k];
    }
  }
  for (int v = 0; v < nodes; ++v) {
    if (s->node_cnt[v] == 0) {
      s->km[v] = 0;
      continue;
    }
    double *M = s->kapa + 9 * v, Kv[3][3];
    for (int k = 0; k < 9; ++k) M[k] /= s->node_cnt[v];
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l < 3; ++l) Kv[k][l] = 0.5 * (M[3 * k + l] + M[3 * l + k]);
    s->km[v] = 0.5 * (Kv[0][0] + Kv[1][1] + Kv[2][2]);
  }
  for (int i = 0; i < n; ++i) {
    s->fb[3 * i] = s->fb[3 * i + 1] = s->fb[3 * i + 2] = 0;
    if (!s->ring_ok[i]) continue;
    double sum, Ai;
    ring_laplace(s, i, s->km, &sum, &Ai);
    const double c = (double)s->p.kbend / Ai * sum;
    for (int d = 0; d < 3; ++d) s->fb[3 * i + d] = c * s->node_n[3 * i + d];
  }
}

int orc_bend(orc_state *s, double *km, double *fb, double *normals) {
  if (!s->bend_ok) return 0;
  bend_chain(s);
  if (km) memcpy(km, s->km, sizeof(double) * s->n);
  if (fb) memcpy(fb, s->fb, sizeof(double) * 3 * s->n);
  if (normals) memcpy(normals, s->node_n, sizeof(double) * 3 * s->n);
  return 1;
}

int orc_ring_laplacian(orc_state *s, const double *f, double *out, double *area) {
  if (!s->bend_ok) return 0;
  for (int i = 0; i < s->n; ++i) {
    double sum = 0, Ai = 0;
    if (s->ring_ok[i]) ring_laplace(s, i, f, &sum, &Ai);
    out[i] = Ai > 0 ? sum / Ai : 0;
    if (area) area[i] = Ai;
  }
  return 1;
}

static void stage_spread(orc_state *s) {
  orc_ib_spread(&s->L, s->n, s->mk[0], s->mk[1], s->mk[2], s->mk[6], s->mk[7], s->mk[8],
                s->coupled, s->p.ib_force_scale, s->F);
}

static void stage_lb(orc_state *s) {
  if (s->order == 0) {
    orc_lb_step(&s->L, s->f, s->f2, s->F, s->F, s->u);
  } else {
    /* reference order: collision velocity from the previous step's force
     * (semLB.cu:374-424); IB velocity from the streamed state and F(n) */
    orc_lb_step(&s->L, s->f, s->f2, s->Fprev, s->F, NULL);
    const int64_t N = s->N;
    pull_moments(&s->L, s->f2, s->F, NULL, s->u, s->u + N, s->u + 2 * N);
    memcpy(s->Fprev, s->F, sizeof(double) * 3 * (size_t)N);
  }
  double *t = s->f;
  s->f = s->f2;
  s->f2 = t;
}

static int bbk_element(const orc_state *s, int i) {
  return s->kind[i] == PFG_MARKER_ELEMENT && s->p.sem_integrator == PFG_SEM_BBK;
}

static void stage_interp(orc_state *s) {
  const int n = s->n;
  if (s->p.sem_integrator == PFG_SEM_BBK) {
    /* elements keep their own velocity; the fluid's goes to uib */
    double *v = (double *)xcalloc((size_t)3 * n + 1, sizeof(double));
    orc_ib_interp(&s->L, n, s->mk[0], s->mk[1], s->mk[2], s->u, s->coupled, v, v + n,
                  v + 2 * n);
    for (int i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) {
        if (bbk_element(s, i)) s->uib[3 * i + a] = v[(size_t)a * n + i];
        else s->mk[3 + a][i] = v[(size_t)a * n + i];
      }
    free(v);
  } else {
    orc_ib_interp(&s->L, n, s->mk[0], s->mk[1], s->mk[2], s->u, s->coupled, s->mk[3],
                  s->mk[4], s->mk[5]);
  }
  for (int i = 0; i < n; ++i)
    if (s->kind[i] == PFG_MARKER_FIXED) s->mk[3][i] = s->mk[4][i] = s->mk[5][i] = 0;
}

static void stage_update(orc_state *s) {
  const pfg_params *p = &s->p;
  const double dt = p->marker_dt;
  s->step++; /* the step index advances with the marker update */
  for (int i = 0; i < s->n; ++i) {
    if (s->kind[i] == PFG_MARKER_FIXED) continue;
    if (s->kind[i] == PFG_MARKER_FIBER && p->fiber_move == PFG_FIBER_KICK) {
      /* network::solverkick (network.cpp:839-881) then x += v dt */
      const double g = p->fiber_gamma;
      const double vdt = exp(-g * dt);
      const double fdt = g > 0 ? (1.0 - vdt) / g : dt;
      /* thermal noise of the exact Ornstein-Uhlenbeck step, unit mass (the
       * BBK form of sem_kernel.cu.orig:251-266), drawn for the step being
       * completed */
      double z[3] = {0, 0, 0};
      const double sig = p->kick_temp > 0 ? sqrt((double)p->kick_temp * (1.0 - exp(-2.0 * g * dt)))
                                          : 0.0;
      if (sig > 0) pfg_normal3(p->seed, s->step - 1, (uint32_t)i, PFG_RNG_KICK, z);
      for (int a = 0; a < 3; ++a) {
        s->mk[3 + a][i] = vdt * s->mk[3 + a][i] + s->mk[6 + a][i] * fdt + sig * z[a];
        s->mk[a][i] += s->mk[3 + a][i] * dt;
      }
      continue;
    }
    if (bbk_element(s, i)) {
      /* Langevin leapfrog (verlet_1 / verlet_3, sem_kernel.cu.orig:251-266,
       * 547-560): two exact Ornstein-Uhlenbeck half steps of length dt/2
       * with this step's force F,
       *   v <- e^{-g dt/2} v + F (1 - e^{-g dt/2}) / (g m) + sqrt(T/m (1 - e^{-g dt})) xi,
       * then x += dt ((1 - w) v + w u) with u the fluid velocity here */
      const double g = p->sem_gamma, m = p->sem_mass, h = 0.5 * dt;
      const double vdt = exp(-g * h);
      const double fdt = g > 0 ? (1.0 - vdt) / g : h;
      double za[3] = {0, 0, 0}, zb[3] = {0, 0, 0}, sig = 0;
      if (p->sem_temp > 0 && g > 0) {
        sig = sqrt((double)p->sem_temp / m * (1.0 - exp(-g * dt)));
        pfg_normal3(p->seed, s->step - 1, (uint32_t)i, PFG_RNG_BBK_A, za);
        pfg_normal3(p->seed, s->step - 1, (uint32_t)i, PFG_RNG_BBK_B, zb);
      }
      const double w = p->sem_fluid_weight;
      for (int a = 0; a < 3; ++a) {
        double v = s->mk[3 + a][i];
        v = vdt * v + fdt * s->mk[6 + a][i] / m + sig * za[a];
        v = vdt * v + fdt * s->mk[6 + a][i] / m + sig * zb[a];
        s->mk[3 + a][i] = v;
        s->mk[a][i] += dt * ((1.0 - w) * v + w * s->uib[3 * i + a]);
      }
      continue;
    }
    /* sem_update_position_kernel (sem_kernel.cu:2112-2138) */
    for (int a = 0; a < 3; ++a) s->mk[a][i] += s->mk[3 + a][i] * dt;
  }
}

static void stage_clear(orc_state *s) {
  memset(s->F, 0, sizeof(double) * 3 * (size_t)s->N);
}

void orc_stage(orc_state *s, int stage) {
  switch (stage) {
    case PFG_STAGE_CELLS: stage_cells(s); break;
    case PFG_STAGE_MEMBRANE: stage_membrane(s); break;
    case PFG_STAGE_BONDS: stage_bonds(s); break;
    case PFG_STAGE_FORCES: stage_forces(s); break;
    case PFG_STAGE_SPREAD: stage_spread(s); break;
    case PFG_STAGE_LB: stage_lb(s); break;
    case PFG_STAGE_INTERP: stage_interp(s); break;
    case PFG_STAGE_UPDATE: stage_update(s); break;
    case PFG_STAGE_CLEAR: stage_clear(s); break;
    default: break;
  }
}

void orc_step(orc_state *s) {
  if (s->n == 0) {
    stage_lb(s);
    return;
  }
  for (int st = 0; st < PFG_STAGE_COUNT; ++st) orc_stage(s, st);
}

double *orc_f(orc_state *s) { return s->f; }
double *orc_force_field(orc_state *s) { return s->F; }
double *orc_vel_field(orc_state *s) { return s->u; }
double *orc_marker(orc_state *s, int which) { return (which >= 0 && which < 9) ? s->mk[which] : NULL; }
double *orc_membrane(orc_state *s, int which) {
  switch (which) {
    case 0: return s->plt_S;
    case 1: return s->plt_V;
    case 2: return s->tri_S;
    case 3: return s->tri_n;
  }
  return NULL;
}
int *orc_lowest(orc_state *s) { return s->lowest; }
int64_t *orc_plt_near_cutoff(orc_state *s) { return s->plt_near_cut; }
int *orc_rec_bond(orc_state *s) { return s->rec_bond; }
int *orc_lig_owner(orc_state *s) { return s->lig_owner; }
double *orc_rec_force(orc_state *s) { return s->rec_force; }
int *orc_cells(orc_state *s, int which, int *nb3) {
  if (nb3)
    for (int k = 0; k < 3; ++k) nb3[k] = s->nb[k];
  return which == 0 ? s->bin_start : s->items;
}
void orc_counters(orc_state *s, int64_t out[6]) {
  out[0] = s->total_bonds;
  out[1] = s->detached;
  out[2] = s->detach_step;
  out[3] = s->near;
  out[4] = s->events;
  out[5] = s->near_cut;
}
