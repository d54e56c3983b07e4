/*
 * pfg_oracle.h -- FP64 CPU restatement of the coupled platelet-fibrin step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls this
 * code; the tests, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
 * it as the checker.  Single-threaded, deterministic, FP64 throughout except the
 * integer decisions that must match the device bit for bit (cell-list bins are
 * taken from the FP32 positions with floorf, as on the device).
 *
 * Each function names the reference lines it restates.  Parity status of every
 * function (what pins it) is listed in the header comment of pfg_oracle.c.
 */
#ifndef PFG_ORACLE_H
#define PFG_ORACLE_H

#include <stdint.h>

#include "../include/pfg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Lattice description shared by the LB / IB routines. */
typedef struct orc_lattice {
  int nx, ny, nz;
  int periodic[3];
  double wall_u_lo[3][3], wall_u_hi[3][3];
  double tau;
  double g[3];
  const uint8_t *solid; /* NULL: no solid cells */
  int zvel;             /* z faces are Zou/He velocity faces */
} orc_lattice;

void orc_lattice_from_params(const pfg_params *p, const uint8_t *solid,
                             orc_lattice *L);

/* D3Q19 tables of LB_kernel.cu:40-47. */
double orc_w(int d);
int orc_c(int d, int axis);
int orc_bb(int d);

/* f = f_eq(rho, u) (gpuSEM_LB.c:1234-1258); NULL fields mean rho=1, u=0. */
void orc_init_eq(const orc_lattice *L, const double *rho, const double *ux,
                 const double *uy, const double *uz, double *f);

/* One LB update from post-collision fin to post-collision fout:
 * pull-stream (LB_kernel.cu:1312-1359) with halfway bounce-back at walls and
 * solid cells (LB_kernel_old_08102011.cu:1155-1224), moments
 * (LB_kernel.cu:1108-1199) and BGK + Guo collision (LB_kernel.cu:1203-1304).
 * Fu: extra force (3*N, may be NULL) used in u = (sum c f + F/2)/rho;
 * Fs: extra force used in the Guo source.  Fused order: Fu = Fs = F(n);
 * reference order: Fu = F(n-1), Fs = F(n).  The body force L->g is added to
 * both.  u_out (3*N, may be NULL) receives the u used in the collision. */
void orc_lb_step(const orc_lattice *L, const double *fin, double *fout,
                 const double *Fu, const double *Fs, double *u_out);

/* rho and u of the pulled (post-stream) distributions with the body force. */
void orc_fields(const orc_lattice *L, const double *f, double *rho, double *ux,
                double *uy, double *uz);

/* Sum over fluid cells of rho and rho*u of the post-collision state. */
void orc_conservation(const orc_lattice *L, const double *f, double out[4]);

/* Cosine 4-point delta (LB_kernel.cu:1477-1478). */
double orc_phi(double r);

/* IB spreading into F (3*N, accumulated) by marker stencils; coupled[m] != 0
 * marks markers that spread.  orc_ib_spread_scan is the reference's own form:
 * every lattice node visits every marker (LB_kernel.cu:1419-1520). */
void orc_ib_spread(const orc_lattice *L, int n, const double *x, const double *y,
                   const double *z, const double *fx, const double *fy,
                   const double *fz, const int *coupled, double scale, double *F);
void orc_ib_spread_scan(const orc_lattice *L, int n, const double *x,
                        const double *y, const double *z, const double *fx,
                        const double *fy, const double *fz, const int *coupled,
                        double scale, double *F);
void orc_ib_interp(const orc_lattice *L, int n, const double *x, const double *y,
                   const double *z, const double *u, const int *coupled,
                   double *vx, double *vy, double *vz);

/* Cell list: bins of the FP32 positions, counting sort stable in marker
 * index.  nb[3] bins per axis; bin_start[nbins+1], items[n]. */
void orc_cell_list(int n, const float *x, const float *y, const float *z, int nx,
                   int ny, int nz, const int periodic[3], const int nb[3],
                   int *bin_start, int *items);

/* ---- full coupled state ---- */
typedef struct orc_state orc_state;

orc_state *orc_create(const pfg_params *p, const pfg_model *m, const uint8_t *flags);
void orc_destroy(orc_state *s);
/* order 0: fused (device order), 1: reference order (see orc_lb_step). */
void orc_set_order(orc_state *s, int order);
/* OpenMP threads of the lattice loops (n <= 0 keeps the current count);
 * returns the count in use, 1 without OpenMP. */
int orc_set_threads(int n);
void orc_step(orc_state *s);
/* One stage, same numbering as pfg_stage. */
void orc_stage(orc_state *s, int stage);
void orc_set_step(orc_state *s, int64_t step);

/* Array access (pointers into the state; lengths as in pfg_model). */
double *orc_f(orc_state *s);          /* 19*N full distributions            */
double *orc_force_field(orc_state *s);/* 3*N spread force of the step       */
double *orc_vel_field(orc_state *s);  /* 3*N velocity used for IB           */
double *orc_marker(orc_state *s, int which); /* 0..8: x y z vx vy vz fx fy fz */
double *orc_membrane(orc_state *s, int which); /* 0 S_all 1 V 2 tri_S 3 tri_n */
int *orc_lowest(orc_state *s);
/* Morse pair visits within 1e-4 of the cutoff so far, per platelet */
int64_t *orc_plt_near_cutoff(orc_state *s);
int *orc_rec_bond(orc_state *s);
int *orc_lig_owner(orc_state *s);
double *orc_rec_force(orc_state *s);
int *orc_cells(orc_state *s, int which, int *nb3); /* 0 bin_start 1 items */
/* counters: [0] bonds [1] detached [2] detach step [3] near-threshold
 * [4] events [5] Morse pair visits within 1e-4 of the cutoff */
void orc_counters(orc_state *s, int64_t out[6]);
/* Bending chain on the current positions: mean curvature km (n), bending
 * force fb (3n, with the kbend of the parameters) and node normals (3n) of
 * the vertices; 0 when the meshes are not closed (then nothing is written). */
int orc_bend(orc_state *s, double *km, double *fb, double *normals);
/* Cotangent ring Laplacian sum/A of per-vertex values f and the areas A. */
int orc_ring_laplacian(orc_state *s, const double *f, double *out, double *area);

#ifdef __cplusplus
}
#endif

#endif
