/*
 * pfg.h -- C ABI of the MI355X coupled platelet-fibrin step.
 *
 * One handle owns one lattice (D3Q19, populations in HBM as 64-cell x-blocks
 * of 19 direction chunks), one marker set
 * (platelet SEM elements and fibrin nodes), their topology and one HIP stream.
 * Every call returns a pfg_status; nothing calls exit().  Handles are
 * thread-compatible (one thread per handle at a time), not thread-safe.
 *
 * The entry points replace the reference's CUDA glue layer
 * (semLB.cu, `extern "C"` block at semLB.cu:48-1036):
 *
 *   pfg_create / pfg_destroy          fluid_allocGPUKernel  semLB.cu:111
 *                                     sem_allocGPUKernel    semLB.cu:749
 *                                     fiber_allocGPUKernel  semLB.cu:562
 *                                     *_releaseGPUKernel    semLB.cu:534,708,1005
 *   pfg_init_equilibrium, pfg_upload_f, pfg_upload_flags
 *                                     fluid_initGPUKernel(aFlag=1) semLB.cu:154
 *   pfg_load_model                    sem_initGPUKernel     semLB.cu:865
 *                                     fiber_init_GPUKernel  semLB.cu:609
 *   pfg_step                          sem_invokeGPUKernel_Force semLB.cu:441
 *                                     + fluid_invokeGPUKernel   semLB.cu:303
 *   pfg_download_*                    fluid_initGPUKernel(aFlag=0) semLB.cu:154
 *                                     sem_copyGPUKernel     semLB.cu:979
 *                                     fiber_copy_GPUKernel  semLB.cu:669
 *
 * Units: everything on the device is in lattice units (dx = dt = 1).  The
 * reference's physical-unit parameter block hostSEMparameters[0..19]
 * (gpuSEM_LB.c:1206-1233) is mapped onto pfg_params by pfg_params_from_sem().
 *
 * Array layouts (host side, caller owned):
 *   lattice scalar field  [nz][ny][nx], x fastest (reference: pitched [z][y][x]
 *                          planes, semLB.cu:127-135)
 *   distributions          [19][nz][ny][nx], direction order of LB_kernel.cu:40-47
 *   markers                SoA arrays of length n_markers
 */
#ifndef PFG_H
#define PFG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pfg_status {
  PFG_OK = 0,
  PFG_ERR_ARG = 1,      /* invalid argument (sizes, null pointers, modes)        */
  PFG_ERR_ALLOC = 2,    /* device or host allocation failed                      */
  PFG_ERR_DEVICE = 3,   /* no usable HIP device / HIP runtime error              */
  PFG_ERR_LAUNCH = 4,   /* kernel launch or execution error                      */
  PFG_ERR_STATE = 5,    /* call out of order (e.g. step before pfg_load_model)   */
  PFG_ERR_IO = 6        /* file read/write failure (host driver helpers)          */
} pfg_status;

/* Storage format of the distributions in HBM; collision arithmetic is FP32. */
typedef enum pfg_store {
  PFG_STORE_F32 = 0,    /* (f_i - w_i) as float: 76 B per cell per copy          */
  PFG_STORE_F16 = 1,    /* (f_i - w_i) * 2^scale as IEEE binary16: 38 B/cell     */
  PFG_STORE_BF16 = 2    /* (f_i - w_i) as bfloat16: 38 B/cell                    */
} pfg_store;

/* Boundary treatment along one lattice axis. */
typedef enum pfg_wall {
  PFG_WALL_PERIODIC = 0, /* periodic wrap (LB_kernel.cu:1347-1352)                */
  PFG_WALL_BOUNCE = 1,   /* halfway bounce-back walls half a cell outside cells 0
                            and n-1, each wall with its own velocity
                            (LB_kernel_old_08102011.cu:1155-1224)                 */
  PFG_WALL_VELOCITY = 2  /* z only: inflow/outflow faces z = 0 and z = nz-1 with
                            the velocity wall_u_lo[2] / wall_u_hi[2] times the
                            duct profile 4s(L-s)/L^2 (s = i + 1/2) over each
                            bounded transverse axis, imposed by Zou/He on the
                            face cells; face cells next to a transverse wall
                            are set to equilibrium (rho 1) at that velocity
                            (fluid3d_in_out_flow_boundary_kernel,
                            LB_kernel_old_08102011.cu:586-1090)                   */
} pfg_wall;

/* Force law of the fibrin links. */
typedef enum pfg_fiber_law {
  PFG_FIBER_WLC = 0,     /* worm-like chain, fiber.cpp:710-733 / network.cpp:378-403 */
  PFG_FIBER_LINEAR = 1   /* Hookean, network.cpp:352-374 (#if 0 branch)            */
} pfg_fiber_law;

/* How fibrin nodes move. */
typedef enum pfg_fiber_move {
  PFG_FIBER_IB = 0,      /* advected with the interpolated fluid velocity, like
                            platelet elements (sem_kernel.cu:2087-2148)           */
  PFG_FIBER_KICK = 1     /* damped kick v = e^{-g dt} v + F (1-e^{-g dt})/g, then
                            x += v dt (network.cpp:839-881); not fluid-coupled    */
} pfg_fiber_move;

/* How platelet SEM elements move. */
typedef enum pfg_sem_integrator {
  PFG_SEM_ADVECT = 0,    /* with the interpolated fluid velocity, x += V dt
                            (sem_update_position_kernel, sem_kernel.cu:2087-2148) */
  PFG_SEM_BBK = 1        /* Langevin leapfrog with the BBK exact-friction half
                            kicks of verlet_1 / verlet_3 (sem_kernel.cu.orig:
                            224-388, 519-568): per step, with F the force of
                            this step, two half kicks
                              v <- e^{-g dt/2} v + F (1 - e^{-g dt/2})/(g m)
                                   + sqrt(T/m (1 - e^{-g dt})) xi
                            (the first completes the previous step, the second
                            starts this one; xi standard normals from the
                            counter RNG), then the drift
                              x += dt ((1 - w) v + w u),
                            u the interpolated fluid velocity, w =
                            sem_fluid_weight (verlet_1's coupled drift is
                            w = 1/2, sem_kernel.cu.orig:377-379; w = 0: free
                            Langevin dynamics).  The stored velocity is the
                            half-step one; loaded velocities are taken as the
                            velocity half a step before the first step.      */
} pfg_sem_integrator;

/* Receptor-ligand rate law. */
typedef enum pfg_bond_law {
  PFG_BOND_BELL = 0,     /* formation at k_on within reach; rupture at
                            k_off0 e^{bell_xb_kT max(tension, 0)}              */
  PFG_BOND_REFERENCE = 1 /* the force-dependent rates of
                            sem_platelet_wall_kernel (sem_kernel.cu:1852-1863):
                            with w = |bond_l0 - r| and f = bond_k w,
                            formation k_on e^{f (x_b - w/2) / kT} toward the
                            candidate ligand at distance r, rupture
                            k_off0 e^{x_b f / kT}, x_b / kT = bell_xb_kT,
                            kT = bell_kT                                      */
} pfg_bond_law;

/* Marker kinds. */
enum {
  PFG_MARKER_ELEMENT = 0, /* platelet SEM element                                 */
  PFG_MARKER_FIBER = 1,   /* free fibrin node                                     */
  PFG_MARKER_FIXED = 2    /* anchored fibrin node (fiber.cpp:350-388): never moves,
                             never spread                                          */
};

typedef struct pfg_params {
  /* ---- lattice ---- */
  int32_t nx, ny, nz;          /* cells per axis                                  */
  int32_t store;               /* pfg_store                                       */
  int32_t f16_scale_log2;      /* PFG_STORE_F16: deviations scaled by 2^this      */
  int32_t wall[3];             /* pfg_wall per axis x, y, z                       */
  float wall_u_lo[3][3];       /* [axis][component] velocity of the lower wall    */
  float wall_u_hi[3][3];       /* [axis][component] velocity of the upper wall    */
  float tau;                   /* BGK relaxation time (reference OMEGA, [13])     */
  float body_force[3];         /* uniform body force density, lattice units      */
  int32_t use_flags;           /* 1: a solid-cell flag array is uploaded          */
  /* ---- immersed boundary ---- */
  float ib_force_scale;        /* marker force -> lattice force factor (F0 chain
                                  of LB_kernel.cu:1158, 1298); 1 in lattice units */
  float marker_dt;             /* position update dt (reference dt/T0,
                                  sem_kernel.cu:2111); 1 in lattice units         */
  /* ---- SEM element mechanics ---- */
  /* Morse U = U0 (e^{-2 rho (r-req)} - 2 e^{-rho (r-req)}) between elements of
   * one platelet, multiplied by a smoothstep that falls from 1 at
   * cutoff - taper to 0 at cutoff (force continuous at the cutoff) */
  float morse_u0, morse_rho, morse_req, morse_cutoff, morse_taper;
  float k_spring;              /* mesh-edge Hookean springs (reference k_spring [2],
                                  sem_kernel.cu:356-396)                          */
  float ks, kv;                /* area / volume constraints ([17], [18],
                                  sem_kernel.cu:1640-1749)                        */
  float wall_rep_a, wall_rep_kappa, wall_rep_hc; /* wall repulsion on the lowest
                                  element of each platelet (sem_kernel.cu:409-430):
                                  f = a*kappa*e^{-kappa h}/(1-e^{-kappa h}), h<=hc  */
  float wall_z;                /* z of the ligand-carrying wall (bottom wall plane) */
  /* ---- receptor-ligand bonds (Bell model) ---- */
  float k_on;                  /* formation rate within reach (reference Kf0 [3])  */
  float k_off0;                /* unstressed rupture rate (reference Kr0 [4])      */
  float bond_k, bond_l0;       /* bond spring constant and rest length            */
  float bond_reach;            /* a free receptor can bind ligands within this    */
  float bell_xb_kT;            /* Bell force sensitivity x_beta / kBT              */
  float detach_rate;           /* detach hazard per step when no bond holds
                                  (reference gamma = SHEAR_RATE*dt, semLB.cu:451)  */
  uint64_t seed;               /* counter-RNG key (reference RANDOM_SEED [1])      */
  /* ---- fibrin ---- */
  int32_t fiber_law;           /* pfg_fiber_law                                    */
  int32_t fiber_move;          /* pfg_fiber_move                                   */
  float fiber_k;               /* WLC: kB*T of fiber.cpp:722 (CC = fiber_k / l0, the
                                  persistence length being l0); linear: Y*A       */
  float fiber_xmax;            /* WLC strain clamp x=(l-l0)/l0 <= xmax (< 1)        */
  float fiber_gamma;           /* damping for PFG_FIBER_KICK                       */
  int32_t fiber_ref_sign;      /* 1: apply link forces with the reference's sign
                                  (network.cpp:405-411 / fiber.cpp:693-695 push the
                                  ends of a stretched link apart); 0: restoring    */
  /* ---- cell lists ---- */
  int32_t bin_cells;           /* bin edge in lattice cells (>= every cutoff)      */
  /* ---- reference parameter block, kept for output/provenance ---- */
  float sem[20];               /* hostSEMparameters[0..19], gpuSEM_LB.c:1206-1233  */
  /* ---- membrane bending ---- */
  float kbend;                 /* Kbend / F0 of the curvature chain
                                  (sem_kernel.cu:874-1128, 1499-1638): each
                                  element gets kbend (1/A_i) (Laplacian of the
                                  mean curvature) n_i; 0: off. Needs closed,
                                  consistently wound platelet meshes.          */
  /* ---- platelet-platelet contact ---- */
  float contact_a, contact_kappa; /* between every two platelets, the closest pair
                                  of elements (lowest indices on ties) repels
                                  with a kappa e^{-kappa h} / (1 - e^{-kappa h}),
                                  h = max(distance, contact_hc / 100), when the
                                  distance is below contact_hc (the non-bonded
                                  term of sem_Force_kernel, sem_kernel.cu:436-509,
                                  on the cell lists); contact_a = 0: off       */
  float contact_hc;            /* <= bin_cells                                   */
  /* ---- thermal noise ---- */
  float kick_temp;             /* fibrin nodes moved by the kick: BBK Langevin
                                  noise xi sqrt(2 T gamma (1 - e^{-2 gamma dt})
                                  / (2 gamma)) per velocity component and step
                                  (sem_kernel.cu.orig:251-266, unit mass), from
                                  the counter RNG; 0: off                       */
  /* ---- SEM element integration ---- */
  int32_t sem_integrator;      /* pfg_sem_integrator                              */
  float sem_mass;              /* element mass (PFG_SEM_BBK)                       */
  float sem_gamma;             /* Langevin friction rate g (PFG_SEM_BBK)           */
  float sem_temp;              /* k_B T of the Langevin noise (PFG_SEM_BBK), 0: off */
  float sem_fluid_weight;      /* w of the BBK drift                              */
  /* ---- bond law ---- */
  int32_t bond_law;            /* pfg_bond_law                                     */
  float bell_kT;               /* kT of PFG_BOND_REFERENCE                         */
} pfg_params;

/* Host description of the immersed structures; all arrays caller-owned and only
 * read during pfg_load_model.  Counts may be zero (then pointers may be NULL). */
typedef struct pfg_model {
  int32_t n_markers;
  const float *x, *y, *z;          /* unwrapped positions                         */
  const int32_t *kind;             /* PFG_MARKER_*                                 */
  const int32_t *body;             /* platelet index for elements, -1 otherwise    */
  /* platelets: elements and triangles of platelet p are contiguous */
  int32_t n_platelets;
  const int32_t *plt_elem_start;   /* n_platelets + 1 marker offsets               */
  const int32_t *plt_tri_start;    /* n_platelets + 1 triangle offsets             */
  int32_t n_tri;
  const int32_t *tri;              /* 3*n_tri marker indices, outward-wound        */
  const double *tri_S0;            /* rest area per triangle                       */
  const double *plt_S0;            /* rest total area per platelet                 */
  const double *plt_V0;            /* rest volume per platelet (reference sign)    */
  int32_t n_edges;
  const int32_t *edge;             /* 2*n_edges marker indices                     */
  const float *edge_l0;            /* rest lengths                                 */
  /* receptors on triangles, position (1-r1)x0 + r1(1-r2)x1 + r1 r2 x2 */
  int32_t n_receptors;
  const int32_t *rec_tri;
  const float *rec_r1, *rec_r2;
  /* static ligand sites (e.g. vWF on the wall); lig_z NULL means z = wall_z */
  int32_t n_wall_ligands;
  const float *lig_x, *lig_y, *lig_z;
  int32_t fibrin_ligands;          /* 1: free fibrin nodes are ligands as well     */
  /* fibrin links */
  int32_t n_links;
  const int32_t *link;             /* 2*n_links marker indices                     */
  const float *link_l0;
} pfg_model;

/* Per-call step report. */
typedef struct pfg_step_info {
  int64_t step;                    /* steps done since pfg_load_model               */
  int32_t total_bonds;             /* bonds held after the last step                */
  int32_t detached;                /* 1 once the detach test fired                  */
  int64_t detach_step;             /* step index at which it fired, else -1         */
  int64_t near_threshold;          /* bond decisions within eps of their threshold  */
  int64_t bond_events;             /* formations + ruptures so far                  */
  int64_t near_cutoff;             /* Morse pair visits (each ordered pair (i, j))
                                      whose distance lay within 1e-4 of the
                                      cutoff, so far: rounding can switch them     */
  int64_t contact_overflow;        /* contact candidates dropped because an element
                                      already had 4 other platelets within
                                      contact_hc, so far (0 means the closest
                                      pairs are exact)                            */
} pfg_step_info;

/* Stages of one coupled step, in execution order (pfg_run_stage). */
typedef enum pfg_stage {
  PFG_STAGE_CELLS = 0,     /* marker bins, counting sort, per-bin index order     */
  PFG_STAGE_MEMBRANE = 1,  /* triangle areas/normals, FP64 area/volume sums       */
  PFG_STAGE_BONDS = 2,     /* Bell rupture/formation decisions and bond forces    */
  PFG_STAGE_FORCES = 3,    /* per-marker force gather (Morse, springs, area/volume,
                              wall, bonds, fibrin links)                          */
  PFG_STAGE_SPREAD = 4,    /* IB force spreading to the lattice force field       */
  PFG_STAGE_LB = 5,        /* fused stream + bounce-back + moments + BGK/Guo      */
  PFG_STAGE_INTERP = 6,    /* IB velocity interpolation to markers                */
  PFG_STAGE_UPDATE = 7,    /* marker position update                              */
  PFG_STAGE_CLEAR = 8,     /* drop a spread no LB stage consumed: zero the force
                              field tiles this step touched (a coupled step never
                              needs it: the LB stage zeroes the force it reads)  */
  PFG_STAGE_COUNT = 9
} pfg_stage;

typedef struct pfg_handle pfg_handle;

/* ---- library ---- */
const char *pfg_version(void);
const char *pfg_status_string(pfg_status s);
pfg_status pfg_device_count(int *count);
/* sizeof of the ABI structs, for binding checks: 0 pfg_params, 1 pfg_model,
 * 2 pfg_step_info, 3 pfg_scene_opts, 4 pfg_tuning; 0 for an unknown id. */
size_t pfg_struct_size(int which);

/* Default parameters: 64x32x32, tau 0.8, all walls periodic, FP32 storage. */
void pfg_params_default(pfg_params *p);

/* Map the reference's FP32 parameter block hostSEMparameters[0..19]
 * (gpuSEM_LB.c:1206-1233; index names sem_kernel.cu:58-76) and lattice
 * geometry onto p.  Fields the reference does not carry keep their values. */
pfg_status pfg_params_from_sem(const float *host_sem_parameters, float dx,
                               float dt, int lx, int ly, int lz, pfg_params *p);

/* ---- lifetime ---- */
pfg_status pfg_create(const pfg_params *p, int device, pfg_handle **out);
pfg_status pfg_destroy(pfg_handle *h);
pfg_status pfg_get_params(const pfg_handle *h, pfg_params *out);
/* Bytes of device memory the handle holds. */
pfg_status pfg_device_bytes(const pfg_handle *h, size_t *bytes);

/* ---- lattice state ---- */
/* f = f_eq(rho, u) per cell (gpuSEM_LB.c:1234-1258).  NULL rho means 1,
 * NULL velocity means 0. */
pfg_status pfg_init_equilibrium(pfg_handle *h, const float *rho,
                                const float *ux, const float *uy,
                                const float *uz);
/* f = f_eq(1, u(z)) for a velocity that varies along z only, one value per
 * z-plane in each array (nz values; NULL means 0): the reference's linear
 * shear start u_y = shear rate * k * dx / UMAX (gpuSEM_LB.c:1234-1258)
 * without a full-lattice host array. */
pfg_status pfg_init_equilibrium_z(pfg_handle *h, const float *ux_z,
                                  const float *uy_z, const float *uz_z);
pfg_status pfg_upload_f(pfg_handle *h, const float *f);      /* [19][N] */
pfg_status pfg_download_f(pfg_handle *h, float *f);          /* [19][N] */
/* Solid-cell flags, 1 byte per cell, nonzero = solid (needs use_flags). */
pfg_status pfg_upload_flags(pfg_handle *h, const uint8_t *flags);
/* rho and u = (sum c f + F/2)/rho from the current distributions, F the body
 * force (LB_kernel.cu:1108-1199).  Any pointer may be NULL. */
pfg_status pfg_download_fields(pfg_handle *h, float *rho, float *ux, float *uy,
                               float *uz);
/* FP64 totals over fluid cells: out[0] = sum rho, out[1..3] = sum rho u. */
pfg_status pfg_conservation(pfg_handle *h, double out[4]);

/* No counterpart in the reference.  pfg_download_fields for callers that
 * cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc): the same values, staged through a fixed device array of
 * PFG_FIELDS_SCRATCH_FLOATS floats (32 MiB, one per device, held until the
 * process ends) a window of whole z-planes at a time; nothing is allocated
 * or freed.  Limit: nx * ny * (number of non-NULL outputs) must not exceed
 * PFG_FIELDS_SCRATCH_FLOATS, else PFG_ERR_ARG.  All calls on one device
 * share the array: calls that overlap on one device (from several host
 * threads) race and must be serialised by the caller. */
#define PFG_FIELDS_SCRATCH_FLOATS (8 * 1024 * 1024)
pfg_status pfg_download_fields_noalloc(pfg_handle *h, float *rho, float *ux, float *uy,
                                       float *uz);

/* ---- immersed structures ---- */
pfg_status pfg_load_model(pfg_handle *h, const pfg_model *m);
pfg_status pfg_upload_markers(pfg_handle *h, const float *x, const float *y,
                              const float *z);
/* Any pointer may be NULL. */
pfg_status pfg_download_markers(pfg_handle *h, float *x, float *y, float *z,
                                float *vx, float *vy, float *vz, float *fx,
                                float *fy, float *fz);
/* Host-side marker forces/velocities, for stage-level tests. */
pfg_status pfg_upload_marker_forces(pfg_handle *h, const float *fx,
                                    const float *fy, const float *fz);
pfg_status pfg_upload_marker_velocities(pfg_handle *h, const float *vx,
                                        const float *vy, const float *vz);
/* Per-platelet FP64 area and volume sums of the last membrane stage. */
pfg_status pfg_download_membrane(pfg_handle *h, double *S_all, double *V,
                                 double *tri_S);
/* Bond state: ligand index per receptor (-1 free; wall ligands 0..n_wall-1,
 * fibrin-node ligand m is n_wall + m), owner receptor per ligand slot. */
pfg_status pfg_download_bonds(pfg_handle *h, int32_t *rec_bond,
                              int32_t *lig_owner, float *rec_force);
pfg_status pfg_upload_bonds(pfg_handle *h, const int32_t *rec_bond);
/* Cell list of the last CELLS stage: bin_start[n_bins+1], bin_items[n_markers].
 * With the force pass by body on (pfg_set_forces_by_body), every element in
 * such a body and no contact search or tile-gather spreading, nothing reads
 * the elements' entries, and the lists hold the other markers only
 * (bin_start[n_bins] is then their number). */
pfg_status pfg_cell_grid(const pfg_handle *h, int32_t dims[3]);
pfg_status pfg_download_cells(pfg_handle *h, int32_t *bin_start,
                              int32_t *bin_items);
/* Lattice force field [3][N] as the LB stage will read it: valid between the
 * SPREAD and LB stages (the LB stage reads it and leaves zeros). */
pfg_status pfg_download_aux(pfg_handle *h, float *a3n);

/* ---- stepping ---- */
/* Run nsteps coupled steps (one hipGraph launch per step when enabled). */
pfg_status pfg_step(pfg_handle *h, int nsteps, pfg_step_info *info);
pfg_status pfg_run_stage(pfg_handle *h, int stage);
pfg_status pfg_set_step(pfg_handle *h, int64_t step);
pfg_status pfg_set_graph(pfg_handle *h, int enable);
/* Skip the marker pipeline (pure LB stepping). */
pfg_status pfg_set_fluid_only(pfg_handle *h, int fluid_only);
/* enable = 1: the run ends at detach, as the reference's loop does
 * (gpuSEM_LB.c:1433-1446 stops when sem_invokeGPUKernel_Force reports done,
 * semLB.cu:451-455).  Steps launched after the detach test fired change
 * neither the lattice, the markers nor the bonds, and pfg_step_info.step
 * reports the steps that ran.  Default 0: stepping continues after detach. */
pfg_status pfg_set_halt_on_detach(pfg_handle *h, int enable);
/* IB spreading form (LB_kernel.cu:1419-1520 restated four ways): 0 automatic
 * (4 from 32768 markers up when the model has compact platelets, else 1
 * there; 2 below), 1 always staged (a workgroup per run
 * of up to 128 markers of one platelet accumulates them in LDS first),
 * 2 always direct (FP32 atomics per marker and node), 3 tile gather (a
 * wavefront per 4x4x4 tile that this step's stencils meet sums the
 * contributions of the markers on the cell lists near it in LDS and writes
 * the tile; no device-wide atomics; needs the cell lists of the current
 * positions, i.e. PFG_STAGE_CELLS before PFG_STAGE_SPREAD when stepping by
 * stages, and falls back to 1 when a periodic axis is shorter than twice
 * the tile's extent along it), 4 by body (a workgroup per platelet sums
 * all its stencils in an LDS box and stores the nodes of the 4x4x4 tiles
 * no other marker's stencil meets this step, which the force pass claims
 * for it, with plain writes, the others with FP32 atomics; the markers
 * outside such platelets spread as in 1; needs the claims of the current
 * positions, i.e. PFG_STAGE_FORCES before PFG_STAGE_SPREAD when stepping
 * by stages, and without them adds every node with atomics).  Same result
 * up to FP32 summation order. */
pfg_status pfg_set_ib_staging(pfg_handle *h, int mode);
/* Lanes per marker in the SEM/fibrin force pass (sem_Force_kernel,
 * sem_kernel.cu:320-513, with the Morse pairs over the cell lists): 0
 * automatic (64 below 32768 markers, else 4), or 1, 2, 4, 8, 16, 32 or 64.  With
 * fewer than 64 the markers are taken in cell-list order.  Same result up
 * to FP32 summation order. */
pfg_status pfg_set_forces_lanes(pfg_handle *h, int lanes);
/* Force pass by body (sem_Force_kernel's Morse term, sem_kernel.cu:320-513,
 * acts between the elements of one platelet): 0 automatic (on from 32768
 * markers), 1 on, 2 off.  On: a workgroup per platelet of at most 2048
 * elements sorts its positions into a cell list of its own in LDS and sums
 * each element's Morse pairs there, then its other terms; the other
 * markers take the per-marker pass (pfg_set_forces_lanes).  It applies
 * when each platelet's elements are one run of indices; otherwise the
 * per-marker pass takes every marker.  Same forces up to FP32 summation
 * order. */
pfg_status pfg_set_forces_by_body(pfg_handle *h, int mode);
/* enable = 1: within a step, the LB update of the lattice rows that no IB
 * stencil touches runs on a second HIP stream concurrently with the marker
 * pipeline, and the touched rows update after the spreading (semLB.cu:303-438
 * runs these phases one after the other); 0 (default): one LB launch after
 * the spreading.  Same result either way; on MI355X the concurrent LB launch
 * slows the latency-bound marker kernels as much as it hides them. */
pfg_status pfg_set_overlap(pfg_handle *h, int enable);
pfg_status pfg_synchronize(pfg_handle *h);

/* ---- batches: many lattices stepped together ----
 * The reference runs one lattice per process (fluid_allocGPUKernel,
 * semLB.cu:111, a run per seed in gpuSEM_LB.c:1614); a batch holds
 * n_lattices independent lattices (vessel segments) of their own size, walls,
 * wall velocities, tau and body force, with their markers, and advances all
 * of them with one set of launches per step (pfg_step, one hipGraph).
 * Storage format, marker mechanics, bond and fibrin parameters come from
 * params[0].  Lattice-state calls of a single-lattice handle
 * (pfg_init_equilibrium, pfg_upload_f, pfg_download_f, pfg_upload_flags,
 * pfg_download_fields, pfg_conservation, pfg_load_model) return
 * PFG_ERR_STATE on a batch; use the pfg_batch_* forms with a lattice index.
 * Marker arrays of the other calls cover all lattices, lattice after lattice.
 * Step counters (pfg_step_info) and the detach test are batch-wide, and the
 * counter RNG is keyed by the batch-wide receptor index. */
pfg_status pfg_batch_create(const pfg_params *params, int n_lattices, int device,
                            pfg_handle **out);
pfg_status pfg_batch_size(const pfg_handle *h, int *n_lattices);
pfg_status pfg_batch_init_equilibrium(pfg_handle *h, int k, const float *rho,
                                      const float *ux, const float *uy,
                                      const float *uz);
pfg_status pfg_batch_init_equilibrium_z(pfg_handle *h, int k, const float *ux_z,
                                        const float *uy_z, const float *uz_z);
pfg_status pfg_batch_upload_flags(pfg_handle *h, int k, const uint8_t *flags);
/* models[0..n_lattices-1], positions in each lattice's own coordinates */
pfg_status pfg_batch_load_models(pfg_handle *h, const pfg_model *models);
pfg_status pfg_batch_download_fields(pfg_handle *h, int k, float *rho, float *ux,
                                     float *uy, float *uz);
pfg_status pfg_batch_download_markers(pfg_handle *h, int k, float *x, float *y,
                                      float *z, float *vx, float *vy, float *vz,
                                      float *fx, float *fy, float *fz);
pfg_status pfg_batch_conservation(pfg_handle *h, int k, double out[4]);

/* ---- measurement ---- */
/* Time nsteps steps with HIP events on the handle's stream.  ms_total: the
 * whole steps; ms_lb: sum of the LB-kernel launches (events around each LB
 * launch, same stream).  Either pointer may be NULL. */
pfg_status pfg_profile(pfg_handle *h, int nsteps, double *ms_total,
                       double *ms_lb);
/* Mean duration of the fused LB kernel: nlaunch back-to-back LB launches (no
 * other kernels) between two HIP events on the handle's stream; advances the
 * lattice by nlaunch steps. */
pfg_status pfg_time_lb(pfg_handle *h, int nlaunch, double *ms_per_launch);

/* Streaming-copy probe on `device`: copies `bytes` (rounded down to 16 B)
 * between two fresh buffers `reps` times with a float4 kernel and reports the
 * read+write rate in GB/s.  Used to calibrate HBM counters and as the
 * practical bandwidth reference; allocates its own buffers. */
pfg_status pfg_copy_probe(int device, size_t bytes, int reps, double *gbs);

/* LB-kernel launch tuning: lattice cells each lane updates along z (1, 2 or
 * 4; a wavefront updates one 64-cell x-block of that many rows).  Applies to
 * batches too; the overlap mode of a batch runs with 1 only. */
pfg_status pfg_set_lb_tuning(pfg_handle *h, int cells_per_lane);
/* Batches: enable = 1 runs the batched LB update as one persistent launch of
 * the workgroups the CUs hold at once, which take the x-block-row tasks of
 * all lattices from a device work queue (in the order of the workgroup
 * table) until it is empty; 0 (default) launches one workgroup per task.
 * Same results either way.  The queue resets itself at the end of each
 * launch, so captured steps carry no clearing node.  PFG_ERR_STATE when
 * enabled on a single-lattice handle. */
pfg_status pfg_set_lb_persistent(pfg_handle *h, int enable);
/* LB-kernel workgroup order: 0 as the hardware deals workgroups round-robin
 * to the 8 XCDs, 1 each XCD a contiguous range of rows, so that
 * neighbouring rows share its L2.  Same results either way.  Default: 1
 * from 2^27 lattice cells up (single lattice), else 0.  In a batch the
 * ranges are taken from its workgroup table (the persistent launch has
 * per-XCD ranges of its own). */
pfg_status pfg_set_lb_xcd_rows(pfg_handle *h, int enable);

/* Launch shapes chosen by pfg_autotune. */
typedef struct pfg_tuning {
  int lb_cells_per_lane; /* pfg_set_lb_tuning                                */
  int lb_xcd_rows;       /* pfg_set_lb_xcd_rows (0 for batches)              */
  int forces_lanes;      /* pfg_set_forces_lanes: 0 automatic, or 1 .. 64     */
  int ib_staging;        /* pfg_set_ib_staging: 0 automatic, 1, 2, 3, 4      */
  int overlap;           /* pfg_set_overlap                                  */
  double ms_per_step;    /* the chosen shapes                                */
  double ms_default;     /* the automatic shapes                             */
  int shapes_timed;
  int lb_persistent;     /* pfg_set_lb_persistent (batches)                  */
  int forces_by_body;    /* pfg_set_forces_by_body: 0 automatic, 1, 2        */
} pfg_tuning;
/* Times `steps` coupled steps of the handle's own lattices and markers for
 * each launch shape, one knob at a time from the automatic choices (LB cells
 * per lane, LB workgroup order, lanes per marker in the force pass, the force
 * pass by body, the spreading form, the LB overlap, the persistent batched
 * launch), keeps the fastest value of each and sets it on the handle.  The
 * state is saved on the device before and restored after (the handle
 * continues from where it was; the trial steps leave no trace), at the cost
 * of one extra copy of the populations and of the marker state while it
 * runs.  The reference has no counterpart (its launch shapes are fixed,
 * semLB.cu:56-78, sem_kernel.cu:131-193).  PFG_ERR_STATE with halt at
 * detach on. */
pfg_status pfg_autotune(pfg_handle *h, int steps, pfg_tuning *out);

#ifdef __cplusplus
}
#endif

#endif /* PFG_H */
