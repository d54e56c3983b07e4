/*
 * pfg_internal.h -- device-side state of a pfg_handle and the launchers each
 * translation unit provides.  Not part of the C ABI.
 */
#ifndef PFG_INTERNAL_H
#define PFG_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pfg.h"
#include "pfg_lattice.h"

namespace pfg {

/* Steps captured per chunk graph (even: a chunk ends on the copy of f it
 * started from). */
constexpr int GRAPH_CHUNK = 8;

/* IB work units: at most IB_CHUNK consecutive markers of one body, staged
 * in LDS; when their stencils span at most IB_BOX_CELLS lattice nodes the
 * chunk is summed node by node. */
constexpr int SCAN_SPAN = 8192; /* bins per block of the bin-count scan */
/* When the cell lists leave the elements out (only the fibrin ligand search
 * reads them), their bins are COARSE_BINS bins of the full grid along each
 * axis: a scan over 64 times fewer bins for a few more candidates. */
constexpr int COARSE_BINS = 4;
constexpr int IB_CHUNK = 128;
/* workgroups of the tile-gather spreading launch (each takes listed tiles in
 * turn) */
constexpr int SPREAD_TILE_GRID = 4096;
constexpr int IB_BOX_CELLS = 1024;
/* Body spreading (spreading form 4): the chunks of a compact body (a
 * platelet of at most BODY_MAX_MARKERS markers, its stencils at most
 * BODY_BOX_DIM nodes along each axis) write their LDS sums to a scratch box
 * each; a workgroup per body adds its chunks' boxes over the body's box
 * (at most BODY_BOX_TILES tiles for the tile claims, at most
 * BODY_MAX_CHUNKS chunks) */
constexpr int BODY_MAX_MARKERS = 4096;
constexpr int BODY_BOX_DIM = 256;
constexpr int BODY_BOX_TILES = 1024;
constexpr int BODY_MAX_CHUNKS = 256;
/* Force pass by body: a workgroup per body of at most FB_MAX markers keeps
 * their positions and a cell list of at most FB_BINS bins in LDS for the
 * Morse pairs (which act within a body only) */
constexpr int FB_MAX = 1024;
constexpr int FB_BINS = 2048;
/* its neighbour lists: at most FB_NBR_MAX partners per element within the
 * cutoff plus a skin of FB_SKIN cutoffs */
constexpr int FB_NBR_MAX = 64;
constexpr float FB_SKIN = 0.25f;
/* lanes of the per-body workgroups (force pass, second spreading pass) */
constexpr int BODY_THREADS = 1024;
/* Marker count from which spreading is staged through LDS by default. */
constexpr int IB_STAGED_MIN = 32768;
/* Marker count from which the forces take 4 lanes per marker in bin order
 * (below it, a wavefront per marker). */
constexpr int FORCES_SORTED_MIN = 32768;
/* Receptor count up to which the cell-list/bond launch resolves the bonds
 * itself (its last workgroup takes them all). */
constexpr int RESOLVE_IN_GATHER_MAX = 2048;
/* Marker count below which the counting launch adds the scan's span totals
 * (above it the counting wavefronts meet on few spans and a reduction
 * launch is cheaper). */
constexpr int SPAN_COUNT_MAX = 32768;
/* Lattice cells from which a single-lattice LB launch gives each XCD a
 * contiguous range of rows by default (faster at 268 M cells, slower at
 * 4 M; the autotuner tries both). */
constexpr int64_t XCD_ROWS_MIN_CELLS = (int64_t)1 << 27;
/* Device counters of a handle (Markers::counters). */
constexpr int N_COUNTERS = 16;
/* Morse pairs whose distance lies within this of the cutoff are counted: a
 * rounding difference can switch them. */
constexpr float MORSE_EPS_R = 1e-4f;

/* Everything the fused LB kernel needs, passed by value. */
struct LBArgs {
  const void *fin;        /* post-collision state of the previous step      */
  void *fout;             /* post-collision state of this step              */
  float *aux;             /* [3][plane]: IB force in, zeroed once read      */
  float4 *vel;            /* per cell (ux, uy, uz, 0): IB velocity out       */
  int64_t plane;          /* stride between the components of aux            */
  const int32_t *tile_stamp; /* per tile: the last step that spread into it */
  const uint8_t *solid;   /* nullptr when there are no solid cells          */
  const uint32_t *bbmask; /* per cell: bit d = upstream of d is solid,
                             bit 31 = cell is solid; nullptr without flags  */
  const int64_t *halt;    /* nonzero *halt: leave the lattice unchanged
                             (run halted at detach); nullptr: never          */
  const int32_t *row_stamp; /* per x-block row: the last step whose IB
                             stencils touch the row                          */
  const int64_t *step;    /* the current step (device counter)              */
  int nx, ny, nz;
  int nxb;                /* x-blocks of XBLOCK cells per row                 */
  int64_t block0;         /* first workgroup of this lattice in a batched
                             launch                                          */
  int tnx, tny;           /* tiles per axis                                   */
  int periodic[3];
  float omega;            /* 1/tau                                           */
  float guo;              /* 1 - 1/(2 tau)                                   */
  float g[3];             /* body force                                      */
  float scale, inv_scale; /* FP16 deviation scale                            */
  /* collision constants for the three weight classes (rest, axis, diagonal),
   * in storage units: ws = w * scale, gB = 9 guo ws, gD = 3 guo ws */
  float keep;             /* 1 - omega                                       */
  float om15, om45, om3, guo3; /* 1.5 omega, 4.5 omega, 3 omega, 3 guo        */
  float ws[3], gB[3], gD[3];
  /* 6 w_d (c_d . u_wall) for [axis][0 lower, 1 upper][d] */
  float wall_term[3][2][Q];
  int zvel;               /* z faces are Zou/He velocity faces                */
  int xcd_rows;           /* 1: each XCD takes a contiguous range of rows     */
  float zvel_u[2][3];     /* their velocities (z = 0, z = nz-1)               */
};

/* Lattice side of the IB coupling for one lattice of a batch (a handle
 * with one lattice has one). */
struct MLat {
  float *aux;             /* the lattice's cells of the [3][plane] force      */
  const float4 *vel;      /* the lattice's cells of the velocity field        */
  int32_t *tile_stamp;    /* the lattice's tiles                              */
  uint64_t *tile_owner;   /* the lattice's tiles (body spreading claims)       */
  int tnx, tny, tnz;
  int nx, ny, nz;
  int periodic[3];
  int nb[3];              /* cell-list bins per axis                          */
  int bin0;               /* first global bin of the lattice                  */
  int nbc[3];             /* coarse bins per axis (lists without elements)    */
  int binc0;              /* first global coarse bin of the lattice           */
  int64_t tile0;          /* first global tile of the lattice                 */
  int nxb;                /* x-blocks per row                                 */
  int32_t *row_stamp;     /* the lattice's x-block rows (nullptr: unused)     */
  /* counter-RNG key of the lattice's own draws (bonds, noise): its seed and
   * lattice-local receptor and marker indices, so that a lattice of a batch
   * draws what it draws in a handle of its own */
  uint64_t seed;
  int32_t rec0;           /* first receptor of the lattice                    */
  int32_t marker0;        /* first marker of the lattice                      */
};

/* Tile-gather spreading: a marker's first stencil node (w = 1: IB-coupled),
 * its kernel weights along x, y, z and its force, per cell-list slot. */
struct SpreadRec {
  int4 o;
  float4 wx, wy, wz, f;
};

/* Marker-side device arrays. */
struct Markers {
  int n = 0;
  int32_t *minst = nullptr;      /* lattice of each marker (batches only)    */
  float *x = nullptr, *y = nullptr, *z = nullptr;
  float4 *pos4 = nullptr;        /* x, y, z, Morse tag per marker, this step */
  float *vx = nullptr, *vy = nullptr, *vz = nullptr;
  float *fx = nullptr, *fy = nullptr, *fz = nullptr;
  /* PFG_SEM_BBK: the fluid velocity interpolated at each marker (the
   * element's own velocity stays in vx, vy, vz) */
  float4 *uib = nullptr;
  int32_t *kind = nullptr, *body = nullptr;
  /* platelets */
  int n_plt = 0;
  int32_t *plt_elem_start = nullptr, *plt_tri_start = nullptr;
  int n_tri = 0;
  int32_t *tri = nullptr;        /* 3*n_tri */
  double *tri_S0 = nullptr, *tri_S = nullptr, *tri_V = nullptr;
  /* per (triangle, vertex slot) 3t+s: area gradient -(L_s x n)/2 and the
   * local area strain (S - S0)/S0; per triangle: (x1-x0) x (x2-x0), the
   * volume gradient at any of its vertices (written each step) */
  float4 *tri_g = nullptr, *tri_a2 = nullptr;
  double *plt_S0 = nullptr, *plt_V0 = nullptr, *plt_S = nullptr, *plt_V = nullptr;
  int32_t *plt_lowest = nullptr; /* lowest element per platelet             */
  /* CSR marker -> incident triangles (value = 3 * triangle + vertex slot) */
  int32_t *m2t_start = nullptr, *m2t = nullptr;
  /* membrane bending: triangles as 6-node elements (3 vertices, 3 edge
   * midpoints), ordered vertex rings, node normals and mean curvature */
  int bend = 0;                  /* the chain runs each step                 */
  int n_tedges = 0;              /* edges of the triangles (midpoint nodes)  */
  int32_t *tri_edge = nullptr;   /* 3 per triangle: edges v0v1, v0v2, v1v2   */
  int32_t *edge_tri = nullptr;   /* 2 per edge: 3 * triangle + slot, or -1   */
  int32_t *ring_start = nullptr, *ring = nullptr; /* empty ring: open vertex */
  double *vn = nullptr, *en = nullptr; /* unit normals of vertices, edges     */
  double *km = nullptr;          /* mean curvature per vertex                */
  /* CSR marker -> mesh-edge partners with rest length */
  int n_edges = 0;
  int32_t *m2e_start = nullptr, *m2e_nbr = nullptr;
  float *m2e_l0 = nullptr;
  /* receptors */
  int n_rec = 0;
  int32_t *rec_tri = nullptr;
  float *rec_r1 = nullptr, *rec_r2 = nullptr;
  int32_t *rec_bond = nullptr;   /* ligand id or -1                          */
  int32_t *rec_cand = nullptr;   /* candidate ligand this step or -1         */
  int32_t *rec_release = nullptr;/* 1 if the bond ruptures this step         */
  float *rec_force = nullptr;    /* 3*n_rec force on the receptor            */
  float *rec_w = nullptr;        /* 3*n_rec share of that force per vertex   */
  int32_t *m2r_start = nullptr, *m2r = nullptr; /* marker -> (receptor*3+slot) */
  /* ligands */
  int n_wall_lig = 0;
  int fibrin_lig = 0;
  int n_lig = 0;                 /* n_wall_lig + n (fibrin node ids)         */
  float *lig_x = nullptr, *lig_y = nullptr, *lig_z = nullptr;
  int32_t *lig_owner = nullptr;  /* receptor or -1                           */
  int32_t *lig_claim = nullptr;  /* lowest claiming receptor or INT_MAX      */
  /* static ligands binned on the cell-list grid: lg_start[nbins+1] */
  int32_t *lg_start = nullptr, *lg_items = nullptr;
  float4 *lg_pos = nullptr;       /* per lg_items slot: x, y, z, ligand index */
  float4 *rec_pos = nullptr;      /* receptor positions of this step, lattice
                                     index in .w (int bits)                   */
  /* fibrin links, CSR marker -> (partner, rest length) */
  int n_links = 0;
  int32_t *m2l_start = nullptr, *m2l_nbr = nullptr;
  float *m2l_l0 = nullptr;
  /* cell list */
  int nb[3] = {1, 1, 1};
  int64_t nbins = 0;
  int64_t nbinsc = 0, scan_blocks_c = 0; /* the coarse grid                  */
  int32_t *bin_count = nullptr, *bin_start = nullptr, *bin_items = nullptr,
          *marker_bin = nullptr, *marker_slot = nullptr;
  int32_t *bin_scatter = nullptr; /* items in scatter (atomic slot) order    */
  float4 *sorted_pos = nullptr;   /* per bin_items slot: x, y, z, Morse tag  */
  int32_t *scan_partials = nullptr; /* sums of SCAN_SPAN bins */
  int64_t scan_blocks = 0;
  /* IB chunks: chunk k holds the markers chunk_items[chunk_start[k] ..
   * chunk_start[k+1]-1], at most IB_CHUNK markers of one body (a platelet,
   * or the markers outside platelets) of one lattice that lay close together
   * at load time; the n_direct markers chunk_items[chunk_start[n_chunks] ..]
   * were too far apart and are spread directly */
  int n_chunks = 0, n_direct = 0;
  int n_chunk_items = 0;         /* markers in the chunks (the direct list follows) */
  int32_t *chunk_start = nullptr;
  int32_t *chunk_items = nullptr;
  /* per chunk the marker whose periodic image places its stencils: the
   * first marker of its platelet (body spreading), else its own first */
  int32_t *chunk_ref = nullptr;
  /* tile-gather spreading: the tiles this step's stencils meet, listed once
   * each (global tile index); tile_count[0] tiles listed, [1] workgroups of
   * the gather launch done (the last one empties the list).  Usable when
   * every periodic axis is at least twice the tile's extent along it */
  int32_t *tile_list = nullptr;
  int32_t *tile_count = nullptr;
  /* body spreading: bodies (compact platelets: body_range[4g .. 4g+3] =
   * their markers m0..m1-1 and chunks c0..c1-1; every marker of a body is
   * in one of its chunks), chunk_body (the body of each chunk, -1: none),
   * the chunks' scratch boxes (chunk_sum: 6 x IB_BOX_CELLS floats per chunk;
   * chunk_info: the box's first node and extents per chunk, extent 0: not
   * written this step), and the chunks and direct list of every marker
   * outside the bodies (the chunk lists' format); per tile of all
   * lattices, the step and the body that claimed it alone this step
   * (OWNER_SHARED: more than one contributor), written by the force launch
   * and read by spreading */
  int n_bodies = 0;
  int32_t *body_range = nullptr;
  int32_t *chunk_body = nullptr;
  float *chunk_sum = nullptr;
  int4 *chunk_info = nullptr;
  int n_rchunks = 0, n_rdirect = 0, n_ritems = 0;
  int32_t *rchunk_start = nullptr, *rchunk_items = nullptr;
  uint64_t *tile_owner = nullptr;
  int32_t *body_spread = nullptr;   /* per body: the last step it spread in  */
  int32_t *body_fallback = nullptr; /* per body: the last step one of its
                                       chunks spread with atomics            */
  /* force pass by body: bodies of at most FB_MAX markers (marker ranges
   * fbody_range[2g], fbody_range[2g+1]) when every body's elements form one
   * run of indices, and every other marker (frest_items) */
  int n_fbodies = 0, n_frest = 0;
  int fb_all_elements = 0;       /* every element is in one of those bodies  */
  int32_t *fbody_range = nullptr, *frest_items = nullptr;
  /* per marker: the positions of its body's last list build (NaN: build),
   * its list length and its partners (body-local indices, partner q of
   * marker i at fb_nbr[q * n + i]) */
  float4 *fb_ref = nullptr;
  int16_t *fb_nlen = nullptr;
  int16_t *fb_nbr = nullptr;
  int32_t *fb_done = nullptr;    /* per body: the step whose forces the build
                                    kernel summed (its lists overflowed)     */
  SpreadRec *spread_rec = nullptr; /* per cell-list slot: the marker's record */
  int tile_gather_ok = 0;
  /* platelet-platelet contact: per platelet pair (lo * n_plt + hi) the best
   * proposal (distance bits << 32 | element of the lower platelet), and the
   * pairs proposed this step */
  uint64_t *contact_slot = nullptr;
  int32_t *contact_list = nullptr;
  /* counters (N_COUNTERS): [0] total bonds, [1] detached, [2] near-threshold
   * lo, [3] hi, [4] events lo, [5] events hi, [6] the step index recorded by
   * the first launch of the step, [7] platelet pairs in contact this step,
   * [8] Morse pair visits within MORSE_EPS_R of the cutoff, [9] contact
   * candidates beyond the per-element platelet slots */
  int64_t *counters = nullptr;
  int32_t *ticket = nullptr;      /* workgroups of the gather launch done       */
  int64_t *detach_step = nullptr;
  int64_t *d_step = nullptr;
};

/* Host-side description of one lattice of a handle. */
struct Inst {
  pfg_params p{};          /* lattice and fluid fields apply per lattice     */
  int64_t N = 0;           /* cells                                          */
  int64_t cell0 = 0;       /* first cell in the concatenated cell fields     */
  int nxb = 0;
  int64_t fblk0 = 0;       /* first x-block in the population arrays          */
  int tnx = 0, tny = 0, tnz = 0;
  int64_t ntiles = 0, tile0 = 0;
  int nb[3] = {1, 1, 1};
  int64_t nbins = 0, bin0 = 0;
  int nbc[3] = {1, 1, 1};
  int64_t nbinsc = 0, binc0 = 0;
  int64_t lb_blocks = 0, lb_block0 = 0; /* workgroups in the batched LB launch */
  int32_t rec0 = 0, marker0 = 0;  /* first receptor and marker of the lattice */
};

}  // namespace pfg

struct pfg_handle {
  pfg_params p{};
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t N = 0, plane = 0;
  int nxb = 0;                 /* x-blocks per row in the population layout */
  int esize = 4;
  /* populations, blocked: element (d,x,y,z) at
   * ((x/64 + nxb*(y + ny*z))*19 + d)*64 + x%64 */
  void *f[2] = {nullptr, nullptr};
  int cur = 0;
  /* IB coupling fields: spreading adds the force of a step into aux and
   * stamps the 4x4x4 tiles it touches with the step; the LB update reads
   * the force of stamped tiles, zeroes it there for the next step and
   * writes the velocity the interpolation reads into vel.  The force is
   * kept as three planes, so the three atomic adds of a node go to three
   * cache lines (atomics to one line are serialised); the velocity is one
   * float4 per cell, so a gather of a node reads one line */
  float *aux = nullptr;
  float4 *vel = nullptr;
  int32_t *tile_stamp = nullptr;
  int tnx = 0, tny = 0, tnz = 0;
  int64_t ntiles = 0;
  uint8_t *solid = nullptr;
  uint32_t *bbmask = nullptr;  /* per-cell bounce mask built from solid        */
  double *red = nullptr;       /* reduction partials                        */
  int64_t red_blocks = 0;
  int lb_cpt = 1;              /* LB launch: z-cells per lane               */
  int lb_xcd_rows = 0;         /* LB launch: contiguous rows per XCD        */
  pfg::Markers m;
  bool model_loaded = false;
  bool fluid_only = false;
  bool halt_on_detach = false; /* steps after the detach test fired change nothing */
  int ib_staging = 0;          /* spreading: 0 auto, 1 LDS-staged chunks, 2 direct */
  int forces_lanes = 0;        /* lanes per marker in the forces: 0 auto, 1, 4, 16, 64 */
  int forces_body = 0;         /* force pass by body: 0 auto, 1 on, 2 off     */
  /* LB rows away from IB stencils update on stream2 while the marker
   * pipeline runs; the rows the stencils touch update after spreading.  Off
   * by default: on MI355X the LB launch saturates HBM and the CU slots, and
   * the latency-bound marker kernels slow down as much as they overlap. */
  bool overlap = false;
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int32_t *row_stamp = nullptr;    /* per x-block row of every lattice      */
  int64_t h_step = 0;
  /* lattices: one, or a batch stepped together */
  int K = 1;
  std::vector<pfg::Inst> inst;
  pfg::LBArgs *d_lbargs[2] = {nullptr, nullptr}; /* per starting copy of f, [K] */
  int32_t *lb_block_inst = nullptr;              /* batched LB: lattice per block */
  int64_t lb_blocks = 0;
  /* batched LB: 1 = one persistent launch of resident workgroups taking the
   * table's tasks from a work queue (lb_queue: next task, workgroups done) */
  int lb_persistent = 0;
  uint32_t *lb_queue = nullptr; /* LB_QUEUES range counters, then the done count */
  pfg::MLat *d_lats = nullptr;                   /* [K] */
  int64_t *d_tile0 = nullptr;                    /* [K+1] first tile per lattice */
  std::vector<int32_t> marker0;                  /* [K+1] first marker per lattice */
  bool use_graph = true;
  /* [starting copy of f][0: one step, 1: GRAPH_CHUNK steps] */
  hipGraphExec_t graphs[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  bool graph_valid = false;
  size_t dev_bytes = 0;
};

namespace pfg {

/* lb.hip; k selects a lattice of the handle */
void make_lb_args(const pfg_handle *h, int k, int cur, LBArgs *a);
hipError_t launch_lb(pfg_handle *h);
/* part 1: rows no IB stencil touches this step (on stream s); part 2: the
 * rows they touch, with the IB force (then the copies of f alternate) */
hipError_t launch_lb_part(pfg_handle *h, int part, hipStream_t s);
hipError_t launch_bbmask(pfg_handle *h, int k);
hipError_t launch_lb_fields(pfg_handle *h, int k, float *rho, float *ux, float *uy,
                            float *uz);
/* per_z: the inputs hold one value per z-plane of lattice k */
hipError_t launch_init_eq(pfg_handle *h, int k, const float *rho, const float *ux,
                          const float *uy, const float *uz, bool per_z = false);
hipError_t launch_convert_f(pfg_handle *h, float *f_out_or_null,
                            const float *f_in_or_null);
hipError_t launch_conservation(pfg_handle *h, int k, double out[4]);
/* persistent batched LB launch: row ranges (one per XCD) and x-block rows per
 * grab */
constexpr int LB_QUEUES = 8;
constexpr int LB_QUEUE_GRAB = 16;
/* pfg_download_fields_noalloc: staging in a fixed device array of
 * FIELDS_SCRATCH_FLOATS floats per device (no allocation) */
constexpr int64_t FIELDS_SCRATCH_FLOATS = PFG_FIELDS_SCRATCH_FLOATS;
constexpr int FIELDS_SCRATCH_MAX_DEVICES = 64;
/* the array's device address on `device` (looked up on the first call) */
hipError_t fields_scratch_setup(int device, float **out_or_null);
/* rho, ux, uy, uz of a single lattice to host (any may be NULL), window of
 * whole z-planes at a time through the fixed array */
hipError_t download_fields_noalloc(pfg_handle *h, float *const host[4]);

/* markers.hip: fused marker-side launches, each with the roles it runs */
void fill_mlat(const pfg_handle *h, int k, MLat *L);
hipError_t launch_begin(pfg_handle *h, bool clear, bool count);
hipError_t launch_scan(pfg_handle *h);
hipError_t launch_place(pfg_handle *h, bool scatter, bool geometry);
/* resolved != nullptr: resolve the bonds in the same launch when there are
 * few receptors (*resolved says whether it did) */
hipError_t launch_gather(pfg_handle *h, bool sort, bool sums, bool decide,
                         bool *resolved = nullptr);
hipError_t launch_resolve(pfg_handle *h);
/* spread = true: the force pass also spreads (only when forces_can_spread) */
hipError_t launch_forces(pfg_handle *h, bool spread = false);
/* whether the step's cell lists hold the elements (else only the other
 * markers, and the step needs no bin-sorted order) */
bool bin_elements(const pfg_handle *h);
bool forces_can_spread(const pfg_handle *h);
hipError_t launch_spread(pfg_handle *h);
hipError_t launch_move(pfg_handle *h, bool interp, bool update);

}  // namespace pfg

#endif
