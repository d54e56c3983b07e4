/*
 * api.hip -- the C ABI (include/pfg.h): handle lifetime, uploads/downloads,
 * model loading (host-side CSR construction) and the step driver.
 *
 * The reference's glue layer (semLB.cu) allocates pitched FP64 arrays, copies
 * the whole SEM struct host->device and back every step (semLB.cu:435, 507) and
 * refills random numbers on the host (semLB.cu:348-365).  Here all state lives
 * on the device between calls; a step is one hipGraph launch whose kernels read
 * the step index from device memory, and the host reads back only counters at
 * the end of a pfg_step call.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <set>
#include <vector>

#include "pfg_internal.h"

using namespace pfg;

namespace {

pfg_status from_hip(hipError_t e) {
  if (e == hipSuccess) return PFG_OK;
  if (e == hipErrorOutOfMemory) return PFG_ERR_ALLOC;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return PFG_ERR_DEVICE;
  return PFG_ERR_LAUNCH;
}

#define PFG_TRY(expr)                         \
  do {                                        \
    hipError_t e_ = (expr);                   \
    if (e_ != hipSuccess) return from_hip(e_); \
  } while (0)

template <typename T>
hipError_t dalloc(pfg_handle *h, T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
  if (e == hipSuccess) h->dev_bytes += sizeof(T) * count;
  return e;
}

template <typename T>
void dfree(T *&p) {
  if (p) hipFree((void *)p);
  p = nullptr;
}

template <typename T>
hipError_t upload(pfg_handle *h, T *dst, const T *src, size_t count) {
  if (count == 0) return hipSuccess;
  hipError_t e = hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyHostToDevice,
                                h->stream);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(h->stream);
}

template <typename T>
hipError_t download(pfg_handle *h, T *dst, const T *src, size_t count) {
  if (count == 0 || dst == nullptr) return hipSuccess;
  hipError_t e = hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyDeviceToHost,
                                h->stream);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(h->stream);
}

template <typename T>
hipError_t fill(pfg_handle *h, T *dst, T value, size_t count) {
  std::vector<T> v(count, value);
  return upload(h, dst, v.data(), count);
}

void free_markers(Markers &m) {
  dfree(m.minst);
  dfree(m.x); dfree(m.y); dfree(m.z);
  dfree(m.vx); dfree(m.vy); dfree(m.vz);
  dfree(m.fx); dfree(m.fy); dfree(m.fz); dfree(m.uib);
  dfree(m.kind); dfree(m.body);
  dfree(m.plt_elem_start); dfree(m.plt_tri_start);
  dfree(m.tri); dfree(m.tri_S0); dfree(m.tri_S); dfree(m.tri_V);
  dfree(m.tri_g); dfree(m.tri_a2); dfree(m.pos4);
  dfree(m.plt_S0); dfree(m.plt_V0); dfree(m.plt_S); dfree(m.plt_V);
  dfree(m.plt_lowest);
  dfree(m.m2t_start); dfree(m.m2t);
  dfree(m.tri_edge); dfree(m.edge_tri); dfree(m.ring_start); dfree(m.ring);
  dfree(m.vn); dfree(m.en); dfree(m.km);
  dfree(m.contact_slot); dfree(m.contact_list);
  dfree(m.m2e_start); dfree(m.m2e_nbr); dfree(m.m2e_l0);
  dfree(m.rec_tri); dfree(m.rec_r1); dfree(m.rec_r2);
  dfree(m.rec_bond); dfree(m.rec_cand); dfree(m.rec_release);
  dfree(m.rec_force); dfree(m.rec_w);
  dfree(m.m2r_start); dfree(m.m2r);
  dfree(m.lig_x); dfree(m.lig_y); dfree(m.lig_z); dfree(m.lig_owner); dfree(m.lig_claim);
  dfree(m.lg_start); dfree(m.lg_items); dfree(m.lg_pos); dfree(m.rec_pos);
  dfree(m.m2l_start); dfree(m.m2l_nbr); dfree(m.m2l_l0);
  dfree(m.bin_count); dfree(m.bin_start); dfree(m.bin_items);
  dfree(m.marker_bin); dfree(m.marker_slot); dfree(m.scan_partials);
  dfree(m.chunk_start); dfree(m.chunk_items); dfree(m.chunk_ref); dfree(m.bin_scatter); dfree(m.sorted_pos);
  dfree(m.tile_list); dfree(m.tile_count); dfree(m.spread_rec);
  dfree(m.body_range); dfree(m.rchunk_start); dfree(m.rchunk_items); dfree(m.tile_owner);
  dfree(m.body_spread); dfree(m.body_fallback); dfree(m.chunk_body); dfree(m.chunk_sum);
  dfree(m.fbody_range); dfree(m.frest_items); dfree(m.fb_ref); dfree(m.fb_nlen); dfree(m.fb_nbr);
  dfree(m.fb_done);
  dfree(m.chunk_info);
  Markers fresh;
  fresh.counters = m.counters;
  fresh.ticket = m.ticket;
  fresh.detach_step = m.detach_step;
  fresh.d_step = m.d_step;
  m = fresh;
}

void drop_graph(pfg_handle *h) {
  for (int g = 0; g < 2; ++g)
    for (int c = 0; c < 2; ++c)
      if (h->graphs[g][c]) {
        hipGraphExecDestroy(h->graphs[g][c]);
        h->graphs[g][c] = nullptr;
      }
  h->graph_valid = false;
}

/* Build a CSR (start[n+1], items) from (key, value) pairs; values keep their
 * input order inside a key. */
void build_csr(int n, const std::vector<std::pair<int, int>> &kv,
               std::vector<int32_t> &start, std::vector<int32_t> &items) {
  start.assign((size_t)n + 1, 0);
  for (const auto &p : kv) start[(size_t)p.first + 1]++;
  for (int i = 0; i < n; ++i) start[(size_t)i + 1] += start[(size_t)i];
  items.assign(kv.size(), 0);
  std::vector<int32_t> fillp(start.begin(), start.end() - 1);
  for (const auto &p : kv) items[(size_t)fillp[(size_t)p.first]++] = p.second;
}

/* one float4 per thread: the streaming pattern that reaches the highest HBM
 * rate on MI355X */
__global__ void copy_probe_kernel(const float4 *__restrict__ a, float4 *__restrict__ b,
                                  size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

bool is_coupled_kind(int k) {
  return k == PFG_MARKER_ELEMENT || k == PFG_MARKER_FIBER || k == PFG_MARKER_FIXED;
}

}  // namespace

extern "C" {

const char *pfg_version(void) { return "platelet_fiber_gpu_amd 0.1 (gfx950)"; }

const char *pfg_status_string(pfg_status s) {
  switch (s) {
    case PFG_OK: return "ok";
    case PFG_ERR_ARG: return "invalid argument";
    case PFG_ERR_ALLOC: return "allocation failed";
    case PFG_ERR_DEVICE: return "no usable HIP device";
    case PFG_ERR_LAUNCH: return "kernel launch or HIP runtime error";
    case PFG_ERR_STATE: return "call out of order";
    case PFG_ERR_IO: return "file I/O error";
  }
  return "unknown status";
}

pfg_status pfg_device_count(int *count) {
  if (!count) return PFG_ERR_ARG;
  *count = 0;
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    return PFG_ERR_DEVICE;
  }
  return PFG_OK;
}

void pfg_params_default(pfg_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->nx = 64;
  p->ny = 32;
  p->nz = 32;
  p->store = PFG_STORE_F32;
  p->f16_scale_log2 = 10;
  p->tau = 0.8f;
  p->ib_force_scale = 1.0f;
  p->marker_dt = 1.0f;
  p->morse_u0 = 0.0f;
  p->morse_rho = 2.0f;
  p->morse_req = 0.8f;
  p->morse_cutoff = 2.5f;
  p->morse_taper = 0.0f;
  p->wall_rep_kappa = 20.0f;
  p->wall_rep_hc = 0.5f;
  p->wall_z = -0.5f;
  p->bond_reach = 1.0f;
  p->bell_xb_kT = 1.0f;
  p->seed = 1101;
  p->fiber_law = PFG_FIBER_WLC;
  p->fiber_move = PFG_FIBER_IB;
  p->fiber_xmax = 0.9f;
  p->fiber_gamma = 1.0f;
  p->bin_cells = 3;
  p->sem_integrator = PFG_SEM_ADVECT;
  p->sem_mass = 1.0f;
  p->sem_gamma = 1.0f;
  p->sem_fluid_weight = 0.5f;
  p->bond_law = PFG_BOND_BELL;
  p->bell_kT = 1.0f;
}

/* hostSEMparameters index map (sem_kernel.cu:58-76, filled at
 * gpuSEM_LB.c:1206-1233): [0] receptors [1] seed [2] k_spring [3] Kf0 [4] Kr0
 * [5] shear rate [6] rho0 [7] L0 [8] T0 [9] M0 [10] F0 [11] UMAX [12] nu
 * [13] tau [14] P0 [15] C [16] UNIT_FACTOR [17] Ks [18] Kv [19] Kbend.
 *
 * The device works in lattice units (length dx, time dt) with forces in
 * units of F0, as the reference's SEM kernels divide every force by F0
 * (sem_kernel.cu:383-389, 428, 1632-1637, 1720-1724).  A stiffness per
 * micrometre becomes one per lattice length (x dx), an area gradient
 * carries one more dx, a volume gradient two; the bending Laplacian of the
 * mean curvature three fewer.  The constants the reference keeps in its
 * kernel source (the vWF bond: BONDTHRESHOLD 0.078 um, k_receptvWF
 * 10 nN/um, the Bell length 0.00071 and KbT 4.05 of sem_kernel.cu:79-85,
 * 1852-1863; the wall repulsion 5e-4 * 2000 e^{-2000 h}/(1 - e^{-2000 h})
 * nN below 0.02 um, sem_kernel.cu:424-427) are mapped the same way. */
constexpr double REF_BOND_THRESHOLD_UM = 0.078;
constexpr double REF_K_RECEPT_VWF = 10.0;
constexpr double REF_BELL_XB = 0.00071;
constexpr double REF_KBT = 4.05;
constexpr double REF_WALL_A = 5e-4, REF_WALL_KAPPA = 2000.0, REF_WALL_HC_UM = 0.02;

pfg_status pfg_params_from_sem(const float *s, float dx, float dt, int lx, int ly, int lz,
                               pfg_params *p) {
  if (!s || !p || dx <= 0.f || dt <= 0.f || lx <= 0 || ly <= 0 || lz <= 0)
    return PFG_ERR_ARG;
  for (int i = 0; i < 20; ++i) p->sem[i] = s[i];
  p->nx = lx;
  p->ny = ly;
  p->nz = lz;
  p->seed = (uint64_t)s[1];
  p->tau = s[13];
  const double F0 = s[10] > 0.f ? s[10] : 1.0, rho0 = s[6], umax = s[11], C = s[15],
               uf = s[16], L0 = s[7] > 0.f ? s[7] : 1.0, T0 = s[8], h = dx;
  /* marker force (F0) -> lattice force density: F0 UNIT_FACTOR dt /
   * (rho0 UMAX) in units of C, as the Guo half force of the moments
   * (LB_kernel.cu:1158-1163) */
  if (rho0 > 0.0 && umax > 0.0 && C > 0.0) p->ib_force_scale = (float)(F0 * uf * dt / (rho0 * umax) / C);
  /* x += V dt / T0 with V in UMAX (sem_kernel.cu:2111-2137): a lattice
   * velocity u = V / C moves C dt L0 / (T0 dx) lattice lengths per step
   * (1 for the reference's own block) */
  p->marker_dt = (T0 > 0.0 && C > 0.0) ? (float)(C * dt * L0 / (T0 * h)) : 1.0f;
  p->k_spring = (float)(s[2] * h / F0);
  p->ks = (float)(s[17] * h / F0);
  p->kv = (float)(s[18] * h * h / F0);
  p->kbend = (float)(s[19] / F0 / (h * h * h));
  /* rates per second -> per step; Kf0 is set from the shear rate by the
   * reference's host before the block is handed over (gpuSEM_LB.c:1209) */
  p->k_on = (float)(s[3] * dt);
  p->k_off0 = (float)(s[4] * dt);
  p->detach_rate = (float)(s[5] * dt); /* gama = SHEAR_RATE dt, gpuSEM_LB.c:1433 */
  p->bond_law = PFG_BOND_REFERENCE;
  p->bond_k = (float)(REF_K_RECEPT_VWF * h / F0);
  p->bond_l0 = (float)(REF_BOND_THRESHOLD_UM / h);
  p->bell_xb_kT = (float)(REF_BELL_XB / REF_KBT * F0);
  p->bell_kT = (float)(REF_KBT / F0 / h);
  p->wall_rep_kappa = (float)(REF_WALL_KAPPA * h);
  p->wall_rep_a = (float)(REF_WALL_A / F0 / h);
  p->wall_rep_hc = (float)(REF_WALL_HC_UM / h);
  return PFG_OK;
}

}  /* extern "C" */

namespace {

pfg_status check_lattice_params(const pfg_params &p) {
  if (p.nx <= 0 || p.ny <= 0 || p.nz <= 0) return PFG_ERR_ARG;
  if ((int64_t)p.nx * p.ny * p.nz > (int64_t)INT_MAX) return PFG_ERR_ARG;
  if (!(p.tau > 0.5f)) return PFG_ERR_ARG;
  for (int k = 0; k < 3; ++k)
    if (p.wall[k] != PFG_WALL_PERIODIC && p.wall[k] != PFG_WALL_BOUNCE &&
        !(k == 2 && p.wall[k] == PFG_WALL_VELOCITY))
      return PFG_ERR_ARG;
  return PFG_OK;
}

/* Batched LB launch layout for h->lb_cpt z-cells per lane: each lattice's
 * workgroups (four wavefronts, one x-block row of lb_cpt z-cells each) and
 * the workgroup -> lattice table. */
hipError_t batch_layout(pfg_handle *h) {
  int64_t blk = 0;
  for (Inst &I : h->inst) {
    const int64_t zrows = (I.p.nz + h->lb_cpt - 1) / h->lb_cpt;
    I.lb_blocks = ((int64_t)I.nxb * I.p.ny * zrows + 3) / 4;
    I.lb_block0 = blk;
    blk += I.lb_blocks;
  }
  h->lb_blocks = blk;
  std::vector<int32_t> table((size_t)blk);
  for (int k = 0; k < h->K; ++k) {
    const Inst &I = h->inst[(size_t)k];
    for (int64_t b = 0; b < I.lb_blocks; ++b) table[(size_t)(I.lb_block0 + b)] = k;
  }
  dfree(h->lb_block_inst);
  hipError_t e = dalloc(h, &h->lb_block_inst, table.size());
  if (e == hipSuccess) e = upload(h, h->lb_block_inst, table.data(), table.size());
  return e;
}

/* Batched LB arguments (both starting copies) and the workgroup table. */
hipError_t upload_batch_args(pfg_handle *h) {
  if (h->K <= 1) return hipSuccess;
  std::vector<LBArgs> args((size_t)h->K);
  for (int cur = 0; cur < 2; ++cur) {
    for (int k = 0; k < h->K; ++k) make_lb_args(h, k, cur, &args[(size_t)k]);
    hipError_t e = upload(h, h->d_lbargs[cur], args.data(), args.size());
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

/* One or K lattices; shared fields (storage, markers, bonds, fibrin) come
 * from ps[0], lattice and fluid fields from each ps[k]. */
pfg_status create_impl(const pfg_params *ps, int K, int device, pfg_handle **out) {
  if (!ps || !out || K < 1) return PFG_ERR_ARG;
  *out = nullptr;
  const pfg_params &p = ps[0];
  if (p.store < PFG_STORE_F32 || p.store > PFG_STORE_BF16) return PFG_ERR_ARG;
  if (p.bin_cells <= 0) return PFG_ERR_ARG;
  if (p.f16_scale_log2 < 0 || p.f16_scale_log2 > 20) return PFG_ERR_ARG;
  /* the contact search covers the neighbouring bins only */
  if (p.contact_a != 0.f &&
      !(p.contact_hc > 0.f && p.contact_hc <= (float)p.bin_cells && p.contact_kappa > 0.f))
    return PFG_ERR_ARG;
  if (p.kick_temp < 0.f) return PFG_ERR_ARG;
  if (p.sem_integrator != PFG_SEM_ADVECT && p.sem_integrator != PFG_SEM_BBK) return PFG_ERR_ARG;
  if (p.sem_integrator == PFG_SEM_BBK &&
      !(p.sem_mass > 0.f && p.sem_gamma >= 0.f && p.sem_temp >= 0.f &&
        p.sem_fluid_weight >= 0.f && p.sem_fluid_weight <= 1.f))
    return PFG_ERR_ARG;
  if (p.bond_law != PFG_BOND_BELL && p.bond_law != PFG_BOND_REFERENCE) return PFG_ERR_ARG;
  if (p.bond_law == PFG_BOND_REFERENCE && !(p.bell_kT > 0.f)) return PFG_ERR_ARG;
  int64_t ncells = 0;
  for (int k = 0; k < K; ++k) {
    const pfg_status st = check_lattice_params(ps[k]);
    if (st != PFG_OK) return st;
    ncells += (int64_t)ps[k].nx * ps[k].ny * ps[k].nz;
  }
  if (ncells > (int64_t)INT_MAX) return PFG_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PFG_ERR_DEVICE;
  if (device < 0 || device >= ndev) return PFG_ERR_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return PFG_ERR_DEVICE;
  /* device address of the fixed staging array, once per device */
  if (fields_scratch_setup(device, nullptr) != hipSuccess) return PFG_ERR_DEVICE;

  pfg_handle *h = new (std::nothrow) pfg_handle();
  if (!h) return PFG_ERR_ALLOC;
  h->p = p;
  h->device = device;
  h->K = K;
  h->esize = p.store == PFG_STORE_F32 ? 4 : 2;
  bool flags = false;
  /* lattice k occupies cells [cell0, cell0 + N), x-blocks [fblk0, ...),
   * tiles [tile0, ...) and LB workgroups [lb_block0, ...) */
  h->inst.resize((size_t)K);
  int64_t cell = 0, fblk = 0, tile = 0;
  for (int k = 0; k < K; ++k) {
    Inst &I = h->inst[(size_t)k];
    I.p = ps[k];
    I.p.store = p.store;
    I.p.f16_scale_log2 = p.f16_scale_log2;
    I.N = (int64_t)I.p.nx * I.p.ny * I.p.nz;
    I.cell0 = cell;
    I.nxb = (I.p.nx + XBLOCK - 1) / XBLOCK;
    I.fblk0 = fblk;
    I.tnx = (I.p.nx + TILE_X - 1) / TILE_X;
    I.tny = (I.p.ny + TILE_Y - 1) / TILE_Y;
    I.tnz = (I.p.nz + TILE_Z - 1) / TILE_Z;
    I.ntiles = (int64_t)I.tnx * I.tny * I.tnz;
    I.tile0 = tile;
    cell += I.N;
    fblk += (int64_t)I.nxb * I.p.ny * I.p.nz;
    tile += I.ntiles;
    flags = flags || I.p.use_flags;
  }
  h->p.use_flags = flags ? 1 : 0;
  pfg_status st = PFG_OK;
  do {
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming) != hipSuccess) {
      st = PFG_ERR_DEVICE;
      break;
    }
    h->N = cell;
    h->plane = (h->N + 255) / 256 * 256;
    h->nxb = h->inst[0].nxb;
    const size_t fbytes = (size_t)h->esize * Q * XBLOCK * (size_t)fblk;
    for (int c = 0; c < 2; ++c) {
      if (hipMalloc(&h->f[c], fbytes) != hipSuccess) {
        st = PFG_ERR_ALLOC;
        break;
      }
      h->dev_bytes += fbytes;
    }
    if (st != PFG_OK) break;
    if (dalloc(h, &h->aux, (size_t)3 * h->plane) != hipSuccess) { st = PFG_ERR_ALLOC; break; }
    if (dalloc(h, &h->vel, (size_t)h->plane) != hipSuccess) { st = PFG_ERR_ALLOC; break; }
    /* row stamps start at -1: no row is touched before the first step */
    if (dalloc(h, &h->row_stamp, (size_t)fblk) != hipSuccess) { st = PFG_ERR_ALLOC; break; }
    if (hipMemset(h->row_stamp, 0xff, sizeof(int32_t) * (size_t)fblk) != hipSuccess) {
      st = PFG_ERR_DEVICE;
      break;
    }
    if (hipMemset(h->aux, 0, sizeof(float) * 3 * h->plane) != hipSuccess ||
        hipMemset(h->vel, 0, sizeof(float4) * h->plane) != hipSuccess) {
      st = PFG_ERR_DEVICE;
      break;
    }
    h->tnx = h->inst[0].tnx;
    h->tny = h->inst[0].tny;
    h->tnz = h->inst[0].tnz;
    h->ntiles = tile;
    /* tile stamps start at -1: no tile is touched before the first step */
    if (dalloc(h, &h->tile_stamp, (size_t)h->ntiles) != hipSuccess) { st = PFG_ERR_ALLOC; break; }
    if (hipMemset(h->tile_stamp, 0xff, sizeof(int32_t) * (size_t)h->ntiles) != hipSuccess) {
      st = PFG_ERR_DEVICE;
      break;
    }
    int64_t nmax = 0;
    for (const Inst &I : h->inst) nmax = std::max(nmax, I.N);
    h->red_blocks = std::min<int64_t>(2048, (nmax + 255) / 256);
    if (dalloc(h, &h->red, (size_t)4 * h->red_blocks) != hipSuccess) { st = PFG_ERR_ALLOC; break; }
    /* z-cells per lane of the LB kernel */
    h->lb_cpt = 1;
    h->lb_xcd_rows = (K == 1 && h->N >= XCD_ROWS_MIN_CELLS) ? 1 : 0;
    if (dalloc(h, &h->m.counters, N_COUNTERS) != hipSuccess) { st = PFG_ERR_ALLOC; break; }
    if (hipMemset(h->m.counters, 0, sizeof(int64_t) * N_COUNTERS) != hipSuccess) { st = PFG_ERR_DEVICE; break; }
    if (dalloc(h, &h->m.ticket, 1) != hipSuccess) { st = PFG_ERR_ALLOC; break; }
    if (hipMemset(h->m.ticket, 0, sizeof(int32_t)) != hipSuccess) { st = PFG_ERR_DEVICE; break; }
    if (dalloc(h, &h->m.detach_step, 1) != hipSuccess) { st = PFG_ERR_ALLOC; break; }
    if (dalloc(h, &h->m.d_step, 1) != hipSuccess) { st = PFG_ERR_ALLOC; break; }
    const int64_t minus1 = -1, zero = 0;
    if (hipMemcpy(h->m.detach_step, &minus1, sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->m.d_step, &zero, sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess) {
      st = PFG_ERR_DEVICE;
      break;
    }
    if (flags) {
      if (dalloc(h, &h->solid, (size_t)h->N) != hipSuccess) { st = PFG_ERR_ALLOC; break; }
      if (hipMemset(h->solid, 0, (size_t)h->N) != hipSuccess) { st = PFG_ERR_DEVICE; break; }
      if (dalloc(h, &h->bbmask, (size_t)h->N) != hipSuccess) { st = PFG_ERR_ALLOC; break; }
      if (hipMemset(h->bbmask, 0, sizeof(uint32_t) * h->N) != hipSuccess) { st = PFG_ERR_DEVICE; break; }
    }
    if (K > 1) {
      std::vector<int64_t> tile0((size_t)K + 1);
      for (int k = 0; k < K; ++k) tile0[(size_t)k] = h->inst[(size_t)k].tile0;
      tile0[(size_t)K] = tile;
      hipError_t e = batch_layout(h);
      if (e == hipSuccess) e = dalloc(h, &h->d_tile0, tile0.size());
      if (e == hipSuccess) e = upload(h, h->d_tile0, tile0.data(), tile0.size());
      for (int c = 0; c < 2 && e == hipSuccess; ++c) e = dalloc(h, &h->d_lbargs[c], (size_t)K);
      if (e == hipSuccess) e = dalloc(h, &h->lb_queue, LB_QUEUES + 1);
      if (e == hipSuccess) e = hipMemset(h->lb_queue, 0, (LB_QUEUES + 1) * sizeof(uint32_t));
      if (e == hipSuccess) e = dalloc(h, &h->d_lats, (size_t)K);
      if (e == hipSuccess) e = upload_batch_args(h);
      if (e != hipSuccess) {
        st = from_hip(e);
        break;
      }
    }
    /* f = equilibrium at rest until the caller says otherwise */
    for (int k = 0; k < K && st == PFG_OK; ++k)
      if (launch_init_eq(h, k, nullptr, nullptr, nullptr, nullptr) != hipSuccess) st = PFG_ERR_LAUNCH;
    if (st == PFG_OK && hipStreamSynchronize(h->stream) != hipSuccess) st = PFG_ERR_LAUNCH;
  } while (0);
  if (st != PFG_OK) {
    pfg_destroy(h);
    return st;
  }
  *out = h;
  return PFG_OK;
}

pfg_status init_eq_impl(pfg_handle *h, int k, const float *rho, const float *ux,
                        const float *uy, const float *uz) {
  hipSetDevice(h->device);
  const int64_t n = h->inst[(size_t)k].N;
  const float *src[4] = {rho, ux, uy, uz};
  float *dev[4] = {nullptr, nullptr, nullptr, nullptr};
  pfg_status st = PFG_OK;
  for (int c = 0; c < 4 && st == PFG_OK; ++c) {
    if (!src[c]) continue;
    if (hipMalloc(&dev[c], sizeof(float) * n) != hipSuccess) {
      st = PFG_ERR_ALLOC;
      break;
    }
    st = from_hip(upload(h, dev[c], src[c], (size_t)n));
  }
  if (st == PFG_OK) {
    st = from_hip(launch_init_eq(h, k, dev[0], dev[1], dev[2], dev[3]));
    if (st == PFG_OK) st = from_hip(hipStreamSynchronize(h->stream));
  }
  for (int c = 0; c < 4; ++c)
    if (dev[c]) hipFree(dev[c]);
  return st;
}

/* f = f_eq(1, u(z)) on lattice k from one velocity value per z-plane. */
pfg_status init_eq_z_impl(pfg_handle *h, int k, const float *ux, const float *uy,
                          const float *uz) {
  hipSetDevice(h->device);
  const int nz = h->inst[(size_t)k].p.nz;
  const float *src[3] = {ux, uy, uz};
  float *dev = nullptr;
  pfg_status st = from_hip(hipMalloc(&dev, sizeof(float) * 3 * (size_t)nz));
  for (int c = 0; c < 3 && st == PFG_OK; ++c)
    if (src[c]) st = from_hip(upload(h, dev + (size_t)c * nz, src[c], (size_t)nz));
  if (st == PFG_OK) {
    st = from_hip(launch_init_eq(h, k, nullptr, ux ? dev : nullptr, uy ? dev + nz : nullptr,
                                 uz ? dev + 2 * (size_t)nz : nullptr, true));
    if (st == PFG_OK) st = from_hip(hipStreamSynchronize(h->stream));
  }
  if (dev) hipFree(dev);
  return st;
}

pfg_status upload_flags_impl(pfg_handle *h, int k, const uint8_t *flags) {
  if (!h->solid || !h->inst[(size_t)k].p.use_flags) return PFG_ERR_STATE;
  hipSetDevice(h->device);
  const Inst &I = h->inst[(size_t)k];
  PFG_TRY(upload(h, h->solid + I.cell0, flags, (size_t)I.N));
  drop_graph(h);
  PFG_TRY(launch_bbmask(h, k));
  PFG_TRY(hipStreamSynchronize(h->stream));
  return PFG_OK;
}

pfg_status fields_impl(pfg_handle *h, int k, float *rho, float *ux, float *uy, float *uz) {
  hipSetDevice(h->device);
  const int64_t n = h->inst[(size_t)k].N;
  float *dev[4] = {nullptr, nullptr, nullptr, nullptr};
  float *host[4] = {rho, ux, uy, uz};
  pfg_status st = PFG_OK;
  for (int c = 0; c < 4; ++c) {
    if (!host[c]) continue;
    if (hipMalloc(&dev[c], sizeof(float) * n) != hipSuccess) {
      st = PFG_ERR_ALLOC;
      break;
    }
    /* solid cells are not written by the fields pass: they read 0 */
    st = from_hip(hipMemsetAsync(dev[c], 0, sizeof(float) * n, h->stream));
    if (st != PFG_OK) break;
  }
  if (st == PFG_OK) st = from_hip(launch_lb_fields(h, k, dev[0], dev[1], dev[2], dev[3]));
  for (int c = 0; c < 4 && st == PFG_OK; ++c)
    if (host[c]) st = from_hip(download(h, host[c], dev[c], (size_t)n));
  for (int c = 0; c < 4; ++c)
    if (dev[c]) hipFree(dev[c]);
  return st;
}

}  // namespace

extern "C" {

pfg_status pfg_create(const pfg_params *pin, int device, pfg_handle **out) {
  return create_impl(pin, 1, device, out);
}

pfg_status pfg_destroy(pfg_handle *h) {
  if (!h) return PFG_ERR_ARG;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  drop_graph(h);
  free_markers(h->m);
  dfree(h->m.counters);
  dfree(h->m.ticket);
  dfree(h->m.detach_step);
  dfree(h->m.d_step);
  for (int k = 0; k < 2; ++k) {
    if (h->f[k]) hipFree(h->f[k]);
    h->f[k] = nullptr;
    dfree(h->d_lbargs[k]);
  }
  dfree(h->aux);
  dfree(h->vel);
  dfree(h->tile_stamp);
  dfree(h->solid);
  dfree(h->bbmask);
  dfree(h->red);
  dfree(h->lb_block_inst);
  dfree(h->lb_queue);
  dfree(h->d_lats);
  dfree(h->d_tile0);
  dfree(h->row_stamp);
  if (h->ev_fork) hipEventDestroy(h->ev_fork);
  if (h->ev_join) hipEventDestroy(h->ev_join);
  if (h->stream2) hipStreamDestroy(h->stream2);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return PFG_OK;
}

pfg_status pfg_get_params(const pfg_handle *h, pfg_params *out) {
  if (!h || !out) return PFG_ERR_ARG;
  *out = h->p;
  return PFG_OK;
}

pfg_status pfg_device_bytes(const pfg_handle *h, size_t *bytes) {
  if (!h || !bytes) return PFG_ERR_ARG;
  *bytes = h->dev_bytes;
  return PFG_OK;
}

/* ---------------- lattice state (single-lattice handles) ---------------- */

pfg_status pfg_init_equilibrium(pfg_handle *h, const float *rho, const float *ux,
                                const float *uy, const float *uz) {
  if (!h) return PFG_ERR_ARG;
  if (h->K != 1) return PFG_ERR_STATE;
  return init_eq_impl(h, 0, rho, ux, uy, uz);
}

pfg_status pfg_init_equilibrium_z(pfg_handle *h, const float *ux_z, const float *uy_z,
                                  const float *uz_z) {
  if (!h) return PFG_ERR_ARG;
  if (h->K != 1) return PFG_ERR_STATE;
  return init_eq_z_impl(h, 0, ux_z, uy_z, uz_z);
}

pfg_status pfg_upload_f(pfg_handle *h, const float *f) {
  if (!h || !f) return PFG_ERR_ARG;
  if (h->K != 1) return PFG_ERR_STATE;
  hipSetDevice(h->device);
  return from_hip(launch_convert_f(h, nullptr, f));
}

pfg_status pfg_download_f(pfg_handle *h, float *f) {
  if (!h || !f) return PFG_ERR_ARG;
  if (h->K != 1) return PFG_ERR_STATE;
  hipSetDevice(h->device);
  return from_hip(launch_convert_f(h, f, nullptr));
}

pfg_status pfg_upload_flags(pfg_handle *h, const uint8_t *flags) {
  if (!h || !flags) return PFG_ERR_ARG;
  if (h->K != 1) return PFG_ERR_STATE;
  return upload_flags_impl(h, 0, flags);
}

pfg_status pfg_download_fields(pfg_handle *h, float *rho, float *ux, float *uy,
                               float *uz) {
  if (!h) return PFG_ERR_ARG;
  if (h->K != 1) return PFG_ERR_STATE;
  return fields_impl(h, 0, rho, ux, uy, uz);
}

pfg_status pfg_download_fields_noalloc(pfg_handle *h, float *rho, float *ux, float *uy,
                                       float *uz) {
  if (!h) return PFG_ERR_ARG;
  if (h->K != 1) return PFG_ERR_STATE;
  float *const host[4] = {rho, ux, uy, uz};
  int64_t nf = 0;
  for (int c = 0; c < 4; ++c) nf += host[c] != nullptr;
  if ((int64_t)h->p.nx * h->p.ny * nf > FIELDS_SCRATCH_FLOATS) {
    fprintf(stderr,
            "pfg_download_fields_noalloc: a z-plane of %d x %d cells x %d fields exceeds the "
            "%lld-float staging array\n",
            h->p.nx, h->p.ny, (int)nf, (long long)FIELDS_SCRATCH_FLOATS);
    return PFG_ERR_ARG;
  }
  hipSetDevice(h->device);
  return from_hip(download_fields_noalloc(h, host));
}

pfg_status pfg_conservation(pfg_handle *h, double out[4]) {
  if (!h || !out) return PFG_ERR_ARG;
  if (h->K != 1) return PFG_ERR_STATE;
  hipSetDevice(h->device);
  return from_hip(launch_conservation(h, 0, out));
}

/* ---------------- immersed structures ---------------- */

}  /* extern "C" */

namespace {

/* Load the markers of all lattices.  minst[i] is the lattice of marker i and
 * linst[l] that of wall ligand l (both NULL: a single lattice). */
pfg_status load_impl(pfg_handle *h, const pfg_model &M, const int32_t *minst,
                     const int32_t *linst) {
  hipSetDevice(h->device);
  const int n = M.n_markers;
  /* ---- validation ---- */
  if (n < 0 || M.n_platelets < 0 || M.n_tri < 0 || M.n_edges < 0 ||
      M.n_receptors < 0 || M.n_wall_ligands < 0 || M.n_links < 0)
    return PFG_ERR_ARG;
  if (n > 0 && (!M.x || !M.y || !M.z || !M.kind || !M.body)) return PFG_ERR_ARG;
  for (int i = 0; i < n; ++i) {
    if (!is_coupled_kind(M.kind[i])) return PFG_ERR_ARG;
    if (M.kind[i] == PFG_MARKER_ELEMENT && (M.body[i] < 0 || M.body[i] >= M.n_platelets))
      return PFG_ERR_ARG;
    if (!std::isfinite(M.x[i]) || !std::isfinite(M.y[i]) || !std::isfinite(M.z[i]))
      return PFG_ERR_ARG;
  }
  if (M.n_platelets > 0) {
    if (!M.plt_elem_start || !M.plt_tri_start || !M.plt_S0 || !M.plt_V0) return PFG_ERR_ARG;
    if (M.plt_elem_start[0] != 0 || M.plt_tri_start[0] != 0) return PFG_ERR_ARG;
    for (int q = 0; q < M.n_platelets; ++q) {
      if (M.plt_elem_start[q + 1] < M.plt_elem_start[q]) return PFG_ERR_ARG;
      if (M.plt_tri_start[q + 1] < M.plt_tri_start[q]) return PFG_ERR_ARG;
      for (int e = M.plt_elem_start[q]; e < M.plt_elem_start[q + 1]; ++e)
        if (e >= n || M.kind[e] != PFG_MARKER_ELEMENT || M.body[e] != q) return PFG_ERR_ARG;
    }
    if (M.plt_tri_start[M.n_platelets] != M.n_tri) return PFG_ERR_ARG;
  } else if (M.n_tri > 0) {
    return PFG_ERR_ARG;
  }
  if (M.n_tri > 0 && (!M.tri || !M.tri_S0)) return PFG_ERR_ARG;
  for (int q = 0; q < M.n_platelets; ++q)
    for (int t = M.plt_tri_start[q]; t < M.plt_tri_start[q + 1]; ++t)
      for (int k = 0; k < 3; ++k) {
        const int v = M.tri[3 * t + k];
        if (v < M.plt_elem_start[q] || v >= M.plt_elem_start[q + 1]) return PFG_ERR_ARG;
      }
  if (M.n_edges > 0 && (!M.edge || !M.edge_l0)) return PFG_ERR_ARG;
  for (int e = 0; e < M.n_edges; ++e)
    for (int k = 0; k < 2; ++k) {
      const int v = M.edge[2 * e + k];
      if (v < 0 || v >= n || M.kind[v] != PFG_MARKER_ELEMENT) return PFG_ERR_ARG;
    }
  if (M.n_receptors > 0 && (!M.rec_tri || !M.rec_r1 || !M.rec_r2)) return PFG_ERR_ARG;
  for (int r = 0; r < M.n_receptors; ++r)
    if (M.rec_tri[r] < 0 || M.rec_tri[r] >= M.n_tri) return PFG_ERR_ARG;
  if (M.n_wall_ligands > 0 && (!M.lig_x || !M.lig_y)) return PFG_ERR_ARG;
  if (M.n_links > 0 && (!M.link || !M.link_l0)) return PFG_ERR_ARG;
  for (int l = 0; l < M.n_links; ++l)
    for (int k = 0; k < 2; ++k) {
      const int v = M.link[2 * l + k];
      if (v < 0 || v >= n || M.kind[v] == PFG_MARKER_ELEMENT) return PFG_ERR_ARG;
    }

  drop_graph(h);
  free_markers(h->m);
  h->model_loaded = false;
  Markers &m = h->m;
  const pfg_params &p = h->p;
  m.n = n;
  m.n_plt = M.n_platelets;
  m.n_tri = M.n_tri;
  m.n_edges = M.n_edges;
  m.n_rec = M.n_receptors;
  m.n_wall_lig = M.n_wall_ligands;
  m.fibrin_lig = M.fibrin_ligands ? 1 : 0;
  m.n_lig = m.n_wall_lig + (m.fibrin_lig ? n : 0);
  m.n_links = M.n_links;

  /* ---- IB chunks.  Markers are grouped by (lattice, body) over runs of
   * consecutive indices; each group is cut by recursive bisection of its
   * starting positions (along the longest extent, at a multiple of the
   * chunk length) into compact chunks, so that a chunk covers a patch of a
   * membrane (or a cluster of fibrin nodes) and its stencils a small box.
   * Chunks whose box already exceeds IB_BOX_CELLS nodes at load (markers
   * far apart, as sparse fibrin nodes are) are listed for direct spreading
   * instead.  The chunk length (16..IB_CHUNK, a power of two) aims at ~2048
   * chunks so that a single platelet still spreads over many workgroups ---- */
  int chunk_len = 16;
  while (chunk_len < IB_CHUNK && (int64_t)chunk_len * 2 * 2048 <= n) chunk_len *= 2;
  std::vector<int32_t> chunk_start, chunk_items, direct_items;
  /* body spreading: the groups that are compact platelets (elements only, at
   * most BODY_MAX_MARKERS, stencils spanning at most BODY_BOX_DIM nodes
   * along each axis at load) are bodies; their leaves are cut further until
   * every chunk's box fits IB_BOX_CELLS, so that no marker of a body is
   * listed for direct spreading; the chunks and direct markers of every
   * other group are listed again for the chunk launch */
  std::vector<int32_t> body_range, chunk_body, rchunk_start, rchunk_items, rdirect_items;
  {
    const float *P[3] = {M.x, M.y, M.z};
    std::vector<int32_t> idx((size_t)n);
    for (int i = 0; i < n; ++i) idx[(size_t)i] = i;
    bool body = false; /* the group being cut is a body */
    auto box_volume = [&](const int32_t *b, int cnt) {
      int64_t vol = 1;
      for (int k = 0; k < 3; ++k) {
        float lo = P[k][b[0]], hi = lo;
        for (int j = 1; j < cnt; ++j) {
          lo = std::min(lo, P[k][b[j]]);
          hi = std::max(hi, P[k][b[j]]);
        }
        vol *= (int64_t)(std::floor(hi) - std::floor(lo)) + 4;
      }
      return vol;
    };
    auto leaf = [&](int32_t *b, int cnt) {
      std::sort(b, b + cnt);
      if (box_volume(b, cnt) <= IB_BOX_CELLS) {
        chunk_start.push_back((int32_t)chunk_items.size());
        chunk_items.insert(chunk_items.end(), b, b + cnt);
        chunk_body.push_back(body ? (int32_t)(body_range.size() / 4) : -1);
      } else {
        direct_items.insert(direct_items.end(), b, b + cnt);
        if (!body) rdirect_items.insert(rdirect_items.end(), b, b + cnt);
      }
    };
    std::vector<std::pair<int32_t *, int>> todo;
    for (int g0 = 0; g0 < n;) {
      int g1 = g0 + 1;
      while (g1 < n && M.body[g1] == M.body[g0] && !(minst && minst[g1] != minst[g0])) ++g1;
      body = g1 - g0 <= BODY_MAX_MARKERS;
      for (int k = 0; k < 3 && body; ++k) {
        float lo = P[k][g0], hi = lo;
        for (int i = g0; i < g1; ++i) {
          body = body && M.kind[i] == PFG_MARKER_ELEMENT;
          lo = std::min(lo, P[k][i]);
          hi = std::max(hi, P[k][i]);
        }
        body = body && (int64_t)(std::floor(hi) - std::floor(lo)) + 4 <= BODY_BOX_DIM;
      }
      const size_t c0 = chunk_start.size();
      todo.push_back({idx.data() + g0, g1 - g0});
      while (!todo.empty()) {
        int32_t *b = todo.back().first;
        const int cnt = todo.back().second;
        todo.pop_back();
        /* a body's leaf is cut in halves until its box fits */
        const bool split = cnt > chunk_len ||
                           (body && cnt > 1 && box_volume(b, cnt) > IB_BOX_CELLS);
        if (!split) {
          leaf(b, cnt);
          continue;
        }
        int axis = 0;
        float ext = -1.f;
        for (int k = 0; k < 3; ++k) {
          float lo = P[k][b[0]], hi = lo;
          for (int j = 1; j < cnt; ++j) {
            lo = std::min(lo, P[k][b[j]]);
            hi = std::max(hi, P[k][b[j]]);
          }
          if (hi - lo > ext) {
            ext = hi - lo;
            axis = k;
          }
        }
        const int left = cnt > chunk_len ? ((cnt + chunk_len - 1) / chunk_len / 2) * chunk_len
                                         : cnt / 2;
        const float *c = P[axis];
        std::nth_element(b, b + left, b + cnt, [c](int32_t u, int32_t v) {
          return c[u] < c[v] || (c[u] == c[v] && u < v);
        });
        todo.push_back({b + left, cnt - left}); /* the left half is cut first */
        todo.push_back({b, left});
      }
      const size_t c1 = chunk_start.size();
      if (body && c1 - c0 > (size_t)BODY_MAX_CHUNKS) { /* too many chunks: not a body */
        body = false;
        for (size_t c = c0; c < c1; ++c) chunk_body[c] = -1;
      }
      if (body) {
        body_range.push_back(g0);
        body_range.push_back(g1);
        body_range.push_back((int32_t)c0);
        body_range.push_back((int32_t)c1);
      } else {
        for (size_t c = c0; c < c1; ++c) {
          const int32_t end_c = c + 1 < c1 ? chunk_start[c + 1] : (int32_t)chunk_items.size();
          rchunk_start.push_back((int32_t)rchunk_items.size());
          rchunk_items.insert(rchunk_items.end(), chunk_items.begin() + chunk_start[c],
                              chunk_items.begin() + end_c);
        }
      }
      g0 = g1;
    }
  }
  m.n_chunks = (int)chunk_start.size();
  m.n_direct = (int)direct_items.size();
  m.n_chunk_items = (int)chunk_items.size();
  chunk_start.push_back((int32_t)chunk_items.size());
  chunk_items.insert(chunk_items.end(), direct_items.begin(), direct_items.end());
  m.n_bodies = (int)(body_range.size() / 4);
  std::vector<int32_t> chunk_ref((size_t)std::max(1, m.n_chunks));
  for (int c = 0; c < m.n_chunks; ++c)
    chunk_ref[(size_t)c] = chunk_body[(size_t)c] >= 0 ? body_range[4 * (size_t)chunk_body[(size_t)c]]
                                                      : chunk_items[(size_t)chunk_start[(size_t)c]];
  m.n_rchunks = (int)rchunk_start.size();
  m.n_rdirect = (int)rdirect_items.size();
  m.n_ritems = (int)(rchunk_items.size() + rdirect_items.size());
  rchunk_start.push_back((int32_t)rchunk_items.size());
  rchunk_items.insert(rchunk_items.end(), rdirect_items.begin(), rdirect_items.end());

  /* ---- force pass by body: the bodies of at most FB_MAX markers, when no
   * body's elements are split over two runs of indices (Morse pairs act
   * between the elements of one body, and a workgroup sees one run) ---- */
  std::vector<int32_t> fbody_range, frest_items;
  {
    std::set<std::pair<int, int>> seen;
    bool ok = true;
    for (int g0 = 0; g0 < n && ok;) {
      int g1 = g0 + 1;
      while (g1 < n && M.body[g1] == M.body[g0] && !(minst && minst[g1] != minst[g0])) ++g1;
      bool elements = false;
      for (int i = g0; i < g1; ++i) elements = elements || M.kind[i] == PFG_MARKER_ELEMENT;
      if (elements) ok = seen.insert({minst ? minst[g0] : 0, M.body[g0]}).second;
      g0 = g1;
    }
    std::vector<char> in_body((size_t)n, 0);
    for (size_t g = 0; ok && g < body_range.size(); g += 4) {
      const int m0 = body_range[g], m1 = body_range[g + 1];
      if (m1 - m0 > FB_MAX) continue;
      fbody_range.push_back(m0);
      fbody_range.push_back(m1);
      for (int i = m0; i < m1; ++i) in_body[(size_t)i] = 1;
    }
    if (!fbody_range.empty())
      for (int k = 0; k < n; ++k) {
        /* the other markers in the spreading's spatial order */
        const int i = chunk_items[(size_t)k];
        if (!in_body[(size_t)i]) frest_items.push_back(i);
      }
  }
  m.n_fbodies = (int)(fbody_range.size() / 2);
  m.n_frest = (int)frest_items.size();
  m.fb_all_elements = m.n_fbodies > 0;
  for (int32_t i : frest_items) m.fb_all_elements = m.fb_all_elements && M.kind[i] != PFG_MARKER_ELEMENT;

  /* ---- host-side CSR construction ---- */
  std::vector<std::pair<int, int>> kv;
  std::vector<int32_t> m2t_start, m2t, m2e_start, m2e_nbr, m2r_start, m2r, m2l_start,
      m2l_nbr;
  std::vector<float> m2e_l0, m2l_l0;
  kv.clear();
  for (int t = 0; t < M.n_tri; ++t)
    for (int k = 0; k < 3; ++k) kv.push_back({M.tri[3 * t + k], 3 * t + k});
  build_csr(n, kv, m2t_start, m2t);
  /* bending topology (gpuSEM_LB.c:846-936): midpoint node of each triangle
   * edge (slots v0v1, v0v2, v1v2), and per vertex the ring of neighbours
   * walked as the reference walks it: the vertex after v in its first
   * triangle, then, in the first unused triangle of v holding the current
   * neighbour, the vertex before it.  A ring that does not close (a vertex
   * on the border of an open mesh) is left empty. */
  std::vector<int32_t> tri_edge((size_t)3 * M.n_tri), edge_tri, ring_start((size_t)n + 1, 0),
      ring;
  {
    static const int sl[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    std::vector<std::pair<int64_t, int32_t>> ek;
    ek.reserve((size_t)3 * M.n_tri);
    for (int t = 0; t < M.n_tri; ++t)
      for (int k = 0; k < 3; ++k) {
        int a = M.tri[3 * t + sl[k][0]], b = M.tri[3 * t + sl[k][1]];
        if (a > b) std::swap(a, b);
        ek.push_back({(int64_t)a * n + b, 3 * t + k});
      }
    std::sort(ek.begin(), ek.end());
    int ne = 0;
    for (size_t i = 0; i < ek.size();) {
      size_t j = i;
      while (j < ek.size() && ek[j].first == ek[i].first) ++j;
      edge_tri.push_back(ek[i].second);
      edge_tri.push_back(j - i > 1 ? ek[i + 1].second : -1);
      for (size_t q = i; q < j; ++q) tri_edge[(size_t)ek[q].second] = ne;
      ++ne;
      i = j;
    }
    m.n_tedges = ne;
    for (int v = 0; v < n; ++v) {
      const int c = m2t_start[(size_t)v + 1] - m2t_start[(size_t)v];
      std::vector<int> el((size_t)c), rg((size_t)c);
      for (int k = 0; k < c; ++k) el[(size_t)k] = m2t[(size_t)m2t_start[(size_t)v] + k] / 3;
      bool ok = c > 2;
      if (ok) {
        int cur = -1;
        for (int j = 0; j < 3; ++j)
          if (M.tri[3 * el[0] + j] == v) cur = M.tri[3 * el[0] + (j == 2 ? 0 : j + 1)];
        rg[0] = cur;
        for (int j = 1; j < c && ok; ++j) {
          bool found = false;
          for (int k = j; k < c && !found; ++k) {
            const int e = el[(size_t)k];
            for (int l = 0; l < 3; ++l)
              if (M.tri[3 * e + l] == cur) {
                found = true;
                cur = M.tri[3 * e + (l == 0 ? 2 : l - 1)];
                rg[(size_t)j] = cur;
                std::swap(el[(size_t)j], el[(size_t)k]);
                break;
              }
          }
          ok = found;
        }
        if (ok) { /* closed: a triangle of v holds the last and the first */
          bool shared = false;
          for (int k = 0; k < c && !shared; ++k) {
            const int e = el[(size_t)k];
            bool hl = false, hf = false;
            for (int l = 0; l < 3; ++l) {
              hl |= M.tri[3 * e + l] == rg[(size_t)c - 1];
              hf |= M.tri[3 * e + l] == rg[0];
            }
            shared = hl && hf;
          }
          ok = shared;
        }
      }
      if (ok) ring.insert(ring.end(), rg.begin(), rg.end());
      ring_start[(size_t)v + 1] = (int32_t)ring.size();
    }
    m.bend = p.kbend != 0.f && M.n_tri > 0;
  }
  kv.clear();
  for (int e = 0; e < M.n_edges; ++e) {
    kv.push_back({M.edge[2 * e], 2 * e});
    kv.push_back({M.edge[2 * e + 1], 2 * e + 1});
  }
  {
    std::vector<int32_t> tmp;
    build_csr(n, kv, m2e_start, tmp);
    m2e_nbr.resize(tmp.size());
    m2e_l0.resize(tmp.size());
    for (size_t k = 0; k < tmp.size(); ++k) {
      const int e = tmp[k] / 2, side = tmp[k] % 2;
      m2e_nbr[k] = M.edge[2 * e + (1 - side)];
      m2e_l0[k] = M.edge_l0[e];
    }
  }
  kv.clear();
  for (int r = 0; r < M.n_receptors; ++r)
    for (int k = 0; k < 3; ++k) kv.push_back({M.tri[3 * M.rec_tri[r] + k], 3 * r + k});
  build_csr(n, kv, m2r_start, m2r);
  kv.clear();
  for (int l = 0; l < M.n_links; ++l) {
    kv.push_back({M.link[2 * l], 2 * l});
    kv.push_back({M.link[2 * l + 1], 2 * l + 1});
  }
  {
    std::vector<int32_t> tmp;
    build_csr(n, kv, m2l_start, tmp);
    m2l_nbr.resize(tmp.size());
    m2l_l0.resize(tmp.size());
    for (size_t k = 0; k < tmp.size(); ++k) {
      const int l = tmp[k] / 2, side = tmp[k] % 2;
      m2l_nbr[k] = M.link[2 * l + (1 - side)];
      m2l_l0[k] = M.link_l0[l];
    }
  }
  /* cell-list geometry: bins of bin_cells cells per lattice, the lattices'
   * bins one after the other */
  int64_t nbins = 0, nbinsc = 0;
  for (Inst &I : h->inst) {
    I.nb[0] = std::max(1, I.p.nx / p.bin_cells);
    I.nb[1] = std::max(1, I.p.ny / p.bin_cells);
    I.nb[2] = std::max(1, I.p.nz / p.bin_cells);
    I.nbins = (int64_t)I.nb[0] * I.nb[1] * I.nb[2];
    I.bin0 = nbins;
    nbins += I.nbins;
    for (int k = 0; k < 3; ++k) I.nbc[k] = std::max(1, I.nb[k] / COARSE_BINS);
    I.nbinsc = (int64_t)I.nbc[0] * I.nbc[1] * I.nbc[2];
    I.binc0 = nbinsc;
    nbinsc += I.nbinsc;
  }
  if (nbins > (int64_t)INT_MAX / 2) return PFG_ERR_ARG;
  for (int k = 0; k < 3; ++k) m.nb[k] = h->inst[0].nb[k];
  m.nbins = nbins;
  m.scan_blocks = (m.nbins + SCAN_SPAN - 1) / SCAN_SPAN;
  m.nbinsc = nbinsc; /* <= nbins: the arrays of the full grid hold it */
  m.scan_blocks_c = (m.nbinsc + SCAN_SPAN - 1) / SCAN_SPAN;
  /* static ligands on the same bin grid, ascending ligand id per bin */
  std::vector<int32_t> lg_start, lg_items;
  std::vector<float> lig_z_host((size_t)M.n_wall_ligands, p.wall_z);
  {
    kv.clear();
    auto cell = [](float v, int nn, int per) {
      int c = (int)std::floor(v);
      if (per) {
        c %= nn;
        if (c < 0) c += nn;
      } else {
        c = std::min(std::max(c, 0), nn - 1);
      }
      return c;
    };
    for (int l = 0; l < M.n_wall_ligands; ++l) {
      if (M.lig_z) lig_z_host[(size_t)l] = M.lig_z[l];
      const Inst &I = h->inst[linst ? (size_t)linst[l] : 0];
      const float c3[3] = {M.lig_x[l], M.lig_y[l], lig_z_host[(size_t)l]};
      const int nn[3] = {I.p.nx, I.p.ny, I.p.nz};
      int b[3];
      for (int k = 0; k < 3; ++k)
        b[k] = cell(c3[k], nn[k], I.p.wall[k] == PFG_WALL_PERIODIC) * I.nb[k] / nn[k];
      kv.push_back({(int)I.bin0 + b[0] + I.nb[0] * (b[1] + I.nb[1] * b[2]), l});
    }
    build_csr((int)m.nbins, kv, lg_start, lg_items);
  }

  /* ---- device allocation + upload ---- */
  hipError_t e = hipSuccess;
#define ALLOC_UP(ptr, src, count)                                   \
  do {                                                              \
    if (e == hipSuccess) e = dalloc(h, &(ptr), (size_t)(count));    \
    if (e == hipSuccess && (src) != nullptr) e = upload(h, (ptr), (src), (size_t)(count)); \
  } while (0)
#define ALLOC_FILL(ptr, value, count)                               \
  do {                                                              \
    if (e == hipSuccess) e = dalloc(h, &(ptr), (size_t)(count));    \
    if (e == hipSuccess) e = fill(h, (ptr), (value), (size_t)(count)); \
  } while (0)
  ALLOC_UP(m.x, M.x, n);
  ALLOC_UP(m.y, M.y, n);
  ALLOC_UP(m.z, M.z, n);
  ALLOC_FILL(m.vx, 0.0f, n);
  ALLOC_FILL(m.vy, 0.0f, n);
  ALLOC_FILL(m.vz, 0.0f, n);
  ALLOC_FILL(m.fx, 0.0f, n);
  if (h->p.sem_integrator == PFG_SEM_BBK)
    ALLOC_FILL(m.uib, make_float4(0.f, 0.f, 0.f, 0.f), n);
  ALLOC_FILL(m.fy, 0.0f, n);
  ALLOC_FILL(m.fz, 0.0f, n);
  ALLOC_UP(m.kind, M.kind, n);
  ALLOC_UP(m.body, M.body, n);
  if (M.n_platelets > 0) {
    ALLOC_UP(m.plt_elem_start, M.plt_elem_start, M.n_platelets + 1);
    ALLOC_UP(m.plt_tri_start, M.plt_tri_start, M.n_platelets + 1);
  } else {
    const int32_t zero2[2] = {0, 0};
    ALLOC_UP(m.plt_elem_start, zero2, 1);
    ALLOC_UP(m.plt_tri_start, zero2, 1);
  }
  ALLOC_UP(m.tri, M.tri, 3 * (size_t)M.n_tri);
  ALLOC_UP(m.tri_S0, M.tri_S0, M.n_tri);
  ALLOC_FILL(m.tri_S, 0.0, M.n_tri);
  ALLOC_FILL(m.tri_V, 0.0, M.n_tri);
  ALLOC_FILL(m.tri_g, make_float4(0.f, 0.f, 0.f, 0.f), 3 * (size_t)M.n_tri);
  ALLOC_FILL(m.tri_a2, make_float4(0.f, 0.f, 0.f, 0.f), (size_t)M.n_tri);
  ALLOC_FILL(m.pos4, make_float4(0.f, 0.f, 0.f, 0.f), (size_t)n);
  ALLOC_UP(m.plt_S0, M.plt_S0, M.n_platelets);
  ALLOC_UP(m.plt_V0, M.plt_V0, M.n_platelets);
  ALLOC_FILL(m.plt_S, 0.0, M.n_platelets);
  ALLOC_FILL(m.plt_V, 0.0, M.n_platelets);
  ALLOC_FILL(m.plt_lowest, -1, M.n_platelets);
  ALLOC_UP(m.m2t_start, m2t_start.data(), m2t_start.size());
  ALLOC_UP(m.m2t, m2t.data(), m2t.size());
  ALLOC_UP(m.tri_edge, tri_edge.data(), tri_edge.size());
  ALLOC_UP(m.edge_tri, edge_tri.data(), edge_tri.size());
  ALLOC_UP(m.ring_start, ring_start.data(), ring_start.size());
  ALLOC_UP(m.ring, ring.data(), ring.size());
  ALLOC_FILL(m.vn, 0.0, m.bend ? (size_t)3 * n : 0);
  ALLOC_FILL(m.en, 0.0, m.bend ? (size_t)3 * m.n_tedges : 0);
  ALLOC_FILL(m.km, 0.0, m.bend ? (size_t)n : 0);
  {
    const bool contact = p.contact_a != 0.f && M.n_platelets >= 2;
    const size_t np = (size_t)M.n_platelets;
    ALLOC_FILL(m.contact_slot, ~(uint64_t)0, contact ? np * np : 0);
    ALLOC_FILL(m.contact_list, 0, contact ? np * (np - 1) / 2 : 0);
  }
  ALLOC_UP(m.m2e_start, m2e_start.data(), m2e_start.size());
  ALLOC_UP(m.m2e_nbr, m2e_nbr.data(), m2e_nbr.size());
  ALLOC_UP(m.m2e_l0, m2e_l0.data(), m2e_l0.size());
  ALLOC_UP(m.rec_tri, M.rec_tri, M.n_receptors);
  ALLOC_UP(m.rec_r1, M.rec_r1, M.n_receptors);
  ALLOC_UP(m.rec_r2, M.rec_r2, M.n_receptors);
  ALLOC_FILL(m.rec_bond, -1, M.n_receptors);
  ALLOC_FILL(m.rec_cand, -1, M.n_receptors);
  ALLOC_FILL(m.rec_release, 0, M.n_receptors);
  ALLOC_FILL(m.rec_force, 0.0f, 3 * (size_t)M.n_receptors);
  ALLOC_FILL(m.rec_w, 0.0f, 3 * (size_t)M.n_receptors);
  ALLOC_UP(m.m2r_start, m2r_start.data(), m2r_start.size());
  ALLOC_UP(m.m2r, m2r.data(), m2r.size());
  ALLOC_UP(m.lig_x, M.lig_x, M.n_wall_ligands);
  ALLOC_UP(m.lig_y, M.lig_y, M.n_wall_ligands);
  ALLOC_UP(m.lig_z, lig_z_host.data(), M.n_wall_ligands);
  ALLOC_FILL(m.lig_owner, -1, m.n_lig);
  ALLOC_FILL(m.lig_claim, (int32_t)INT_MAX, m.n_lig);
  ALLOC_UP(m.lg_start, lg_start.data(), lg_start.size());
  ALLOC_UP(m.lg_items, lg_items.data(), lg_items.size());
  {
    std::vector<float4> lg_pos(lg_items.size());
    for (size_t k = 0; k < lg_items.size(); ++k) {
      const int l = lg_items[k];
      float tag;
      std::memcpy(&tag, &l, sizeof(tag)); /* the index as the bits of .w */
      lg_pos[k] = make_float4(M.lig_x[l], M.lig_y[l], lig_z_host[(size_t)l], tag);
    }
    ALLOC_UP(m.lg_pos, lg_pos.data(), lg_pos.size());
  }
  ALLOC_FILL(m.rec_pos, make_float4(0.f, 0.f, 0.f, 0.f), m.n_rec);
  ALLOC_UP(m.m2l_start, m2l_start.data(), m2l_start.size());
  ALLOC_UP(m.m2l_nbr, m2l_nbr.data(), m2l_nbr.size());
  ALLOC_UP(m.m2l_l0, m2l_l0.data(), m2l_l0.size());
  ALLOC_FILL(m.bin_count, 0, m.scan_blocks * SCAN_SPAN); /* zero padding to whole scan blocks */
  ALLOC_FILL(m.bin_start, 0, m.nbins + 1);
  ALLOC_FILL(m.bin_items, 0, n);
  ALLOC_FILL(m.marker_bin, 0, n);
  ALLOC_FILL(m.marker_slot, 0, n);
  ALLOC_FILL(m.scan_partials, 0, m.scan_blocks);
  ALLOC_UP(m.chunk_start, chunk_start.data(), chunk_start.size());
  ALLOC_UP(m.chunk_items, chunk_items.data(), chunk_items.size());
  ALLOC_UP(m.chunk_ref, chunk_ref.data(), chunk_ref.size());
  if (!fbody_range.empty()) {
    ALLOC_UP(m.fbody_range, fbody_range.data(), fbody_range.size());
    if (!frest_items.empty()) ALLOC_UP(m.frest_items, frest_items.data(), frest_items.size());
    /* no list is built yet */
    ALLOC_FILL(m.fb_ref, make_float4(NAN, NAN, NAN, NAN), n);
    if (e == hipSuccess) e = dalloc(h, &m.fb_nlen, (size_t)n);
    if (e == hipSuccess) e = dalloc(h, &m.fb_nbr, (size_t)FB_NBR_MAX * n);
    ALLOC_FILL(m.fb_done, (int32_t)-1, m.n_fbodies);
  }
  if (!body_range.empty()) {
    ALLOC_UP(m.body_range, body_range.data(), body_range.size());
    ALLOC_UP(m.chunk_body, chunk_body.data(), chunk_body.size());
    if (e == hipSuccess) e = dalloc(h, &m.chunk_sum, (size_t)6 * IB_BOX_CELLS * m.n_chunks);
    if (e == hipSuccess) e = dalloc(h, &m.chunk_info, (size_t)2 * m.n_chunks);
  }
  ALLOC_UP(m.rchunk_start, rchunk_start.data(), rchunk_start.size());
  if (!rchunk_items.empty()) ALLOC_UP(m.rchunk_items, rchunk_items.data(), rchunk_items.size());
  /* no tile is claimed, and no body spread, before the first step */
  ALLOC_FILL(m.tile_owner, ~(uint64_t)0, h->ntiles > 0 ? h->ntiles : 1);
  ALLOC_FILL(m.body_spread, (int32_t)-1, m.n_bodies > 0 ? m.n_bodies : 1);
  ALLOC_FILL(m.body_fallback, (int32_t)-1, m.n_bodies > 0 ? m.n_bodies : 1);
  if (minst) ALLOC_UP(m.minst, minst, n);
  ALLOC_FILL(m.bin_scatter, 0, n);
  ALLOC_FILL(m.sorted_pos, make_float4(0.f, 0.f, 0.f, 0.f), n);
  ALLOC_FILL(m.tile_list, 0, h->ntiles > 0 ? h->ntiles : 1);
  ALLOC_FILL(m.tile_count, 0, 2);
  if (e == hipSuccess) e = dalloc(h, &m.spread_rec, n > 0 ? (size_t)n : 1);
  m.tile_gather_ok = 1;
  for (const Inst &I : h->inst) {
    const int ext[3] = {TILE_X, TILE_Y, TILE_Z}, len[3] = {I.p.nx, I.p.ny, I.p.nz};
    for (int ax = 0; ax < 3; ++ax)
      if (I.p.wall[ax] == PFG_WALL_PERIODIC && len[ax] < 2 * ext[ax]) m.tile_gather_ok = 0;
  }
#undef ALLOC_UP
#undef ALLOC_FILL
  if (e != hipSuccess) {
    free_markers(h->m);
    return from_hip(e);
  }
  /* fresh counters and step index */
  const int64_t zeros[N_COUNTERS] = {}, minus1 = -1;
  PFG_TRY(upload(h, m.counters, zeros, N_COUNTERS));
  PFG_TRY(upload(h, m.detach_step, &minus1, 1));
  PFG_TRY(upload(h, m.d_step, zeros, 1));
  /* the step index starts again: no tile carries a stamp or a force */
  PFG_TRY(hipMemset(h->tile_stamp, 0xff, sizeof(int32_t) * (size_t)h->ntiles));
  PFG_TRY(hipMemset(h->aux, 0, sizeof(float) * 3 * h->plane));
  if (h->K > 1) {
    std::vector<MLat> lats((size_t)h->K);
    for (int k = 0; k < h->K; ++k) fill_mlat(h, k, &lats[(size_t)k]);
    PFG_TRY(upload(h, h->d_lats, lats.data(), lats.size()));
  }
  h->h_step = 0;
  h->model_loaded = true;
  return PFG_OK;
}

}  // namespace

extern "C" {

pfg_status pfg_load_model(pfg_handle *h, const pfg_model *md) {
  if (!h || !md) return PFG_ERR_ARG;
  if (h->K != 1) return PFG_ERR_STATE;
  const pfg_status st = load_impl(h, *md, nullptr, nullptr);
  if (st == PFG_OK) h->marker0 = {0, md->n_markers};
  return st;
}

/* ---------------- batches ---------------- */

pfg_status pfg_batch_create(const pfg_params *params, int n_lattices, int device,
                            pfg_handle **out) {
  if (!params || n_lattices < 1 || !out) return PFG_ERR_ARG;
  return create_impl(params, n_lattices, device, out);
}

pfg_status pfg_batch_size(const pfg_handle *h, int *n_lattices) {
  if (!h || !n_lattices) return PFG_ERR_ARG;
  *n_lattices = h->K;
  return PFG_OK;
}

pfg_status pfg_batch_init_equilibrium(pfg_handle *h, int k, const float *rho, const float *ux,
                                      const float *uy, const float *uz) {
  if (!h || k < 0 || k >= h->K) return PFG_ERR_ARG;
  return init_eq_impl(h, k, rho, ux, uy, uz);
}

pfg_status pfg_batch_init_equilibrium_z(pfg_handle *h, int k, const float *ux_z,
                                        const float *uy_z, const float *uz_z) {
  if (!h || k < 0 || k >= h->K) return PFG_ERR_ARG;
  return init_eq_z_impl(h, k, ux_z, uy_z, uz_z);
}

pfg_status pfg_batch_upload_flags(pfg_handle *h, int k, const uint8_t *flags) {
  if (!h || !flags || k < 0 || k >= h->K) return PFG_ERR_ARG;
  return upload_flags_impl(h, k, flags);
}

pfg_status pfg_batch_download_fields(pfg_handle *h, int k, float *rho, float *ux, float *uy,
                                     float *uz) {
  if (!h || k < 0 || k >= h->K) return PFG_ERR_ARG;
  return fields_impl(h, k, rho, ux, uy, uz);
}

pfg_status pfg_batch_conservation(pfg_handle *h, int k, double out[4]) {
  if (!h || !out || k < 0 || k >= h->K) return PFG_ERR_ARG;
  hipSetDevice(h->device);
  return from_hip(launch_conservation(h, k, out));
}

/* The models of all lattices become one marker set: lattice k's markers,
 * triangles, edges, receptors, ligands and links follow those of k-1, with
 * their indices shifted. */
pfg_status pfg_batch_load_models(pfg_handle *h, const pfg_model *models) {
  if (!h || !models) return PFG_ERR_ARG;
  const int K = h->K;
  int64_t n = 0, npl = 0, ntri = 0, ned = 0, nrec = 0, nlig = 0, nlink = 0;
  int fibrin_lig = -1;
  for (int k = 0; k < K; ++k) {
    const pfg_model &M = models[k];
    if (M.n_markers < 0 || M.n_platelets < 0 || M.n_tri < 0 || M.n_edges < 0 ||
        M.n_receptors < 0 || M.n_wall_ligands < 0 || M.n_links < 0)
      return PFG_ERR_ARG;
    if (M.n_platelets > 0 && (!M.plt_elem_start || !M.plt_tri_start || !M.plt_S0 || !M.plt_V0))
      return PFG_ERR_ARG;
    if (M.n_markers > 0 && (!M.x || !M.y || !M.z || !M.kind || !M.body)) return PFG_ERR_ARG;
    if (M.n_tri > 0 && (!M.tri || !M.tri_S0)) return PFG_ERR_ARG;
    if (M.n_edges > 0 && (!M.edge || !M.edge_l0)) return PFG_ERR_ARG;
    if (M.n_receptors > 0 && (!M.rec_tri || !M.rec_r1 || !M.rec_r2)) return PFG_ERR_ARG;
    if (M.n_wall_ligands > 0 && (!M.lig_x || !M.lig_y)) return PFG_ERR_ARG;
    if (M.n_links > 0 && (!M.link || !M.link_l0)) return PFG_ERR_ARG;
    if (fibrin_lig >= 0 && (M.fibrin_ligands != 0) != (fibrin_lig != 0)) return PFG_ERR_ARG;
    fibrin_lig = M.fibrin_ligands != 0;
    n += M.n_markers;
    npl += M.n_platelets;
    ntri += M.n_tri;
    ned += M.n_edges;
    nrec += M.n_receptors;
    nlig += M.n_wall_ligands;
    nlink += M.n_links;
  }
  if (n > INT_MAX / 4 || ntri > INT_MAX / 4) return PFG_ERR_ARG;
  std::vector<float> x, y, z, edge_l0, r1, r2, lx, ly, lz, link_l0;
  std::vector<int32_t> kind, body, pes(1, 0), pts(1, 0), tri, edge, rec_tri, link, minst, linst;
  std::vector<double> tri_S0, plt_S0, plt_V0;
  int32_t m0 = 0, p0 = 0, t0 = 0;
  for (int k = 0; k < K; ++k) {
    const pfg_model &M = models[k];
    for (int i = 0; i < M.n_markers; ++i) {
      x.push_back(M.x[i]);
      y.push_back(M.y[i]);
      z.push_back(M.z[i]);
      kind.push_back(M.kind[i]);
      body.push_back(M.body[i] >= 0 ? M.body[i] + p0 : M.body[i]);
      minst.push_back(k);
    }
    for (int q = 0; q < M.n_platelets; ++q) {
      pes.push_back(M.plt_elem_start[q + 1] + m0);
      pts.push_back(M.plt_tri_start[q + 1] + t0);
      plt_S0.push_back(M.plt_S0[q]);
      plt_V0.push_back(M.plt_V0[q]);
    }
    for (int t = 0; t < M.n_tri; ++t) {
      for (int c = 0; c < 3; ++c) tri.push_back(M.tri[3 * t + c] + m0);
      tri_S0.push_back(M.tri_S0[t]);
    }
    for (int e = 0; e < M.n_edges; ++e) {
      edge.push_back(M.edge[2 * e] + m0);
      edge.push_back(M.edge[2 * e + 1] + m0);
      edge_l0.push_back(M.edge_l0[e]);
    }
    for (int r = 0; r < M.n_receptors; ++r) {
      rec_tri.push_back(M.rec_tri[r] + t0);
      r1.push_back(M.rec_r1[r]);
      r2.push_back(M.rec_r2[r]);
    }
    for (int l = 0; l < M.n_wall_ligands; ++l) {
      lx.push_back(M.lig_x[l]);
      ly.push_back(M.lig_y[l]);
      lz.push_back(M.lig_z ? M.lig_z[l] : h->inst[(size_t)k].p.wall_z);
      linst.push_back(k);
    }
    for (int l = 0; l < M.n_links; ++l) {
      link.push_back(M.link[2 * l] + m0);
      link.push_back(M.link[2 * l + 1] + m0);
      link_l0.push_back(M.link_l0[l]);
    }
    m0 += M.n_markers;
    p0 += M.n_platelets;
    t0 += M.n_tri;
  }
  /* each lattice's receptors and markers keep their own indices in the
   * lattice's random draws */
  {
    int32_t r0 = 0, mk0 = 0;
    for (int k = 0; k < K; ++k) {
      h->inst[(size_t)k].rec0 = r0;
      h->inst[(size_t)k].marker0 = mk0;
      r0 += models[k].n_receptors;
      mk0 += models[k].n_markers;
    }
  }
  pfg_model all{};
  all.n_markers = (int32_t)n;
  all.x = x.data();
  all.y = y.data();
  all.z = z.data();
  all.kind = kind.data();
  all.body = body.data();
  all.n_platelets = (int32_t)npl;
  all.plt_elem_start = pes.data();
  all.plt_tri_start = pts.data();
  all.n_tri = (int32_t)ntri;
  all.tri = tri.data();
  all.tri_S0 = tri_S0.data();
  all.plt_S0 = plt_S0.data();
  all.plt_V0 = plt_V0.data();
  all.n_edges = (int32_t)ned;
  all.edge = edge.data();
  all.edge_l0 = edge_l0.data();
  all.n_receptors = (int32_t)nrec;
  all.rec_tri = rec_tri.data();
  all.rec_r1 = r1.data();
  all.rec_r2 = r2.data();
  all.n_wall_ligands = (int32_t)nlig;
  all.lig_x = lx.data();
  all.lig_y = ly.data();
  all.lig_z = lz.data();
  all.fibrin_ligands = fibrin_lig > 0 ? 1 : 0;
  all.n_links = (int32_t)nlink;
  all.link = link.data();
  all.link_l0 = link_l0.data();
  const pfg_status st =
      load_impl(h, all, K > 1 ? minst.data() : nullptr, K > 1 ? linst.data() : nullptr);
  if (st == PFG_OK) {
    h->marker0.assign(1, 0);
    for (int k = 0; k < K; ++k) h->marker0.push_back(h->marker0.back() + models[k].n_markers);
  }
  return st;
}

pfg_status pfg_upload_markers(pfg_handle *h, const float *x, const float *y,
                              const float *z) {
  if (!h || !x || !y || !z) return PFG_ERR_ARG;
  if (!h->model_loaded) return PFG_ERR_STATE;
  hipSetDevice(h->device);
  const size_t n = (size_t)h->m.n;
  PFG_TRY(upload(h, h->m.x, x, n));
  PFG_TRY(upload(h, h->m.y, y, n));
  PFG_TRY(upload(h, h->m.z, z, n));
  /* tile claims and neighbour lists made from the old positions no longer
   * hold (all-ones words are NaN references: the lists are built again) */
  if (h->m.tile_owner)
    PFG_TRY(hipMemset(h->m.tile_owner, 0xff, sizeof(uint64_t) * (size_t)h->ntiles));
  if (h->m.fb_ref) PFG_TRY(hipMemset(h->m.fb_ref, 0xff, sizeof(float4) * n));
  return PFG_OK;
}

pfg_status pfg_download_markers(pfg_handle *h, float *x, float *y, float *z,
                                float *vx, float *vy, float *vz, float *fx,
                                float *fy, float *fz) {
  if (!h) return PFG_ERR_ARG;
  if (!h->model_loaded) return PFG_ERR_STATE;
  hipSetDevice(h->device);
  const size_t n = (size_t)h->m.n;
  PFG_TRY(download(h, x, h->m.x, n));
  PFG_TRY(download(h, y, h->m.y, n));
  PFG_TRY(download(h, z, h->m.z, n));
  PFG_TRY(download(h, vx, h->m.vx, n));
  PFG_TRY(download(h, vy, h->m.vy, n));
  PFG_TRY(download(h, vz, h->m.vz, n));
  PFG_TRY(download(h, fx, h->m.fx, n));
  PFG_TRY(download(h, fy, h->m.fy, n));
  PFG_TRY(download(h, fz, h->m.fz, n));
  return PFG_OK;
}

pfg_status pfg_batch_download_markers(pfg_handle *h, int k, float *x, float *y, float *z,
                                      float *vx, float *vy, float *vz, float *fx, float *fy,
                                      float *fz) {
  if (!h || k < 0 || k >= h->K) return PFG_ERR_ARG;
  if (!h->model_loaded || h->marker0.size() != (size_t)h->K + 1) return PFG_ERR_STATE;
  hipSetDevice(h->device);
  const size_t m0 = (size_t)h->marker0[(size_t)k];
  const size_t n = (size_t)h->marker0[(size_t)k + 1] - m0;
  PFG_TRY(download(h, x, h->m.x + m0, n));
  PFG_TRY(download(h, y, h->m.y + m0, n));
  PFG_TRY(download(h, z, h->m.z + m0, n));
  PFG_TRY(download(h, vx, h->m.vx + m0, n));
  PFG_TRY(download(h, vy, h->m.vy + m0, n));
  PFG_TRY(download(h, vz, h->m.vz + m0, n));
  PFG_TRY(download(h, fx, h->m.fx + m0, n));
  PFG_TRY(download(h, fy, h->m.fy + m0, n));
  PFG_TRY(download(h, fz, h->m.fz + m0, n));
  return PFG_OK;
}

pfg_status pfg_upload_marker_forces(pfg_handle *h, const float *fx,
                                    const float *fy, const float *fz) {
  if (!h || !fx || !fy || !fz) return PFG_ERR_ARG;
  if (!h->model_loaded) return PFG_ERR_STATE;
  hipSetDevice(h->device);
  const size_t n = (size_t)h->m.n;
  PFG_TRY(upload(h, h->m.fx, fx, n));
  PFG_TRY(upload(h, h->m.fy, fy, n));
  PFG_TRY(upload(h, h->m.fz, fz, n));
  return PFG_OK;
}

pfg_status pfg_upload_marker_velocities(pfg_handle *h, const float *vx,
                                        const float *vy, const float *vz) {
  if (!h || !vx || !vy || !vz) return PFG_ERR_ARG;
  if (!h->model_loaded) return PFG_ERR_STATE;
  hipSetDevice(h->device);
  const size_t n = (size_t)h->m.n;
  PFG_TRY(upload(h, h->m.vx, vx, n));
  PFG_TRY(upload(h, h->m.vy, vy, n));
  PFG_TRY(upload(h, h->m.vz, vz, n));
  return PFG_OK;
}

pfg_status pfg_download_membrane(pfg_handle *h, double *S_all, double *V,
                                 double *tri_S) {
  if (!h) return PFG_ERR_ARG;
  if (!h->model_loaded) return PFG_ERR_STATE;
  hipSetDevice(h->device);
  PFG_TRY(download(h, S_all, h->m.plt_S, (size_t)h->m.n_plt));
  PFG_TRY(download(h, V, h->m.plt_V, (size_t)h->m.n_plt));
  PFG_TRY(download(h, tri_S, h->m.tri_S, (size_t)h->m.n_tri));
  return PFG_OK;
}

pfg_status pfg_download_bonds(pfg_handle *h, int32_t *rec_bond, int32_t *lig_owner,
                              float *rec_force) {
  if (!h) return PFG_ERR_ARG;
  if (!h->model_loaded) return PFG_ERR_STATE;
  hipSetDevice(h->device);
  PFG_TRY(download(h, rec_bond, h->m.rec_bond, (size_t)h->m.n_rec));
  PFG_TRY(download(h, lig_owner, h->m.lig_owner, (size_t)h->m.n_lig));
  PFG_TRY(download(h, rec_force, h->m.rec_force, 3 * (size_t)h->m.n_rec));
  return PFG_OK;
}

pfg_status pfg_upload_bonds(pfg_handle *h, const int32_t *rec_bond) {
  if (!h || !rec_bond) return PFG_ERR_ARG;
  if (!h->model_loaded) return PFG_ERR_STATE;
  hipSetDevice(h->device);
  const Markers &m = h->m;
  std::vector<int32_t> owner((size_t)m.n_lig, -1);
  for (int r = 0; r < m.n_rec; ++r) {
    const int l = rec_bond[r];
    if (l < -1 || l >= m.n_lig) return PFG_ERR_ARG;
    if (l >= m.n_wall_lig && l >= 0) {
      if (!m.fibrin_lig) return PFG_ERR_ARG;
    }
    if (l >= 0) {
      if (owner[(size_t)l] >= 0) return PFG_ERR_ARG;
      owner[(size_t)l] = r;
    }
  }
  PFG_TRY(upload(h, m.rec_bond, rec_bond, (size_t)m.n_rec));
  PFG_TRY(upload(h, m.lig_owner, owner.data(), owner.size()));
  return PFG_OK;
}

pfg_status pfg_cell_grid(const pfg_handle *h, int32_t dims[3]) {
  if (!h || !dims) return PFG_ERR_ARG;
  if (!h->model_loaded) return PFG_ERR_STATE;
  /* the coarse grid when the lists leave the elements out */
  const bool full = bin_elements(h);
  for (int k = 0; k < 3; ++k) dims[k] = full ? h->m.nb[k] : h->inst[0].nbc[k];
  return PFG_OK;
}

pfg_status pfg_download_cells(pfg_handle *h, int32_t *bin_start, int32_t *bin_items) {
  if (!h) return PFG_ERR_ARG;
  if (!h->model_loaded) return PFG_ERR_STATE;
  hipSetDevice(h->device);
  const int64_t nb = bin_elements(h) ? h->m.nbins : h->m.nbinsc;
  PFG_TRY(download(h, bin_start, h->m.bin_start, (size_t)nb + 1));
  PFG_TRY(download(h, bin_items, h->m.bin_items, (size_t)h->m.n));
  return PFG_OK;
}

pfg_status pfg_download_aux(pfg_handle *h, float *a3n) {
  if (!h || !a3n) return PFG_ERR_ARG;
  hipSetDevice(h->device);
  for (int k = 0; k < 3; ++k)
    PFG_TRY(download(h, a3n + (size_t)k * h->N, h->aux + (size_t)k * h->plane,
                     (size_t)h->N));
  return PFG_OK;
}

/* ---------------- stepping ---------------- */

static hipError_t run_stage(pfg_handle *h, int s) {
  hipError_t e = hipSuccess;
  switch (s) {
    case PFG_STAGE_CELLS:
      e = launch_begin(h, false, true);
      if (e == hipSuccess) e = launch_scan(h);
      if (e == hipSuccess) e = launch_place(h, true, false);
      if (e == hipSuccess) e = launch_gather(h, true, false, false);
      return e;
    case PFG_STAGE_MEMBRANE:
      e = launch_place(h, false, true);
      if (e == hipSuccess) e = launch_gather(h, false, true, false);
      return e;
    case PFG_STAGE_BONDS:
      e = launch_gather(h, false, false, true);
      if (e == hipSuccess) e = launch_resolve(h);
      return e;
    case PFG_STAGE_FORCES: return launch_forces(h);
    case PFG_STAGE_SPREAD: return launch_spread(h);
    case PFG_STAGE_LB: return launch_lb(h);
    case PFG_STAGE_INTERP: return launch_move(h, true, false);
    case PFG_STAGE_UPDATE: return launch_move(h, false, true);
    case PFG_STAGE_CLEAR: return launch_begin(h, true, false);
  }
  return hipErrorInvalidValue;
}

/* One coupled step, in the order of the reference's per-step calls
 * (gpuSEM_LB.c:1441-1442 -> semLB.cu:441-531 then semLB.cu:303-438):
 * forces from X(n), spread, LB with F(n), interpolate, move.  The stages are
 * fused into seven to ten launches (cell lists, scan, scatter, sort + sums +
 * bond decisions, bond resolution, forces, spreading, LB, move; resolution
 * joins the sort launch with few receptors, spreading joins the forces with
 * a wavefront per marker); the LB update zeroes the force it reads, so the
 * next step's spreading adds into zeros without a clearing pass.
 * With overlap, the first marker launch stamps the rows the IB stencils
 * touch; the LB update of all other rows then runs on stream2 while the
 * marker pipeline runs on the main stream, and the stamped rows update
 * after the spreading.  Without overlap (or when the LB launch is timed by
 * lb0/lb1), one LB launch follows the spreading. */
static hipError_t enqueue_step(pfg_handle *h, hipEvent_t lb0, hipEvent_t lb1) {
  hipError_t e = hipSuccess;
  const bool markers = h->model_loaded && h->m.n > 0 && !h->fluid_only;
  /* the row-split LB passes take one z-row per wavefront */
  const bool overlap = markers && h->overlap && h->row_stamp && !lb0 && !lb1 &&
                       (h->K == 1 || h->lb_cpt == 1);
  if (markers) {
    e = launch_begin(h, false, true);
    if (overlap) {
      if (e == hipSuccess) e = hipEventRecord(h->ev_fork, h->stream);
      if (e == hipSuccess) e = hipStreamWaitEvent(h->stream2, h->ev_fork, 0);
      if (e == hipSuccess) e = launch_lb_part(h, 1, h->stream2);
      if (e == hipSuccess) e = hipEventRecord(h->ev_join, h->stream2);
    }
    if (e == hipSuccess) e = launch_scan(h);
    if (e == hipSuccess) e = launch_place(h, true, true);
    bool resolved = false;
    /* the bin-sorted order serves the Morse gathers, the contact search and
     * the tile gather, none of which runs when the lists leave the elements
     * out */
    if (e == hipSuccess) e = launch_gather(h, bin_elements(h), true, true, &resolved);
    if (e == hipSuccess && !resolved) e = launch_resolve(h);
    const bool fused = forces_can_spread(h);
    if (e == hipSuccess) e = launch_forces(h, fused);
    if (e == hipSuccess && !fused) e = launch_spread(h);
  }
  if (overlap) {
    if (e == hipSuccess) e = hipStreamWaitEvent(h->stream, h->ev_join, 0);
    if (e == hipSuccess) e = launch_lb_part(h, 2, h->stream);
  } else {
    if (e == hipSuccess && lb0) e = hipEventRecord(lb0, h->stream);
    if (e == hipSuccess) e = launch_lb(h);
    if (e == hipSuccess && lb1) e = hipEventRecord(lb1, h->stream);
  }
  if (markers && e == hipSuccess) e = launch_move(h, true, true);
  return e;
}

/* Graphs of one step and of GRAPH_CHUNK steps, for each starting copy of the
 * distributions (the copies alternate every step; a chunk has an even number
 * of steps and ends on the copy it started from). */
static hipError_t build_graphs(pfg_handle *h) {
  drop_graph(h);
  const int cur0 = h->cur;
  hipError_t e = hipSuccess;
  for (int g = 0; g < 2 && e == hipSuccess; ++g) {
    for (int chunk = 0; chunk < 2 && e == hipSuccess; ++chunk) {
      h->cur = g;
      hipGraph_t graph = nullptr;
      e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) break;
      hipError_t e2 = hipSuccess;
      const int n = chunk ? GRAPH_CHUNK : 1;
      for (int s = 0; s < n && e2 == hipSuccess; ++s) e2 = enqueue_step(h, nullptr, nullptr);
      e = hipStreamEndCapture(h->stream, &graph);
      if (e2 != hipSuccess) e = e2;
      if (e == hipSuccess) {
        hipGraphExec_t ex = nullptr;
        e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
        if (e == hipSuccess) h->graphs[g][chunk] = ex;
      }
      if (graph) hipGraphDestroy(graph);
    }
  }
  h->cur = cur0;
  if (e == hipSuccess) h->graph_valid = true;
  else drop_graph(h);
  return e;
}

static pfg_status fill_info(pfg_handle *h, pfg_step_info *info) {
  if (!info) return PFG_OK;
  int64_t c[N_COUNTERS] = {};
  int64_t ds = -1;
  PFG_TRY(download(h, c, h->m.counters, N_COUNTERS));
  PFG_TRY(download(h, &ds, h->m.detach_step, 1));
  info->step = h->h_step;
  info->total_bonds = (int32_t)c[0];
  info->detached = (int32_t)c[1];
  info->detach_step = ds;
  info->near_threshold = c[2];
  info->bond_events = c[4];
  info->near_cutoff = c[8];
  info->contact_overflow = c[9];
  return PFG_OK;
}

pfg_status pfg_step(pfg_handle *h, int nsteps, pfg_step_info *info) {
  if (!h || nsteps < 0) return PFG_ERR_ARG;
  hipSetDevice(h->device);
  if (h->use_graph && nsteps > 0 && !h->graph_valid) PFG_TRY(build_graphs(h));
  int s = 0;
  if (h->use_graph) {
    for (; s + GRAPH_CHUNK <= nsteps; s += GRAPH_CHUNK) {
      PFG_TRY(hipGraphLaunch(h->graphs[h->cur][1], h->stream));
      h->h_step += GRAPH_CHUNK;
    }
  }
  for (; s < nsteps; ++s) {
    if (h->use_graph) {
      PFG_TRY(hipGraphLaunch(h->graphs[h->cur][0], h->stream));
      h->cur = 1 - h->cur;
    } else {
      PFG_TRY(enqueue_step(h, nullptr, nullptr));
    }
    h->h_step++;
  }
  PFG_TRY(hipStreamSynchronize(h->stream));
  if (h->halt_on_detach && h->model_loaded && !h->fluid_only) {
    /* steps launched after the detach test fired left the lattice in place
     * while the host alternated the copies: realign the current copy with
     * the steps that actually ran */
    int64_t done = 0;
    PFG_TRY(download(h, &done, h->m.d_step, 1));
    if (h->h_step > done) {
      if ((h->h_step - done) & 1) h->cur = 1 - h->cur;
      h->h_step = done;
    }
  }
  return fill_info(h, info);
}

pfg_status pfg_set_overlap(pfg_handle *h, int enable) {
  if (!h) return PFG_ERR_ARG;
  h->overlap = enable != 0;
  drop_graph(h);
  return PFG_OK;
}

pfg_status pfg_set_ib_staging(pfg_handle *h, int mode) {
  if (!h || mode < 0 || mode > 4) return PFG_ERR_ARG;
  h->ib_staging = mode;
  drop_graph(h);
  return PFG_OK;
}

pfg_status pfg_set_forces_lanes(pfg_handle *h, int lanes) {
  if (!h || !(lanes == 0 || lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16 ||
              lanes == 32 || lanes == 64))
    return PFG_ERR_ARG;
  h->forces_lanes = lanes;
  drop_graph(h);
  return PFG_OK;
}

pfg_status pfg_set_forces_by_body(pfg_handle *h, int mode) {
  if (!h || mode < 0 || mode > 2) return PFG_ERR_ARG;
  h->forces_body = mode;
  drop_graph(h);
  return PFG_OK;
}

pfg_status pfg_set_halt_on_detach(pfg_handle *h, int enable) {
  if (!h) return PFG_ERR_ARG;
  h->halt_on_detach = enable != 0;
  drop_graph(h);
  hipSetDevice(h->device);
  PFG_TRY(upload_batch_args(h)); /* the batched LB arguments carry the flag */
  return PFG_OK;
}

pfg_status pfg_run_stage(pfg_handle *h, int stage) {
  if (!h || stage < 0 || stage >= PFG_STAGE_COUNT) return PFG_ERR_ARG;
  if (stage != PFG_STAGE_LB && !h->model_loaded) return PFG_ERR_STATE;
  hipSetDevice(h->device);
  PFG_TRY(run_stage(h, stage));
  PFG_TRY(hipStreamSynchronize(h->stream));
  return PFG_OK;
}

pfg_status pfg_set_step(pfg_handle *h, int64_t step) {
  if (!h || step < 0) return PFG_ERR_ARG;
  hipSetDevice(h->device);
  PFG_TRY(upload(h, h->m.d_step, &step, 1));
  /* stamps and claims of other steps could equal the new index: no tile or
   * row keeps one (the tile-gather spreading lists a tile when its stamp
   * changes; body spreading stores into a tile its body claimed alone) */
  PFG_TRY(hipMemset(h->tile_stamp, 0xff, sizeof(int32_t) * (size_t)h->ntiles));
  if (h->m.tile_owner)
    PFG_TRY(hipMemset(h->m.tile_owner, 0xff, sizeof(uint64_t) * (size_t)h->ntiles));
  if (h->m.body_spread)
    PFG_TRY(hipMemset(h->m.body_spread, 0xff, sizeof(int32_t) * (size_t)std::max(1, h->m.n_bodies)));
  if (h->m.body_fallback)
    PFG_TRY(hipMemset(h->m.body_fallback, 0xff, sizeof(int32_t) * (size_t)std::max(1, h->m.n_bodies)));
  if (h->m.fb_done) PFG_TRY(hipMemset(h->m.fb_done, 0xff, sizeof(int32_t) * (size_t)h->m.n_fbodies));
  if (h->row_stamp) {
    size_t rows = 0;
    for (const Inst &I : h->inst) rows += (size_t)I.nxb * I.p.ny * I.p.nz;
    PFG_TRY(hipMemset(h->row_stamp, 0xff, sizeof(int32_t) * rows));
  }
  h->h_step = step;
  return PFG_OK;
}

pfg_status pfg_set_graph(pfg_handle *h, int enable) {
  if (!h) return PFG_ERR_ARG;
  h->use_graph = enable != 0;
  drop_graph(h);
  return PFG_OK;
}

pfg_status pfg_set_fluid_only(pfg_handle *h, int fluid_only) {
  if (!h) return PFG_ERR_ARG;
  h->fluid_only = fluid_only != 0;
  drop_graph(h);
  return PFG_OK;
}

pfg_status pfg_synchronize(pfg_handle *h) {
  if (!h) return PFG_ERR_ARG;
  hipSetDevice(h->device);
  PFG_TRY(hipStreamSynchronize(h->stream));
  return PFG_OK;
}

pfg_status pfg_profile(pfg_handle *h, int nsteps, double *ms_total, double *ms_lb) {
  if (!h || nsteps <= 0) return PFG_ERR_ARG;
  hipSetDevice(h->device);
  std::vector<hipEvent_t> ev((size_t)2 * nsteps, nullptr);
  hipEvent_t t0 = nullptr, t1 = nullptr;
  hipError_t e = hipEventCreate(&t0);
  if (e == hipSuccess) e = hipEventCreate(&t1);
  for (size_t k = 0; k < ev.size() && e == hipSuccess; ++k) e = hipEventCreate(&ev[k]);
  if (e == hipSuccess) e = hipEventRecord(t0, h->stream);
  for (int s = 0; s < nsteps && e == hipSuccess; ++s) {
    e = enqueue_step(h, ev[2 * (size_t)s], ev[2 * (size_t)s + 1]);
    h->h_step++;
  }
  if (e == hipSuccess) e = hipEventRecord(t1, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  double lb = 0.0;
  float ms = 0.f;
  for (int s = 0; s < nsteps && e == hipSuccess; ++s) {
    e = hipEventElapsedTime(&ms, ev[2 * (size_t)s], ev[2 * (size_t)s + 1]);
    lb += ms;
  }
  float tot = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&tot, t0, t1);
  for (auto x : ev)
    if (x) hipEventDestroy(x);
  if (t0) hipEventDestroy(t0);
  if (t1) hipEventDestroy(t1);
  if (e != hipSuccess) return from_hip(e);
  if (ms_total) *ms_total = tot;
  if (ms_lb) *ms_lb = lb;
  return PFG_OK;
}

pfg_status pfg_set_lb_tuning(pfg_handle *h, int cells_per_lane) {
  if (!h) return PFG_ERR_ARG;
  if (cells_per_lane != 1 && cells_per_lane != 2 && cells_per_lane != 4) return PFG_ERR_ARG;
  if (cells_per_lane == h->lb_cpt) return PFG_OK;
  hipSetDevice(h->device);
  h->lb_cpt = cells_per_lane;
  drop_graph(h);
  if (h->K > 1) { /* the batched launch's workgroups depend on it */
    PFG_TRY(hipStreamSynchronize(h->stream));
    PFG_TRY(batch_layout(h));
    PFG_TRY(upload_batch_args(h));
  }
  return PFG_OK;
}

/* ---- autotuning of the launch shapes ---- */

namespace {

/* The state a step carries into the next one: the current population copy,
 * marker positions and velocities, bonds and ligand ownership, the counters,
 * the detach step and the step index.  Everything else is rebuilt inside a
 * step (cell lists, membrane sums, forces) or reset here (the force field,
 * the tile and row stamps). */
struct Snapshot {
  void *f = nullptr;
  size_t f_bytes = 0;
  float *mk = nullptr;   /* x, y, z, vx, vy, vz */
  int32_t *bonds = nullptr; /* rec_bond, lig_owner, lig_claim */
  int64_t *ctr = nullptr;   /* counters[N_COUNTERS], detach_step, d_step */
  int cur = 0;
  int64_t h_step = 0;
};

size_t population_bytes(const pfg_handle *h) {
  size_t blocks = 0;
  for (const Inst &I : h->inst) blocks += (size_t)I.nxb * I.p.ny * I.p.nz;
  return (size_t)h->esize * Q * XBLOCK * blocks;
}

void snapshot_free(Snapshot &S) {
  if (S.f) hipFree(S.f);
  if (S.mk) hipFree(S.mk);
  if (S.bonds) hipFree(S.bonds);
  if (S.ctr) hipFree(S.ctr);
  S = Snapshot{};
}

/* copies between the handle and the snapshot on the handle's stream */
hipError_t snapshot_copy(pfg_handle *h, Snapshot &S, bool save) {
  const Markers &m = h->m;
  auto cp = [&](void *hv, void *sv, size_t bytes) {
    if (bytes == 0) return hipSuccess;
    return save ? hipMemcpyAsync(sv, hv, bytes, hipMemcpyDeviceToDevice, h->stream)
                : hipMemcpyAsync(hv, sv, bytes, hipMemcpyDeviceToDevice, h->stream);
  };
  const int cur = save ? h->cur : S.cur;
  hipError_t e = cp(h->f[cur], S.f, S.f_bytes);
  if (h->model_loaded) {
    const size_t n = (size_t)m.n;
    float *const arr[6] = {m.x, m.y, m.z, m.vx, m.vy, m.vz};
    for (int k = 0; k < 6 && e == hipSuccess; ++k) e = cp(arr[k], S.mk + k * n, 4 * n);
    const size_t nr = (size_t)m.n_rec, nl = (size_t)m.n_lig;
    if (e == hipSuccess) e = cp(m.rec_bond, S.bonds, 4 * nr);
    if (e == hipSuccess) e = cp(m.lig_owner, S.bonds + nr, 4 * nl);
    if (e == hipSuccess) e = cp(m.lig_claim, S.bonds + nr + nl, 4 * nl);
  }
  if (e == hipSuccess) e = cp(h->m.counters, S.ctr, N_COUNTERS * sizeof(int64_t));
  if (e == hipSuccess) e = cp(h->m.detach_step, S.ctr + N_COUNTERS, sizeof(int64_t));
  if (e == hipSuccess) e = cp(h->m.d_step, S.ctr + N_COUNTERS + 1, sizeof(int64_t));
  if (!save) {
    /* no tile or row carries a stamp or a force of the steps being undone */
    if (e == hipSuccess) e = hipMemsetAsync(h->tile_stamp, 0xff, 4 * (size_t)h->ntiles, h->stream);
    if (e == hipSuccess && h->m.tile_owner)
      e = hipMemsetAsync(h->m.tile_owner, 0xff, 8 * (size_t)h->ntiles, h->stream);
    if (e == hipSuccess && h->m.body_spread)
      e = hipMemsetAsync(h->m.body_spread, 0xff, 4 * (size_t)std::max(1, h->m.n_bodies), h->stream);
    if (e == hipSuccess && h->m.body_fallback)
      e = hipMemsetAsync(h->m.body_fallback, 0xff, 4 * (size_t)std::max(1, h->m.n_bodies), h->stream);
    if (e == hipSuccess && h->m.fb_done)
      e = hipMemsetAsync(h->m.fb_done, 0xff, 4 * (size_t)h->m.n_fbodies, h->stream);
    if (e == hipSuccess && h->row_stamp) {
      size_t rows = 0;
      for (const Inst &I : h->inst) rows += (size_t)I.nxb * I.p.ny * I.p.nz;
      e = hipMemsetAsync(h->row_stamp, 0xff, 4 * rows, h->stream);
    }
    if (e == hipSuccess) e = hipMemsetAsync(h->aux, 0, sizeof(float) * 3 * (size_t)h->plane, h->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess && save) {
    S.cur = h->cur;
    S.h_step = h->h_step;
  }
  if (e == hipSuccess && !save) {
    h->cur = S.cur;
    h->h_step = S.h_step;
  }
  return e;
}

hipError_t snapshot_take(pfg_handle *h, Snapshot &S) {
  S.f_bytes = population_bytes(h);
  const size_t n = h->model_loaded ? (size_t)h->m.n : 0;
  const size_t nb = h->model_loaded ? (size_t)h->m.n_rec + 2 * (size_t)h->m.n_lig : 0;
  hipError_t e = hipMalloc(&S.f, S.f_bytes);
  if (e == hipSuccess) e = hipMalloc((void **)&S.mk, 6 * 4 * (n > 0 ? n : 1));
  if (e == hipSuccess) e = hipMalloc((void **)&S.bonds, 4 * (nb > 0 ? nb : 1));
  if (e == hipSuccess) e = hipMalloc((void **)&S.ctr, (N_COUNTERS + 2) * sizeof(int64_t));
  if (e == hipSuccess) e = snapshot_copy(h, S, true);
  return e;
}

struct Shape {
  int cpt, lanes, staging, overlap, xcd, persistent, fbody;
};

hipError_t apply_shape(pfg_handle *h, const Shape &s) {
  drop_graph(h);
  h->lb_xcd_rows = s.xcd;
  h->forces_lanes = s.lanes;
  h->forces_body = s.fbody;
  h->ib_staging = s.staging;
  h->overlap = s.overlap != 0;
  h->lb_persistent = h->K > 1 ? s.persistent : 0;
  if (h->lb_cpt == s.cpt) return hipSuccess;
  h->lb_cpt = s.cpt;
  if (h->K == 1) return hipSuccess;
  /* the batched LB launch's workgroups follow the cells per lane */
  hipError_t e = batch_layout(h);
  if (e == hipSuccess) e = upload_batch_args(h);
  return e;
}

/* ms per step of `steps` steps from the snapshot (after `warm` steps from the
 * snapshot that build the graphs and warm the caches) */
pfg_status time_shape(pfg_handle *h, Snapshot &S, int steps, int warm, double *ms) {
  PFG_TRY(snapshot_copy(h, S, false));
  if (warm > 0) {
    pfg_status st = pfg_step(h, warm, nullptr);
    if (st != PFG_OK) return st;
    PFG_TRY(snapshot_copy(h, S, false));
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipEventRecord(e0, h->stream);
  pfg_status st = PFG_OK;
  if (e == hipSuccess) st = pfg_step(h, steps, nullptr);
  if (e == hipSuccess && st == PFG_OK) e = hipEventRecord(e1, h->stream);
  if (e == hipSuccess && st == PFG_OK) e = hipEventSynchronize(e1);
  float t = 0.f;
  if (e == hipSuccess && st == PFG_OK) e = hipEventElapsedTime(&t, e0, e1);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (st != PFG_OK) return st;
  if (e != hipSuccess) return from_hip(e);
  *ms = (double)t / steps;
  return PFG_OK;
}

}  // namespace

pfg_status pfg_autotune(pfg_handle *h, int steps, pfg_tuning *out) {
  if (!h || steps <= 0 || !out) return PFG_ERR_ARG;
  if (h->halt_on_detach) return PFG_ERR_STATE; /* a detach inside a trial ends it early */
  hipSetDevice(h->device);
  const bool markers = h->model_loaded && h->m.n > 0 && !h->fluid_only;
  const Shape start{h->lb_cpt, h->forces_lanes, h->ib_staging, h->overlap ? 1 : 0,
                    h->lb_xcd_rows, h->lb_persistent, h->forces_body};
  Snapshot S;
  hipError_t e = snapshot_take(h, S);
  if (e != hipSuccess) {
    snapshot_free(S);
    return from_hip(e);
  }
  /* knobs, tried one at a time from the automatic shape, each kept at its
   * fastest value before the next is tried (values 0 are the automatic
   * choices); a value replaces the best only when it is 1 % faster */
  const int cpts[3] = {1, 2, 4}, lanes[7] = {64, 32, 16, 8, 4, 2, 1}, ovl[2] = {0, 1},
            xcds[2] = {0, 1}, pers[2] = {0, 1}, fbod[2] = {1, 2};
  /* spreading forms that apply to this model */
  int stag[4] = {1, 2, 0, 0}, nstag = 2;
  if (h->m.n_bodies > 0) stag[nstag++] = 4;
  if (h->m.tile_gather_ok) stag[nstag++] = 3;
  Shape best{1, 0, 0, 0, (h->K == 1 && h->N >= XCD_ROWS_MIN_CELLS) ? 1 : 0, 0, 0};
  const int warm = GRAPH_CHUNK;
  pfg_status st = PFG_OK;
  double best_ms = 0.0;
  int tried = 0;
  e = apply_shape(h, best);
  st = e == hipSuccess ? time_shape(h, S, steps, warm, &best_ms) : from_hip(e);
  out->ms_default = best_ms;
  tried = 1;
  /* knobs: 0 LB cells per lane, 1 LB row order, 5 the
   * persistent batched launch (batches), 6 the force pass by body, 2 force
   * lanes, 3 spreading form, 4 overlap */
  const int order[7] = {0, 1, 5, 6, 2, 3, 4};
  for (int ki = 0; ki < 7 && st == PFG_OK; ++ki) {
    const int knob = order[ki];
    const int *vals = knob == 0 ? cpts : knob == 1 ? xcds : knob == 2 ? lanes
                                       : knob == 3 ? stag : knob == 4 ? ovl
                                                          : knob == 5 ? pers : fbod;
    const int nv = knob == 0 ? 3 : knob == 2 ? 7 : knob == 3 ? nstag : 2;
    if (knob == 5 && h->K == 1) continue;
    if (knob == 6 && (h->m.n_fbodies == 0 || !markers)) continue;
    if ((knob >= 2 && knob <= 4) && !markers) break;
    for (int v = 0; v < nv && st == PFG_OK; ++v) {
      Shape s = best;
      int *slot = knob == 0 ? &s.cpt : knob == 1 ? &s.xcd : knob == 2 ? &s.lanes
                                     : knob == 3 ? &s.staging : knob == 4 ? &s.overlap
                                     : knob == 5 ? &s.persistent : &s.fbody;
      if (*slot == vals[v]) continue;
      *slot = vals[v];
      double ms = 0.0;
      e = apply_shape(h, s);
      st = e == hipSuccess ? time_shape(h, S, steps, warm, &ms) : from_hip(e);
      ++tried;
      if (st == PFG_OK && ms < 0.99 * best_ms) {
        best = s;
        best_ms = ms;
      }
    }
  }
  if (st == PFG_OK) {
    e = apply_shape(h, best);
    if (e == hipSuccess) e = snapshot_copy(h, S, false);
    if (e != hipSuccess) st = from_hip(e);
  } else {
    apply_shape(h, start);
    snapshot_copy(h, S, false);
  }
  snapshot_free(S);
  if (st != PFG_OK) return st;
  out->lb_cells_per_lane = best.cpt;
  out->lb_xcd_rows = best.xcd;
  out->forces_lanes = best.lanes;
  out->ib_staging = best.staging;
  out->overlap = best.overlap;
  out->lb_persistent = best.persistent;
  out->forces_by_body = best.fbody;
  out->ms_per_step = best_ms;
  out->shapes_timed = tried;
  return PFG_OK;
}

pfg_status pfg_set_lb_persistent(pfg_handle *h, int enable) {
  if (!h) return PFG_ERR_ARG;
  if (h->K == 1) return enable ? PFG_ERR_STATE : PFG_OK;
  h->lb_persistent = enable != 0;
  drop_graph(h);
  return PFG_OK;
}

pfg_status pfg_set_lb_xcd_rows(pfg_handle *h, int enable) {
  if (!h) return PFG_ERR_ARG;
  h->lb_xcd_rows = enable != 0;
  drop_graph(h);
  return PFG_OK;
}

pfg_status pfg_time_lb(pfg_handle *h, int nlaunch, double *ms_per_launch) {
  if (!h || nlaunch <= 0 || !ms_per_launch) return PFG_ERR_ARG;
  hipSetDevice(h->device);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipEventRecord(e0, h->stream);
  for (int k = 0; k < nlaunch && e == hipSuccess; ++k) {
    e = launch_lb(h);
    h->h_step++;
  }
  if (e == hipSuccess) e = hipEventRecord(e1, h->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e == hipSuccess) *ms_per_launch = (double)ms / nlaunch;
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  return from_hip(e);
}

pfg_status pfg_copy_probe(int device, size_t bytes, int reps, double *gbs) {
  if (!gbs || reps <= 0 || bytes < 16) return PFG_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return PFG_ERR_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return PFG_ERR_DEVICE;
  const size_t n4 = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, n4 * 16);
  if (e == hipSuccess) e = hipMalloc(&b, n4 * 16);
  if (e == hipSuccess) e = hipMemset(a, 0, n4 * 16);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  const dim3 grid((unsigned)((n4 + 255) / 256));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(copy_probe_kernel, grid, dim3(256), 0, nullptr, a, b, n4);
    e = hipEventRecord(e0, nullptr);
  }
  for (int r = 0; r < reps && e == hipSuccess; ++r) {
    hipLaunchKernelGGL(copy_probe_kernel, grid, dim3(256), 0, nullptr, a, b, n4);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e == hipSuccess) *gbs = 2.0 * (double)n4 * 16.0 * reps / (ms * 1e-3) / 1e9;
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (a) hipFree(a);
  if (b) hipFree(b);
  return from_hip(e);
}

}  // extern "C"
