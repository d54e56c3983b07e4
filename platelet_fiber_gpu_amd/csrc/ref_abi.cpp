/*
 * ref_abi.cpp -- the reference's own C entry points (include/pfg_ref.h)
 * over the pfg ABI.
 *
 * The reference keeps three host structs (fluid, SEM and fibrin grids) that
 * own pitched FP64 device copies of the host arrays and re-copies the SEM
 * struct every step (semLB.cu:111-1034).  Here the fluid grid owns one pfg
 * handle; the SEM and fibrin grids only hold what their init calls handed
 * over, and the handle's model is built from both at the first step.  Host
 * arrays stay in the reference's units and layouts; the conversions are:
 *   position   device = X / dx (z - 1 in the channel mapping, where the
 *              device lattice is the interior planes)
 *   velocity   device = V / C  (the reference's velocities are in UMAX)
 *   force      device = F      (both in units of F0)
 *   populations device = fIN   (the reference's f are lattice-unit ones:
 *              gpuSEM_LB.c:1247-1252 divides c.u by C)
 */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "pfg.h"
#include "pfg_ref.h"

namespace {

pfg_status g_last = PFG_OK;

bool report(const char *where, pfg_status st) {
  g_last = st;
  if (st == PFG_OK) return true;
  std::fprintf(stderr, "pfg: %s: %s\n", where, pfg_status_string(st));
  return false;
}

struct FiberGrid {
  int max_nodes = 0, max_links = 0, max_conn = 0;
  double dt = 0;
  std::vector<int> type;       /* 1: anchored                       */
  std::vector<int> link;       /* 2 per link                        */
  std::vector<double> l0;      /* rest lengths, micrometres         */
  std::vector<double> x, y, z; /* micrometres                       */
  bool loaded = false;
  int marker0 = 0;             /* first marker of the network once the
                                  model is on the device            */
  struct FluidGrid *fluid = nullptr;
};

struct SemGrid {
  int max_cells = 0, max_elems = 0, surf = 0, newnode = 0, nrec_per_tri = 0;
  float dt = 0;
  double S0_all = 0, V0 = 0;
  std::vector<float> params;   /* hostSEMparameters[0..19]          */
  int n_cells = 0;
  std::vector<int> n_elem;
  std::vector<double> X, Y, Z; /* [maxElements][maxCells], um      */
  std::vector<double> bonds;   /* [maxElements][maxElements], um   */
  std::vector<int> tri6;       /* [SurfElem][6]                     */
  std::vector<float> r1, r2;   /* [SurfElem][nrec]                  */
  std::vector<double> S0;
  std::vector<int> rbond;      /* [3][nrec][SurfElem]               */
  bool loaded = false;
  /* marker index of element (m, c), -1 beyond numOfElements */
  std::vector<int> marker_of;
  struct FluidGrid *fluid = nullptr;
};

struct FluidGrid {
  float dt = 0, dx = 0;
  int width = 0, height = 0, depth = 0;
  pfg_handle *h = nullptr;
  pfg_params p{};
  bool channel = false;        /* device lattice = interior z planes */
  double C = 1.0;              /* dx / dt / UMAX                      */
  double u_lo = 0, u_hi = 0;   /* wall velocities along y (UMAX units) */
  double u_plate = 0;          /* the top obstacle plane's velocity    */
  SemGrid *sem = nullptr;
  FiberGrid *fiber = nullptr;
  bool model_on_device = false;
  int64_t n_device = 0;        /* device cells                        */
};

FiberGrid *g_fiber = nullptr;  /* the last fibrin grid: the reference passes
                                  it to no other call (semLB.cu:562-718) */

inline int dz_off(const FluidGrid *F) { return F->channel ? 1 : 0; }

/* Build the device model from the SEM and fibrin grids (first step). */
bool load_model(FluidGrid *F) {
  if (F->model_on_device) return true;
  SemGrid *S = F->sem;
  FiberGrid *B = F->fiber ? F->fiber : g_fiber;
  if (B && !B->loaded) B = nullptr;
  if (!S || !S->loaded) {
    if (!B) return report("model", PFG_ERR_STATE);
  }
  const double dx = F->dx;
  const double zoff = dz_off(F);
  std::vector<float> x, y, z;
  std::vector<int32_t> kind, body, pes(1, 0), pts(1, 0), tri, edge, rec_tri, link;
  std::vector<float> edge_l0, r1, r2, lx, ly, lz, link_l0;
  std::vector<double> tri_S0, plt_S0, plt_V0;
  if (S && S->loaded) {
    S->marker_of.assign((size_t)S->max_elems * S->max_cells, -1);
    for (int c = 0; c < S->n_cells; ++c) {
      const int base = (int)x.size();
      for (int m = 0; m < S->n_elem[(size_t)c]; ++m) {
        const size_t k = (size_t)m * S->max_cells + c;
        S->marker_of[k] = (int)x.size();
        x.push_back((float)(S->X[k] / dx));
        y.push_back((float)(S->Y[k] / dx));
        z.push_back((float)(S->Z[k] / dx - zoff));
        kind.push_back(PFG_MARKER_ELEMENT);
        body.push_back(c);
      }
      /* the surface mesh of the reference's platelet (its kernels take
       * cell 0, sem_kernel.cu:2094); every cell with that element count
       * carries it */
      const int ne = S->n_elem[(size_t)c];
      const bool mesh = S->surf > 0 && (c == 0 || ne == S->n_elem[0]);
      const int t0 = (int)tri_S0.size();
      if (mesh) {
        for (int t = 0; t < S->surf; ++t) {
          for (int v = 0; v < 3; ++v) tri.push_back(base + S->tri6[(size_t)t * 6 + v]);
          tri_S0.push_back(S->S0[(size_t)t] / (dx * dx));
          for (int r = 0; r < S->nrec_per_tri; ++r) {
            rec_tri.push_back(t0 + t);
            r1.push_back(S->r1[(size_t)t * S->nrec_per_tri + r]);
            r2.push_back(S->r2[(size_t)t * S->nrec_per_tri + r]);
          }
        }
      }
      /* springs: the positive entries of the dense rest-length matrix
       * (sem_kernel.cu:356-396) */
      for (int i = 0; i < ne; ++i)
        for (int j = i + 1; j < ne; ++j) {
          const double L = S->bonds[(size_t)i * S->max_elems + j];
          if (L > 0) {
            edge.push_back(base + i);
            edge.push_back(base + j);
            edge_l0.push_back((float)(L / dx));
          }
        }
      pes.push_back((int)x.size());
      pts.push_back((int)tri_S0.size());
      plt_S0.push_back(S->S0_all / (dx * dx));
      plt_V0.push_back(S->V0 / (dx * dx * dx));
    }
  }
  if (B) {
    B->marker0 = (int)x.size();
    for (int i = 0; i < B->max_nodes; ++i) {
      x.push_back((float)(B->x[(size_t)i] / dx));
      y.push_back((float)(B->y[(size_t)i] / dx));
      z.push_back((float)(B->z[(size_t)i] / dx - zoff));
      kind.push_back(B->type[(size_t)i] == 1 ? PFG_MARKER_FIXED : PFG_MARKER_FIBER);
      body.push_back(-1);
    }
    for (int l = 0; l < B->max_links; ++l) {
      link.push_back(B->marker0 + B->link[(size_t)2 * l]);
      link.push_back(B->marker0 + B->link[(size_t)2 * l + 1]);
      link_l0.push_back((float)(B->l0[(size_t)l] / dx));
    }
  }
  /* the vWF grid of the bottom wall: 4 slots per lattice node
   * (vWFbond[4][height][width], semLB.cu:127), as 4 ligands at the node */
  const int W = F->width, H = F->height;
  if (S && S->loaded && !rec_tri.empty()) {
    for (int k = 0; k < 4; ++k)
      for (int j = 0; j < H; ++j)
        for (int i = 0; i < W; ++i) {
          lx.push_back((float)i);
          ly.push_back((float)j);
          lz.push_back((float)(-zoff));
        }
  }
  pfg_model m{};
  m.n_markers = (int32_t)x.size();
  m.x = x.data();
  m.y = y.data();
  m.z = z.data();
  m.kind = kind.data();
  m.body = body.data();
  m.n_platelets = (int32_t)plt_S0.size();
  m.plt_elem_start = pes.data();
  m.plt_tri_start = pts.data();
  m.n_tri = (int32_t)tri_S0.size();
  m.tri = tri.data();
  m.tri_S0 = tri_S0.data();
  m.plt_S0 = plt_S0.data();
  m.plt_V0 = plt_V0.data();
  m.n_edges = (int32_t)edge_l0.size();
  m.edge = edge.data();
  m.edge_l0 = edge_l0.data();
  m.n_receptors = (int32_t)rec_tri.size();
  m.rec_tri = rec_tri.data();
  m.rec_r1 = r1.data();
  m.rec_r2 = r2.data();
  m.n_wall_ligands = (int32_t)lx.size();
  m.lig_x = lx.data();
  m.lig_y = ly.data();
  m.lig_z = lz.data();
  m.fibrin_ligands = 0;
  m.n_links = (int32_t)link_l0.size();
  m.link = link.data();
  m.link_l0 = link_l0.data();
  if (!report("pfg_load_model", pfg_load_model(F->h, &m))) return false;
  /* bonds held at the start (receptBond slot 2 = 1, at grid node (i, j)):
   * the first free slot of the node */
  if (S && S->loaded && m.n_receptors > 0) {
    const int nr = S->nrec_per_tri, T = S->surf;
    std::vector<int32_t> rb((size_t)m.n_receptors, -1);
    std::vector<int> used((size_t)W * H, 0);
    bool any = false;
    for (int t = 0; t < T; ++t)
      for (int r = 0; r < nr; ++r) {
        const size_t k = (size_t)r * T + t;
        if (S->rbond[(size_t)2 * nr * T + k] != 1) continue;
        const int i = S->rbond[k], j = S->rbond[(size_t)nr * T + k];
        if (i < 0 || i >= W || j < 0 || j >= H) continue;
        int &u = used[(size_t)j * W + i];
        if (u >= 4) continue;
        rb[(size_t)t * nr + r] = (u++) * W * H + j * W + i;
        any = true;
      }
    if (any && !report("pfg_upload_bonds", pfg_upload_bonds(F->h, rb.data()))) return false;
  }
  F->model_on_device = true;
  return true;
}

}  // namespace

extern "C" {

pfg_status pfg_ref_last_status(void) { return g_last; }

pfg_handle *pfg_ref_handle(void *g) { return g ? ((FluidGrid *)g)->h : nullptr; }

void *fluid_allocGPUKernel(void *model, float dt, float dx, int width, int height, int depth) {
  (void)model;
  if (!(dt > 0.f) || !(dx > 0.f) || width <= 0 || height <= 0 || depth <= 0) {
    report("fluid_allocGPUKernel", PFG_ERR_ARG);
    return nullptr;
  }
  FluidGrid *F = new (std::nothrow) FluidGrid();
  if (!F) {
    report("fluid_allocGPUKernel", PFG_ERR_ALLOC);
    return nullptr;
  }
  F->dt = dt;
  F->dx = dx;
  F->width = width;
  F->height = height;
  F->depth = depth;
  report("fluid_allocGPUKernel", PFG_OK);
  return F;
}

void fluid_initGPUKernel(void *model, void *g, int aFlag, float *hostParameters, void **fIN,
                         void **fOUT, void *ux, void *uy, void *uz, void *rho, void *obstacle,
                         void *Fx, void *Fy, void *Fz, void *vWFbond) {
  (void)model;
  FluidGrid *F = (FluidGrid *)g;
  if (!F) {
    report("fluid_initGPUKernel", PFG_ERR_ARG);
    return;
  }
  const int W = F->width, H = F->height, D = F->depth;
  const int64_t plane = (int64_t)W * H;
  if (aFlag) {
    if (F->h == nullptr) {
      pfg_params p;
      pfg_params_default(&p);
      /* obstacle cells: the channel's two planes, or a general pattern */
      const float *ob = (const float *)obstacle;
      bool lo = true, hi = true, inner = false;
      for (int64_t c = 0; ob && c < plane * D; ++c) {
        const int k = (int)(c / plane);
        const bool solid = ob[c] > 0.5f;
        if (k == 0) lo = lo && solid;
        else if (k == D - 1) hi = hi && solid;
        else inner = inner || solid;
      }
      F->channel = ob != nullptr && lo && hi && !inner && D > 2;
      const int nz = F->channel ? D - 2 : D;
      if (hostParameters) {
        if (!report("pfg_params_from_sem", pfg_params_from_sem(hostParameters, F->dx, F->dt, W,
                                                               H, nz, &p)))
          return;
        F->C = hostParameters[15] > 0.f ? hostParameters[15] : 1.0;
      } else {
        p.nx = W;
        p.ny = H;
        p.nz = nz;
      }
      for (int a = 0; a < 3; ++a) p.wall[a] = PFG_WALL_PERIODIC;
      if (F->channel) {
        /* halfway walls between the obstacle planes and the interior,
         * moving with the reference's linear shear there (u_y at node k is
         * SHEAR_RATE k dx / UMAX, gpuSEM_LB.c:1239; lattice velocity = that
         * over C) */
        const double rate = hostParameters ? hostParameters[5] : 0.0;
        const double umax = hostParameters && hostParameters[11] > 0.f ? hostParameters[11] : 1.0;
        F->u_lo = rate * 0.5 * F->dx / umax;
        F->u_hi = rate * (D - 1.5) * F->dx / umax;
        F->u_plate = rate * (D - 1) * F->dx / umax; /* LB_kernel.cu:630 */
        p.wall[2] = PFG_WALL_BOUNCE;
        p.wall_u_lo[2][1] = (float)(F->u_lo / F->C);
        p.wall_u_hi[2][1] = (float)(F->u_hi / F->C);
        p.wall_z = -1.0f; /* the reference's wall plane z = 0 (node 0) */
      } else if (ob) {
        p.use_flags = 1;
      }
      F->p = p;
      F->n_device = (int64_t)W * H * nz;
      if (!report("pfg_create", pfg_create(&p, 0, &F->h))) return;
      if (p.use_flags) {
        std::vector<uint8_t> fl((size_t)F->n_device);
        for (int64_t c = 0; c < F->n_device; ++c) fl[(size_t)c] = ob[c] > 0.5f ? 1 : 0;
        if (!report("pfg_upload_flags", pfg_upload_flags(F->h, fl.data()))) return;
      }
    }
    /* populations of the device cells (interior planes in the channel) */
    if (fIN) {
      const int off = dz_off(F);
      std::vector<float> f((size_t)19 * F->n_device);
      for (int d = 0; d < 19; ++d) {
        const double *src = (const double *)fIN[d] + (size_t)off * plane;
        float *dst = f.data() + (size_t)d * F->n_device;
        for (int64_t c = 0; c < F->n_device; ++c) dst[c] = (float)src[c];
      }
      report("pfg_upload_f", pfg_upload_f(F->h, f.data()));
    }
    (void)fOUT;
    (void)ux;
    (void)uy;
    (void)uz;
    (void)rho;
    (void)Fx;
    (void)Fy;
    (void)Fz;
    (void)vWFbond;
    return;
  }
  /* device -> host */
  if (!F->h) {
    report("fluid_initGPUKernel", PFG_ERR_STATE);
    return;
  }
  const int off = dz_off(F);
  const int64_t n = F->n_device;
  if (fIN || fOUT) {
    std::vector<float> f((size_t)19 * n);
    if (!report("pfg_download_f", pfg_download_f(F->h, f.data()))) return;
    for (int d = 0; d < 19; ++d)
      for (int which = 0; which < 2; ++which) {
        void **arr = which ? fOUT : fIN;
        if (!arr || !arr[d]) continue;
        double *dst = (double *)arr[d] + (size_t)off * plane;
        const float *src = f.data() + (size_t)d * n;
        for (int64_t c = 0; c < n; ++c) dst[c] = src[c];
      }
  }
  std::vector<float> r((size_t)n), vx((size_t)n), vy((size_t)n), vz((size_t)n);
  if (!report("pfg_download_fields",
              pfg_download_fields(F->h, r.data(), vx.data(), vy.data(), vz.data())))
    return;
  double *outs[4] = {(double *)rho, (double *)ux, (double *)uy, (double *)uz};
  const float *ins[4] = {r.data(), vx.data(), vy.data(), vz.data()};
  for (int q = 0; q < 4; ++q) {
    if (!outs[q]) continue;
    const double scale = q == 0 ? 1.0 : F->C; /* velocities in UMAX */
    for (int64_t c = 0; c < n; ++c) outs[q][(size_t)off * plane + c] = ins[q][c] * scale;
    if (F->channel) { /* the obstacle planes carry the wall state */
      /* the obstacle planes carry the walls: rest below, the moving
       * plate's u_y on top (LB_kernel.cu:622-633) */
      const double wall_lo = q == 0 ? 1.0 : 0.0;
      const double wall_hi = q == 0 ? 1.0 : (q == 2 ? F->u_plate : 0.0);
      for (int64_t c = 0; c < plane; ++c) {
        outs[q][c] = wall_lo;
        outs[q][(size_t)(F->depth - 1) * plane + c] = wall_hi;
      }
    }
  }
  double *fo[3] = {(double *)Fx, (double *)Fy, (double *)Fz};
  for (int q = 0; q < 3; ++q)
    if (fo[q]) std::memset(fo[q], 0, sizeof(double) * (size_t)plane * F->depth);
  if (vWFbond && F->sem && F->sem->loaded && F->model_on_device) {
    /* slot owners from the bond state: the receptor's flat index, -1 free */
    SemGrid *S = F->sem;
    int32_t *vw = (int32_t *)vWFbond;
    std::fill(vw, vw + (size_t)4 * plane, -1);
    const int nrec = S->surf * S->nrec_per_tri;
    std::vector<int32_t> rb((size_t)nrec);
    if (report("pfg_download_bonds", pfg_download_bonds(F->h, rb.data(), nullptr, nullptr)))
      for (int r = 0; r < nrec; ++r)
        if (rb[(size_t)r] >= 0 && rb[(size_t)r] < 4 * plane) vw[rb[(size_t)r]] = r;
  }
}

void fluid_invokeGPUKernel(void *model, void *g, void *g_SEM, double *randNum, void *r,
                           int timeSteps) {
  (void)model;
  (void)randNum;
  (void)r;
  FluidGrid *F = (FluidGrid *)g;
  if (!F || !F->h || timeSteps < 0) {
    report("fluid_invokeGPUKernel", F && F->h ? PFG_ERR_ARG : PFG_ERR_STATE);
    return;
  }
  if (g_SEM && !F->sem) {
    F->sem = (SemGrid *)g_SEM;
    F->sem->fluid = F;
  }
  if (!F->fiber && g_fiber && g_fiber->loaded) {
    F->fiber = g_fiber;
    g_fiber->fluid = F;
  }
  if (!load_model(F)) return;
  pfg_step_info info;
  report("pfg_step", pfg_step(F->h, timeSteps, &info));
}

void fluid_releaseGPUKernel(void *model, void *g) {
  (void)model;
  FluidGrid *F = (FluidGrid *)g;
  if (!F) return;
  if (F->h) pfg_destroy(F->h);
  if (F->sem) F->sem->fluid = nullptr;
  if (F->fiber) F->fiber->fluid = nullptr;
  delete F;
  report("fluid_releaseGPUKernel", PFG_OK);
}

void *sem_allocGPUKernel(void *model, int maxCells, int maxElements, int SurfElem, int newnode,
                         int numReceptorsPerElem, float dt, double S0_all,
                         float *hostParameters) {
  (void)model;
  if (maxCells <= 0 || maxElements <= 0 || SurfElem < 0 || numReceptorsPerElem < 0) {
    report("sem_allocGPUKernel", PFG_ERR_ARG);
    return nullptr;
  }
  SemGrid *S = new (std::nothrow) SemGrid();
  if (!S) {
    report("sem_allocGPUKernel", PFG_ERR_ALLOC);
    return nullptr;
  }
  S->max_cells = maxCells;
  S->max_elems = maxElements;
  S->surf = SurfElem;
  S->newnode = newnode;
  S->nrec_per_tri = numReceptorsPerElem;
  S->dt = dt;
  S->S0_all = S0_all;
  if (hostParameters) S->params.assign(hostParameters, hostParameters + 20);
  report("sem_allocGPUKernel", PFG_OK);
  return S;
}

void sem_initGPUKernel(void *model, void *g, int numOfCells, int *numOfElements, int SurfElem,
                       int numReceptorsPerElem, void *hostX_Ref, void *hostY_Ref,
                       void *hostZ_Ref, void *hostX, void *hostY, void *hostRY, void *hostZ,
                       void *hostVX, void *hostVY, void *hostVZ, void *hostFX, void *hostFY,
                       void *hostFZ, void *hostType, void *hostBonds, void *triElem,
                       void *receptor_r1, void *receptor_r2, void *node_share_Elem, void *N,
                       void *node_nbrElemNum, void *node_nbr_nodes, void *S0, double V0,
                       void *receptBond, void *randNum) {
  (void)model;
  (void)hostX_Ref;
  (void)hostY_Ref;
  (void)hostZ_Ref;
  (void)hostRY;
  (void)hostVX;
  (void)hostVY;
  (void)hostVZ;
  (void)hostFX;
  (void)hostFY;
  (void)hostFZ;
  (void)hostType;
  (void)node_share_Elem;
  (void)N;
  (void)node_nbrElemNum;
  (void)node_nbr_nodes;
  (void)randNum;
  SemGrid *S = (SemGrid *)g;
  if (!S || numOfCells < 0 || numOfCells > S->max_cells || !numOfElements || !hostX ||
      !hostY || !hostZ || SurfElem != S->surf || numReceptorsPerElem != S->nrec_per_tri) {
    report("sem_initGPUKernel", PFG_ERR_ARG);
    return;
  }
  const size_t ne = (size_t)S->max_elems * S->max_cells;
  S->n_cells = numOfCells;
  S->n_elem.assign(numOfElements, numOfElements + S->max_cells);
  for (int c = 0; c < numOfCells; ++c)
    if (S->n_elem[(size_t)c] < 0 || S->n_elem[(size_t)c] > S->max_elems) {
      report("sem_initGPUKernel", PFG_ERR_ARG);
      return;
    }
  S->X.assign((const double *)hostX, (const double *)hostX + ne);
  S->Y.assign((const double *)hostY, (const double *)hostY + ne);
  S->Z.assign((const double *)hostZ, (const double *)hostZ + ne);
  const size_t nb = (size_t)S->max_elems * S->max_elems;
  if (hostBonds) S->bonds.assign((const double *)hostBonds, (const double *)hostBonds + nb);
  else S->bonds.assign(nb, 0.0);
  const size_t T = (size_t)S->surf, R = (size_t)S->nrec_per_tri;
  if (T > 0) {
    if (!triElem || !S0) {
      report("sem_initGPUKernel", PFG_ERR_ARG);
      return;
    }
    S->tri6.assign((const int *)triElem, (const int *)triElem + 6 * T);
    for (size_t t = 0; t < T; ++t)
      for (int v = 0; v < 3; ++v) {
        const int e = S->tri6[t * 6 + v];
        if (e < 0 || e >= S->max_elems) {
          report("sem_initGPUKernel", PFG_ERR_ARG);
          return;
        }
      }
    S->S0.assign((const double *)S0, (const double *)S0 + T);
  }
  if (T * R > 0) {
    if (!receptor_r1 || !receptor_r2) {
      report("sem_initGPUKernel", PFG_ERR_ARG);
      return;
    }
    S->r1.assign((const float *)receptor_r1, (const float *)receptor_r1 + T * R);
    S->r2.assign((const float *)receptor_r2, (const float *)receptor_r2 + T * R);
    if (receptBond) S->rbond.assign((const int *)receptBond, (const int *)receptBond + 3 * T * R);
    else S->rbond.assign(3 * T * R, -1);
  }
  S->V0 = V0;
  S->loaded = true;
  report("sem_initGPUKernel", PFG_OK);
}

void sem_invokeGPUKernel_Force(void *model, void *g, int timeSteps, int *done, int *totalT,
                               void *r, double gama) {
  (void)model;
  (void)timeSteps;
  (void)r;
  (void)gama;
  SemGrid *S = (SemGrid *)g;
  if (!S || !done || !totalT) {
    report("sem_invokeGPUKernel_Force", PFG_ERR_ARG);
    return;
  }
  FluidGrid *F = S->fluid;
  if (!F || !F->h || !F->model_on_device) {
    /* before the first fluid step: nothing has run yet; the forces of the
     * step are computed inside the next fluid_invokeGPUKernel */
    (*totalT)++;
    report("sem_invokeGPUKernel_Force", PFG_OK);
    return;
  }
  pfg_step_info info;
  if (!report("pfg_step", pfg_step(F->h, 0, &info))) return;
  if (info.detached) {
    *done = 1; /* semLB.cu:451-455 */
    return;
  }
  (*totalT)++;
}

void sem_copyGPUKernel(void *model, void *g, void *hostX, void *hostY, void *hostRY,
                       void *hostZ, void *hostVX, void *hostVY, void *hostVZ, void *hostFX,
                       void *hostFY, void *hostFZ, int timeSteps) {
  (void)model;
  (void)timeSteps;
  SemGrid *S = (SemGrid *)g;
  FluidGrid *F = S ? S->fluid : nullptr;
  if (!S || !F || !F->h || !F->model_on_device) {
    report("sem_copyGPUKernel", PFG_ERR_STATE);
    return;
  }
  int n = 0;
  for (int c = 0; c < S->n_cells; ++c) n += S->n_elem[(size_t)c];
  const FiberGrid *B = F->fiber;
  const int total = n + (B && B->loaded ? B->max_nodes : 0);
  std::vector<float> a[9];
  float *p[9];
  for (int k = 0; k < 9; ++k) {
    a[k].resize((size_t)total);
    p[k] = a[k].data();
  }
  if (!report("pfg_download_markers", pfg_download_markers(F->h, p[0], p[1], p[2], p[3], p[4],
                                                           p[5], p[6], p[7], p[8])))
    return;
  const double dx = F->dx, C = F->C, zoff = dz_off(F);
  const double Ly = F->height * dx;
  double *out[10] = {(double *)hostX, (double *)hostY, (double *)hostRY, (double *)hostZ,
                     (double *)hostVX, (double *)hostVY, (double *)hostVZ, (double *)hostFX,
                     (double *)hostFY, (double *)hostFZ};
  for (int c = 0; c < S->n_cells; ++c)
    for (int m = 0; m < S->n_elem[(size_t)c]; ++m) {
      const size_t k = (size_t)m * S->max_cells + c;
      const int i = S->marker_of[k];
      if (i < 0) continue;
      const double ry = a[1][(size_t)i] * dx;
      double y = std::fmod(ry, Ly);
      if (y < 0) y += Ly;
      const double v[10] = {a[0][(size_t)i] * dx, y, ry, (a[2][(size_t)i] + zoff) * dx,
                            a[3][(size_t)i] * C, a[4][(size_t)i] * C, a[5][(size_t)i] * C,
                            a[6][(size_t)i], a[7][(size_t)i], a[8][(size_t)i]};
      for (int q = 0; q < 10; ++q)
        if (out[q]) out[q][k] = v[q];
    }
  report("sem_copyGPUKernel", PFG_OK);
}

void sem_releaseGPUKernel(void *model, void *g) {
  (void)model;
  SemGrid *S = (SemGrid *)g;
  if (!S) return;
  if (S->fluid) S->fluid->sem = nullptr;
  delete S;
  report("sem_releaseGPUKernel", PFG_OK);
}

void *fiber_allocGPUKernel(void *model, int maxNodes, int maxLinks, int max_N_conn_at_Node,
                           double dt, double *hostParameters) {
  (void)model;
  (void)hostParameters;
  if (maxNodes < 0 || maxLinks < 0 || max_N_conn_at_Node < 0) {
    report("fiber_allocGPUKernel", PFG_ERR_ARG);
    return nullptr;
  }
  FiberGrid *B = new (std::nothrow) FiberGrid();
  if (!B) {
    report("fiber_allocGPUKernel", PFG_ERR_ALLOC);
    return nullptr;
  }
  B->max_nodes = maxNodes;
  B->max_links = maxLinks;
  B->max_conn = max_N_conn_at_Node;
  B->dt = dt;
  g_fiber = B;
  report("fiber_allocGPUKernel", PFG_OK);
  return B;
}

void fiber_init_GPUKernel(void *model, void *fg, void *NodeType, void *N_Conn_at_Node,
                          void *Link_at_Node, void *lAdjVer, void *linkLengths,
                          void *linkLengths0, void *linkThick, void *X, void *V_X, void *X0,
                          void *F_X, void *Y, void *Y0, void *V_Y, void *F_Y, void *Z, void *Z0,
                          void *V_Z, void *F_Z) {
  (void)model;
  (void)N_Conn_at_Node;
  (void)Link_at_Node;
  (void)linkLengths;
  (void)linkThick;
  (void)V_X;
  (void)X0;
  (void)F_X;
  (void)Y0;
  (void)V_Y;
  (void)F_Y;
  (void)Z0;
  (void)V_Z;
  (void)F_Z;
  FiberGrid *B = (FiberGrid *)fg;
  if (!B || !NodeType || !lAdjVer || !linkLengths0 || !X || !Y || !Z) {
    report("fiber_init_GPUKernel", PFG_ERR_ARG);
    return;
  }
  const size_t n = (size_t)B->max_nodes, L = (size_t)B->max_links;
  B->type.assign((const int *)NodeType, (const int *)NodeType + n);
  B->link.assign((const int *)lAdjVer, (const int *)lAdjVer + 2 * L);
  for (size_t l = 0; l < 2 * L; ++l)
    if (B->link[l] < 0 || B->link[l] >= B->max_nodes) {
      report("fiber_init_GPUKernel", PFG_ERR_ARG);
      return;
    }
  B->l0.assign((const double *)linkLengths0, (const double *)linkLengths0 + L);
  B->x.assign((const double *)X, (const double *)X + n);
  B->y.assign((const double *)Y, (const double *)Y + n);
  B->z.assign((const double *)Z, (const double *)Z + n);
  B->loaded = true;
  report("fiber_init_GPUKernel", PFG_OK);
}

void fiber_copy_GPUKernel(void *model, void *fg, void *X, void *V_X, void *X0, void *F_X,
                          void *Y, void *Y0, void *V_Y, void *F_Y, void *Z, void *Z0, void *V_Z,
                          void *F_Z) {
  (void)model;
  (void)X0;
  (void)Y0;
  (void)Z0;
  FiberGrid *B = (FiberGrid *)fg;
  FluidGrid *F = B ? B->fluid : nullptr; /* the fluid grid that took it */
  if (!B || !B->loaded) {
    report("fiber_copy_GPUKernel", PFG_ERR_STATE);
    return;
  }
  const size_t n = (size_t)B->max_nodes;
  double *out[9] = {(double *)X, (double *)Y, (double *)Z, (double *)V_X, (double *)V_Y,
                    (double *)V_Z, (double *)F_X, (double *)F_Y, (double *)F_Z};
  if (!F || !F->h || !F->model_on_device) {
    /* nothing stepped yet: the network as loaded */
    for (size_t i = 0; i < n; ++i) {
      if (out[0]) out[0][i] = B->x[i];
      if (out[1]) out[1][i] = B->y[i];
      if (out[2]) out[2][i] = B->z[i];
      for (int q = 3; q < 9; ++q)
        if (out[q]) out[q][i] = 0.0;
    }
    report("fiber_copy_GPUKernel", PFG_OK);
    return;
  }
  int total = 0;
  if (F->sem && F->sem->loaded)
    for (int c = 0; c < F->sem->n_cells; ++c) total += F->sem->n_elem[(size_t)c];
  total += (int)n;
  std::vector<float> a[9];
  float *p[9];
  for (int k = 0; k < 9; ++k) {
    a[k].resize((size_t)total);
    p[k] = a[k].data();
  }
  if (!report("pfg_download_markers", pfg_download_markers(F->h, p[0], p[1], p[2], p[3], p[4],
                                                           p[5], p[6], p[7], p[8])))
    return;
  const double dx = F->dx, C = F->C, zoff = dz_off(F);
  for (size_t i = 0; i < n; ++i) {
    const size_t j = (size_t)B->marker0 + i;
    const double v[9] = {a[0][j] * dx, a[1][j] * dx, (a[2][j] + zoff) * dx,
                         a[3][j] * C,  a[4][j] * C,  a[5][j] * C,
                         a[6][j],      a[7][j],      a[8][j]};
    for (int q = 0; q < 9; ++q)
      if (out[q]) out[q][i] = v[q];
  }
  report("fiber_copy_GPUKernel", PFG_OK);
}

void fiber_release_GPUKernel(void *model, void *fg) {
  (void)model;
  FiberGrid *B = (FiberGrid *)fg;
  if (!B) return;
  if (B->fluid) B->fluid->fiber = nullptr;
  if (g_fiber == B) g_fiber = nullptr;
  delete B;
  report("fiber_release_GPUKernel", PFG_OK);
}

}  // extern "C"
