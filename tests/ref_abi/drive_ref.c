/*
 * drive_ref.c -- drives libpfg.so through the reference's own entry points
 * in the reference host's call order (gpuSEM_LB.c:1293-1304 allocation and
 * initialisation, 1433-1454 the step loop, 1453-1454 the copies back), on a
 * small sheared channel with one platelet, and checks the result against
 * the same run through the pfg_* ABI with the mapping pfg_ref.h documents.
 *
 * Setup as gpuSEM_LB.c:1168-1258 builds it: obstacle planes z = 0 and
 * z = depth-1, linear shear u_y = SHEAR_RATE k dx / UMAX at equilibrium,
 * the derived parameter block hostSEMparameters[0..19]; the platelet is an
 * icosahedron of 12 elements with springs on its edges and 3 receptors per
 * triangle; a fibrin chain of six nodes (the first anchored) goes in through
 * fiber_allocGPUKernel / fiber_init_GPUKernel and comes back through
 * fiber_copy_GPUKernel.
 *
 * Prints one line "REF_ABI ok ..." and exits 0, or "REF_ABI FAIL ..." and 1.
 * Exit 77: no HIP device (the program then only shows that it links).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pfg.h"
#include "pfg_ref.h"

#define LX 40
#define LY 24
#define LZ 20
#define NE 12
#define NT 20
#define NREC 3
#define STEPS 24
#define NF 6 /* fibrin chain: nodes (the first anchored) and NF - 1 links */

static const double DX = 0.2, DT = 1e-8, RATE = 1000.0;

static void icosahedron(double c[3], double rad, double X[NE][3], int T[NT][3]) {
  const double p = (1.0 + sqrt(5.0)) / 2.0;
  const double v[NE][3] = {{-1, p, 0}, {1, p, 0},  {-1, -p, 0}, {1, -p, 0},
                           {0, -1, p}, {0, 1, p},  {0, -1, -p}, {0, 1, -p},
                           {p, 0, -1}, {p, 0, 1},  {-p, 0, -1}, {-p, 0, 1}};
  const int f[NT][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                        {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                        {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                        {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
  const double n = sqrt(1.0 + p * p);
  for (int i = 0; i < NE; ++i)
    for (int a = 0; a < 3; ++a) X[i][a] = c[a] + rad * v[i][a] / n;
  memcpy(T, f, sizeof(f));
}

static double lcg(unsigned *s) {
  *s = *s * 1664525u + 1013904223u;
  return (*s >> 8) / 16777216.0;
}

int main(void) {
  int ndev = 0;
  if (pfg_device_count(&ndev) != PFG_OK || ndev < 1) {
    printf("REF_ABI skip: no HIP device\n");
    return 77;
  }
  /* ---- host setup (gpuSEM_LB.c:1168-1258) ---- */
  float s[20];
  memset(s, 0, sizeof(s));
  const double umax = RATE * (LZ - 1) * DX, C = DX / DT / umax;
  s[1] = 1101.0f;
  s[2] = 0.05f;                                   /* k_spring, nN/um  */
  s[3] = (float)(0.0165 * exp(0.00065 * RATE));  /* Kf0              */
  s[4] = 3.21f;                                   /* Kr0              */
  s[5] = (float)RATE;
  s[6] = 1.0f;                                    /* rho0             */
  s[7] = 1.0f;                                    /* L0               */
  s[8] = (float)(1.0 / umax);                     /* T0 = L0 / UMAX   */
  s[10] = 1.0f;                                   /* F0               */
  s[11] = (float)umax;
  s[12] = (float)(1.2e6 / s[6]);
  s[13] = (float)(3.0 * s[12] * DT / (DX * DX) + 0.5);
  s[15] = (float)C;
  s[16] = 1e12f;
  const size_t plane = (size_t)LX * LY, N = plane * LZ;
  double *fIN[19], *fOUT[19];
  const double w[19] = {1. / 3,  1. / 18, 1. / 18, 1. / 18, 1. / 18, 1. / 18, 1. / 18,
                        1. / 36, 1. / 36, 1. / 36, 1. / 36, 1. / 36, 1. / 36, 1. / 36,
                        1. / 36, 1. / 36, 1. / 36, 1. / 36, 1. / 36};
  const int cy[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1};
  for (int d = 0; d < 19; ++d) {
    fIN[d] = malloc(sizeof(double) * N);
    fOUT[d] = calloc(N, sizeof(double));
    for (size_t c = 0; c < N; ++c) {
      const int k = (int)(c / plane);
      const double uy = RATE * k * DX / umax; /* gpuSEM_LB.c:1239 */
      const double cu = 3.0 * cy[d] * uy;
      fIN[d][c] = w[d] * (1.0 + cu / C + 0.5 * cu * cu / (C * C) - 1.5 * uy * uy / (C * C));
    }
  }
  double *ux = calloc(N, 8), *uy = calloc(N, 8), *uz = calloc(N, 8), *rho = calloc(N, 8);
  double *Fx = calloc(N, 8), *Fy = calloc(N, 8), *Fz = calloc(N, 8);
  float *obst = calloc(N, 4);
  int *vwf = calloc(4 * plane, sizeof(int));
  for (size_t c = 0; c < plane; ++c) obst[c] = obst[(LZ - 1) * plane + c] = 1.0f;
  for (size_t c = 0; c < 4 * plane; ++c) vwf[c] = -1;
  /* platelet: one cell, NE elements ([maxElements][maxCells] = [NE][1]) */
  double X[NE][3], cen[3] = {3.1, 2.5, 1.0};
  int T[NT][3];
  icosahedron(cen, 0.55, X, T);
  double hx[NE], hy[NE], hz[NE], zero[NE], bonds[NE * NE];
  int etype[NE], tri6[NT * 6], numE[1] = {NE};
  memset(zero, 0, sizeof(zero));
  memset(bonds, 0, sizeof(bonds));
  memset(etype, 0, sizeof(etype));
  for (int i = 0; i < NE; ++i) {
    hx[i] = X[i][0];
    hy[i] = X[i][1];
    hz[i] = X[i][2];
  }
  double S0[NT], S0_all = 0.0, V0 = 0.0;
  for (int t = 0; t < NT; ++t) {
    const int *v = T[t];
    double e1[3], e2[3], n[3], c3[3];
    for (int a = 0; a < 3; ++a) {
      e1[a] = X[v[1]][a] - X[v[0]][a];
      e2[a] = X[v[2]][a] - X[v[0]][a];
      c3[a] = (X[v[0]][a] + X[v[1]][a] + X[v[2]][a]) / 3.0;
    }
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    const double ln = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    S0[t] = 0.5 * ln;
    S0_all += S0[t];
    /* reference sign: n_elem = -(e01 x e02)/|..| (sem_kernel.cu:793-808) */
    V0 += -(c3[0] * n[0] + c3[1] * n[1] + c3[2] * n[2]) / ln * S0[t] / 3.0;
    for (int q = 0; q < 3; ++q) {
      tri6[t * 6 + q] = v[q];
      tri6[t * 6 + 3 + q] = -1;
      const int a = v[q], b = v[(q + 1) % 3];
      double L = 0;
      for (int k = 0; k < 3; ++k) L += (X[a][k] - X[b][k]) * (X[a][k] - X[b][k]);
      bonds[a * NE + b] = bonds[b * NE + a] = sqrt(L);
    }
  }
  float r1[NT * NREC], r2[NT * NREC];
  unsigned seed = 7u;
  for (int k = 0; k < NT * NREC; ++k) {
    r1[k] = (float)sqrt(lcg(&seed));
    r2[k] = (float)lcg(&seed);
  }
  int rbond[3 * NREC * NT];
  for (int k = 0; k < 3 * NREC * NT; ++k) rbond[k] = -1;
  double randNum[2 * NREC * NT];
  for (int k = 0; k < 2 * NREC * NT; ++k) randNum[k] = -1.0;

  /* a fibrin chain along x above the platelet's side, node 0 anchored, links
   * stretched by 10 % (the network arrays of fiber.cpp:145-217) */
  int ntype[NF], nconn[NF], natl[2 * NF], ladj[2 * (NF - 1)];
  double nx0[NF], ny0[NF], nz0[NF], nzero[NF], ll0[NF - 1], lth[NF - 1];
  for (int i = 0; i < NF; ++i) {
    ntype[i] = i == 0 ? 1 : 0;
    nconn[i] = (i == 0 || i == NF - 1) ? 1 : 2;
    natl[2 * i] = i > 0 ? i - 1 : -1;
    natl[2 * i + 1] = i < NF - 1 ? i : -1;
    nx0[i] = 1.0 + 0.55 * i;
    ny0[i] = 3.9;
    nz0[i] = 1.7;
    nzero[i] = 0.0;
  }
  for (int l = 0; l < NF - 1; ++l) {
    ladj[2 * l] = l;
    ladj[2 * l + 1] = l + 1;
    ll0[l] = 0.5;
    lth[l] = 0.05;
  }
  /* ---- the reference's calls (gpuSEM_LB.c:1279-1304; the fibrin network's
   * copies to the device as semLB.cu:609-664 takes them) ---- */
  void *FIBgrid = fiber_allocGPUKernel(NULL, NF, NF - 1, 2, DT, NULL);
  fiber_init_GPUKernel(NULL, FIBgrid, ntype, nconn, natl, ladj, ll0, ll0, lth, nx0, nzero, nx0,
                       nzero, ny0, ny0, nzero, nzero, nz0, nz0, nzero, nzero);
  void *LBgrids = fluid_allocGPUKernel(NULL, (float)DT, (float)DX, LX, LY, LZ);
  fluid_initGPUKernel(NULL, LBgrids, 1, s, (void **)fIN, (void **)fOUT, ux, uy, uz, rho, obst,
                      Fx, Fy, Fz, vwf);
  void *SEMgrids = sem_allocGPUKernel(NULL, 1, NE, NT, 0, NREC, (float)DT, S0_all, s);
  sem_initGPUKernel(NULL, SEMgrids, 1, numE, NT, NREC, hx, hy, hz, hx, hy, hy, hz, zero, zero,
                    zero, zero, zero, zero, etype, bonds, tri6, r1, r2, NULL, NULL, NULL, NULL,
                    S0, V0, rbond, randNum);
  if (pfg_ref_last_status() != PFG_OK) {
    printf("REF_ABI FAIL init: %s\n", pfg_status_string(pfg_ref_last_status()));
    return 1;
  }
  /* the step loop (gpuSEM_LB.c:1433-1454) */
  int done = 0, totalT = 0;
  const double gama = RATE * DT;
  for (int t = 0; t < STEPS && !done; ++t) {
    sem_invokeGPUKernel_Force(NULL, SEMgrids, 1, &done, &totalT, NULL, gama);
    if (!done) fluid_invokeGPUKernel(NULL, LBgrids, SEMgrids, randNum, NULL, 1);
  }
  fluid_initGPUKernel(NULL, LBgrids, 0, s, (void **)fIN, (void **)fOUT, ux, uy, uz, rho, obst,
                      Fx, Fy, Fz, vwf);
  double ox[NE], oy[NE], ory[NE], oz[NE], ovx[NE], ovy[NE], ovz[NE], ofx[NE], ofy[NE],
      ofz[NE];
  sem_copyGPUKernel(NULL, SEMgrids, ox, oy, ory, oz, ovx, ovy, ovz, ofx, ofy, ofz, 1);
  double fbx[NF], fby[NF], fbz[NF], fbvx[NF], fbvy[NF], fbvz[NF], fbfx[NF], fbfy[NF], fbfz[NF];
  fiber_copy_GPUKernel(NULL, FIBgrid, fbx, fbvx, NULL, fbfx, fby, NULL, fbvy, fbfy, fbz, NULL,
                       fbvz, fbfz);
  if (pfg_ref_last_status() != PFG_OK || done) {
    printf("REF_ABI FAIL run: %s done %d\n", pfg_status_string(pfg_ref_last_status()), done);
    return 1;
  }

  /* ---- the same run through pfg_*, mapped as pfg_ref.h documents ---- */
  pfg_params p;
  pfg_params_default(&p);
  const int nz = LZ - 2;
  if (pfg_params_from_sem(s, (float)DX, (float)DT, LX, LY, nz, &p) != PFG_OK) return 1;
  p.wall[0] = p.wall[1] = PFG_WALL_PERIODIC;
  p.wall[2] = PFG_WALL_BOUNCE;
  p.wall_u_lo[2][1] = (float)(RATE * 0.5 * DX / umax / C);
  p.wall_u_hi[2][1] = (float)(RATE * (LZ - 1.5) * DX / umax / C);
  p.wall_z = -1.0f;
  pfg_handle *h = NULL;
  if (pfg_create(&p, 0, &h) != PFG_OK) return 1;
  const size_t n = plane * nz;
  float *f = malloc(sizeof(float) * 19 * n);
  for (int d = 0; d < 19; ++d) {
    /* the fresh host planes (the adapter's download overwrote fIN) */
    for (size_t c = 0; c < n; ++c) {
      const int k = (int)(c / plane) + 1;
      const double uyk = RATE * k * DX / umax, cu = 3.0 * cy[d] * uyk;
      f[d * n + c] = (float)(w[d] * (1.0 + cu / C + 0.5 * cu * cu / (C * C) -
                                     1.5 * uyk * uyk / (C * C)));
    }
  }
  if (pfg_upload_f(h, f) != PFG_OK) return 1;
  float mx[NE + NF], my[NE + NF], mz[NE + NF];
  int32_t kind[NE + NF], body[NE + NF], pes[2] = {0, NE}, pts[2] = {0, NT}, tri[3 * NT],
      edge[2 * 30], link[2 * (NF - 1)];
  float link_l0[NF - 1];
  for (int i = 0; i < NF; ++i) {
    mx[NE + i] = (float)(nx0[i] / DX);
    my[NE + i] = (float)(ny0[i] / DX);
    mz[NE + i] = (float)(nz0[i] / DX - 1.0);
    kind[NE + i] = ntype[i] == 1 ? PFG_MARKER_FIXED : PFG_MARKER_FIBER;
    body[NE + i] = -1;
  }
  for (int l = 0; l < NF - 1; ++l) {
    link[2 * l] = NE + ladj[2 * l];
    link[2 * l + 1] = NE + ladj[2 * l + 1];
    link_l0[l] = (float)(ll0[l] / DX);
  }
  float el0[30];
  double tS0[NT], pS0 = S0_all / (DX * DX), pV0 = V0 / (DX * DX * DX);
  int ne = 0;
  for (int i = 0; i < NE; ++i) {
    mx[i] = (float)(X[i][0] / DX);
    my[i] = (float)(X[i][1] / DX);
    mz[i] = (float)(X[i][2] / DX - 1.0);
    kind[i] = PFG_MARKER_ELEMENT;
    body[i] = 0;
    for (int j = i + 1; j < NE; ++j)
      if (bonds[i * NE + j] > 0) {
        edge[2 * ne] = i;
        edge[2 * ne + 1] = j;
        el0[ne++] = (float)(bonds[i * NE + j] / DX);
      }
  }
  int32_t rt[NT * NREC];
  for (int t = 0; t < NT; ++t) {
    for (int q = 0; q < 3; ++q) tri[3 * t + q] = T[t][q];
    tS0[t] = S0[t] / (DX * DX);
    for (int r = 0; r < NREC; ++r) rt[t * NREC + r] = t;
  }
  float *lx = malloc(sizeof(float) * 4 * plane), *ly = malloc(sizeof(float) * 4 * plane),
        *lz = malloc(sizeof(float) * 4 * plane);
  for (size_t k = 0; k < 4 * plane; ++k) {
    lx[k] = (float)(k % LX);
    ly[k] = (float)((k / LX) % LY);
    lz[k] = -1.0f;
  }
  pfg_model m;
  memset(&m, 0, sizeof(m));
  m.n_markers = NE + NF;
  m.x = mx;
  m.y = my;
  m.z = mz;
  m.kind = kind;
  m.body = body;
  m.n_platelets = 1;
  m.plt_elem_start = pes;
  m.plt_tri_start = pts;
  m.n_tri = NT;
  m.tri = tri;
  m.tri_S0 = tS0;
  m.plt_S0 = &pS0;
  m.plt_V0 = &pV0;
  m.n_edges = ne;
  m.edge = edge;
  m.edge_l0 = el0;
  m.n_receptors = NT * NREC;
  m.rec_tri = rt;
  m.rec_r1 = r1;
  m.rec_r2 = r2;
  m.n_wall_ligands = (int32_t)(4 * plane);
  m.lig_x = lx;
  m.lig_y = ly;
  m.lig_z = lz;
  m.n_links = NF - 1;
  m.link = link;
  m.link_l0 = link_l0;
  if (pfg_load_model(h, &m) != PFG_OK) return 1;
  pfg_step_info info;
  if (pfg_step(h, totalT, &info) != PFG_OK) return 1;
  float dxp[NE + NF], dyp[NE + NF], dzp[NE + NF], dvx[NE + NF], dvy[NE + NF], dvz[NE + NF];
  if (pfg_download_markers(h, dxp, dyp, dzp, dvx, dvy, dvz, NULL, NULL, NULL) != PFG_OK)
    return 1;
  float *fuy = malloc(sizeof(float) * n);
  if (pfg_download_fields(h, NULL, NULL, fuy, NULL) != PFG_OK) return 1;

  /* ---- checks ---- */
  double dpos = 0.0, dvel = 0.0, moved = 0.0;
  for (int i = 0; i < NE; ++i) {
    const double a[3] = {ox[i], ory[i], oz[i]};
    const double b[3] = {dxp[i] * DX, dyp[i] * DX, (dzp[i] + 1.0) * DX};
    for (int k = 0; k < 3; ++k) {
      dpos = fmax(dpos, fabs(a[k] - b[k]));
      moved = fmax(moved, fabs(a[k] - X[i][k]));
    }
    /* Y wraps at height dx, RY does not (sem_kernel.cu:2120-2124) */
    double yw = fmod(ory[i], LY * DX);
    if (yw < 0) yw += LY * DX;
    if (fabs(yw - oy[i]) > 1e-12) dpos = 1.0;
    const double va[3] = {ovx[i], ovy[i], ovz[i]}, vb[3] = {dvx[i] * C, dvy[i] * C, dvz[i] * C};
    for (int k = 0; k < 3; ++k) dvel = fmax(dvel, fabs(va[k] - vb[k]));
  }
  /* the fibrin chain: positions and velocities through fiber_copy_GPUKernel
   * against the direct run; the anchored node stays, the others move */
  double dfib = 0.0, fmoved = 0.0;
  for (int i = 0; i < NF; ++i) {
    const double a[3] = {fbx[i], fby[i], fbz[i]};
    const double b[3] = {dxp[NE + i] * DX, dyp[NE + i] * DX, (dzp[NE + i] + 1.0) * DX};
    const double x0[3] = {nx0[i], ny0[i], nz0[i]};
    const double va[3] = {fbvx[i], fbvy[i], fbvz[i]};
    const double vb[3] = {dvx[NE + i] * C, dvy[NE + i] * C, dvz[NE + i] * C};
    for (int k = 0; k < 3; ++k) {
      dfib = fmax(dfib, fabs(a[k] - b[k]));
      dfib = fmax(dfib, fabs(va[k] - vb[k]));
      if (i == 0) dfib = fmax(dfib, fabs(a[k] - x0[k]) > 1e-6 ? 1.0 : 0.0);
      else fmoved = fmax(fmoved, fabs(a[k] - x0[k]));
    }
  }
  /* the downloaded flow against the direct run, and the linear shear away
   * from the platelet (x within 4 nodes of the far end) */
  double dflow = 0.0, dlin = 0.0;
  for (size_t c = 0; c < n; ++c) {
    const int k = (int)(c / plane) + 1;
    dflow = fmax(dflow, fabs(uy[k * plane + c % plane] - fuy[c] * C));
    const int i = (int)(c % LX);
    if (i >= LX - 4) dlin = fmax(dlin, fabs(uy[k * plane + c % plane] - RATE * k * DX / umax));
  }
  /* the top obstacle plane holds the plate velocity SHEAR_RATE (depth-1) dx / UMAX
   * with the FP32 dx of fluid_allocGPUKernel (LB_kernel.cu:647) */
  const double u_plate = RATE * (LZ - 1) * (double)(float)DX / umax;
  const int ok = totalT == STEPS && dpos < 1e-6 && dvel < 1e-6 && dflow < 1e-6 && dlin < 1e-3 &&
                 dfib < 1e-6 && fmoved > 1e-5 &&
                 moved > 1e-4 && fabs(uy[(LZ - 1) * plane] - u_plate) < 1e-12 && uy[0] == 0.0;
  printf("REF_ABI %s steps %d moved %.3e um, |dpos| %.2e um, |dvel| %.2e, |dflow| %.2e, "
         "linear shear %.2e (UMAX units), fibrin moved %.3e um, |dfibrin| %.2e\n",
         ok ? "ok" : "FAIL", totalT, moved, dpos, dvel, dflow, dlin, fmoved, dfib);
  pfg_destroy(h);
  sem_releaseGPUKernel(NULL, SEMgrids);
  fiber_release_GPUKernel(NULL, FIBgrid);
  fluid_releaseGPUKernel(NULL, LBgrids);
  return ok ? 0 : 1;
}
