/*
 * pfg_driver.cpp -- host driver of the coupled platelet step over the C ABI.
 *
 * Keeps the reference's run structure:
 *   main            gpuSEM_LB.c:1582-1615  (-d device, a run per seed)
 *   ParReader       gpuSEM_LB.c:148-344    (keyword file, fixed order)
 *   runGPUSimulation gpuSEM_LB.c:649-1540  (parameter block, step loop that
 *                   ends at detach, output frames, pause time)
 * and its output files (gpuSEM_LB.c:1330-1336, 1384-1405, 1464-1494, 1533):
 *   Output/cellsPos.data       binary FP64 x, y(unwrapped), z, z per element,
 *                              initial frame then one frame per OUT_FREQ steps
 *   Output/cellsVel.data       binary FP64 vx, vy, vz, vz, initial frame
 *   Output/cellsForce.data     binary FP64 fx, fy, fz, fz, initial frame
 *   Output/cells{Pos,Vel,Force}ASCII.dat  the same as text, every frame
 *   Output/cellsPauseTime.dat  appended "pause time, seed" at detach
 *
 * The platelets are the NEU mesh named by --mesh, else by the file's
 * NEU_PATH when that file exists (as written, next to the parameter file or
 * in its parent directory), read as NEUReader reads it (gpuSEM_LB.c:346-424)
 * and scaled by 1 / DELTA_X (mesh units to lattice units; --mesh-scale
 * overrides); without a mesh file they come from the seeded generator.
 * Positions, velocities and forces are written in lattice units; the pause
 * time is steps x DELTA_T.
 *
 * Physics: the file's parameter block (with its derived entries) sets the
 * run's physics through pfg_params_from_sem -- tau, the IB force scale, the
 * marker step, the spring, area, volume and bending stiffnesses, the bond
 * rates and the reference's bond and wall-repulsion constants -- and the
 * shear: the top wall moves at SHEAR_RATE across the lattice depth, the flow
 * starting from that linear profile.  The platelets carry no Morse term, as
 * in the reference (NO_NB, sem_kernel.cu:78).  --scene-physics keeps the
 * scene's own lattice-unit physics instead (with --rates-from-file taking
 * only the rates from the file).
 *
 * Usage: pfg_driver [-d device] [-p parameter file] [-o output dir]
 *                   [--scene shear|channel|thrombus|large] [--lattice nx ny nz]
 *                   [--store f32|f16|bf16] [--seeds first last] [--max-steps n]
 *                   [--mesh file.neu] [--mesh-scale s] [--no-mesh]
 *                   [--fibrin network.txt] [--fibrin-scale s]
 *                   [--rates-from-file] [--scene-physics] [--scene-counts]
 *                   [--parse-only]
 * --fibrin places the network of a fibrin text file (the format
 * init_fibrin_network reads, fiber.cpp:145-217) in the scene's fibrin box.
 * Exit status: 0 done, 1 parameter or usage error, 2 I/O error, 11 device
 * error (the reference's exit code for an invalid device), 3 other failure.
 */
#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pfg.h"
#include "pfg_geometry.h"

namespace {

/* ---------------- parameter file ---------------- */

struct ParFile {
  std::string pdb, psf, neu;
  float dx = 0, dt = 0, sim_time = 0;
  int out_freq = 0, max_cell = 0, max_elem = 0, surf_elem = 0;
  float host[100] = {0};  /* hostSEMparameters */
};

/* Reads the keywords in the reference's fixed order; a keyword out of place
 * is an error, reported as the reference reports it. */
bool read_par_file(const char *path, ParFile *pf, std::string *err) {
  FILE *in = std::fopen(path, "r");
  if (!in) {
    *err = std::string("Parameter File was not open: ") + path;
    return false;
  }
  char key[64], sval[256];
  auto expect = [&](const char *want) -> bool {
    if (std::fscanf(in, "%63s", key) != 1 || std::strcmp(key, want) != 0) {
      *err = std::string("Parameter Read Failed for ") + want;
      return false;
    }
    return true;
  };
  auto read_str = [&](const char *want, std::string *out) -> bool {
    if (!expect(want) || std::fscanf(in, "%255s", sval) != 1) {
      *err = std::string("Parameter Read Failed for ") + want;
      return false;
    }
    *out = sval;
    return true;
  };
  auto read_f = [&](const char *want, float *out) -> bool {
    if (!expect(want) || std::fscanf(in, "%f", out) != 1) {
      *err = std::string("Parameter Read Failed for ") + want;
      return false;
    }
    return true;
  };
  auto read_i = [&](const char *want, int *out) -> bool {
    if (!expect(want) || std::fscanf(in, "%d", out) != 1) {
      *err = std::string("Parameter Read Failed for ") + want;
      return false;
    }
    return true;
  };
  const bool ok = read_str("PDB_PATH", &pf->pdb) && read_str("PSF_PATH", &pf->psf) &&
                  read_str("NEU_PATH", &pf->neu) && read_f("DELTA_X", &pf->dx) &&
                  read_f("DELTA_T", &pf->dt) && read_f("SIM_TIME", &pf->sim_time) &&
                  read_i("OUT_FREQ", &pf->out_freq) && read_i("MAX_CELL", &pf->max_cell) &&
                  read_i("MAX_ELEM", &pf->max_elem) && read_i("SURF_ELEM", &pf->surf_elem) &&
                  read_f("RECEPTOR", &pf->host[0]) && read_f("RANDOM_SEED", &pf->host[1]) &&
                  read_f("K_SPRING", &pf->host[2]) && read_f("Kr0", &pf->host[4]) &&
                  read_f("SHEAR_RATE", &pf->host[5]) && read_f("INIT_RHO", &pf->host[6]);
  std::fclose(in);
  if (ok && (pf->dx <= 0.f || pf->dt <= 0.f || pf->out_freq <= 0)) {
    *err = "DELTA_X, DELTA_T and OUT_FREQ must be positive";
    return false;
  }
  return ok;
}

/* The derived entries of hostSEMparameters, as runGPUSimulation fills them
 * (gpuSEM_LB.c:1206-1233): lz is the lattice depth, Lm the mean rest length
 * of the membrane edges (gpuSEM_LB.c:834-844) in the same length unit as
 * DELTA_X. */
void fill_derived(ParFile *pf, int lz, double Lm) {
  float *h = pf->host;
  const double shear = h[5], rho0 = h[6], dx = pf->dx, dt = pf->dt;
  const double umax = shear * (lz - 1) * dx;
  h[3] = (float)(0.0165 * std::exp(0.00065 * shear)); /* Kf0 */
  const double L0 = 1.0;
  h[7] = (float)L0;
  h[8] = (float)(L0 / umax);
  h[9] = (float)(rho0 * L0 * L0 * L0);
  h[10] = 1.0f;
  h[11] = (float)umax;
  h[12] = (float)(1.2e6 / rho0);
  h[13] = (float)(3.0 * (double)h[12] * dt / (dx * dx) + 0.5);
  h[14] = (float)(rho0 * umax * umax);
  h[15] = (float)(dx / dt / umax);
  h[16] = 1e12f;
  const double kbt = 1.3806488e-8 * 300.0;
  h[17] = (float)(600.0 * kbt / 2.0 / Lm / Lm);
  h[18] = (float)(600.0 * kbt / 2.0 / Lm / Lm / Lm);
  h[19] = (float)(500.0 * kbt * 1e-6);
}

/* ---------------- command line ---------------- */

struct Options {
  std::string par = "Input/parameter.dat";
  std::string out = ".";
  int device = 0;
  long long seed_first = -1, seed_last = -1;
  int scene = PFG_SCENE_SHEAR;
  int nx = 0, ny = 0, nz = 0;
  int store = -1;
  long long max_steps = -1;
  bool rates_from_file = false;
  bool scene_physics = false; /* --scene-physics: keep the scene's physics */
  bool scene_counts = false;
  bool parse_only = false;
  std::string mesh;      /* --mesh */
  float mesh_scale = 0;  /* --mesh-scale; 0: 1 / DELTA_X */
  bool no_mesh = false;  /* --no-mesh: generated platelets even when NEU_PATH exists */
  std::string fibrin;    /* --fibrin */
  float fibrin_scale = 0;
};

int usage(const char *msg) {
  std::fprintf(stderr, "pfg_driver: %s\n", msg);
  std::fprintf(stderr,
               "usage: pfg_driver [-d device] [-p parameter file] [-o output dir]\n"
               "       [--scene shear|channel|thrombus|large] [--lattice nx ny nz]\n"
               "       [--store f32|f16|bf16] [--seeds first last] [--max-steps n]\n"
               "       [--mesh file.neu] [--mesh-scale s] [--no-mesh]\n"
               "       [--fibrin network.txt] [--fibrin-scale s]\n"
               "       [--rates-from-file] [--scene-physics] [--scene-counts]\n"
               "       [--parse-only]\n");
  return 1;
}

bool parse_args(int argc, char **argv, Options *o, std::string *err) {
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto need = [&](int n) -> bool {
      if (i + n >= argc) {
        *err = "missing value after " + a;
        return false;
      }
      return true;
    };
    if (a == "-d") {
      if (!need(1)) return false;
      o->device = std::atoi(argv[++i]);
    } else if (a == "-p") {
      if (!need(1)) return false;
      o->par = argv[++i];
    } else if (a == "-o") {
      if (!need(1)) return false;
      o->out = argv[++i];
    } else if (a == "--scene") {
      if (!need(1)) return false;
      const std::string s = argv[++i];
      if (s == "shear") o->scene = PFG_SCENE_SHEAR;
      else if (s == "channel") o->scene = PFG_SCENE_CHANNEL;
      else if (s == "thrombus") o->scene = PFG_SCENE_THROMBUS;
      else if (s == "large") o->scene = PFG_SCENE_LARGE;
      else {
        *err = "unknown scene " + s;
        return false;
      }
    } else if (a == "--lattice") {
      if (!need(3)) return false;
      o->nx = std::atoi(argv[++i]);
      o->ny = std::atoi(argv[++i]);
      o->nz = std::atoi(argv[++i]);
      if (o->nx <= 0 || o->ny <= 0 || o->nz <= 0) {
        *err = "--lattice needs three positive sizes";
        return false;
      }
    } else if (a == "--store") {
      if (!need(1)) return false;
      const std::string s = argv[++i];
      if (s == "f32") o->store = PFG_STORE_F32;
      else if (s == "f16") o->store = PFG_STORE_F16;
      else if (s == "bf16") o->store = PFG_STORE_BF16;
      else {
        *err = "unknown store " + s;
        return false;
      }
    } else if (a == "--seeds") {
      if (!need(2)) return false;
      o->seed_first = std::atoll(argv[++i]);
      o->seed_last = std::atoll(argv[++i]);
      if (o->seed_first <= 0 || o->seed_last < o->seed_first) {
        *err = "--seeds needs 0 < first <= last";
        return false;
      }
    } else if (a == "--max-steps") {
      if (!need(1)) return false;
      o->max_steps = std::atoll(argv[++i]);
      if (o->max_steps <= 0) {
        *err = "--max-steps must be positive";
        return false;
      }
    } else if (a == "--mesh") {
      if (!need(1)) return false;
      o->mesh = argv[++i];
    } else if (a == "--mesh-scale") {
      if (!need(1)) return false;
      o->mesh_scale = (float)std::atof(argv[++i]);
      if (!(o->mesh_scale > 0)) {
        *err = "--mesh-scale must be positive";
        return false;
      }
    } else if (a == "--no-mesh") {
      o->no_mesh = true;
    } else if (a == "--fibrin") {
      if (!need(1)) return false;
      o->fibrin = argv[++i];
    } else if (a == "--fibrin-scale") {
      if (!need(1)) return false;
      o->fibrin_scale = (float)std::atof(argv[++i]);
      if (!(o->fibrin_scale > 0)) {
        *err = "--fibrin-scale must be positive";
        return false;
      }
    } else if (a == "--rates-from-file") {
      o->rates_from_file = true;
    } else if (a == "--scene-physics") {
      o->scene_physics = true;
    } else if (a == "--scene-counts") {
      o->scene_counts = true;
    } else if (a == "--parse-only") {
      o->parse_only = true;
    } else {
      *err = "unknown argument " + a;
      return false;
    }
  }
  return true;
}

/* ---------------- output files ---------------- */

struct Frame {
  std::vector<float> x, y, z, vx, vy, vz, fx, fy, fz;
  explicit Frame(int n) : x(n), y(n), z(n), vx(n), vy(n), vz(n), fx(n), fy(n), fz(n) {}
};

struct Outputs {
  FILE *f[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  ~Outputs() {
    for (FILE *p : f)
      if (p) std::fclose(p);
  }
  bool open(const std::string &dir) {
    const std::string od = dir + "/Output";
    if (::mkdir(od.c_str(), 0777) != 0 && errno != EEXIST) return false;
    static const char *names[7] = {"cellsPos.data",      "cellsVel.data",      "cellsForce.data",
                                   "cellsPosASCII.dat",  "cellsVelASCII.dat",  "cellsForceASCII.dat",
                                   "cellsPauseTime.dat"};
    for (int k = 0; k < 7; ++k) {
      f[k] = std::fopen((od + "/" + names[k]).c_str(), k == 6 ? "a+" : "wb+");
      if (!f[k]) return false;
    }
    for (int k = 3; k < 6; ++k)
      std::fprintf(f[k], "Timestep       ElemID        X         Y         Z\n");
    return true;
  }
};

void put4(FILE *fp, float a, float b, float c) {
  const double v[4] = {a, b, c, c}; /* the reference repeats z for alignment */
  std::fwrite(v, sizeof(double), 4, fp);
}

/* Element order: cell by cell, element by element (the reference's
 * [element][cell] arrays read with the cell loop outside). */
void write_frame(Outputs &o, const Frame &fr, const std::vector<int32_t> &start, long long t,
                 bool initial) {
  const int ncell = (int)start.size() - 1;
  for (int j = 0; j < ncell; ++j) {
    for (int i = 0; i < start[j + 1] - start[j]; ++i) {
      const int m = start[j] + i;
      put4(o.f[0], fr.x[m], fr.y[m], fr.z[m]);
      if (initial) {
        put4(o.f[1], fr.vx[m], fr.vy[m], fr.vz[m]);
        put4(o.f[2], fr.fx[m], fr.fy[m], fr.fz[m]);
        std::fprintf(o.f[3], "0      %d    %e   %e   %e\n ", i, fr.x[m], fr.y[m], fr.z[m]);
        std::fprintf(o.f[4], "0      %d    %e   %e   %e\n ", i, fr.vx[m], fr.vy[m], fr.vz[m]);
        std::fprintf(o.f[5], "0      %d    %e   %e   %e\n ", i, fr.fx[m], fr.fy[m], fr.fz[m]);
      } else {
        std::fprintf(o.f[3], "%lld      %d  %e  %e  %e\n ", t, i, fr.x[m], fr.y[m], fr.z[m]);
        std::fprintf(o.f[4], "%lld      %d  %e  %e  %e\n ", t, i, fr.vx[m], fr.vy[m], fr.vz[m]);
        std::fprintf(o.f[5], "%lld      %d  %e  %e  %e\n ", t, i, fr.fx[m], fr.fy[m], fr.fz[m]);
      }
    }
  }
}

/* ---------------- one run (runGPUSimulation) ---------------- */

struct SceneGuard {
  pfg_scene *s = nullptr;
  ~SceneGuard() {
    if (s) pfg_scene_destroy(s);
  }
};
struct HandleGuard {
  pfg_handle *h = nullptr;
  ~HandleGuard() {
    if (h) pfg_destroy(h);
  }
};

double mean_edge_length(const pfg_model &m) {
  if (m.n_edges <= 0) return 1.0;
  double s = 0.0;
  for (int e = 0; e < m.n_edges; ++e) s += m.edge_l0[e];
  return s / m.n_edges;
}

bool readable(const std::string &path) {
  FILE *f = std::fopen(path.c_str(), "r");
  if (!f) return false;
  std::fclose(f);
  return true;
}

/* The mesh file of the run: --mesh, else NEU_PATH as written, next to the
 * parameter file, or in its parent directory; "" when there is none. */
std::string mesh_file(const Options &o, const ParFile &pf) {
  if (o.no_mesh) return "";
  if (!o.mesh.empty()) return o.mesh;
  if (pf.neu.empty()) return "";
  if (readable(pf.neu)) return pf.neu;
  const size_t cut = o.par.find_last_of('/');
  const std::string dir = cut == std::string::npos ? "." : o.par.substr(0, cut);
  const size_t b = pf.neu.find_last_of('/');
  const std::string base = b == std::string::npos ? pf.neu : pf.neu.substr(b + 1);
  if (readable(dir + "/" + base)) return dir + "/" + base;
  if (readable(dir + "/../" + pf.neu)) return dir + "/../" + pf.neu;
  return "";
}

int build_scene(const Options &o, ParFile *pf, long long seed, SceneGuard *sg, pfg_params *p,
                pfg_model *model) {
  pfg_scene_opts so;
  pfg_scene_opts_init(&so, o.scene);
  const std::string mesh = mesh_file(o, *pf);
  if (!mesh.empty()) {
    so.mesh_path = mesh.c_str();
    so.mesh_scale = o.mesh_scale > 0 ? o.mesh_scale : (pf->dx > 0 ? 1.0f / pf->dx : 1.0f);
  }
  if (!o.fibrin.empty()) {
    so.fibrin_path = o.fibrin.c_str();
    so.fibrin_scale = o.fibrin_scale;
  }
  so.nx = o.nx;
  so.ny = o.ny;
  so.nz = o.nz;
  so.store = o.store;
  so.seed = (uint64_t)seed;
  if (!o.scene_counts) {
    /* MAX_CELL platelets of MAX_ELEM elements; receptors per triangle as
     * gpuSEM_LB.c:723 (RECEPTOR / SURF_ELEM / 2, truncated) over the 2n-4
     * triangles of the generated mesh */
    so.n_platelets = pf->max_cell > 0 ? pf->max_cell : -1;
    if (pf->max_elem >= 4) so.elements = pf->max_elem;
    const int per_tri = pf->surf_elem > 0 ? (int)(pf->host[0] / pf->surf_elem / 2) : 0;
    int ntri = 2 * (so.elements > 0 ? so.elements : pf->max_elem) - 4;
    if (!mesh.empty()) {
      int32_t nn = 0, nt = 0;
      if (pfg_geom_read_neu(mesh.c_str(), 0, 0, nullptr, nullptr, nullptr, nullptr, &nn, &nt) !=
          PFG_OK) {
        std::fprintf(stderr, "pfg_driver: cannot read mesh %s\n", mesh.c_str());
        return 2;
      }
      ntri = nt;
    }
    so.receptors = per_tri > 0 ? per_tri * ntri : -1;
  }
  const pfg_status bs = pfg_scene_build(&so, p, &sg->s);
  if (bs == PFG_ERR_IO) return 2;
  if (bs != PFG_OK) return 3;
  if (pfg_scene_model(sg->s, model) != PFG_OK) return 3;
  fill_derived(pf, p->nz, mean_edge_length(*model) * pf->dx);
  for (int k = 0; k < 20; ++k) p->sem[k] = pf->host[k];
  p->seed = (uint64_t)seed;
  if (!o.scene_physics) {
    pfg_params q = *p;
    if (pfg_params_from_sem(pf->host, pf->dx, pf->dt, p->nx, p->ny, p->nz, &q) != PFG_OK)
      return 1;
    p->tau = q.tau;
    p->ib_force_scale = q.ib_force_scale;
    p->marker_dt = q.marker_dt;
    p->k_spring = q.k_spring;
    p->ks = q.ks;
    p->kv = q.kv;
    p->kbend = q.kbend;
    p->k_on = q.k_on;
    p->k_off0 = q.k_off0;
    p->detach_rate = q.detach_rate;
    p->bond_law = q.bond_law;
    p->bond_k = q.bond_k;
    p->bond_l0 = q.bond_l0;
    p->bell_xb_kT = q.bell_xb_kT;
    p->bell_kT = q.bell_kT;
    p->wall_rep_a = q.wall_rep_a;
    p->wall_rep_kappa = q.wall_rep_kappa;
    p->wall_rep_hc = q.wall_rep_hc;
    p->morse_u0 = 0.f; /* NO_NB: the reference's platelets have no Morse term */
    /* the shear: SHEAR_RATE x dt per lattice length, over the nz cells
     * between the two z walls, along the scene's flow direction */
    if (p->wall[2] == PFG_WALL_BOUNCE) {
      const float u_top = (float)((double)pf->host[5] * pf->dt * p->nz);
      for (int a = 0; a < 3; ++a)
        if (p->wall_u_hi[2][a] != 0.f) p->wall_u_hi[2][a] = u_top;
    }
  } else if (o.rates_from_file) {
    p->k_on = pf->host[3] * pf->dt;
    p->k_off0 = pf->host[4] * pf->dt;
    p->detach_rate = pf->host[5] * pf->dt; /* gama = SHEAR_RATE * dt */
  }
  return 0;
}

/* The initial flow per z-plane: the scene's profile, rescaled to the top
 * wall velocity of the run (a linear Couette profile in the shear scenes). */
pfg_status initial_flow(const pfg_scene *s, const pfg_params &p, std::vector<float> *u) {
  for (int a = 0; a < 3; ++a) u[a].assign((size_t)p.nz, 0.f);
  pfg_status st = pfg_scene_velocity_z(s, u[0].data(), u[1].data(), u[2].data());
  if (st != PFG_OK) return st;
  /* the scene's flow was built for its own wall velocity */
  pfg_params sp;
  if ((st = pfg_scene_params(s, &sp)) != PFG_OK) return st;
  for (int a = 0; a < 3; ++a)
    if (sp.wall_u_hi[2][a] != 0.f)
      for (float &v : u[a]) v *= p.wall_u_hi[2][a] / sp.wall_u_hi[2][a];
  return PFG_OK;
}

int run_seed(const Options &o, ParFile pf, long long seed, long long max_steps) {
  SceneGuard sg;
  pfg_params p;
  pfg_model model;
  int rc = build_scene(o, &pf, seed, &sg, &p, &model);
  if (rc) {
    std::fprintf(stderr, "pfg_driver: scene build failed\n");
    return rc;
  }
  HandleGuard hg;
  pfg_status st = pfg_create(&p, o.device, &hg.h);
  if (st != PFG_OK) {
    std::fprintf(stderr, "pfg_driver: pfg_create: %s\n", pfg_status_string(st));
    return st == PFG_ERR_DEVICE ? 11 : 3;
  }
  pfg_handle *h = hg.h;
  std::vector<float> u0[3];
  const uint8_t *flags = pfg_scene_flags(sg.s);
  if ((flags && p.use_flags && (st = pfg_upload_flags(h, flags)) != PFG_OK) ||
      (st = initial_flow(sg.s, p, u0)) != PFG_OK ||
      (st = pfg_init_equilibrium_z(h, u0[0].data(), u0[1].data(), u0[2].data())) != PFG_OK ||
      (st = pfg_load_model(h, &model)) != PFG_OK || (st = pfg_set_halt_on_detach(h, 1)) != PFG_OK) {
    std::fprintf(stderr, "pfg_driver: set-up: %s\n", pfg_status_string(st));
    return 3;
  }
  std::vector<int32_t> start(model.plt_elem_start, model.plt_elem_start + model.n_platelets + 1);
  if (model.n_platelets == 0) start.assign(1, 0);
  Outputs out;
  if (!out.open(o.out)) {
    std::fprintf(stderr, "pfg_driver: cannot open the files under %s/Output\n", o.out.c_str());
    return 2;
  }
  std::printf("seed %lld: lattice %dx%dx%d, %d platelets, %d markers, %d receptors, tau %.4f\n",
              seed, p.nx, p.ny, p.nz, model.n_platelets, model.n_markers, model.n_receptors,
              p.tau);
  Frame fr(model.n_markers);
  auto fetch = [&]() {
    return pfg_download_markers(h, fr.x.data(), fr.y.data(), fr.z.data(), fr.vx.data(),
                                fr.vy.data(), fr.vz.data(), fr.fx.data(), fr.fy.data(),
                                fr.fz.data());
  };
  if ((st = fetch()) != PFG_OK) return 3;
  write_frame(out, fr, start, 0, true);

  long long totalT = 0;
  bool detached = false;
  while (!detached && totalT < max_steps) {
    const long long n = std::min<long long>(pf.out_freq, max_steps - totalT);
    pfg_step_info info;
    if ((st = pfg_step(h, (int)n, &info)) != PFG_OK) {
      std::fprintf(stderr, "pfg_driver: pfg_step: %s\n", pfg_status_string(st));
      return 3;
    }
    totalT = info.step;
    detached = info.detached != 0;
    if ((st = fetch()) != PFG_OK) return 3;
    std::printf("Time Step = %lld, time = %e, bonds %d\n", totalT, totalT * (double)pf.dt,
                info.total_bonds);
    write_frame(out, fr, start, totalT, false);
  }
  if (detached) {
    std::fprintf(out.f[6], "%e, %lld\n", totalT * (double)pf.dt, seed);
    std::printf("seed %lld: detached after %lld steps (%e)\n", seed, totalT,
                totalT * (double)pf.dt);
  } else {
    std::printf("seed %lld: no detach within %lld steps\n", seed, max_steps);
  }
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  Options o;
  std::string err;
  if (!parse_args(argc, argv, &o, &err)) return usage(err.c_str());
  ParFile pf;
  if (!read_par_file(o.par.c_str(), &pf, &err)) {
    std::printf("%s \n", err.c_str());
    return 1;
  }
  const long long seed0 = (long long)pf.host[1];
  const long long first = o.seed_first > 0 ? o.seed_first : (seed0 > 0 ? seed0 : 1101);
  const long long last = o.seed_first > 0 ? o.seed_last : first;
  /* SIM_TIME / DELTA_T steps (gpuSEM_LB.c:701) cap each run */
  long long max_steps = o.max_steps > 0 ? o.max_steps
                                        : (long long)std::floor(pf.sim_time / pf.dt + 0.5);
  if (max_steps < 1) max_steps = 1;

  if (o.parse_only) {
    SceneGuard sg;
    pfg_params p;
    pfg_model model;
    const int rc = build_scene(o, &pf, first, &sg, &p, &model);
    if (rc) return rc;
    std::printf("{\"mesh\": \"%s\", \"triangles\": %d, \"links\": %d, ",
                mesh_file(o, pf).c_str(), model.n_tri, model.n_links);
    std::printf("\"max_steps\": %lld, \"out_freq\": %d, \"seeds\": [%lld, %lld], "
                "\"lattice\": [%d, %d, %d], \"platelets\": %d, \"markers\": %d, "
                "\"receptors\": %d, \"tau\": %.9g, \"k_on\": %.9g, \"k_off0\": %.9g, "
                "\"detach_rate\": %.9g, \"seed\": %llu, ",
                max_steps, pf.out_freq, first, last, p.nx, p.ny, p.nz, model.n_platelets,
                model.n_markers, model.n_receptors, p.tau, p.k_on, p.k_off0, p.detach_rate,
                (unsigned long long)p.seed);
    std::printf("\"physics\": \"%s\", \"ib_force_scale\": %.9g, \"marker_dt\": %.9g, "
                "\"k_spring\": %.9g, \"ks\": %.9g, \"kv\": %.9g, \"kbend\": %.9g, "
                "\"bond_law\": %d, \"bond_k\": %.9g, \"bond_l0\": %.9g, "
                "\"wall_rep_a\": %.9g, \"morse_u0\": %.9g, \"wall_u_top\": %.9g, \"sem\": [",
                o.scene_physics ? "scene" : "file", p.ib_force_scale, p.marker_dt, p.k_spring,
                p.ks, p.kv, p.kbend, p.bond_law, p.bond_k, p.bond_l0, p.wall_rep_a, p.morse_u0,
                p.wall_u_hi[2][0] + p.wall_u_hi[2][1]);
    for (int k = 0; k < 20; ++k) std::printf("%s%.9g", k ? ", " : "", p.sem[k]);
    std::printf("]}\n");
    return 0;
  }
  int ndev = 0;
  if (pfg_device_count(&ndev) != PFG_OK || o.device < 0 || o.device >= ndev) {
    std::printf("invalid device number %d (%d visible)\n", o.device, ndev);
    return 11;
  }
  for (long long seed = first; seed <= last; ++seed) {
    const int rc = run_seed(o, pf, seed, max_steps);
    if (rc) return rc;
  }
  return 0;
}
