/*
 * pfg_geometry.h -- host-side (CPU only) construction of the seeded scenes the
 * solver runs: platelet meshes, receptors, ligands, fibrin networks, solid
 * flags and initial flow.  Replaces the reference's setup in
 * runGPUSimulation (gpuSEM_LB.c:649-1258: NEU mesh, topology, S0/V0,
 * receptor placement, vWF grid, equilibrium init) and init_fibrin_network
 * (fiber.cpp:109-498) with generators driven by the counter RNG, so the same
 * seed gives the same scene on every machine.
 */
#ifndef PFG_GEOMETRY_H
#define PFG_GEOMETRY_H

#include "pfg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Workloads.  Sizes are those of the named configuration unless overridden
 * by a nonzero field of pfg_scene_opts. */
enum {
  PFG_SCENE_CHANNEL = 1,  /* 64x32x32 body-force channel, one 256-element platelet  */
  PFG_SCENE_SHEAR = 2,    /* 256x128x128 Couette, 2048 elements, 512 receptors,
                             4096 wall ligands                                      */
  PFG_SCENE_THROMBUS = 3, /* 512x256x256 tube R=120, 500x1000 elements, fibrin
                             20k nodes / 30k links, wall + fibrin ligands          */
  PFG_SCENE_LARGE = 4,    /* shear physics at 1024x512x512, FP16 storage, 2000x1000
                             elements, fibrin 80k / 120k                            */
  PFG_SCENE_BATCH = 5     /* one instance of the batched workload                  */
};

typedef struct pfg_scene_opts {
  int32_t config;                  /* PFG_SCENE_*                                  */
  int32_t nx, ny, nz;              /* lattice override (0: config default)         */
  int32_t n_platelets;             /* override (0: default; -1: none)              */
  int32_t elements;                /* elements per platelet                        */
  int32_t receptors;               /* receptors per platelet (-1: none)            */
  int32_t ligands;                 /* static wall ligands (-1: none)               */
  int32_t fibrin_nodes;            /* -1: none                                     */
  int32_t fibrin_links;
  float tube_radius;               /* 0: config default                            */
  int32_t store;                   /* pfg_store, -1: config default                */
  uint64_t seed;                   /* 0: 1101 (gpuSEM_LB.c:1614)                   */
  const char *mesh_path;           /* NULL: generated ellipsoids; else every
                                      platelet is this Gambit .neu mesh, read as
                                      NEUReader reads it (gpuSEM_LB.c:346-424),
                                      centred on its placement and scaled       */
  float mesh_scale;                /* lattice units per mesh unit; 0: the scale
                                      that makes the mean mesh edge 0.75        */
  const char *fibrin_path;         /* NULL: generated network (when the scene
                                      has one); else the network of this file
                                      (pfg_geom_read_fibrin), centred in the
                                      scene's fibrin box and scaled, replacing
                                      fibrin_nodes / fibrin_links              */
  float fibrin_scale;              /* lattice units per file unit; 0: fit the
                                      file's bounding box into the fibrin box   */
} pfg_scene_opts;

typedef struct pfg_scene pfg_scene;

/* Zeroed options for a configuration (store = -1). */
void pfg_scene_opts_init(pfg_scene_opts *o, int config);
/* Build the scene and the parameters it runs with. */
pfg_status pfg_scene_build(const pfg_scene_opts *o, pfg_params *params,
                           pfg_scene **out);
pfg_status pfg_scene_destroy(pfg_scene *s);
/* Pointers into scene-owned arrays, valid until pfg_scene_destroy. */
pfg_status pfg_scene_model(const pfg_scene *s, pfg_model *out);
/* Solid flags [nz][ny][nx] or NULL when the scene has none. */
const uint8_t *pfg_scene_flags(const pfg_scene *s);
/* Initial velocity field [nz][ny][nx] per component (equilibrium init). */
pfg_status pfg_scene_velocity(const pfg_scene *s, float *ux, float *uy, float *uz);
/* The same initial velocity as one value per z-plane (it varies along z
 * only): nz values per component. */
pfg_status pfg_scene_velocity_z(const pfg_scene *s, float *ux_z, float *uy_z,
                                float *uz_z);
/* The parameters the scene was built with (as pfg_scene_build returned). */
pfg_status pfg_scene_params(const pfg_scene *s, pfg_params *out);
/* Fluid-cell count (cells not flagged solid). */
int64_t pfg_scene_fluid_cells(const pfg_scene *s);

/* ---- primitives (exposed for tests) ---- */

/* n points on the ellipsoid (a,b,c) about (cx,cy,cz): Fibonacci spiral on the
 * unit sphere, convex-hull triangulation (2n-4 triangles, wound so that
 * e01 x e02 points outward), each coordinate jittered by a seeded uniform in
 * [-jitter, jitter].  tri must hold 3*(2n-4) ints. */
pfg_status pfg_geom_ellipsoid(int n, float a, float b, float c, float cx, float cy,
                              float cz, float jitter, uint64_t seed, float *x,
                              float *y, float *z, int32_t *tri, int32_t *n_tri);

/* Gambit .neu surface mesh as NEUReader reads it (gpuSEM_LB.c:346-424): the
 * node and triangle counts, then one line per node "index x y z" and one
 * per triangle "index type nodes n1 n2 n3", indices from 1.  With x == NULL
 * only the counts are returned; otherwise the buffers must hold max_nodes
 * nodes and max_tri triangles (3 ints each, from 0).  PFG_ERR_IO: the file
 * cannot be read; PFG_ERR_ARG: a count, an index or a value is missing or
 * out of range, or a buffer is too small. */
pfg_status pfg_geom_read_neu(const char *path, int max_nodes, int max_tri, float *x,
                             float *y, float *z, int32_t *tri, int32_t *n_nodes,
                             int32_t *n_tri);

/* Fibrin network text file as init_fibrin_network reads it
 * (fiber.cpp:145-217): one line per node, "index x y z n c_1 .. c_n
 * t_1 .. t_n" with indices from 1, neighbours c_k and fibre thicknesses t_k.
 * Links are the pairs (i, c) with c > i of each node's list, repeats kept
 * (the reference counts fibres so, fiber.cpp:288-306); self links are
 * dropped (fiber.cpp:190-196).  link holds 2 ints per link (from 0),
 * thickness one float.  With x == NULL only the counts are returned.
 * PFG_ERR_IO: the file cannot be read; PFG_ERR_ARG: a malformed line, an
 * index out of range, a node listed twice or missing, a neighbour list that
 * is not symmetric, or a buffer too small. */
pfg_status pfg_geom_read_fibrin(const char *path, int max_nodes, int max_links, float *x,
                                float *y, float *z, int32_t *link, float *thickness,
                                int32_t *n_nodes, int32_t *n_links);

/* Triangle areas, total area and the reference-signed volume
 * V = sum (c . n_elem) S / 3 with n_elem = -(e01 x e02)/|e01 x e02|
 * (sem_kernel.cu:793-808), in FP64. */
pfg_status pfg_geom_area_volume(int n_tri, const int32_t *tri, const float *x,
                                const float *y, const float *z, double *tri_S,
                                double *S_all, double *V);

/* Random geometric graph: n nodes uniform in the box [lo, hi), links between
 * pairs closer than radius (network.cpp:197-218 with i < j), at most
 * max_links of them chosen by seeded hash, sorted by (i, j). Returns the
 * link count through n_links; link must hold 2*max_links ints. */
pfg_status pfg_geom_fibrin(int n, const float lo[3], const float hi[3],
                           float radius, int max_links, uint64_t seed, float *x,
                           float *y, float *z, int32_t *link, float *link_l0,
                           int32_t *n_links);

#ifdef __cplusplus
}
#endif

#endif
