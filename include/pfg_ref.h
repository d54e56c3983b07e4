/*
 * pfg_ref.h -- the reference's own C entry points, exported by libpfg.so.
 *
 * The reference's host program (gpuSEM_LB.c) declares and calls these names
 * (prototypes at gpuSEM_LB.c:57-84 and GPUDefines.h:211-225); they are the
 * `extern "C"` glue of semLB.cu:48-1036.  libpfg.so exports every one of
 * them with the same signature, so that program links against this library
 * unchanged.  Each is a thin adapter over the pfg_* ABI (include/pfg.h):
 * it takes the reference's host layouts, converts units and layouts, and
 * drives a pfg handle.
 *
 * Host layouts (caller owned, as the reference's host code builds them):
 *   lattice fields       FP64 [depth][height][width] (x fastest); fIN/fOUT
 *                        are 19 such planes in the direction order of
 *                        LB_kernel.cu:40-47; obstacle is FP32, > 0.5 solid
 *   SEM element arrays   FP64 [maxElements][maxCells] (element row, platelet
 *                        column); positions in micrometres, velocities in
 *                        units of UMAX, forces in units of F0
 *   triElem              int [SurfElem][6]: the triangle's three elements,
 *                        then its three edge-midpoint nodes
 *   receptor_r1/_r2      float [SurfElem][numReceptorsPerElem]
 *   receptBond           int [3][numReceptorsPerElem][SurfElem]: vWF grid
 *                        column and row of the bond, and 1 when bonded
 *   bondLengths          FP64 [maxElements][maxElements] dense rest lengths
 *                        (micrometres); a positive entry is a spring
 *   vWFbond              int [4][height][width] ligand slots of the wall
 *
 * Units: the lattice step dx and time step dt come from
 * fluid_allocGPUKernel; C = dx/dt/UMAX, UMAX = hostParameters[11]; the
 * physical parameter block hostParameters[0..19] (gpuSEM_LB.c:1206-1233) is
 * mapped by pfg_params_from_sem.
 *
 * Geometry: the reference's channel (gpuSEM_LB.c:1182-1199) has obstacle
 * planes at z = 0 and z = depth-1 and a top plate moving along y
 * (fluid3d_moving_plate_boundary_kernel, LB_kernel.cu:602-780).  When the
 * obstacle cells are exactly those two planes, the device lattice is the
 * depth-2 interior planes with halfway bounce-back walls between the
 * obstacle and the first interior plane, each moving with the linear shear
 * profile's velocity there (so the reference's linear shear is an exact
 * steady state of the device lattice); lattice node k of the reference is
 * device cell k-1.  Any other obstacle pattern is uploaded as solid cells of
 * a fully periodic lattice (LB_kernel.cu:1347-1352 wraps every axis).
 *
 * The gsl_rng* arguments are accepted and not used: every random draw comes
 * from the counter RNG keyed by hostParameters[1] (the seed).  The reference
 * passes gama = SHEAR_RATE * dt to sem_invokeGPUKernel_Force; the detach
 * hazard is hostParameters[5] * dt, the same number.
 *
 * Errors: the reference prints and calls exit() on a CUDA error (semLB.cu:
 * 86-94).  These adapters print the pfg status to stderr and keep the
 * handle unchanged; pfg_ref_last_status() returns the last status.
 */
#ifndef PFG_REF_H
#define PFG_REF_H

#include "pfg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- fluid (semLB.cu:111-557) ---- */
/* semLB.cu:111 */
void *fluid_allocGPUKernel(void *model, float dt, float dx, int width, int height,
                           int depth);
/* semLB.cu:154; aFlag 1: host -> device, 0: device -> host.  The reference
 * passes its hostSEMparameters block as hostParameters (gpuSEM_LB.c:1294);
 * the handle is created at the first aFlag = 1 call from it. */
void fluid_initGPUKernel(void *model, void *g, int aFlag, float *hostParameters,
                         void **fIN, void **fOUT, void *ux, void *uy, void *uz,
                         void *rho, void *obstacle, void *Fx, void *Fy, void *Fz,
                         void *vWFbond);
/* semLB.cu:303; r is the reference's gsl_rng* (not used) */
void fluid_invokeGPUKernel(void *model, void *g, void *g_SEM, double *randNum, void *r,
                           int timeSteps);
/* semLB.cu:534 */
void fluid_releaseGPUKernel(void *model, void *g);

/* ---- platelets (semLB.cu:749-1034) ---- */
/* semLB.cu:749 */
void *sem_allocGPUKernel(void *model, int maxCells, int maxElements, int SurfElem,
                         int newnode, int numReceptorsPerElem, float dt, double S0_all,
                         float *hostParameters);
/* semLB.cu:865 */
void sem_initGPUKernel(void *model, void *g, int numOfCells, int *numOfElements,
                       int SurfElem, int numReceptorsPerElem, void *hostX_Ref,
                       void *hostY_Ref, void *hostZ_Ref, void *hostX, void *hostY,
                       void *hostRY, void *hostZ, void *hostVX, void *hostVY, void *hostVZ,
                       void *hostFX, void *hostFY, void *hostFZ, void *hostType,
                       void *hostBonds, void *triElem, void *receptor_r1,
                       void *receptor_r2, void *node_share_Elem, void *N,
                       void *node_nbrElemNum, void *node_nbr_nodes, void *S0, double V0,
                       void *receptBond, void *randNum);
/* semLB.cu:441; r is the reference's gsl_rng* (not used).  *done becomes 1
 * once the detach test has fired (no bond held, hazard gama per step). */
void sem_invokeGPUKernel_Force(void *model, void *g, int timeSteps, int *done,
                               int *totalT, void *r, double gama);
/* semLB.cu:979 */
void sem_copyGPUKernel(void *model, void *g, void *hostX, void *hostY, void *hostRY,
                       void *hostZ, void *hostVX, void *hostVY, void *hostVZ, void *hostFX,
                       void *hostFY, void *hostFZ, int timeSteps);
/* semLB.cu:1005 */
void sem_releaseGPUKernel(void *model, void *g);

/* ---- fibrin network (semLB.cu:562-718) ---- */
/* semLB.cu:562 */
void *fiber_allocGPUKernel(void *model, int maxNodes, int maxLinks, int max_N_conn_at_Node,
                           double dt, double *hostParameters);
/* semLB.cu:609; NodeType 1 = anchored node; positions in micrometres */
void fiber_init_GPUKernel(void *model, void *fg, void *NodeType, void *N_Conn_at_Node,
                          void *Link_at_Node, void *lAdjVer, void *linkLengths,
                          void *linkLengths0, void *linkThick, void *X, void *V_X, void *X0,
                          void *F_X, void *Y, void *Y0, void *V_Y, void *F_Y, void *Z,
                          void *Z0, void *V_Z, void *F_Z);
/* semLB.cu:669 */
void fiber_copy_GPUKernel(void *model, void *fg, void *X, void *V_X, void *X0, void *F_X,
                          void *Y, void *Y0, void *V_Y, void *F_Y, void *Z, void *Z0,
                          void *V_Z, void *F_Z);
/* semLB.cu:708 */
void fiber_release_GPUKernel(void *model, void *fg);

/* ---- additions (no counterpart in the reference) ---- */
/* Status of the last adapter call. */
pfg_status pfg_ref_last_status(void);
/* The pfg handle behind a fluid grid of fluid_allocGPUKernel (NULL before
 * the first fluid_initGPUKernel with aFlag = 1), for callers that mix both
 * interfaces. */
pfg_handle *pfg_ref_handle(void *fluid_grid);

#ifdef __cplusplus
}
#endif

#endif /* PFG_REF_H */
