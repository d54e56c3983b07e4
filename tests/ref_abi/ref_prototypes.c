/*
 * ref_prototypes.c -- the entry-point declarations as the reference's host
 * program writes them (gpuSEM_LB.c:57-84 and GPUDefines.h:211-225; the
 * fiber_* calls as semLB.cu:562-718 defines them), with the GSL generator
 * as an opaque type as in the reference host.  Linked against libpfg.so
 * without pfg_ref.h: the reference host's object code resolves every name
 * it calls to this library.  Prints the number of entry points resolved.
 */
#include <stdio.h>

typedef struct gsl_rng gsl_rng; /* opaque: the library never dereferences it */

extern void *fiber_allocGPUKernel(void *, int, int, int, double, double *);
extern void *sem_allocGPUKernel(void *model, int maxCells, int maxElements, int SurfElem,
                                int newnode, int numReceptorsPerElem, float dt, double S0_all,
                                float *hostSEMparameters);
extern void sem_initGPUKernel(void *model, void *g, int numOfCells, int *numOfElements,
                              int SurfElem, int numReceptorsPerElem, void *hostX_Ref,
                              void *hostY_Ref, void *hostZ_Ref, void *hostX, void *hostY,
                              void *hostRY, void *hostZ, void *hostVX, void *hostVY,
                              void *hostVZ, void *hostFX, void *hostFY, void *hostFZ,
                              void *hostType, void *hostBonds, void *triElem,
                              void *receptor_r1, void *receptor_r2, void *node_share_Elem,
                              void *N, void *node_nbrElemNum, void *node_nbr_nodes, void *S0,
                              double V0, void *receptBond, void *randNum);
extern void sem_invokeGPUKernel_Force(void *model, void *g, int timeSteps, int *done,
                                      int *totalT, gsl_rng *r, double gama);
extern void sem_copyGPUKernel(void *model, void *g, void *hostX, void *hostY, void *hostRY,
                              void *hostZ, void *hostVX, void *hostVY, void *hostVZ,
                              void *hostFX, void *hostFY, void *hostFZ, int timeSteps);
extern void sem_releaseGPUKernel(void *model, void *g);
void *fluid_allocGPUKernel(void *model, float dt, float dx, int width, int height, int depth);
void fluid_initGPUKernel(void *model, void *g, int aFlag, float *hostLBparameters, void **fIN,
                         void **fOUT, void *ux, void *uy, void *uz, void *rho, void *obstacle,
                         void *Fx, void *Fy, void *Fz, void *vWFbond);
void fluid_invokeGPUKernel(void *model, void *g, void *g_SEM, double *randNum, gsl_rng *r,
                           int timeSteps);
void fluid_releaseGPUKernel(void *model, void *g);
extern void fiber_init_GPUKernel(void *model, void *fg, void *NodeType, void *N_Conn_at_Node,
                                 void *Link_at_Node, void *lAdjVer, void *linkLengths,
                                 void *linkLengths0, void *linkThick, void *X, void *V_X,
                                 void *X0, void *F_X, void *Y, void *Y0, void *V_Y, void *F_Y,
                                 void *Z, void *Z0, void *V_Z, void *F_Z);
extern void fiber_copy_GPUKernel(void *model, void *fg, void *X, void *V_X, void *X0,
                                 void *F_X, void *Y, void *Y0, void *V_Y, void *F_Y, void *Z,
                                 void *Z0, void *V_Z, void *F_Z);
extern void fiber_release_GPUKernel(void *model, void *fg);

int main(void) {
  void *fns[] = {
      (void *)fiber_allocGPUKernel,   (void *)sem_allocGPUKernel,
      (void *)sem_initGPUKernel,      (void *)sem_invokeGPUKernel_Force,
      (void *)sem_copyGPUKernel,      (void *)sem_releaseGPUKernel,
      (void *)fluid_allocGPUKernel,   (void *)fluid_initGPUKernel,
      (void *)fluid_invokeGPUKernel,  (void *)fluid_releaseGPUKernel,
      (void *)fiber_init_GPUKernel,   (void *)fiber_copy_GPUKernel,
      (void *)fiber_release_GPUKernel};
  int n = 0;
  for (unsigned k = 0; k < sizeof(fns) / sizeof(fns[0]); ++k) n += fns[k] != NULL;
  printf("resolved %d\n", n);
  return n == 13 ? 0 : 1;
}
