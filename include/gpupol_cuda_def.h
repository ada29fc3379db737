/*
 * gpupol_cuda_def.h -- device-side state of the MI355X backend behind the
 * reference's GPULib* entry points.
 *
 * Replaces gpupol2/gpupol_cuda_def.h:8-20 and gpupol/gpupol_cuda_def.h:8-16,
 * the two headers of the reference that belong to its CUDA backend (they
 * include <cuda.h> and hold raw device buffers).  A host program of the
 * reference (gpupol2/gpupol2.c, gpupol/gpupol.c) compiled against this header
 * in their place declares its `GPULibContext gpuContext` and
 * `GPUDeviceState devices[N_MAX_GPU]` with the layout the libraries
 * libgpupol3_ring.so and libgpupol_lin.so expect; nothing in it is a CUDA or
 * HIP type.  The reference's plain-C structs (SimProperties, SimState,
 * Polymer in gpupol_def.h) are unchanged and byte-identical in
 * gpupol_ring.h / gpupol_lin.h.
 */
#ifndef GPUPOL_CUDA_DEF_H_INC
#define GPUPOL_CUDA_DEF_H_INC
#include <stdio.h>
#include <stdlib.h>

/* gpupol2/gpupol_cuda_def.h:8-10 */
typedef struct GPULibContext {
    int devIds[8];
} GPULibContext;

/* One device of the backend.  The reference keeps two lattice buffers, the
   trans table and the per-lane seed table here (gpupol2/gpupol_cuda_def.h:12-20);
   the MI355X engine keeps one in-place lattice and no RNG state, behind an
   opaque handle of libelp.so (include/elp_gpu.h). */
typedef struct GPUDeviceState {
    int devId;
    void* engine;                 /* elp_engine* of this device */
    int curBuf;                   /* always 0: one lattice, updated in place */
    long cumDt, cumDu, cumDv;     /* rings: cumulative tile-grid shift (gpupol2/gpupol_cuda.cu:137-139) */
    int dt, du, dv;               /* rings: last drawn shift */
    /* schedule of the engine (defaults reproduce the reference's: 32^3 tiles,
       one site per 4^3 cell, nSteps iterations per time unit) */
    int tileT, tileU, tileV, sublattice, nIter;
    double lastKernelMs;          /* device time of the last GPULibRun */
} GPUDeviceState;

#endif
