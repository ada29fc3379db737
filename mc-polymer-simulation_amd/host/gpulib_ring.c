/*
 * gpulib_ring.c -- the GPULib* boundary of the ring drop-in, with the
 * reference signatures (gpupol2/gpupol_cuda.h:6-12), implemented over the
 * C ABI of libelp.so (include/elp_gpu.h).
 *
 * Host RNG stream (reference seed semantics): 3 LFSR113 draws at load
 * (gpupol2/gpupol_cuda.cu:36-38) and 3 per time unit (:133-135), each scaled
 * to [0, tile edge).  The draws accumulate into the tile-grid origin; time
 * unit k moves in the frame of the shifts drawn before it, exactly as the
 * reference's relayout chain does.  The device lattice stays in one fixed
 * row-major frame, so no relayout kernel and no host un-shift are needed.
 *
 * Failures print and exit, as gpuErrchk does (gpupol2/gpupol_cuda_def.h:22-28).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../include/elp_gpu.h"
#include "gpupol_ring.h"
#include "elp_host_common.h"

#define ELPCHK(call)                                                                  \
    do {                                                                              \
        int rc_ = (call);                                                             \
        if (rc_) {                                                                    \
            fprintf(stderr, "GPUassert: %s (%s) %s %d\n", elp_last_error(), #call, __FILE__, __LINE__); \
            exit(rc_ < 0 ? -rc_ : rc_);                                               \
        }                                                                             \
    } while (0)

static int parse_tile(const char* s, int* t, int* u, int* v) {
    int a, b, c;
    if (sscanf(s, "%dx%dx%d", &a, &b, &c) == 3) { *t = a; *u = b; *v = c; return 0; }
    if (sscanf(s, "%d", &a) == 1) { *t = *u = *v = a; return 0; }
    return -1;
}

void ElpRingScheduleFromEnv(SimProperties* sp, GPUDeviceState* dev) {
    dev->tileT = dev->tileU = dev->tileV = ELP_RING_TILE;
    dev->sublattice = ELP_SUBLAT_CUBIC;
    dev->nIter = (int)sp->nSteps;
    const char* s;
    if ((s = getenv("ELP_TILE")) && *s) {
        if (parse_tile(s, &dev->tileT, &dev->tileU, &dev->tileV)) { fprintf(stderr, "bad ELP_TILE=%s\n", s); exit(2); }
    }
    if ((s = getenv("ELP_SUBLATTICE")) && *s) {
        if (!strcmp(s, "dense")) dev->sublattice = ELP_SUBLAT_DENSE;
        else if (!strcmp(s, "cubic")) dev->sublattice = ELP_SUBLAT_CUBIC;
        else { fprintf(stderr, "bad ELP_SUBLATTICE=%s\n", s); exit(2); }
    }
    if ((s = getenv("ELP_NITER")) && *s) dev->nIter = atoi(s);
}

/* gpupol2/gpupol_cuda.cu:6-27.  One MI355X per process: the first visible
   device is used (the reference's multi-device replication is broken beyond
   one device, gpupol2/gpupol_cuda.cu:146). */
int GPULibInit(SimProperties* sp, GPUDeviceState* devStates, GPULibContext* ctx) {
    (void)devStates;
    sp->nDevices = 1;
    ctx->devIds[0] = 0;
    return 0;
}

/* gpupol2/gpupol_cuda.cu:29-50 */
int GPULoadLattices(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx) {
    (void)ctx;
    for (int d = 0; d < sp->nDevices; d++) {
        GPUDeviceState* dev = devStates + d;
        ELPCHK(elp_upload((elp_engine*)dev->engine, ss[d].lattice, sp->latSize));
        dev->dt = elp_draw_shift(&ss[d].rngState, dev->tileT);
        dev->du = elp_draw_shift(&ss[d].rngState, dev->tileU);
        dev->dv = elp_draw_shift(&ss[d].rngState, dev->tileV);
        dev->cumDt = dev->dt; dev->cumDu = dev->du; dev->cumDv = dev->dv;
        dev->curBuf = 0;
    }
    return 0;
}

/* gpupol2/gpupol_cuda.cu:52-76 */
int GPULatticeToCPU(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx) {
    (void)ctx;
    for (int d = 0; d < sp->nDevices; d++)
        ELPCHK(elp_download((elp_engine*)devStates[d].engine, ss[d].lattice, sp->latSize));
    if (GetAllRingPolymers(sp, ss)) exit(193);
    for (int d = 0; d < sp->nDevices; d++)
        if (UpdatePolymerWithLabels(ss + d)) exit(123);
    return 0;
}

int ElpGpuTuvEnabled(void) {
    const char* env = getenv("ELP_GPU_TUV");
    return !(env && env[0] == '0');
}

int WriteLatticeFileDevice(SimProperties* sp, SimState* ss, GPUDeviceState* dev, const char* file) {
    elp_engine* e = (elp_engine*)dev->engine;
    uint64_t len = 0;
    const int maxLen = sp->polLength > 0 ? sp->polLength : (int)sp->maxPolLength;
    ELPCHK(elp_ring_tuv(e, ss->label2Id, ss->nPol, maxLen, (int)sp->maxPolLength, NULL, 0, &len));
    char* text = (char*)malloc(len ? len : 1);
    ELPCHK(elp_ring_tuv(e, ss->label2Id, ss->nPol, maxLen, (int)sp->maxPolLength, text, len, &len));
    FILE* f = fopen(file, "w");
    if (!f) { printf("Error opening file %s\n", file); free(text); return -1; }
    fwrite(text, 1, len, f);
    fclose(f);
    free(text);
    return 0;
}

/* gpupol2/gpupol_cuda.cu:78-99 */
int GPULibLoadBuffers(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx) {
    printf("Loading GPU buffers\n");
    for (int d = 0; d < sp->nDevices; d++) {
        GPUDeviceState* dev = devStates + d;
        ElpRingScheduleFromEnv(sp, dev);
        elp_engine* e = NULL;
        ELPCHK(elp_ring_create(&e, ctx->devIds[d], (int)sp->LT, (int)sp->LU, (int)sp->LV, dev->tileT, dev->tileU,
                               dev->tileV, dev->sublattice, dev->nIter));
        dev->engine = e;
        dev->devId = ctx->devIds[d];
    }
    GPULoadLattices(sp, ss, devStates, ctx);
    printf("Finished loading buffers\n");
    return 0;
}

/* gpupol2/gpupol_cuda.cu:101-112 */
int GPULibRelease(SimProperties* sp, GPUDeviceState* devStates, GPULibContext* ctx) {
    (void)ctx;
    for (int d = 0; d < sp->nDevices; d++) {
        elp_destroy((elp_engine*)devStates[d].engine);
        devStates[d].engine = NULL;
    }
    return 0;
}

/* Draw the shifts of nTime time units and run them; origins of time unit k
   are the accumulated shifts before its own draw. */
static void run_units(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, int nTime, long tau0) {
    if (nTime <= 0) return;
    int32_t* org = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)nTime);
    for (int d = 0; d < sp->nDevices; d++) {
        GPUDeviceState* dev = devStates + d;
        for (int i = 0; i < nTime; i++) {
            org[3 * i + 0] = (int32_t)(dev->cumDt % dev->tileT);
            org[3 * i + 1] = (int32_t)(dev->cumDu % dev->tileU);
            org[3 * i + 2] = (int32_t)(dev->cumDv % dev->tileV);
            dev->dt = elp_draw_shift(&ss[d].rngState, dev->tileT);
            dev->du = elp_draw_shift(&ss[d].rngState, dev->tileU);
            dev->dv = elp_draw_shift(&ss[d].rngState, dev->tileV);
            dev->cumDt += dev->dt; dev->cumDu += dev->du; dev->cumDv += dev->dv;
        }
        elp_engine* e = (elp_engine*)dev->engine;
        ELPCHK(elp_run(e, nTime, org, (uint64_t)tau0, sp->seed, (uint32_t)d));
        ELPCHK(elp_sync(e));
        float ms = 0;
        if (!elp_last_run_ms(e, &ms)) dev->lastKernelMs = ms;
    }
    free(org);
}

/* gpupol2/gpupol_cuda.cu:114-152 */
int GPULibRun(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx, int nTime) {
    run_units(sp, ss, devStates, nTime, sp->curT);
    GPULatticeToCPU(sp, ss, devStates, ctx);
    return 0;
}

/* GPULibRun without the copy of the lattice and the chains to the host (the
   caller writes the tuv file from the device, WriteLatticeFileDevice). */
int GPULibRunDevice(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx, int nTime) {
    (void)ctx;
    run_units(sp, ss, devStates, nTime, sp->curT);
    return 0;
}

/* gpupol2/gpupol_cuda.cu:154-198: runs of fastEq time units separated by a
   uniform reshuffle of stored length on the host. */
int GPULibEqRun(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx, int nTime) {
    long dt = sp->fastEq;
    long nEq = nTime / dt;
    long done = 0;
    for (long t = 0; t < nTime; t += dt) {
        long cur = dt + (((t / dt) < (nEq % dt)) ? 1 : 0);
        run_units(sp, ss, devStates, (int)cur, sp->curT + done);
        done += cur;
        GPULatticeToCPU(sp, ss, devStates, ctx);
        RedistribSL(ss, sp);
        GPULoadLattices(sp, ss, devStates, ctx);
    }
    return 0;
}
