/*
 * gpulib_lin.c -- GPULib* boundary of the linear-chain drop-in with the
 * reference signatures (gpupol/gpupol_cuda.h:6-9) over libelp.so, plus the
 * handle API used by bench.py and the tests.
 *
 * Host RNG stream: 3 LFSR113 draws per time unit, each scaled to [0, tile
 * edge), used directly as that time unit's tile-grid origin
 * (gpupol/gpupol_cuda.cu:99-105; the linear kernel reads its tile at that
 * offset with periodic wrap, in place).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../include/elp_gpu.h"
#include "gpupol_lin.h"
#include "elp_host_common.h"

#define ELPCHK(call)                                                                                  \
    do {                                                                                              \
        int rc_ = (call);                                                                             \
        if (rc_) {                                                                                    \
            fprintf(stderr, "GPUassert: %s (%s) %s %d\n", elp_last_error(), #call, __FILE__, __LINE__); \
            exit(rc_ < 0 ? -rc_ : rc_);                                                               \
        }                                                                                             \
    } while (0)

void ElpLinScheduleFromEnv(SimProperties* sp, GPUDeviceState* dev) {
    dev->tileT = dev->tileU = dev->tileV = ELP_LIN_TILE;
    dev->sublattice = ELP_SUBLAT_CUBIC;
    dev->nIter = (int)sp->nSteps;
    const char* s;
    int a, b, c;
    if ((s = getenv("ELP_TILE")) && *s) {
        if (sscanf(s, "%dx%dx%d", &a, &b, &c) == 3) { dev->tileT = a; dev->tileU = b; dev->tileV = c; }
        else if (sscanf(s, "%d", &a) == 1) dev->tileT = dev->tileU = dev->tileV = a;
    }
    if ((s = getenv("ELP_SUBLATTICE")) && *s) dev->sublattice = !strcmp(s, "dense") ? ELP_SUBLAT_DENSE : ELP_SUBLAT_CUBIC;
    if ((s = getenv("ELP_NITER")) && *s) dev->nIter = atoi(s);
}

/* gpupol/gpupol_cuda.cu:6-27 (one GPU per process) */
int GPULibInit(SimProperties* sp, GPUDeviceState* devStates, GPULibContext* ctx) {
    (void)devStates;
    sp->nDevices = 1;
    ctx->devIds[0] = 0;
    return 0;
}

/* gpupol/gpupol_cuda.cu:29-48 */
int GPULibLoadBuffers(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx) {
    for (int d = 0; d < sp->nDevices; d++) {
        GPUDeviceState* dev = devStates + d;
        ElpLinScheduleFromEnv(sp, dev);
        elp_engine* e = NULL;
        ELPCHK(elp_lin_create(&e, ctx->devIds[d], (int)sp->LT, (int)sp->LU, (int)sp->LV, dev->tileT, dev->tileU,
                              dev->tileV, dev->sublattice, dev->nIter));
        dev->engine = e;
        dev->devId = ctx->devIds[d];
        ELPCHK(elp_upload(e, ss[d].lattice, sizeof(uint16_t) * (uint64_t)sp->L));
    }
    return 0;
}

/* gpupol/gpupol_cuda.cu:50-60 */
int GPULibRelease(SimProperties* sp, GPUDeviceState* devStates, GPULibContext* ctx) {
    (void)ctx;
    for (int d = 0; d < sp->nDevices; d++) { elp_destroy((elp_engine*)devStates[d].engine); devStates[d].engine = NULL; }
    return 0;
}

/* gpupol/gpupol_cuda.cu:62-82 */
int GPULatticeToCPU(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx) {
    (void)ctx;
    for (int d = 0; d < sp->nDevices; d++)
        ELPCHK(elp_download((elp_engine*)devStates[d].engine, ss[d].lattice, sizeof(uint16_t) * (uint64_t)sp->L));
    if (GetAllRingPolymers(sp, ss)) exit(193);
    for (int d = 0; d < sp->nDevices; d++)
        if (UpdatePolymerWithLabels(ss + d)) exit(123);
    return 0;
}

/* gpupol/gpupol_cuda.cu:84-112 */
static void run_units(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, int nTime) {
    int32_t* org = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)(nTime > 0 ? nTime : 1));
    for (int d = 0; d < sp->nDevices; d++) {
        GPUDeviceState* dev = devStates + d;
        for (int i = 0; i < nTime; i++) {
            org[3 * i + 0] = elp_draw_shift(&ss[d].rngState, dev->tileT);
            org[3 * i + 1] = elp_draw_shift(&ss[d].rngState, dev->tileU);
            org[3 * i + 2] = elp_draw_shift(&ss[d].rngState, dev->tileV);
        }
        elp_engine* e = (elp_engine*)dev->engine;
        if (nTime > 0) {
            ELPCHK(elp_run(e, nTime, org, (uint64_t)sp->curT, sp->seed, (uint32_t)d));
            ELPCHK(elp_sync(e));
            float ms = 0;
            if (!elp_last_run_ms(e, &ms)) dev->lastKernelMs = ms;
        }
    }
    free(org);
}

int GPULibRun(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx, int nTime) {
    run_units(sp, ss, devStates, nTime);
    GPULatticeToCPU(sp, ss, devStates, ctx);
    return 0;
}

/* As GPULibRun without the copy back and host recovery: the caller writes
   the tuv file from the device (WriteLatticeFileDevice). */
int GPULibRunDevice(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx, int nTime) {
    (void)ctx;
    run_units(sp, ss, devStates, nTime);
    return 0;
}

/* Device-side chain recovery and tuv text unless ELP_GPU_TUV=0. */
int ElpGpuTuvEnabled(void) {
    const char* env = getenv("ELP_GPU_TUV");
    return !(env && env[0] == '0');
}

int WriteLatticeFileDevice(SimProperties* sp, SimState* ss, GPUDeviceState* dev, const char* file) {
    elp_engine* e = (elp_engine*)dev->engine;
    uint64_t len = 0;
    const int maxLen = sp->polLength > 0 ? sp->polLength : (int)sp->maxPolLength;
    ELPCHK(elp_lin_tuv(e, ss->label2Id, ss->nPol, maxLen, (int)sp->maxPolLength, NULL, 0, &len));
    char* text = (char*)malloc(len ? len : 1);
    ELPCHK(elp_lin_tuv(e, ss->label2Id, ss->nPol, maxLen, (int)sp->maxPolLength, text, len, &len));
    FILE* f = fopen(file, "w");
    if (!f) { printf("Error opening file %s\n", file); free(text); return -1; }
    fwrite(text, 1, len, f);
    fclose(f);
    free(text);
    return 0;
}

/* ---------------- handle API (bench.py, tests, other drivers) ---------------- */
struct elp_lin_host { SimProperties sp; SimState ss; };

/* Fresh start with box edges truncated to multiples of box_unit (0: 32, as the
   reference; a smaller unit, e.g. 16, allows boxes such as 48 for engines with
   16^3 tiles, with a seed table sized by that unit). */
elp_lin_host* elp_lin_host_new_unit(int N, int LT, int LU, int LV, double density, uint32_t seed, int box_unit) {
    if (box_unit < 0 || (box_unit && box_unit % 4)) return NULL;
    elp_lin_host* h = (elp_lin_host*)calloc(1, sizeof *h);
    SimPropDefaultInit(&h->sp);
    h->sp.polLength = N; h->sp.density = density; h->sp.seed = seed; h->sp.nDevices = 1;
    SetBoxDimensionUnit(&h->sp, LT, LU, LV, box_unit ? box_unit : ELP_LIN_TILE);
    if (!h->sp.L) { elp_lin_host_free(h); return NULL; }
    SimStateInit(&h->ss, &h->sp);
    if (GetBestPolymerPartitioning(&h->sp, &h->ss, N, density)) { elp_lin_host_free(h); return NULL; }
    RedistribSL(&h->ss, &h->sp);
    if (GetAllRingPolymers(&h->sp, &h->ss) || UpdatePolymerWithLabels(&h->ss)) { elp_lin_host_free(h); return NULL; }
    return h;
}

elp_lin_host* elp_lin_host_new(int N, int LT, int LU, int LV, double density, uint32_t seed) {
    return elp_lin_host_new_unit(N, LT, LU, LV, density, seed, 0);
}

/* Fresh start whose chains are random self-avoiding walks covering about
   site_fraction of the box (lin_saw.c); site_fraction <= 0 takes
   ELP_SAW_SITE_FRACTION.  Stored length is then spread by RedistribSL. */
elp_lin_host* elp_lin_host_new_saw(int N, int LT, int LU, int LV, double density, uint32_t seed, int box_unit,
                                   double site_fraction) {
    if (box_unit < 0 || (box_unit && box_unit % 4)) return NULL;
    elp_lin_host* h = (elp_lin_host*)calloc(1, sizeof *h);
    SimPropDefaultInit(&h->sp);
    h->sp.polLength = N; h->sp.density = density; h->sp.seed = seed; h->sp.nDevices = 1;
    SetBoxDimensionUnit(&h->sp, LT, LU, LV, box_unit ? box_unit : ELP_LIN_TILE);
    if (!h->sp.L) { elp_lin_host_free(h); return NULL; }
    SimStateInit(&h->ss, &h->sp);
    const double frac = site_fraction > 0 ? site_fraction : ELP_SAW_SITE_FRACTION;
    if (BuildSAWPolymers(&h->sp, &h->ss, N, density, frac)) { elp_lin_host_free(h); return NULL; }
    RedistribSL(&h->ss, &h->sp);
    if (GetAllRingPolymers(&h->sp, &h->ss) || UpdatePolymerWithLabels(&h->ss)) { elp_lin_host_free(h); return NULL; }
    return h;
}

/* Start from a tuv file whose box edges are multiples of box_unit (a multiple
   of 4); NULL on a bad or unreadable file. */
elp_lin_host* elp_lin_host_from_file_unit(const char* path, uint32_t seed, int box_unit) {
    if (box_unit <= 0 || box_unit % 4) return NULL;
    elp_lin_host* h = (elp_lin_host*)calloc(1, sizeof *h);
    SimPropDefaultInit(&h->sp);
    h->sp.seed = seed; h->sp.nDevices = 1;
    if (ReadLatticeFileUnit(&h->ss, &h->sp, path, box_unit) || UpdatePolymerWithLabels(&h->ss) ||
        GetAllRingPolymers(&h->sp, &h->ss) || UpdatePolymerWithLabels(&h->ss) || !h->ss.nPol) {
        elp_lin_host_free(h);
        return NULL;
    }
    return h;
}

/* elp_lin_host_from_file_unit with the largest unit of 32, 16, 8, 4 dividing
   the file's box edges (32 for every box the reference writes) */
elp_lin_host* elp_lin_host_from_file(const char* path, uint32_t seed) {
    int L[3] = {0, 0, 0};
    FILE* f = fopen(path, "r");
    if (!f) return NULL;
    int ok = fscanf(f, "%*s %i", L) == 1 && fscanf(f, "%*s %i", L + 1) == 1 && fscanf(f, "%*s %i", L + 2) == 1;
    fclose(f);
    if (!ok) return NULL;
    for (int unit = ELP_LIN_TILE; unit >= 4; unit /= 2)
        if (L[0] > 0 && L[1] > 0 && L[2] > 0 && !(L[0] % unit) && !(L[1] % unit) && !(L[2] % unit))
            return elp_lin_host_from_file_unit(path, seed, unit);
    return NULL;
}

void elp_lin_host_free(elp_lin_host* h) {
    if (!h) return;
    SimStateFree(&h->ss);
    free(h->sp.trans);
    free(h);
}

void elp_lin_host_info(const elp_lin_host* h, int* LT, int* LU, int* LV, int* nPol, int* N, int* maxPolLength) {
    if (LT) *LT = (int)h->sp.LT;
    if (LU) *LU = (int)h->sp.LU;
    if (LV) *LV = (int)h->sp.LV;
    if (nPol) *nPol = h->ss.nPol;
    if (N) *N = h->ss.nPol ? h->ss.pol[0].length : 0;
    if (maxPolLength) *maxPolLength = (int)h->sp.maxPolLength;
}

int elp_lin_host_get_lattice(const elp_lin_host* h, uint16_t* out, uint64_t nbytes) {
    if (nbytes != sizeof(uint16_t) * (uint64_t)h->sp.L) return -1;
    memcpy(out, h->ss.lattice, nbytes);
    return 0;
}

int elp_lin_host_set_lattice(elp_lin_host* h, const uint16_t* in, uint64_t nbytes) {
    if (nbytes != sizeof(uint16_t) * (uint64_t)h->sp.L) return -1;
    memcpy(h->ss.lattice, in, nbytes);
    if (GetAllRingPolymers(&h->sp, &h->ss)) return -1;
    return UpdatePolymerWithLabels(&h->ss);
}

int elp_lin_host_write_tuv(elp_lin_host* h, const char* path) { return WriteLatticeFile(&h->sp, &h->ss, path); }

int elp_lin_host_polymer(const elp_lin_host* h, int id, uint32_t* bonds, int cap, uint32_t* start3) {
    for (int p = 0; p < h->ss.nPol; p++) {
        if (h->ss.polNumTranslate[p] != id) continue;
        const Polymer* pol = h->ss.pol + p;
        if (pol->length > cap) return -1;
        for (int i = 0; i < pol->length; i++) bonds[i] = pol->bonds[i];
        start3[0] = pol->startTUV.t; start3[1] = pol->startTUV.u; start3[2] = pol->startTUV.v;
        return pol->length;
    }
    return -1;
}

int elp_lin_host_label2id(const elp_lin_host* h, int32_t* out, uint64_t n) {
    if (n != ((uint64_t)1 << h->ss.labelBit) || !h->ss.label2Id) return -1;
    for (uint64_t i = 0; i < n; i++) out[i] = h->ss.label2Id[i];
    return 0;
}

/* the linear driver draws no load-time shift */
void elp_lin_host_load_shift(elp_lin_host* h, int tt, int tu, int tv) { (void)h; (void)tt; (void)tu; (void)tv; }

void elp_lin_host_origins(elp_lin_host* h, int tt, int tu, int tv, int n, int32_t* out) {
    for (int i = 0; i < n; i++) {
        out[3 * i + 0] = elp_draw_shift(&h->ss.rngState, tt);
        out[3 * i + 1] = elp_draw_shift(&h->ss.rngState, tu);
        out[3 * i + 2] = elp_draw_shift(&h->ss.rngState, tv);
    }
}

uint32_t elp_lin_host_seed(const elp_lin_host* h) { return h->sp.seed; }

/* per-lane seed table of SimStateInit (gpupol/gpupol_init.c:23-34); see
   elp_ring_host_seeds */
int elp_lin_host_seeds(const elp_lin_host* h, uint32_t* out, uint64_t n) {
    uint64_t need = (uint64_t)h->sp.nwg * h->sp.ws * 2 * h->sp.R;
    if (n != need) return -1;
    memcpy(out, h->ss.seeds, need * sizeof(uint32_t));
    return 0;
}
