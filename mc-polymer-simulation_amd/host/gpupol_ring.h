/*
 * gpupol_ring.h -- host model/state of the ring-polymer drop-in
 * (counterpart of gpupol2/gpupol_def.h, gpupol_init.h, oclpol.h, gpupol_io.h
 * and gpupol_cuda.h for the gpupol3 ring executable).
 *
 * The structs keep the reference's names and fields so that the GPULib*
 * entry points keep their reference signatures (gpupol2/gpupol_cuda.h:6-12).
 * Differences, all internal: SimState.lattice is row-major (t fastest,
 * t + LT*(u + LU*v)) instead of the reference tile-major GetGpuSite layout,
 * and GPUDeviceState holds an opaque engine handle instead of raw buffers.
 */
#ifndef GPUPOL_RING_H
#define GPUPOL_RING_H
#include <stdio.h>
#include <stdint.h>
#include "elp_host_common.h"

#define ELP_RING_LCELL 4            /* gpupol3: LCELL=4 (gpupol2/Makefile:26) */
#define ELP_RING_WS_EDGE 8          /* cells per tile edge (gpupol2/gpupol.h:3-5) */
#define ELP_RING_TILE (ELP_RING_LCELL * ELP_RING_WS_EDGE)   /* 32 */
#define ELP_N_MAX_GPU 16

typedef struct TUVCoor { uint t, u, v; } TUVCoor;

/* gpupol2/gpupol_def.h:23-32 */
typedef struct Polymer {
    uint* bonds;          /* per monomer: bond code to the next monomer, 0 = same site */
    TUVCoor startTUV;
    int startLeftSL;
    uint* label;          /* per monomer label bit */
    uint labelStart;
    int length;
} Polymer;

/* gpupol2/gpupol_def.h:34-47 */
typedef struct SimState {
    uint8_t* lattice;     /* row-major site bytes (see include/elp_gpu.h) */
    uint8_t* gpuLattice;  /* download buffer */
    uint* seeds;          /* reference per-lane seed table (stream position only) */
    int nPol;
    Polymer* pol;
    int* label2Id;
    int* id2Label;
    int* polNumTranslate;
    int labelHead;
    int headBit;
    int labelBit;
    uint4_t rngState;
} SimState;

/* gpupol2/gpupol_def.h:50-76 */
typedef struct SimProperties {
    size_t nwg;
    size_t ws;
    int double_step;
    long tMax;
    long curT;
    uint R;
    long nSteps;
    long writeInterval;
    uint nwt, nwu, nwv;
    uint LT, LU, LV, L;
    uint latSize;
    uint seed;
    int polLength;
    long fastEq;
    int equilibrated;
    char* dirOut;
    int nDevices;
    uint* trans;
    uint maxPolLength;
    double density;
    int boxUnit;       /* box edges are truncated to multiples of this; 0 = the reference's 32 */
} SimProperties;

/* gpupol2/gpupol_cuda_def.h:8-20 (device buffers replaced by an engine) */
typedef struct GPULibContext { int devIds[8]; } GPULibContext;
typedef struct GPUDeviceState {
    int devId;
    void* engine;                 /* elp_engine* */
    int curBuf;
    long cumDt, cumDu, cumDv;     /* cumulative tile-grid shift */
    int dt, du, dv;               /* last drawn shift */
    /* schedule of the engine (defaults reproduce ringpol3: 32^3 tiles,
       cubic sub-lattice, nSteps iterations) */
    int tileT, tileU, tileV, sublattice, nIter;
    double lastKernelMs;
} GPUDeviceState;

#ifdef __cplusplus
extern "C" {
#endif

/* ---- model / state (gpupol2/gpupol_init.c) ---- */
int  SimPropDefaultInit(SimProperties* sp);
void SetBoxDimension(SimProperties* sp, int LT, int LU, int LV);
int  SimStateInit(SimState* ss, SimProperties* sp);
void ClearLattice(SimState* ss, SimProperties* sp);
void GetBestPolymerPartitioning(SimProperties* sp, SimState* ss, int polLength, double density);
void ConstructCubeRingPolymer(SimState* ss, SimProperties* sp, uint tStart, uint uStart, uint vStart,
                              uint dt, uint du, uint dv, uint sl);
void AddPolymerToSim(SimState* ss, SimProperties* sp, uint t, uint u, uint v, int length, uint* bonds);
int  UniqueLabel(int startLabel, int nLabelBit, int head, int headBit);
void RedistribSL(SimState* ss, SimProperties* sp);
int  SetGPUTrans(uint* trans);
void LoadPolymers(SimProperties* sp, SimState* ss, GPUDeviceState* devices, GPULibContext* ctx);
void SimStateFree(SimState* ss);

/* ---- lattice encoding + chain recovery (gpupol2/oclpol.c) ---- */
uint SiteIndex(uint t, uint u, uint v, const SimProperties* sp);
void AddUnitTUV(uint unit, uint* t, uint* u, uint* v, const SimProperties* sp);
void SetBondVecs(uint8_t* lattice, uint t, uint u, uint v, uint* bonds, uint polSize, uint* labels,
                 SimProperties* sp);
void SetLatticeFromSS(SimState* ss, SimProperties* sp);
int  GetAllRingPolymers(SimProperties* sp, SimState* ss);
int  UpdatePolymerWithLabels(SimState* ss);

/* ---- tuv + settings IO (gpupol2/gpupol_io.c) ---- */
int  ReadLatticeFile(SimState* ss, SimProperties* sp, const char* file);
int  WriteLatticeFile(SimProperties* sp, SimState* ss, const char* file);
void WriteRingPolymer(FILE* pFile, Polymer* pol);
void WriteSimulationSettings(SimProperties* sp, SimState* ss);

/* ---- GPU boundary with the reference signatures (gpupol2/gpupol_cuda.h:6-12) ---- */
int GPULibInit(SimProperties* sp, GPUDeviceState* devStates, GPULibContext* ctx);
int GPULibLoadBuffers(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx);
int GPULibRelease(SimProperties* sp, GPUDeviceState* devStates, GPULibContext* ctx);
int GPULibRun(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx, int nTime);
int GPULoadLattices(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx);
int GPULatticeToCPU(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx);

/* tuv file of a device's lattice, written from the device-side ring recovery
   (elp_ring_tuv); same bytes as GPULatticeToCPU + WriteLatticeFile.  Used by
   the drop-in unless ELP_GPU_TUV=0. */
int WriteLatticeFileDevice(SimProperties* sp, SimState* ss, GPUDeviceState* dev, const char* file);
int ElpGpuTuvEnabled(void);
int GPULibRunDevice(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx, int nTime);
int GPULibEqRun(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx, int nTime);

/* Schedule of the device engine, read from the environment:
   ELP_TILE ("32" or "16" or "32x32x16"), ELP_SUBLATTICE ("cubic"|"dense"),
   ELP_NITER (iterations per time unit).  Defaults reproduce ringpol3. */
void ElpRingScheduleFromEnv(SimProperties* sp, GPUDeviceState* dev);

#ifdef __cplusplus
}
#endif
#endif
