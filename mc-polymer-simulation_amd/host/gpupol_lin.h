/*
 * gpupol_lin.h -- host model/state of the linear-chain drop-in
 * (counterpart of gpupol/gpupol_def.h, gpupol_init.h, oclpol.h, gpupol_io.h
 * and gpupol_cuda.h for bin/gpupol_cuda_lin).
 *
 * Reference structs and field names are kept so the GPULib* entry points keep
 * their signatures (gpupol/gpupol_cuda.h:6-9).  Internal difference: the
 * host lattice is one uint16 word per site, row-major (t + LT*(u + LU*v)),
 * in the device format (bits 0-3 next bond, 4-7 prev bond, 8 stored length,
 * 9 label of the bond carrier, 10 label of the stored-length monomer) instead
 * of the reference's three packed arrays.
 */
#ifndef GPUPOL_LIN_H
#define GPUPOL_LIN_H
#include <stdio.h>
#include <stdint.h>
#include "elp_host_common.h"

#define ELP_LIN_TILE 32
#define ELP_N_MAX_GPU 16

#define LW_NEXT(w) ((uint)(w) & 0xfu)
#define LW_PREV(w) (((uint)(w) >> 4) & 0xfu)
#define LW_SL(w) (((uint)(w) >> 8) & 1u)
#define LW_LABELS(w) (((uint)(w) >> 9) & 3u)

typedef struct TUVCoor { uint t, u, v; } TUVCoor;

/* gpupol/gpupol_def.h:23-33 (polCoor dropped: coordinates are not kept) */
typedef struct Polymer {
    uint* bonds;          /* per monomer: bond to the next monomer, 0 = stored length, 0xf = chain end */
    TUVCoor startTUV;
    int startLeftSL;
    uint* label;
    uint labelStart;
    int length;
} Polymer;

/* gpupol/gpupol_def.h:35-47 */
typedef struct SimState {
    uint16_t* lattice;
    uint16_t* gpuLattice;
    uint* seeds;
    int nPol;
    Polymer* pol;
    int* label2Id;
    int* id2Label;
    int* polNumTranslate;
    int labelHead, headBit, labelBit;
    uint4_t rngState;
} SimState;

/* gpupol/gpupol_def.h:50-77 */
typedef struct SimProperties {
    size_t nwg;
    size_t ws;
    long tMax;
    long curT;
    uint R;
    long nSteps;
    long writeInterval;
    uint nwt, nwu, nwv;
    uint LT, LU, LV, L;
    uint latSize;
    int equilibrated;
    int double_step;
    uint seed;
    uint gpuLatSize;
    int polLength;
    long fastEq;
    char* dirOut;
    int nDevices;
    uint* trans;
    uint maxPolLength;
    double density;
    int boxUnit;       /* box edges are truncated to multiples of this; 0 = the reference's 32 */
    double sawSiteFraction; /* > 0: fresh starts grow self-avoiding walks covering this site fraction (lin_saw.c) */
} SimProperties;

typedef struct GPULibContext { int devIds[8]; } GPULibContext;
typedef struct GPUDeviceState {
    int devId;
    void* engine;
    int tileT, tileU, tileV, sublattice, nIter;
    double lastKernelMs;
} GPUDeviceState;

#ifdef __cplusplus
extern "C" {
#endif
int  SimPropDefaultInit(SimProperties* sp);
void SetBoxDimension(SimProperties* sp, int LT, int LU, int LV);
int  SimStateInit(SimState* ss, SimProperties* sp);
void SimStateFree(SimState* ss);
void GetBestPolymerPartitioning(SimProperties* sp, SimState* ss, int polLength, double density);
void AddPolymerToSim(SimState* ss, SimProperties* sp, uint t, uint u, uint v, int length, uint* bonds);
void RedistribSL(SimState* ss, SimProperties* sp);
#define ELP_SAW_SITE_FRACTION 0.7
int  BuildSAWPolymers(SimProperties* sp, SimState* ss, int polLength, double density, double siteFraction);
void SetBondVecs(SimState* ss, SimProperties* sp, uint t, uint u, uint v, uint* bonds, uint polSize, uint* labels);
int  GetAllRingPolymers(SimProperties* sp, SimState* ss);
int  UpdatePolymerWithLabels(SimState* ss);
int  ReadLatticeFile(SimState* ss, SimProperties* sp, const char* file);
int  WriteLatticeFile(SimProperties* sp, SimState* ss, const char* file);
void WriteSimulationSettings(SimProperties* sp, SimState* ss);
void LoadPolymers(SimProperties* sp, SimState* ss, GPUDeviceState* devices, GPULibContext* ctx);
void AddUnitTUV(uint unit, uint* t, uint* u, uint* v, const SimProperties* sp);
uint SiteIndex(uint t, uint u, uint v, const SimProperties* sp);

int GPULibInit(SimProperties* sp, GPUDeviceState* devStates, GPULibContext* ctx);
int GPULibLoadBuffers(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx);
int GPULibRelease(SimProperties* sp, GPUDeviceState* devStates, GPULibContext* ctx);
int GPULibRun(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx, int nTime);
/* GPULibRun without the copy back; the tuv file is then written from the
   device (device-side chain recovery, unless ELP_GPU_TUV=0). */
int GPULibRunDevice(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx, int nTime);
int ElpGpuTuvEnabled(void);
int WriteLatticeFileDevice(SimProperties* sp, SimState* ss, GPUDeviceState* dev, const char* file);
int GPULatticeToCPU(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx);
void ElpLinScheduleFromEnv(SimProperties* sp, GPUDeviceState* dev);
#ifdef __cplusplus
}
#endif
#endif
