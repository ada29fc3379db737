/*
 * gpupol_lin.h -- host model and GPU boundary of the linear-chain drop-in
 * (libgpupol_lin.so): counterpart of gpupol/gpupol_def.h, gpupol_init.h,
 * oclpol.h, gpupol_io.h and gpupol_cuda.h for bin/gpupol_cuda_lin.
 *
 * SimProperties, SimState and Polymer are byte-identical to
 * gpupol/gpupol_def.h:23-78 (tests/test_cseam.py compares sizeof and every
 * offsetof), so a reference host program compiled against that header and
 * linked with this library shares them; the device state is the backend's own
 * (gpupol_cuda_def.h in this directory).  Internal differences: the host
 * lattice is one uint16 word per site, row-major (t + LT*(u + LU*v)), in the
 * device format (bits 0-3 next bond, 4-7 prev bond, 8 stored length, 9 label
 * of the bond carrier, 10 label of the stored-length monomer) instead of the
 * reference's three packed arrays, so slLattice, labLattice and polCoor stay
 * NULL.
 */
#ifndef GPUPOL_LIN_H
#define GPUPOL_LIN_H
#include <stdio.h>
#include <stdint.h>
#include <stddef.h>
#include "gpupol_cuda_def.h"

#ifndef UINT_DEF
#define UINT_DEF
typedef unsigned int uint;
#endif
#ifndef ELP_UINT4_T_DEF
#define ELP_UINT4_T_DEF
typedef struct uint4_t { uint x, y, z, w; } uint4_t;     /* gpupol/gpupol_def.h:8-10 */
#endif

#define ELP_LIN_TILE 32
#define ELP_N_MAX_GPU 16

#define LW_NEXT(w) ((uint)(w) & 0xfu)
#define LW_PREV(w) (((uint)(w) >> 4) & 0xfu)
#define LW_SL(w) (((uint)(w) >> 8) & 1u)
#define LW_LABELS(w) (((uint)(w) >> 9) & 3u)

typedef struct TUVCoor { uint t, u, v; } TUVCoor;

/* gpupol/gpupol_def.h:23-33 */
typedef struct Polymer {
    uint* bonds;          /* per monomer: bond to the next monomer, 0 = stored length, 0xf = chain end */
    struct Coor* polCoor; /* not kept (NULL) */
    TUVCoor startTUV;
    int startLeftSL;
    uint* label;
    uint labelStart;
    int length;
} Polymer;

/* gpupol/gpupol_def.h:35-47 */
typedef struct SimState {
    uint16_t* lattice;    /* row-major site words */
    uint16_t* gpuLattice; /* download buffer */
    uint* slLattice;      /* not used (NULL) */
    uint* labLattice;     /* not used (NULL) */
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
    uint slSize, labSize;
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
} SimProperties;

#ifdef __cplusplus
extern "C" {
#endif
/* gpupol/gpupol_init.h:7 */
int  SimPropDefaultInit(SimProperties* sp);
/* gpupol/gpupol_init.h:17: edges truncated to multiples of 32, the tile edge */
void SetBoxDimension(SimProperties* sp, int LT, int LU, int LV);
/* the same with another box unit (a multiple of 4); no reference counterpart */
void SetBoxDimensionUnit(SimProperties* sp, int LT, int LU, int LV, int unit);
int  SimStateInit(SimState* ss, SimProperties* sp);
void SimStateFree(SimState* ss);
/* gpupol/gpupol_init.h:8 (void there); returns 0, or 192 / 120 when the
   chains do not fit the box, after printing the reference's message */
int  GetBestPolymerPartitioning(SimProperties* sp, SimState* ss, int polLength, double density);
/* gpupol/gpupol_init.h:11 (void there); returns 0, or 1 when a site is taken */
int  AddPolymerToSim(SimState* ss, SimProperties* sp, uint t, uint u, uint v, int length, uint* bonds);
/* gpupol/gpupol_init.h:15 */
void RedistribSL(SimState* ss, SimProperties* sp);
#define ELP_SAW_SITE_FRACTION 0.7
int  BuildSAWPolymers(SimProperties* sp, SimState* ss, int polLength, double density, double siteFraction);
/* returns 0, or 1 when a site is already occupied */
int  SetBondVecs(SimState* ss, SimProperties* sp, uint t, uint u, uint v, uint* bonds, uint polSize, uint* labels);
int  GetAllRingPolymers(SimProperties* sp, SimState* ss);
int  UpdatePolymerWithLabels(SimState* ss);
/* gpupol/gpupol_io.h:7; returns 0, 199 on a bad file, 202 when a box edge is
   not a multiple of 32 */
int  ReadLatticeFile(SimState* ss, SimProperties* sp, const char* file);
int  ReadLatticeFileUnit(SimState* ss, SimProperties* sp, const char* file, int unit);
/* gpupol/gpupol_io.h:8-10 */
int  WriteLatticeFile(SimProperties* sp, SimState* ss, const char* file);
void WriteSimulationSettings(SimProperties* sp, SimState* ss);
/* gpupol/gpupol_init.h:5, gpupol/gpupol_init.c:86-150: resume from the highest
   t= file or build the reference start (ELP_LIN_INIT=saw[:FRACTION]: random
   self-avoiding walks instead); exits on failure, as the reference does */
void LoadPolymers(SimProperties* sp, SimState* ss);
void AddUnitTUV(uint unit, uint* t, uint* u, uint* v, const SimProperties* sp);
uint SiteIndex(uint t, uint u, uint v, const SimProperties* sp);

/* ---- GPU boundary with the reference signatures (gpupol/gpupol_cuda.h:6-9) ---- */
/* gpupol/gpupol_cuda.h:6 -- one device per process */
int GPULibInit(SimProperties* sp, GPUDeviceState* devStates, GPULibContext* ctx);
/* gpupol/gpupol_cuda.h:7 -- create the engine and upload the lattice */
int GPULibLoadBuffers(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx);
/* gpupol/gpupol_cuda.h:8 */
int GPULibRelease(SimProperties* sp, GPUDeviceState* devStates, GPULibContext* ctx);
/* gpupol/gpupol_cuda.h:9 -- nTime time units (3 host draws each), then the
   lattice and the chains back to the host */
int GPULibRun(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx, int nTime);
/* GPULibRun without the copy back; the tuv file is then written from the
   device (device-side chain recovery, unless ELP_GPU_TUV=0). */
int GPULibRunDevice(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx, int nTime);
int ElpGpuTuvEnabled(void);
int WriteLatticeFileDevice(SimProperties* sp, SimState* ss, GPUDeviceState* dev, const char* file);
int GPULatticeToCPU(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx);
void ElpLinScheduleFromEnv(SimProperties* sp, GPUDeviceState* dev);

/* ---- handle API over the host model (no reference counterpart) ---- */
typedef struct elp_lin_host elp_lin_host;
/* fresh start, box edges truncated to multiples of box_unit (0: 32); NULL when
   the chains do not fit */
elp_lin_host* elp_lin_host_new_unit(int N, int LT, int LU, int LV, double density, uint32_t seed, int box_unit);
elp_lin_host* elp_lin_host_new(int N, int LT, int LU, int LV, double density, uint32_t seed);
/* fresh start of random self-avoiding walks covering site_fraction of the box
   (<= 0: ELP_SAW_SITE_FRACTION) */
elp_lin_host* elp_lin_host_new_saw(int N, int LT, int LU, int LV, double density, uint32_t seed, int box_unit,
                                   double site_fraction);
/* start from a tuv file (box edges multiples of box_unit); NULL on a bad file */
elp_lin_host* elp_lin_host_from_file_unit(const char* path, uint32_t seed, int box_unit);
/* the same with the largest of 32, 16, 8, 4 dividing the file's box */
elp_lin_host* elp_lin_host_from_file(const char* path, uint32_t seed);
void elp_lin_host_free(elp_lin_host* h);
void elp_lin_host_info(const elp_lin_host* h, int* LT, int* LU, int* LV, int* nPol, int* N, int* maxPolLength);
int  elp_lin_host_get_lattice(const elp_lin_host* h, uint16_t* out, uint64_t nbytes);
int  elp_lin_host_set_lattice(elp_lin_host* h, const uint16_t* in, uint64_t nbytes);
int  elp_lin_host_write_tuv(elp_lin_host* h, const char* path);
int  elp_lin_host_polymer(const elp_lin_host* h, int id, uint32_t* bonds, int cap, uint32_t* start3);
void elp_lin_host_load_shift(elp_lin_host* h, int tt, int tu, int tv);
void elp_lin_host_origins(elp_lin_host* h, int tt, int tu, int tv, int n, int32_t* out);
uint32_t elp_lin_host_seed(const elp_lin_host* h);
int  elp_lin_host_label2id(const elp_lin_host* h, int32_t* out, uint64_t n);
int  elp_lin_host_seeds(const elp_lin_host* h, uint32_t* out, uint64_t n);
#ifdef __cplusplus
}
#endif
#endif
