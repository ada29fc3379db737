/*
 * gpupol_ring.h -- host model and GPU boundary of the ring-polymer drop-in
 * (libgpupol3_ring.so): counterpart of gpupol2/gpupol_def.h, gpupol_init.h,
 * oclpol.h, gpupol_io.h and gpupol_cuda.h for the gpupol3 ring executable.
 *
 * SimProperties, SimState and Polymer are byte-identical to
 * gpupol2/gpupol_def.h:23-76 (same fields, types of the same size, same
 * order), so a reference host program compiled against that header and
 * linked with this library shares them (tests/test_cseam.py compares sizeof
 * and every offsetof).  The device state is the backend's own
 * (gpupol_cuda_def.h in this directory, replacing the reference's CUDA one).
 * Internal differences: SimState.lattice holds one byte per site in
 * row-major order (t + LT*(u + LU*v)) instead of the reference's tile-major
 * GetGpuSite layout, and the seed table is drawn only for the stream
 * position of the host generator (the engine's random numbers are
 * counter-based).
 */
#ifndef GPUPOL_RING_H
#define GPUPOL_RING_H
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
typedef struct uint4_t { uint x, y, z, w; } uint4_t;     /* gpupol2/gpupol_def.h:8-10 */
#endif

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
    uint8_t* lattice;     /* row-major site bytes (site encoding in elp_gpu.h) */
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
} SimProperties;

#ifdef __cplusplus
extern "C" {
#endif

/* ---- model / state (gpupol2/gpupol_init.c) ---- */
/* gpupol2/gpupol_init.h:7 */
int  SimPropDefaultInit(SimProperties* sp);
/* gpupol2/gpupol_init.h:16: edges truncated to multiples of 32, the tile edge */
void SetBoxDimension(SimProperties* sp, int LT, int LU, int LV);
/* the same with another box unit (a multiple of 4; e.g. 16 allows boxes of 48
   for engines with 16^3 tiles); no reference counterpart */
void SetBoxDimensionUnit(SimProperties* sp, int LT, int LU, int LV, int unit);
int  SimStateInit(SimState* ss, SimProperties* sp);
void ClearLattice(SimState* ss, SimProperties* sp);
/* gpupol2/gpupol_init.h:8 (void there); returns 0, or 192 when the rings do
   not fit the box, after printing the reference's message */
int  GetBestPolymerPartitioning(SimProperties* sp, SimState* ss, int polLength, double density);
int  ConstructCubeRingPolymer(SimState* ss, SimProperties* sp, uint tStart, uint uStart, uint vStart,
                              uint dt, uint du, uint dv, uint sl);
/* gpupol2/gpupol_init.h:11 (void there); returns 0, or 1 when a site is
   already occupied */
int  AddPolymerToSim(SimState* ss, SimProperties* sp, uint t, uint u, uint v, int length, uint* bonds);
int  UniqueLabel(int startLabel, int nLabelBit, int head, int headBit);
void RedistribSL(SimState* ss, SimProperties* sp);
int  SetGPUTrans(uint* trans);
/* gpupol2/gpupol_init.h:5, gpupol2/gpupol_init.c:83-152: resume from the
   highest t=<T>_dev=<i>.res files in sp->dirOut or build the reference start;
   exits on failure, as the reference does */
void LoadPolymers(SimProperties* sp, SimState* ss);
void SimStateFree(SimState* ss);

/* ---- lattice encoding + chain recovery (gpupol2/oclpol.c) ---- */
uint SiteIndex(uint t, uint u, uint v, const SimProperties* sp);
void AddUnitTUV(uint unit, uint* t, uint* u, uint* v, const SimProperties* sp);
/* returns 0, or 1 when a site is already occupied */
int  SetBondVecs(uint8_t* lattice, uint t, uint u, uint v, uint* bonds, uint polSize, uint* labels,
                 SimProperties* sp);
void SetLatticeFromSS(SimState* ss, SimProperties* sp);
int  GetAllRingPolymers(SimProperties* sp, SimState* ss);
int  UpdatePolymerWithLabels(SimState* ss);

/* ---- tuv + settings IO (gpupol2/gpupol_io.c) ---- */
/* gpupol2/gpupol_io.h:7; returns 0, 199 on a bad file, 202 when a box edge
   is not a multiple of 32 */
int  ReadLatticeFile(SimState* ss, SimProperties* sp, const char* file);
/* the same with another box unit (a multiple of 4); no reference counterpart */
int  ReadLatticeFileUnit(SimState* ss, SimProperties* sp, const char* file, int unit);
int  WriteLatticeFile(SimProperties* sp, SimState* ss, const char* file);
void WriteRingPolymer(FILE* pFile, Polymer* pol);
void WriteSimulationSettings(SimProperties* sp, SimState* ss);

/* ---- GPU boundary with the reference signatures (gpupol2/gpupol_cuda.h:6-12) ---- */
/* gpupol2/gpupol_cuda.h:6 -- one device per process (sp->nDevices = 1) */
int GPULibInit(SimProperties* sp, GPUDeviceState* devStates, GPULibContext* ctx);
/* gpupol2/gpupol_cuda.h:7 -- create the engine, upload the lattice, draw the first shift */
int GPULibLoadBuffers(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx);
/* gpupol2/gpupol_cuda.h:8 */
int GPULibRelease(SimProperties* sp, GPUDeviceState* devStates, GPULibContext* ctx);
/* gpupol2/gpupol_cuda.h:9 -- nTime time units, then GPULatticeToCPU */
int GPULibRun(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx, int nTime);
/* gpupol2/gpupol_cuda.h:10 */
int GPULoadLattices(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx);
/* gpupol2/gpupol_cuda.h:11 -- lattice to ss[i].lattice, rings recovered into ss[i].pol */
int GPULatticeToCPU(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx);

/* tuv file of a device's lattice, written from the device-side ring recovery
   (elp_ring_tuv); same bytes as GPULatticeToCPU + WriteLatticeFile.  Used by
   the drop-in unless ELP_GPU_TUV=0. */
int WriteLatticeFileDevice(SimProperties* sp, SimState* ss, GPUDeviceState* dev, const char* file);
int ElpGpuTuvEnabled(void);
int GPULibRunDevice(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx, int nTime);
/* gpupol2/gpupol_cuda.h:12 -- runs of fastEq time units, stored length reshuffled between them */
int GPULibEqRun(SimProperties* sp, SimState* ss, GPUDeviceState* devStates, GPULibContext* ctx, int nTime);

/* ---- handle API over the host model (no reference counterpart), for drivers
   other than the drop-in executable; the Python package binds it ---- */
typedef struct elp_ring_host elp_ring_host;
/* fresh start: rings of N monomers at the given density and seed, box edges
   truncated to multiples of box_unit (0: 32); NULL when the rings do not fit */
elp_ring_host* elp_ring_host_new_unit(int N, int LT, int LU, int LV, double density, uint32_t seed, int box_unit);
elp_ring_host* elp_ring_host_new(int N, int LT, int LU, int LV, double density, uint32_t seed);
/* start from a tuv file (box edges multiples of box_unit); NULL on a bad file */
elp_ring_host* elp_ring_host_from_file_unit(const char* path, uint32_t seed, int box_unit);
/* the same with the largest of 32, 16, 8, 4 dividing the file's box */
elp_ring_host* elp_ring_host_from_file(const char* path, uint32_t seed);
void elp_ring_host_free(elp_ring_host* h);
void elp_ring_host_info(const elp_ring_host* h, int* LT, int* LU, int* LV, int* nPol, int* N, int* maxPolLength);
int  elp_ring_host_get_lattice(const elp_ring_host* h, uint8_t* out, uint64_t nbytes);
/* replace the lattice and recover the rings; -1 unless it holds this state's rings */
int  elp_ring_host_set_lattice(elp_ring_host* h, const uint8_t* in, uint64_t nbytes);
int  elp_ring_host_write_tuv(elp_ring_host* h, const char* path);
int  elp_ring_host_polymer(const elp_ring_host* h, int id, uint32_t* bonds, int cap, uint32_t* start3);
/* host shift stream (gpupol2/gpupol_cuda.cu:36-38,133-139) */
void elp_ring_host_load_shift(elp_ring_host* h, int tt, int tu, int tv);
void elp_ring_host_origins(elp_ring_host* h, int tt, int tu, int tv, int n, int32_t* out);
uint32_t elp_ring_host_seed(const elp_ring_host* h);
int  elp_ring_host_label2id(const elp_ring_host* h, int32_t* out, uint64_t n);
int  elp_ring_host_seeds(const elp_ring_host* h, uint32_t* out, uint64_t n);
int  elp_host_lfsr113(uint32_t state[4], uint32_t* out, int n);

/* Schedule of the device engine, read from the environment:
   ELP_TILE ("32" or "16" or "32x32x16"), ELP_SUBLATTICE ("cubic"|"dense"),
   ELP_NITER (iterations per time unit).  Defaults reproduce ringpol3. */
void ElpRingScheduleFromEnv(SimProperties* sp, GPUDeviceState* dev);

#ifdef __cplusplus
}
#endif
#endif
