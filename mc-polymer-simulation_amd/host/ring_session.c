/*
 * ring_session.c -- a handle-based C API over the ring host model, for
 * callers that are not the drop-in executable (bench.py, the tests, other
 * drivers).  It runs exactly the drop-in's start-up (SetBoxDimension,
 * SimStateInit, GetBestPolymerPartitioning, RedistribSL: gpupol2/gpupol_init.c
 * :83-152) and exposes the lattice, the host shift stream and the tuv writer.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "gpupol_ring.h"
#include "elp_host_common.h"

struct elp_ring_host {
    SimProperties sp;
    SimState ss;
    long cumT, cumU, cumV;
};

/* Fresh start (no files): box edges truncated to multiples of box_unit (0: 32,
   as the reference; a smaller unit, e.g. 16, allows boxes such as 48 for
   engines with 16^3 tiles, with a seed table sized by that unit), rings of N
   monomers at the given density and seed.  Returns NULL on failure. */
elp_ring_host* elp_ring_host_new_unit(int N, int LT, int LU, int LV, double density, uint32_t seed, int box_unit) {
    if (box_unit < 0 || (box_unit && box_unit % 4)) return NULL;
    elp_ring_host* h = (elp_ring_host*)calloc(1, sizeof *h);
    SimPropDefaultInit(&h->sp);
    h->sp.polLength = N;
    h->sp.density = density;
    h->sp.seed = seed;
    h->sp.nDevices = 1;
    SetBoxDimensionUnit(&h->sp, LT, LU, LV, box_unit ? box_unit : ELP_RING_TILE);
    if (!h->sp.L) { elp_ring_host_free(h); return NULL; }
    SimStateInit(&h->ss, &h->sp);
    if (GetBestPolymerPartitioning(&h->sp, &h->ss, N, density)) { elp_ring_host_free(h); return NULL; }
    RedistribSL(&h->ss, &h->sp);
    if (GetAllRingPolymers(&h->sp, &h->ss) || UpdatePolymerWithLabels(&h->ss)) { elp_ring_host_free(h); return NULL; }
    return h;
}

/* Fresh start with the reference's box rule (edges truncated to multiples of 32). */
elp_ring_host* elp_ring_host_new(int N, int LT, int LU, int LV, double density, uint32_t seed) {
    return elp_ring_host_new_unit(N, LT, LU, LV, density, seed, 0);
}

/* Start from a tuv file (resume path of LoadPolymers, without the seed bump);
   box edges must be multiples of box_unit (a multiple of 4).  Returns NULL on
   a bad or unreadable file. */
elp_ring_host* elp_ring_host_from_file_unit(const char* path, uint32_t seed, int box_unit) {
    if (box_unit <= 0 || box_unit % 4) return NULL;
    elp_ring_host* h = (elp_ring_host*)calloc(1, sizeof *h);
    SimPropDefaultInit(&h->sp);
    h->sp.seed = seed;
    h->sp.nDevices = 1;
    if (ReadLatticeFileUnit(&h->ss, &h->sp, path, box_unit) || UpdatePolymerWithLabels(&h->ss) ||
        GetAllRingPolymers(&h->sp, &h->ss) || UpdatePolymerWithLabels(&h->ss) || !h->ss.nPol) {
        elp_ring_host_free(h);
        return NULL;
    }
    h->sp.polLength = h->ss.pol[0].length;
    return h;
}

/* elp_ring_host_from_file_unit with the largest unit of 32, 16, 8, 4 that
   divides the file's box edges (32 for every box the reference writes, which
   keeps its seed-table size and so the host stream position) */
elp_ring_host* elp_ring_host_from_file(const char* path, uint32_t seed) {
    int L[3] = {0, 0, 0};
    FILE* f = fopen(path, "r");
    if (!f) return NULL;
    int ok = fscanf(f, "%*s %i", L) == 1 && fscanf(f, "%*s %i", L + 1) == 1 && fscanf(f, "%*s %i", L + 2) == 1;
    fclose(f);
    if (!ok) return NULL;
    for (int unit = ELP_RING_TILE; unit >= 4; unit /= 2)
        if (L[0] > 0 && L[1] > 0 && L[2] > 0 && !(L[0] % unit) && !(L[1] % unit) && !(L[2] % unit))
            return elp_ring_host_from_file_unit(path, seed, unit);
    return NULL;
}

void elp_ring_host_free(elp_ring_host* h) {
    if (!h) return;
    SimStateFree(&h->ss);
    free(h->sp.trans);
    free(h);
}

/* box edges, number of rings, monomers per ring, maxPolLength */
void elp_ring_host_info(const elp_ring_host* h, int* LT, int* LU, int* LV, int* nPol, int* N, int* maxPolLength) {
    if (LT) *LT = (int)h->sp.LT;
    if (LU) *LU = (int)h->sp.LU;
    if (LV) *LV = (int)h->sp.LV;
    if (nPol) *nPol = h->ss.nPol;
    if (N) *N = h->ss.nPol ? h->ss.pol[0].length : 0;
    if (maxPolLength) *maxPolLength = (int)h->sp.maxPolLength;
}

int elp_ring_host_get_lattice(const elp_ring_host* h, uint8_t* out, uint64_t nbytes) {
    if (nbytes != h->sp.latSize) return -1;
    memcpy(out, h->ss.lattice, nbytes);
    return 0;
}

/* Replace the lattice and recover the rings from it (GetAllRingPolymers +
   UpdatePolymerWithLabels, as after every GPU interval).  -1 if the lattice
   does not hold exactly the labelled rings of this state. */
int elp_ring_host_set_lattice(elp_ring_host* h, const uint8_t* in, uint64_t nbytes) {
    if (nbytes != h->sp.latSize) return -1;
    memcpy(h->ss.lattice, in, nbytes);
    if (GetAllRingPolymers(&h->sp, &h->ss)) return -1;
    return UpdatePolymerWithLabels(&h->ss);
}

int elp_ring_host_write_tuv(elp_ring_host* h, const char* path) { return WriteLatticeFile(&h->sp, &h->ss, path); }

/* Bonds (incl. 0 = stored length) and start of ring `id` in tuv order.
   Returns the ring length, or -1. */
int elp_ring_host_polymer(const elp_ring_host* h, int id, uint32_t* bonds, int cap, uint32_t* start3) {
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

/* Host shift stream of the drop-in (gpupol2/gpupol_cuda.cu:36-38,133-139):
   elp_ring_host_load_shift draws the load-time shift; each following
   elp_ring_host_origins call emits the tile-grid origins of n time units
   (mod the tile edges) and advances the accumulated shift. */
void elp_ring_host_load_shift(elp_ring_host* h, int tt, int tu, int tv) {
    h->cumT = elp_draw_shift(&h->ss.rngState, tt);
    h->cumU = elp_draw_shift(&h->ss.rngState, tu);
    h->cumV = elp_draw_shift(&h->ss.rngState, tv);
}

void elp_ring_host_origins(elp_ring_host* h, int tt, int tu, int tv, int n, int32_t* out) {
    for (int i = 0; i < n; i++) {
        out[3 * i + 0] = (int32_t)(h->cumT % tt);
        out[3 * i + 1] = (int32_t)(h->cumU % tu);
        out[3 * i + 2] = (int32_t)(h->cumV % tv);
        h->cumT += elp_draw_shift(&h->ss.rngState, tt);
        h->cumU += elp_draw_shift(&h->ss.rngState, tu);
        h->cumV += elp_draw_shift(&h->ss.rngState, tv);
    }
}

uint32_t elp_ring_host_seed(const elp_ring_host* h) { return h->sp.seed; }

/* The table of 2^18 entries mapping a ring's 18 id label bits to its polymer
   id (-1: unused), as elp_ring_chains / elp_ring_tuv take it. */
int elp_ring_host_label2id(const elp_ring_host* h, int32_t* out, uint64_t n) {
    if (n != ((uint64_t)1 << h->ss.labelBit) || !h->ss.label2Id) return -1;
    for (uint64_t i = 0; i < n; i++) out[i] = h->ss.label2Id[i];
    return 0;
}

/* The per-lane seed table drawn by SimStateInit (gpupol2/gpupol_init.c:22-33):
   (nwg * 512 * 8) uint32, lane l of work-group w at 8*(l + 512*w): 4 words of
   its own state, then 4 words shared by the work-group.  The counter-based
   engine does not use it; drawing it keeps the host stream position. */
int elp_ring_host_seeds(const elp_ring_host* h, uint32_t* out, uint64_t n) {
    uint64_t need = (uint64_t)h->sp.nwg * h->sp.ws * 2 * h->sp.R;
    if (n != need) return -1;
    memcpy(out, h->ss.seeds, need * sizeof(uint32_t));
    return 0;
}

/* The host generator (gpupol2/gpupol.h:84-95, the LFSR113 step of util/rng.h)
   from an explicit state: n outputs into out, the state advanced in place.
   Exposed so that tests can compare it with the efpol restatement's. */
int elp_host_lfsr113(uint32_t state[4], uint32_t* out, int n) {
    if (!state || !out || n < 0) return -1;
    uint4_t s = {state[0], state[1], state[2], state[3]};
    for (int i = 0; i < n; i++) out[i] = elp_lfsr113(&s);
    state[0] = s.x; state[1] = s.y; state[2] = s.z; state[3] = s.w;
    return 0;
}
