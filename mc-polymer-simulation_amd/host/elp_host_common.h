/*
 * elp_host_common.h -- host-side helpers shared by the ring and linear
 * drop-in libraries: the LFSR113 host generator with the reference's seed
 * semantics, FCC bond-code algebra and t=<T> file-name parsing.
 */
#ifndef ELP_HOST_COMMON_H
#define ELP_HOST_COMMON_H
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef UINT_DEF
#define UINT_DEF
typedef unsigned int uint;
#endif

/* 2^-32: scales a 32-bit draw to [0,1) (gpupol2/gpupol.h:15, util/rng.h:27) */
#define ELP_RNG_FAC 2.3283064365386962890625E-10

#ifndef ELP_UINT4_T_DEF
#define ELP_UINT4_T_DEF
typedef struct uint4_t { uint x, y, z, w; } uint4_t;
#endif

/* L'Ecuyer LFSR113 (gpupol2/gpupol.h:84-95, identical to util/rng.h:10-21). */
static inline uint elp_lfsr113(uint4_t* s) {
    uint b;
    b = ((s->x << 6) ^ s->x) >> 13;  s->x = ((s->x & 0xFFFFFFFEu) << 18) ^ b;
    b = ((s->y << 2) ^ s->y) >> 27;  s->y = ((s->y & 0xFFFFFFF8u) << 2) ^ b;
    b = ((s->z << 13) ^ s->z) >> 21; s->z = ((s->z & 0xFFFFFFF0u) << 7) ^ b;
    b = ((s->w << 3) ^ s->w) >> 12;  s->w = ((s->w & 0xFFFFFF80u) << 13) ^ b;
    return s->x ^ s->y ^ s->z ^ s->w;
}

/* Per-device seeding of the GPU drivers (gpupol2/gpupol_init.c:14-17). */
static inline void elp_lfsr113_seed_gpu(uint4_t* s, uint seed, int dev) {
    s->x = seed ^ (85127698u + (uint)dev * 716576218u);
    s->y = seed ^ (12369123u + (uint)dev * 612737112u);
    s->z = seed ^ (81261212u + (uint)dev * 172635798u);
    s->w = seed ^ (51203016u + (uint)dev * 410287361u);
}

/* Tile-grid shift drawn per time unit: (int)(edge * 2^-32 * r)
   (gpupol2/gpupol_cuda.cu:133-135). */
static inline int elp_draw_shift(uint4_t* s, int edge) {
    return (int)(edge * (double)ELP_RNG_FAC * elp_lfsr113(s));
}

/* ---- FCC bonds: 4-bit code (t,u,v,w), displacement (t-w, u-w, v-w) ---- */
static inline int elp_bond_valid(uint a) { return a != 0x0 && a != 0x3 && a != 0xc && a != 0xf; }
static inline uint elp_bond_add(uint a, uint b) { return ((a | b) == 0xf) ? (a & b) : (a | b); }
/* 1 iff a+b is a single bond; *c receives the sum when the codes combine */
static inline int elp_bond_add_valid(uint a, uint b, uint* c) {
    if ((a | b) != 0xf && (a & b)) return 0;
    uint r = elp_bond_add(a, b);
    *c = r;
    return elp_bond_valid(r);
}
static inline void elp_bond_delta(uint b, int* dt, int* du, int* dv) {
    int w = (b >> 3) & 1;
    *dt = (int)(b & 1) - w; *du = (int)((b >> 1) & 1) - w; *dv = (int)((b >> 2) & 1) - w;
}

/* t=<T> from a file name (util/file.c:37-53, gpupol2/util.c GetTFromName) */
long elp_t_from_name(const char* file);

/* trans table of the two decompositions of every valid bond
   (gpupol2/gpupol_init.c:552-584); returns 0 on success */
int elp_set_gpu_trans(uint trans[4]);

#endif
