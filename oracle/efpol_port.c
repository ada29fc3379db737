/*
 * efpol_port.c -- ORACLE / TEST INFRASTRUCTURE ONLY.
 *
 * A single-threaded C restatement of the reference CPU ELP engine
 * (cpupol/efpol*, "efpol").  It is used (a) by the tests as the statistical
 * reference for the GPU engines and (b) by bench.py as the CPU baseline timed
 * on the host cores.  Nothing in the product links or calls this file.
 *
 * Pinned by: byte-for-byte equality of every tuv file it writes with golden
 * files written by the reference efpol binaries (tests/golden/efpol/, made by
 * tests/golden/make_efpol_golden.sh from oracle/_ref), for ring and linear
 * chains and all three stored-length models (tests/test_efpol_port.py).
 *
 * Followed, function by function:
 *   LFSR113 generator, Seed, DRng ......... util/rng.h:10-35
 *   lattice constants / coordinates ....... cpupol/efpol_init.c:209-227,
 *                                           cpupol/efpol_lib.c:1-64
 *   transition table (AddTrans, SetTransTable) cpupol/efpol_init.c:285-424
 *   initial two-site loops (CreatePolymers)  cpupol/efpol_init.c:239-283
 *   one MC step (DoStep) / DoMCStep ........ cpupol/efpol_lib.c:122-207
 *   tuv writer (WriteLatticeFile) .......... cpupol/efpol_io.c:70-97
 *   settings writer ........................ cpupol/efpol_io.c:131-156
 *   driver (argv, output times) ............ cpupol/efpol.c:4-115
 * Not restated: lattice doubling (DoubleLattice, dblStep > 0 is rejected).
 *
 * Usage:
 *   efpol_port SEED DIR DENSITY TIME INTERVAL NMONO DBL_STEP L MODEL
 *       same argv as the reference (cpupol/efpol.c:22-38); writes the same
 *       files DIR/N_<NMONO>/t=<T>_dev=0.res and simulation_settings.txt.
 *   efpol_port from TUV SEED DIR TIME INTERVAL MAXSTACK [MODEL]
 *       efpol's dynamics (MODEL sl_equal, the default, sl_double or sl_quad)
 *       from the chains of a tuv file; MAXSTACK 2 rejects every move that
 *       would put a third monomer on a site, so the chain samples the state
 *       space of the GPU engines (at most one stored-length monomer per
 *       site); 0 = efpol itself.
 *   efpol_port table MODEL
 *       prints the transition table, one slot per line.
 *   efpol_port bench L NMONO DENSITY SEED NTIME [NWARM]
 *       no files; runs NWARM untimed time units, then times NTIME units of
 *       the DoMCStep region only (cpupol/efpol.c:92-95) and prints
 *       "attempts=<n> seconds=<s>" (attempts = NTIME*NMONO*nPol).
 *
 * Compile-time: -DPOL_RING (default) or -DPOL_LIN.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/time.h>
#include <errno.h>

#if !defined(POL_RING) && !defined(POL_LIN)
#define POL_RING 1
#endif

enum { MODEL_EQUAL = 1, MODEL_DOUBLE = 2, MODEL_QUAD = 4 };
#define NSLOT 16   /* transition slots per (src1,src2) pair */
#define BLOCK 8    /* polymers moved per step */

/* ---------------- RNG: LFSR113 (util/rng.h:10-35) ---------------- */
static uint32_t lfsr113(uint32_t s[4]) {
    uint32_t b;
    b = ((s[0] << 6) ^ s[0]) >> 13;  s[0] = ((s[0] & 0xFFFFFFFEu) << 18) ^ b;
    b = ((s[1] << 2) ^ s[1]) >> 27;  s[1] = ((s[1] & 0xFFFFFFF8u) << 2) ^ b;
    b = ((s[2] << 13) ^ s[2]) >> 21; s[2] = ((s[2] & 0xFFFFFFF0u) << 7) ^ b;
    b = ((s[3] << 3) ^ s[3]) >> 12;  s[3] = ((s[3] & 0xFFFFFF80u) << 13) ^ b;
    return s[0] ^ s[1] ^ s[2] ^ s[3];
}
static const double RNG_SCALE = 2.3283064365386962890625E-10; /* 2^-32 */
static double lfsr113_unit(uint32_t s[4]) { return RNG_SCALE * lfsr113(s); }
static void lfsr113_seed_efpol(uint32_t s[4], uint32_t seed) {
    s[0] = seed; s[1] = seed ^ 0xa741f123u; s[2] = seed ^ 0x213910abu; s[3] = seed ^ 0xcd123e14u;
}

/* ---------------- FCC bond algebra ---------------- */
static int bond_ok(uint32_t a) { return a != 0x0 && a != 0x3 && a != 0xc && a != 0xf; }
/* sum of two bond codes; returns 0 if the sum is not a single bond (the sum
   code is still stored when the operands are compatible) */
static int bond_sum(uint32_t a, uint32_t b, uint32_t* c) {
    if ((a | b) != 0xf && (a & b)) return 0;
    uint32_t r = ((a | b) == 0xf) ? (a & b) : (a | b);
    *c = r;
    return bond_ok(r);
}

/* ---------------- lattice geometry ----------------
 * Sites are indexed t + L*(u + L*v).  The reference packs coordinates with
 * guard bits; only the neighbour relation and the periodic wrap matter. */
typedef struct {
    int BL, L;
    int nPol, polSize;
    int wordsPerMono;      /* uint32 words holding 4-bit units of 8 polymers */
    uint32_t* occ;         /* 1 bit per site */
    uint32_t* unit;        /* [mono][pol/8] nibble (pol%8) */
    uint32_t* pos;         /* [mono][pol] site index */
    uint32_t rng[4];
    /* transition table, flat over (src1, src2, slot) like the reference */
    uint32_t trans[16 * 16 * NSLOT];
    int ntrans[16 * 16];
    int add_occ[16 * 16 * NSLOT];
    int rem_occ[16 * 16 * NSLOT];
    int maxStack;          /* 2: reject moves that put a third monomer on a site (0: no limit, efpol itself) */
} Sim;

static uint32_t site_of(const Sim* S, int t, int u, int v) { return (uint32_t)(t + S->L * (u + S->L * v)); }
static uint32_t site_step(const Sim* S, uint32_t site, uint32_t b) {
    int L = S->L;
    int t = site % L, u = (site / L) % L, v = site / (L * L);
    int w = (b >> 3) & 1;
    t = (t + (b & 1) - w + L) % L;
    u = (u + ((b >> 1) & 1) - w + L) % L;
    v = (v + ((b >> 2) & 1) - w + L) % L;
    return site_of(S, t, u, v);
}
static int occ_get(const Sim* S, uint32_t s) { return (S->occ[s >> 5] >> (s & 31)) & 1; }
static void occ_set(Sim* S, uint32_t s) { S->occ[s >> 5] |= 1u << (s & 31); }
static void occ_clr(Sim* S, uint32_t s) { S->occ[s >> 5] &= ~(1u << (s & 31)); }
static uint32_t unit_get(const Sim* S, int mono, int pol) {
    return (S->unit[pol / 8 + mono * S->wordsPerMono] >> (4 * (pol % 8))) & 0xf;
}
static void unit_or(Sim* S, int mono, int pol, uint32_t u) {
    S->unit[pol / 8 + mono * S->wordsPerMono] |= u << (4 * (pol % 8));
}

/* ---------------- transition table (cpupol/efpol_init.c:285-424) ---------------- */
static void add_trans(Sim* S, uint32_t s1, uint32_t s2, uint32_t d1, uint32_t d2) {
    int pair = 16 * (int)s1 + (int)s2;
    int id = NSLOT * pair + S->ntrans[pair];   /* no overflow guard, as in the reference */
    S->trans[id] = d1 | (d2 << 4);
    if (d1 || d2) {
        uint32_t mv = 0;
        int ok = (d1 != 0xf) ? bond_sum(d1, (~s1) & 0xf, &mv) : bond_sum(s2, (~d2) & 0xf, &mv);
        if (!ok) { fprintf(stderr, "bad transition [%x,%x]=>[%x,%x]\n", s1, s2, d1, d2); exit(1); }
        S->trans[id] |= mv << 8;   /* displacement of the moved monomer */
    }
    S->rem_occ[id] = (s1 && s2);
    S->add_occ[id] = (d1 && d2);
    S->ntrans[pair]++;
}

/* "large unit": bit k of a 4-bit code weighted so that a sum of four codes
   is zero modulo the all-axes unit iff the net displacement vanishes */
static uint64_t large_unit(const Sim* S, uint32_t b) {
    uint64_t L = (uint64_t)S->L;
    return (uint64_t)(b & 1) + (uint64_t)((b >> 1) & 1) * 2 * L + (uint64_t)((b >> 2) & 1) * 4 * L * L
         + (uint64_t)((b >> 3) & 1) * 8 * L * L * L;
}

static void build_trans(Sim* S, int model) {
    memset(S->ntrans, 0, sizeof(S->ntrans));
    uint32_t r;
    /* insert / remove a stored-length unit next to a bond pair */
    for (uint32_t a = 1; a < 0xf; a++) {
        if (!bond_ok(a)) continue;
        for (uint32_t b = a + 1; b < 0xf; b++) {
            if (!bond_ok(b)) continue;
            if (!bond_sum(a, b, &r)) continue;
            add_trans(S, a, b, r, 0); add_trans(S, a, b, 0, r);
            add_trans(S, b, a, r, 0); add_trans(S, b, a, 0, r);
            add_trans(S, r, 0, a, b); add_trans(S, r, 0, b, a);
            add_trans(S, 0, r, a, b); add_trans(S, 0, r, b, a);
        }
    }
    /* move a stored-length unit along the chain */
    for (uint32_t b = 1; b < 0xf; b++) {
        if (!bond_ok(b)) continue;
        for (int k = 0; k < 4; k++) { add_trans(S, 0, b, b, 0); add_trans(S, b, 0, 0, b); }
    }
    /* kink / crankshaft: replace (a,c) by (b,d) with a+c = b+d */
    uint64_t all_axes = large_unit(S, 0xf);
    for (uint32_t a = 1; a < 0xf; a++) {
        if (!bond_ok(a)) continue;
        for (uint32_t b = 1; b < 0xf; b++) {
            if (!bond_ok(b)) continue;
            if (!bond_sum(a, (~b) & 0xf, &r)) continue;
            for (uint32_t c = 1; c < 0xf; c++) {
                if (!bond_ok(c)) continue;
                if (model == MODEL_QUAD && bond_sum(a, c, &r)) continue;
                for (uint32_t d = 1; d < 0xf; d++) {
                    if (!bond_ok(d)) continue;
                    uint64_t net = large_unit(S, a) + large_unit(S, c) + large_unit(S, (~b) & 0xf)
                                 + large_unit(S, (~d) & 0xf);
                    if (net % all_axes) continue;
                    add_trans(S, a, c, b, d); add_trans(S, a, c, b, d);
                    if (model == MODEL_EQUAL && bond_sum(a, c, &r)) {
                        for (int k = 0; k < 2; k++) { add_trans(S, a, c, b, d); add_trans(S, a, c, b, d); }
                    }
                }
            }
        }
    }
    /* replicate partially filled pairs up to 16 slots (sequential tests, as in the reference) */
    for (int p = 0; p < 256; p++) {
        if (S->ntrans[p] == 2) for (int i = 0; i < 2; i++) { uint32_t e = S->trans[p * NSLOT + i]; add_trans(S, p / 16, p % 16, e & 0xf, (e >> 4) & 0xf); }
        if (S->ntrans[p] == 4) for (int i = 0; i < 4; i++) { uint32_t e = S->trans[p * NSLOT + i]; add_trans(S, p / 16, p % 16, e & 0xf, (e >> 4) & 0xf); }
        if (S->ntrans[p] == 8) for (int i = 0; i < 8; i++) { uint32_t e = S->trans[p * NSLOT + i]; add_trans(S, p / 16, p % 16, e & 0xf, (e >> 4) & 0xf); }
    }
    /* chain-end moves (code 0xf marks the end of a linear chain) */
    for (uint32_t a = 0; a < 0xf; a++) {
        if (!bond_ok(a)) continue;
        for (int k = 0; k < model; k++) { add_trans(S, a, 0xf, 0x0, 0xf); add_trans(S, 0xf, a, 0xf, 0x0); }
        add_trans(S, 0x0, 0xf, a, 0xf); add_trans(S, 0xf, 0x0, 0xf, a);
    }
    for (int p = 0; p < 256; p++)
        while (S->ntrans[p] < NSLOT) add_trans(S, p / 16, p % 16, 0, 0);
}

/* ---------------- state ---------------- */
static void sim_alloc(Sim* S, int BL, int polSize, int nPol) {
    S->BL = BL; S->L = 1 << BL; S->polSize = polSize; S->nPol = nPol;
    S->wordsPerMono = nPol / 8 + ((nPol % 8) ? 1 : 0);
    size_t nsite = (size_t)S->L * S->L * S->L;
    S->occ = calloc(nsite / 32 + 1, sizeof(uint32_t));
    S->unit = calloc((size_t)S->wordsPerMono * polSize, sizeof(uint32_t));
    S->pos = calloc((size_t)polSize * nPol, sizeof(uint32_t));
}

/* two-site loops on a regular grid (cpupol/efpol_init.c:239-283) */
static void create_polymers(Sim* S) {
    int L = S->L, delta, nTot;
    for (delta = 1; delta < L; delta++) {
        nTot = (L / delta) * (L / delta) * (L / (delta + 1));
        if (nTot < S->nPol) { delta--; break; }
    }
    int nT = L / delta, nV = L / (delta + 1);
    for (int p = 0; p < S->nPol; p++) {
        int t = delta * (p % nT);
        int v = (delta + 1) * ((p / nT) % nV);
        int u = delta * (p / nT / nV);
        uint32_t a = site_of(S, t, u, v), b = site_of(S, t, u, v + 1);
        S->pos[p] = a;
        unit_or(S, 0, p, 0x4);
        occ_set(S, a); occ_set(S, b);
        for (int m = 1; m < S->polSize; m++) S->pos[p + S->nPol * m] = b;
#ifdef POL_LIN
        unit_or(S, S->polSize - 1, p, 0xf);
#else
        unit_or(S, S->polSize - 1, p, 0xb);
#endif
    }
}

/* ---------------- one MC step (cpupol/efpol_lib.c:132-207) ---------------- */
static long mc_step(Sim* S) {
    int N = S->polSize;
    int mono = (int)(lfsr113_unit(S->rng) * N);
#ifdef POL_LIN
    int prev = (mono == 0) ? (N - 1) : (mono - 1);
#else
    int prev = (mono - 1 + N) % N;
#endif
    int nBlocks = (S->nPol - 1) / BLOCK + 1;
    int p0 = (int)(lfsr113_unit(S->rng) * nBlocks) * BLOCK;
    int p1 = (S->nPol < p0 + BLOCK) ? S->nPol : p0 + BLOCK;
    uint32_t bits = 0;
    long changed = 0;
    uint32_t* w1 = &S->unit[p0 / 8 + prev * S->wordsPerMono];
    uint32_t* w2 = &S->unit[p0 / 8 + mono * S->wordsPerMono];
    for (int p = p0; p < p1; p++) {
        int sh = 4 * ((p - p0) % 8);
        if (sh == 0) bits = lfsr113(S->rng);
        uint32_t s1 = (*w1 >> sh) & 0xf, s2 = (*w2 >> sh) & 0xf;
        int opt = 16 * NSLOT * (int)s1 + NSLOT * (int)s2 + (int)((bits >> sh) & 0xf);
        uint32_t e = S->trans[opt];
        uint32_t d1 = e & 0xf, d2 = (e >> 4) & 0xf;
        if (!d1 && !d2) continue;
        if (S->maxStack == 2) {
            /* a site holds three monomers iff two consecutive bonds are both 0 */
#ifdef POL_LIN
            int pp = prev - 1, nn = mono + 1;
            uint32_t bpp = pp >= 0 ? unit_get(S, pp, p) : 0xf, bnn = nn < N ? unit_get(S, nn, p) : 0xf;
#else
            uint32_t bpp = unit_get(S, (prev - 1 + N) % N, p), bnn = unit_get(S, (mono + 1) % N, p);
#endif
            if ((d1 == 0 && (bpp == 0 || d2 == 0)) || (d2 == 0 && bnn == 0)) continue;
        }
        uint32_t* ps = &S->pos[p + mono * S->nPol];
        uint32_t to = site_step(S, *ps, e >> 8);
        if (S->add_occ[opt]) {
            if (occ_get(S, to)) continue;
            occ_set(S, to);
        }
        if (S->rem_occ[opt]) occ_clr(S, *ps);
        *ps = to;
        *w1 = (*w1 & ~(0xfu << sh)) | (d1 << sh);
        *w2 = (*w2 & ~(0xfu << sh)) | (d2 << sh);
        changed++;
    }
    return changed;
}

static void mc_run(Sim* S, long nTime) {
    long n = nTime * (long)S->polSize * ((S->nPol - 1) / BLOCK + 1);
    for (long j = 0; j < n; j++) mc_step(S);
}

/* ---------------- output (cpupol/efpol_io.c) ---------------- */
static void write_tuv(const Sim* S, const char* file) {
    FILE* f = fopen(file, "w");
    if (!f) { fprintf(stderr, "cannot open %s\n", file); exit(1); }
    fprintf(f, "LT= %i\nLU= %i\nLV= %i\nnp= %i\nmaxPolLength= %i\n", S->L, S->L, S->L, S->nPol, S->polSize);
    for (int p = 0; p < S->nPol; p++) {
        uint32_t s = S->pos[p];
        fprintf(f, "len= %i\n", S->polSize);
        fprintf(f, "%u  %u  %u\t", s % S->L, (s / S->L) % S->L, s / (S->L * S->L));
        for (int m = 0; m < S->polSize; m++) fprintf(f, "%x", unit_get(S, m, p));
        fprintf(f, "\n");
    }
    fclose(f);
}

static int bl_for(int L);

/* Start from a tuv file (box edge a power of two, every chain of the same
   length): bonds as given, monomer positions walked from each start.
   Returns 0, or -1 on a file this reader does not take. */
static int read_tuv(Sim* S, const char* file) {
    FILE* f = fopen(file, "r");
    if (!f) return -1;
    int LT, LU, LV, np, maxLen;
    if (fscanf(f, "%*s %d %*s %d %*s %d %*s %d %*s %d", &LT, &LU, &LV, &np, &maxLen) != 5 || LT != LU || LT != LV ||
        np <= 0 || maxLen <= 0 || (LT & (LT - 1))) { fclose(f); return -1; }
    char* text = malloc((size_t)maxLen + 2);
    char fmt[32];
    snprintf(fmt, sizeof fmt, "%%%ds", maxLen + 1);
    int rc = 0;
    for (int p = 0; p < np && !rc; p++) {
        int len; unsigned t, u, v;
        if (fscanf(f, "%*s %d %u %u %u", &len, &t, &u, &v) != 4 || fscanf(f, fmt, text) != 1 || (int)strlen(text) != len) {
            rc = -1; break;
        }
        if (p == 0) sim_alloc(S, bl_for(LT), len, np);
        if (len != S->polSize) { rc = -1; break; }
        uint32_t site = site_of(S, (int)t, (int)u, (int)v);
        for (int m = 0; m < len; m++) {
            char c = text[m];
            uint32_t b = (c >= '0' && c <= '9') ? (uint32_t)(c - '0') : (uint32_t)(10 + (c | 0x20) - 'a');
            if (b > 0xf) { rc = -1; break; }
            S->pos[p + S->nPol * m] = site;
            occ_set(S, site);
            unit_or(S, m, p, b);
            if (b && b != 0xf) site = site_step(S, site, b);
        }
    }
    free(text);
    fclose(f);
    return rc;
}

static void write_settings(const Sim* S, const char* dir, uint32_t seed, double density,
                           int initPolSize, long interval, int model) {
    char file[4096];
    snprintf(file, sizeof file, "%s/simulation_settings.txt", dir);
    FILE* f = fopen(file, "w");
    if (!f) return;
    fprintf(f, "Start_seed = %u\nLength = %i\n", seed, S->polSize);
#ifdef POL_LIN
    fprintf(f, "Polytype = lin\n");
#else
    fprintf(f, "Polytype = ring\n");
#endif
    fprintf(f, "Density = %lf\nLatsize = %i\nStart_polysize = %i\nDouble_step = %i\n", density, S->L, initPolSize, 0);
    fprintf(f, "Interval = %li\nEquilibrated = 0\nNpol = %i\n", interval, S->nPol);
    fprintf(f, "Polymodel = %s\n", model == MODEL_EQUAL ? "sl_equal" : model == MODEL_DOUBLE ? "sl_double" : "sl_quad");
    fprintf(f, "Executable = efpol\n");
    fclose(f);
}

static int mkdir_p(const char* path) {
    char buf[4096];
    snprintf(buf, sizeof buf, "%s", path);
    for (char* p = buf + 1; *p; p++) {
        if (*p == '/') { *p = 0; mkdir(buf, 0777); *p = '/'; }
    }
    if (mkdir(buf, 0777) && errno != EEXIST) return -1;
    return 0;
}

static int n_pol_for(double density, int polSize, int nsite) {
    int n = (int)(density * nsite / (double)polSize + 0.5);
    return n < 1 ? 1 : n;
}

static int bl_for(int L) {
    int b;
    for (b = 0; b <= 9; b++) if ((1 << b) >= L) break;
    return b;
}

static double now_s(void) { struct timeval tv; gettimeofday(&tv, NULL); return tv.tv_sec + 1e-6 * tv.tv_usec; }

int main(int argc, char** argv) {
    if (argc >= 2 && !strcmp(argv[1], "rng")) {
        /* rng X Y Z W N: N outputs of the generator from the state (X,Y,Z,W), one per line */
        if (argc < 7) { fprintf(stderr, "usage: %s rng X Y Z W N\n", argv[0]); return 2; }
        uint32_t s[4];
        for (int k = 0; k < 4; k++) s[k] = (uint32_t)strtoul(argv[2 + k], NULL, 10);
        const long n = atol(argv[6]);
        for (long i = 0; i < n; i++) printf("%u\n", lfsr113(s));
        return 0;
    }
    if (argc >= 2 && !strcmp(argv[1], "from")) {
        /* from TUV SEED DIR TIME INTERVAL MAXSTACK [MODEL]: efpol's dynamics from the chains of a tuv
           file, with at most MAXSTACK (2, or 0 = no limit) monomers per site; writes
           DIR/N_<len>/t=<T>_dev=0.res every INTERVAL time units like the reference driver */
        if (argc < 8) {
            fprintf(stderr, "usage: %s from TUV SEED DIR TIME INTERVAL MAXSTACK [sl_equal|sl_double|sl_quad]\n", argv[0]);
            return 2;
        }
        Sim S; memset(&S, 0, sizeof S);
        if (read_tuv(&S, argv[2])) { fprintf(stderr, "cannot read %s\n", argv[2]); return 199; }
        uint32_t seed = (uint32_t)atol(argv[3]);
        long tMax = atol(argv[5]), interval = atol(argv[6]);
        S.maxStack = atoi(argv[7]);
        lfsr113_seed_efpol(S.rng, seed);
        const char* mname = argc > 8 ? argv[8] : "sl_equal";
        build_trans(&S, !strcmp(mname, "sl_double") ? MODEL_DOUBLE : !strcmp(mname, "sl_quad") ? MODEL_QUAD : MODEL_EQUAL);
        char sub[4096], file[8192];
        snprintf(sub, sizeof sub, "%s/N_%i", argv[4], S.polSize);
        mkdir_p(sub);
        for (long t = 0; t <= tMax; t += interval) {
            snprintf(file, sizeof file, "%s/t=%li_dev=%i.res", sub, t, 0);
            write_tuv(&S, file);
            if (t < tMax) mc_run(&S, interval);
        }
        return 0;
    }
    if (argc >= 2 && !strcmp(argv[1], "table")) {
        /* table MODEL: every transition slot, "s1 s2 slot d1 d2" in hex (16 slots per bond pair) */
        if (argc < 3) { fprintf(stderr, "usage: %s table sl_equal|sl_double|sl_quad\n", argv[0]); return 2; }
        int model = !strcmp(argv[2], "sl_double") ? MODEL_DOUBLE : !strcmp(argv[2], "sl_quad") ? MODEL_QUAD : MODEL_EQUAL;
        Sim S; memset(&S, 0, sizeof S);
        S.BL = 5; S.L = 32;
        build_trans(&S, model);
        for (int pair = 0; pair < 256; pair++)
            for (int k = 0; k < NSLOT; k++) {
                uint32_t e = S.trans[NSLOT * pair + k];
                printf("%x %x %d %x %x\n", pair >> 4, pair & 15, k, e & 0xf, (e >> 4) & 0xf);
            }
        return 0;
    }
    if (argc >= 2 && !strcmp(argv[1], "bench")) {
        if (argc < 7) { fprintf(stderr, "usage: %s bench L NMONO DENSITY SEED NTIME [NWARM]\n", argv[0]); return 2; }
        int L = atoi(argv[2]), N = atoi(argv[3]);
        double rho = atof(argv[4]);
        uint32_t seed = (uint32_t)atol(argv[5]);
        long nTime = atol(argv[6]);
        long nWarm = argc > 7 ? atol(argv[7]) : 0;
        Sim S; memset(&S, 0, sizeof S);
        int BL = bl_for(L);
        int nPol = n_pol_for(rho, N, (1 << BL) * (1 << BL) * (1 << BL));
        sim_alloc(&S, BL, N, nPol);
        lfsr113_seed_efpol(S.rng, seed);
        build_trans(&S, MODEL_EQUAL);
        create_polymers(&S);
        mc_run(&S, nWarm);   /* untimed relaxation from the two-site loops */
        double t0 = now_s();
        mc_run(&S, nTime);
        double t1 = now_s();
        printf("attempts=%ld seconds=%.6f L=%d nPol=%d N=%d\n", nTime * (long)N * nPol, t1 - t0, S.L, nPol, N);
        return 0;
    }
    if (argc <= 9) { fprintf(stderr, "usage: %s SEED DIR DENSITY TIME INTERVAL NMONO DBL_STEP L MODEL\n", argv[0]); return 2; }
    uint32_t seed = (uint32_t)atol(argv[1]);
    const char* dir = argv[2];
    double density = atof(argv[3]);
    long tMax = (long)atof(argv[4]);
    long interval = (long)atof(argv[5]);
    int initPolSize = atoi(argv[6]);
    int dbl = atoi(argv[7]);
    int Lreq = atoi(argv[8]);
    int model;
    if (!strcmp(argv[9], "sl_equal")) model = MODEL_EQUAL;
    else if (!strcmp(argv[9], "sl_double")) model = MODEL_DOUBLE;
    else if (!strcmp(argv[9], "sl_quad")) model = MODEL_QUAD;
    else { printf("Unknown polymer model, aborting!\n"); return 193; }
    if (dbl != 0) { fprintf(stderr, "lattice doubling is not restated\n"); return 2; }
    if (Lreq < 2 || initPolSize < 2) { printf("Error polymer or lattice too small\n"); return 212; }

    Sim S; memset(&S, 0, sizeof S);
    int BL = bl_for(Lreq);
    int nPol = n_pol_for(density, initPolSize, (1 << BL) * (1 << BL) * (1 << BL));
    sim_alloc(&S, BL, initPolSize, nPol);
    lfsr113_seed_efpol(S.rng, seed);
    char sub[4096], file[8192];
    snprintf(sub, sizeof sub, "%s/N_%i", dir, initPolSize);
    mkdir_p(sub);
    build_trans(&S, model);
    create_polymers(&S);
    for (long t = 0; t < tMax; t += interval) {
        snprintf(file, sizeof file, "%s/t=%li_dev=%i.res", sub, t, 0);
        write_tuv(&S, file);
        double t0 = now_s();
        mc_run(&S, interval);
        double dt = now_s() - t0;
        printf("\rL=%i at %.2lf M/s", S.L, 1e-6 * interval * S.polSize * S.nPol / (dt > 0 ? dt : 1e-9));
        fflush(stdout);
    }
    snprintf(file, sizeof file, "%s/t=%li_dev=%i.res", sub, ((tMax - 1) / interval + 1) * interval, 0);
    write_tuv(&S, file);
    write_settings(&S, sub, seed, density, initPolSize, interval, model);
    printf("\n");
    return 0;
}
