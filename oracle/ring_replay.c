/*
 * ring_replay.c -- ORACLE / TEST INFRASTRUCTURE ONLY.
 *
 * Single-threaded C replay of the ring engine's schedule: the bit-exact
 * reference for the gfx950 kernel (mc-polymer-simulation_amd/csrc/elp_ring.hip).
 * Written independently of the kernel: tiles are copied into a plain
 * t-fastest buffer, slots are visited one at a time in a selectable order,
 * and the moves are restated directly from the reference kernel:
 *   TransForw  gpupol2/ringpol3.cu:92-145
 *   TransBack  gpupol2/ringpol3.cu:147-192
 *   DiffuseSL  gpupol2/ringpol3.cu:194-218
 *   OOB reject gpupol2/ringpol3.cu:58-76 (touched site outside the tile)
 *   substep order per iteration TF, DSL, TF, TB  gpupol2/ringpol3.cu:278-292
 *   trans table gpupol2/gpupol_init.c:552-584
 * Schedule (the gfx950 engine's, documented in include/elp_gpu.h): per time
 * unit i the tile grid origin is origins[3i..3i+2]; Philox4x32-10 with key
 * (seed, stream) gives the tile-uniform coset per substep (counter
 * (tau, tile, 0xFFFFFFFF, iter/4), word iter%4, byte substep) and the move
 * bits per slot (counter (tau, tile, slot, iter/4), same word/byte); under
 * balanced rates bits 6-7 of the slot's byte (the move drawn) are those of
 * the tile's coset byte.
 *
 * Pinned by: Philox known-answer vectors (Random123 kat_vectors); slot-order
 * independence (ascending, descending and shuffled orders give identical
 * lattices -- a brute-force check that slots of one substep commute);
 * invariants after every run (monomer count, ring closure, <= 1 stored
 * length per site, cyclic label sequence of every ring unchanged); exact
 * stationary-distribution checks on tiny systems (tests/test_ring_balance.py:
 * every move has a reverse and the exact forward/reverse rate ratios;
 * tests/test_learned_gpu.py: the six-monomer ring's exact measure).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---------------- Philox4x32-10 ---------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static uint32_t philox_word(uint32_t tau, uint32_t tile, uint32_t lane, uint32_t iter, uint32_t seed, uint32_t stream) {
    uint32_t ctr[4] = {tau, tile, lane, iter >> 2}, key[2] = {seed, stream}, out[4];
    orc_philox4x32_10(ctr, key, out);
    return out[iter & 3];
}

/* ---------------- bond algebra ---------------- */
static int dT(uint32_t b) { return (int)(b & 1) - (int)(b >> 3); }
static int dU(uint32_t b) { return (int)((b >> 1) & 1) - (int)(b >> 3); }
static int dV(uint32_t b) { return (int)((b >> 2) & 1) - (int)(b >> 3); }
static uint32_t badd(uint32_t a, uint32_t b) { return ((a | b) == 0xf) ? (a & b) : (a | b); }

static void gpu_trans(uint32_t tr[4]) {
    int nf[16] = {0};
    memset(tr, 0, 4 * sizeof(uint32_t));
    for (uint32_t a = 1; a < 15; a++) {
        if (a == 3 || a == 0xc) continue;
        for (uint32_t b = a + 1; b < 15; b++) {
            if (b == 3 || b == 0xc) continue;
            if ((a | b) != 0xf && (a & b)) continue;
            uint32_t s = badd(a, b);
            if (s == 0 || s == 3 || s == 0xc || s == 0xf) continue;
            tr[s / 4] |= a << (8 * (s % 4) + 4 * nf[s]);
            nf[s]++;
        }
    }
}

/* ---------------- per-move random bits ----------------
 * The slot's byte for this substep, fixed in advance (where the reference
 * kernel draws from its per-lane xorshift128 state, gpupol2/ringpol3.cu:108,
 * :171, the engine reads this byte). */
typedef struct { uint32_t fixed; } Draw;
static uint32_t draw(Draw* d) { return d->fixed; }

/* ---------------- moves on a tile buffer ---------------- */
typedef struct { uint8_t* b; int T, U, V; uint32_t tr[4]; int rules; } Tile;   /* rules: 0 reference, 1 balanced */

/* index of (t,u,v)+bond, or -1 if it leaves the tile */
static int step(const Tile* X, int t, int u, int v, uint32_t bond, int* nt, int* nu, int* nv) {
    *nt = t + dT(bond); *nu = u + dU(bond); *nv = v + dV(bond);
    if (*nt < 0 || *nt >= X->T || *nu < 0 || *nu >= X->U || *nv < 0 || *nv >= X->V) return -1;
    return *nt + X->T * (*nu + X->U * *nv);
}

static int trans_forw(Tile* X, int t, int u, int v, Draw* rd) {
    int x = t + X->T * (u + X->U * v), at, au, av;
    uint32_t s = X->b[x];
    if (!s) return 0;
    uint32_t nxt = s & 0xf, lab = s & 0x30, sl = s & 0x40;
    int n = step(X, t, u, v, nxt, &at, &au, &av);
    if (n < 0) return 0;
    uint32_t ns = X->b[n], nsl = ns & 0x40;
    if (sl + nsl == 0) return 0;
    uint32_t r = draw(rd);
    uint32_t b1 = (X->tr[nxt / 4] >> (4 * (2 * (nxt % 4) + (r & 1)))) & 0xf;
    uint32_t b2 = badd((~b1) & 0xf, nxt);
    uint32_t first = (r & 2) ? b1 : b2, second = (r & 2) ? b2 : b1;
    int d = step(X, t, u, v, first, &at, &au, &av);
    if (d < 0) return 0;
    if (X->b[d]) return 0;
    int moveFirst = (sl + nsl == 0x80) ? (int)((r & 4) >> 2) : (sl ? 1 : 0);
    uint32_t dn = second;
    if (moveFirst) {
        s = first | ((lab >> 1) & 0x10);
        dn |= lab & 0x10;
    } else {
        s = first | lab | sl;
        dn |= (ns & 0x20) >> 1;
        ns &= 0x1f;
    }
    X->b[x] = (uint8_t)s;
    X->b[d] = (uint8_t)dn;
    if (!moveFirst) X->b[n] = (uint8_t)ns;
    return 1;
}

static int trans_back(Tile* X, int t, int u, int v, Draw* rd) {
    int x = t + X->T * (u + X->U * v), at, au, av;
    uint32_t s = X->b[x];
    uint32_t nxt = s & 0xf, lab = s & 0x30, sl = s & 0x40;
    if (!s) return 0;
    int src = step(X, t, u, v, nxt, &at, &au, &av);
    if (src < 0) return 0;
    uint32_t ss = X->b[src];
    uint32_t srcNext = ss & 0xf, srcLab = ss & 0x30;
    if (ss & 0x40) return 0;
    if ((nxt | srcNext) != 0xf && (nxt & srcNext)) return 0;
    uint32_t nn = badd(nxt, srcNext);
    if (nn == 0x3 || nn == 0xc) return 0;
    int w = step(X, t, u, v, nn, &at, &au, &av);
    if (w < 0) return 0;
    uint32_t ws = X->b[w], wsl = ws & 0x40;
    if (sl + wsl == 0x80) return 0;
    uint32_t r = draw(rd);
    /* balanced rates: a fold forced by stored length at one end is accepted
       with probability 1/4, the rate of the TransForw that undoes it */
    if (X->rules && (sl | wsl) && (r & 6)) return 0;
    int moveFirst = (sl + wsl == 0) ? (int)(r & 1) : (sl == 0x40 ? 0 : 1);
    if (moveFirst) {
        s = nn | (lab << 1) | srcLab | 0x40;
    } else {
        s = nn | lab | sl;
        ws = (ws & 0x3f) | (srcLab << 1) | 0x40;
    }
    X->b[src] = 0;
    X->b[x] = (uint8_t)s;
    X->b[w] = (uint8_t)ws;
    return 1;
}

static int diffuse_sl(Tile* X, int t, int u, int v) {
    int x = t + X->T * (u + X->U * v), at, au, av;
    uint32_t s = X->b[x];
    uint32_t nxt = s & 0xf, lab = s & 0x30, sl = s & 0x40;
    if (!s) return 0;
    int n = step(X, t, u, v, nxt, &at, &au, &av);
    if (n < 0) return 0;
    uint32_t ns = X->b[n], nlab = ns & 0x30, nsl = ns & 0x40;
    if (nsl + sl != 0x40) return 0;
    if (sl) {
        ns = ns | ((lab & 0x10) << 1) | 0x40;
        s = nxt | ((lab >> 1) & 0x10);
    } else {
        s = nxt | (lab << 1) | ((nlab >> 1) & 0x10) | 0x40;
        ns = ns & 0x1f;
    }
    X->b[x] = (uint8_t)s;
    X->b[n] = (uint8_t)ns;
    return 1;
}

/* ---------------- move counting ----------------
 * Outcome counts of a replay (orc_ring_counts), the quantities the engine's
 * counting kernels report: attempts at an empty site; attempts whose move
 * would touch a site outside the tile -- TransForw: the bond partner or the
 * target common neighbour (first bond of the pair the random bits pick);
 * DiffuseSL: the partner; TransBack: the source or the fold target beyond it
 * -- whatever the move's other conditions; accepted moves (the move wrote). */
static int g_counting = 0;
static uint64_t g_counts[3];

static int leaves_tile(const Tile* X, int t, int u, int v, int type, uint32_t r) {
    uint32_t s = X->b[t + X->T * (u + X->U * v)];
    uint32_t nxt = s & 0xf;
    int at, au, av, bt, bu, bv;
    int n = step(X, t, u, v, nxt, &at, &au, &av);
    if (n < 0) return 1;
    if (type == 1) return 0;
    if (type == 0) {
        uint32_t b1 = (X->tr[nxt / 4] >> (4 * (2 * (nxt % 4) + (r & 1)))) & 0xf;
        uint32_t first = (r & 2) ? b1 : badd((~b1) & 0xf, nxt);
        return step(X, t, u, v, first, &bt, &bu, &bv) < 0;
    }
    uint32_t srcNext = X->b[n] & 0xf;
    return step(X, at, au, av, srcNext, &bt, &bu, &bv) < 0;
}

/* 1: start counting (counts cleared); 0: stop.  out (may be NULL): empty,
   leaving the tile, accepted since the last start. */
int orc_ring_counts(int enable, uint64_t out[3]) {
    if (out) { out[0] = g_counts[0]; out[1] = g_counts[1]; out[2] = g_counts[2]; }
    if (enable) g_counts[0] = g_counts[1] = g_counts[2] = 0;
    g_counting = enable;
    return 0;
}

/*
 * One move on a T x U x V tile buffer (t fastest) with the random byte r:
 * type 0 TransForw, 1 DiffuseSL, 2 TransBack; rules 0 reference, 1 balanced.  For enumerating transition
 * probabilities of the move set (detailed-balance tests).  Returns 0, or -1
 * on a bad type / site.
 */
int orc_ring_move(uint8_t* tile, int T, int U, int V, int type, int t, int u, int v, uint32_t r, int rules) {
    if (t < 0 || t >= T || u < 0 || u >= U || v < 0 || v >= V) return -1;
    Tile X = {tile, T, U, V, {0, 0, 0, 0}, rules};
    gpu_trans(X.tr);
    Draw d = {r};
    if (type == 0) trans_forw(&X, t, u, v, &d);
    else if (type == 1) diffuse_sl(&X, t, u, v);
    else if (type == 2) trans_back(&X, t, u, v, &d);
    else return -1;
    return 0;
}

/* site of slot `slot` in coset c of the given sub-lattice */
static void slot_site(int sublat, int T, int U, int slot, uint32_t c, int* t, int* u, int* v) {
    int ncT = T / 4, ncU = U / 4;
    int cell = sublat ? slot / 4 : slot, j = sublat ? slot % 4 : 0;
    int bt = 4 * (cell % ncT), bu = 4 * ((cell / ncT) % ncU), bv = 4 * (cell / (ncT * ncU));
    if (sublat) {       /* 4Z^3 + <(1,3,2)>: rep (0, c&3, c>>2) + j*(1,3,2) mod 4 */
        *t = bt + j;
        *u = bu + (int)((c + 3u * j) % 4);
        *v = bv + (int)(((c >> 2) + 2u * j) % 4);
    } else {            /* 4Z^3: offset (c&3, (c>>2)&3, (c>>4)&3) */
        *t = bt + (int)(c & 3);
        *u = bu + (int)((c >> 2) & 3);
        *v = bv + (int)((c >> 4) & 3);
    }
}

/*
 * order: 0 ascending slots, 1 descending, 2 shuffled (fixed LCG per substep).
 * Returns 0, or -1 on bad arguments.
 */
int orc_ring_run_rules(uint8_t* lat, int LT, int LU, int LV, int TT, int TU, int TV, int sublat, int n_iter,
                       int n_time, const int32_t* origins, uint64_t tau0, uint32_t seed, uint32_t stream, int order,
                       int rules) {
    if (TT % 4 || TU % 4 || TV % 4 || LT % TT || LU % TU || LV % TV) return -1;
    Tile X;
    X.T = TT; X.U = TU; X.V = TV;
    X.rules = rules;
    X.b = (uint8_t*)malloc((size_t)TT * TU * TV);
    gpu_trans(X.tr);
    int nT = LT / TT, nU = LU / TU, nV = LV / TV;
    int nslot = (TT / 4) * (TU / 4) * (TV / 4) * (sublat ? 4 : 1);
    int* perm = (int*)malloc(sizeof(int) * nslot);
    uint32_t cmask = sublat ? 0xf : 0x3f;
    for (int i = 0; i < n_time; i++) {
        int ot = origins[3 * i], ou = origins[3 * i + 1], ov = origins[3 * i + 2];
        uint32_t tau = (uint32_t)(tau0 + (uint64_t)i);
        for (int kv = 0; kv < nV; kv++)
            for (int ku = 0; ku < nU; ku++)
                for (int kt = 0; kt < nT; kt++) {
                    uint32_t tile = (uint32_t)(kt + nT * (ku + nU * kv));
                    for (int v = 0; v < TV; v++)
                        for (int u = 0; u < TU; u++)
                            for (int t = 0; t < TT; t++) {
                                size_t g = (size_t)((ot + kt * TT + t) % LT) +
                                           (size_t)LT * ((size_t)((ou + ku * TU + u) % LU) + (size_t)LU * ((ov + kv * TV + v) % LV));
                                X.b[t + TT * (u + TU * v)] = lat[g];
                            }
                    for (int it = 0; it < n_iter; it++) {
                        uint32_t ow = philox_word(tau, tile, 0xFFFFFFFFu, (uint32_t)it, seed, stream);
                        for (int s = 0; s < 4; s++) {
                            uint32_t c = (ow >> (8 * s)) & cmask;
                            for (int k = 0; k < nslot; k++) perm[k] = k;
                            if (order == 1) for (int k = 0; k < nslot; k++) perm[k] = nslot - 1 - k;
                            if (order == 2) {
                                uint32_t x = 2654435761u * (uint32_t)(it * 4 + s + 1) + tile;
                                for (int k = nslot - 1; k > 0; k--) {
                                    x = x * 1664525u + 1013904223u;
                                    int j = (int)((x >> 8) % (uint32_t)(k + 1));
                                    int tmp = perm[k]; perm[k] = perm[j]; perm[j] = tmp;
                                }
                            }
                            for (int kk = 0; kk < nslot; kk++) {
                                int slot = perm[kk], t, u, v;
                                slot_site(sublat, TT, TU, slot, c, &t, &u, &v);
                                Draw rd = {(philox_word(tau, tile, (uint32_t)slot, (uint32_t)it, seed, stream) >> (8 * s)) & 0xff};
                                /* move of this slot: the substep's (TF, DSL, TF, TB), or under
                                   balanced rates bits 6-7 of the byte: 0,1 TF, 2 DSL, 3 TB, those two
                                   bits taken from the tile's offset byte (the same for every slot) */
                                if (rules) rd.fixed = (rd.fixed & 0x3fu) | ((ow >> (8 * s)) & 0xc0u);
                                int mv = s == 1 ? 1 : (s == 3 ? 2 : 0);
                                if (rules) mv = (rd.fixed >> 6) < 2 ? 0 : (int)(rd.fixed >> 6) - 1;
                                const int empty = X.b[t + TT * (u + TU * v)] == 0;
                                const int face = !empty && leaves_tile(&X, t, u, v, mv == 0 ? 0 : (mv == 1 ? 1 : 2), rd.fixed);
                                int acc;
                                if (mv == 1) acc = diffuse_sl(&X, t, u, v);
                                else if (mv == 2) acc = trans_back(&X, t, u, v, &rd);
                                else acc = trans_forw(&X, t, u, v, &rd);
                                if (g_counting) {
                                    g_counts[0] += (uint64_t)empty;
                                    g_counts[1] += (uint64_t)face;
                                    g_counts[2] += (uint64_t)(!empty && !face && acc);
                                }
                            }
                        }
                    }
                    for (int v = 0; v < TV; v++)
                        for (int u = 0; u < TU; u++)
                            for (int t = 0; t < TT; t++) {
                                size_t g = (size_t)((ot + kt * TT + t) % LT) +
                                           (size_t)LT * ((size_t)((ou + ku * TU + u) % LU) + (size_t)LU * ((ov + kv * TV + v) % LV));
                                lat[g] = X.b[t + TT * (u + TU * v)];
                            }
                }
    }
    free(perm);
    free(X.b);
    return 0;
}

/* orc_ring_run_rules with the reference's move rates */
int orc_ring_run(uint8_t* lat, int LT, int LU, int LV, int TT, int TU, int TV, int sublat, int n_iter,
                 int n_time, const int32_t* origins, uint64_t tau0, uint32_t seed, uint32_t stream, int order) {
    return orc_ring_run_rules(lat, LT, LU, LV, TT, TU, TV, sublat, n_iter, n_time, origins, tau0, seed, stream, order, 0);
}
