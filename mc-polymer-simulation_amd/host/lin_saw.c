/*
 * lin_saw.c -- random self-avoiding-walk start for the linear-chain melt.
 *
 * The reference starts linear chains as compact folded bars or cubes
 * (gpupol/gpupol_init.c:180-360, restated in lin_host.c).  This builder
 * instead grows every chain as a random self-avoiding walk on the FCC lattice
 * from the host LFSR113 stream of SimStateInit, so a seed fixes the start.
 *
 * Chain count and length follow GetBestPolymerPartitioning
 * (gpupol/gpupol_init.c:182-185): nPol = round(density * L^3 / N).  Each chain
 * occupies m distinct sites, m = round(siteFraction * L^3 / nPol) clipped to
 * [ceil((N+1)/2), N]; the other N - m monomers are stored length.  Stored
 * length sits in front of the first N - m bonds and is then spread uniformly
 * over the chain's bonds by RedistribSL, as after the reference builders
 * (gpupol/gpupol_init.c:484-532), so a site holds at most one stored-length
 * unit and the last site none.
 *
 * Growth: chains one after another; a chain starts at a uniformly drawn free
 * site and each step moves to a uniformly drawn free, untried neighbour among
 * the 12 FCC directions.  A dead end pops the last site (depth-first
 * backtracking); after 64*m pops the chain is abandoned, its sites freed and a
 * new start drawn.  More than 64 + nPol/4 abandoned chains in all end the
 * build with an error (the box is too full for this site fraction; at
 * density 1.2, L=128, N=500 fractions up to 0.8 build, 0.85 jams).  The walks
 * are an initial condition, not an equilibrium self-avoiding-walk ensemble:
 * the move engine equilibrates from them.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "gpupol_lin.h"
#include "elp_host_common.h"

static const uint kSawDirs[12] = {0x1, 0x2, 0x4, 0x5, 0x6, 0x7, 0x8, 0x9, 0xa, 0xb, 0xd, 0xe};

/* uniform integer in [0, n) from one host draw */
static uint saw_uniform(uint4_t* s, uint n) {
    uint k = (uint)((double)n * (double)ELP_RNG_FAC * elp_lfsr113(s));
    return k < n ? k : n - 1;
}

static uint saw_step(uint site, uint dir, const SimProperties* sp) {
    uint t = site % sp->LT, u = (site / sp->LT) % sp->LU, v = site / (sp->LT * sp->LU);
    AddUnitTUV(dir, &t, &u, &v, sp);
    return SiteIndex(t, u, v, sp);
}

int BuildSAWPolymers(SimProperties* sp, SimState* ss, int polLength, double density, double siteFraction) {
    const uint L3 = sp->L;
    const int N = polLength;
    const int nPol = (int)(density * L3 / (double)N + 0.5);
    if (N < 2 || nPol < 1) { printf("Error: no polymer to place\n"); return -1; }
    int m = (int)(siteFraction * L3 / (double)nPol + 0.5);
    if (m > N) m = N;
    if (m < (N + 2) / 2) m = (N + 2) / 2;
    if ((double)m * nPol > 0.95 * L3) {
        printf("Error: %i chains of %i sites do not fit into %u sites\n", nPol, m, L3);
        return -1;
    }
    printf("Self-avoiding walk start: %i chains, %i monomers on %i sites each (stored length %i)\n", nPol, N, m, N - m);

    uint8_t* occ = (uint8_t*)calloc(L3, 1);
    uint* path = (uint*)malloc(sizeof(uint) * (size_t)m);
    uint* dirs = (uint*)malloc(sizeof(uint) * (size_t)m);     /* bond from path[i] to path[i+1] */
    uint* tried = (uint*)malloc(sizeof(uint) * (size_t)m);    /* directions already tried from path[i] */
    uint* bonds = (uint*)malloc(sizeof(uint) * ((size_t)N + 1));
    uint4_t* rng = &ss->rngState;
    int rc = 0;
    sp->maxPolLength = (uint)polLength + 1;   /* sizes each chain's arrays in AddPolymerToSim */
    long restarts = 0;

    for (int p = 0; p < nPol && !rc; p++) {
        for (;;) {
            uint s0;
            do s0 = saw_uniform(rng, L3); while (occ[s0]);
            int len = 1;
            long pops = 0;
            path[0] = s0; tried[0] = 0; occ[s0] = 1;
            while (len < m && pops <= 64L * m) {
                uint cand[12], nc = 0;
                const uint cur = path[len - 1];
                for (uint d = 0; d < 12; d++) {
                    if (tried[len - 1] & (1u << d)) continue;
                    if (!occ[saw_step(cur, kSawDirs[d], sp)]) cand[nc++] = d;
                }
                if (!nc) {                         /* dead end: back up one site */
                    if (len == 1) break;
                    occ[cur] = 0;
                    len--;
                    pops++;
                    continue;
                }
                const uint d = cand[saw_uniform(rng, nc)];
                tried[len - 1] |= 1u << d;
                dirs[len - 1] = kSawDirs[d];
                path[len] = saw_step(cur, kSawDirs[d], sp);
                occ[path[len]] = 1;
                tried[len] = 0;
                len++;
            }
            if (len == m) break;
            for (int i = 0; i < len; i++) occ[path[i]] = 0;
            if (++restarts > 64L + nPol / 4) {
                printf("Error: self-avoiding walk start jammed at chain %i of %i\n", p, nPol);
                rc = -1;
                break;
            }
        }
        if (rc) break;
        /* stored length in front of the first N - m bonds, then the walk, then the chain end */
        int k = 0;
        for (int i = 0; i < m - 1; i++) {
            if (i < N - m) bonds[k++] = 0x0;
            bonds[k++] = dirs[i];
        }
        bonds[k++] = 0xf;
        if (k != N) { printf("Internal error: chain of %i monomers built as %i\n", N, k); rc = -1; break; }
        const uint t = path[0] % sp->LT, u = (path[0] / sp->LT) % sp->LU, v = path[0] / (sp->LT * sp->LU);
        for (int d = 0; d < sp->nDevices; d++) AddPolymerToSim(ss + d, sp, t, u, v, N, bonds);
    }
    if (restarts) printf("Self-avoiding walk start: %li chain restarts\n", restarts);
    free(occ); free(path); free(dirs); free(tried); free(bonds);
    if (rc) return rc;
    sp->maxPolLength = (uint)polLength + 1;
    if (GetAllRingPolymers(sp, ss)) return -1;
    for (int d = 0; d < sp->nDevices; d++)
        if (UpdatePolymerWithLabels(ss + d)) return -1;
    printf("------------------------------------------\n");
    return 0;
}
