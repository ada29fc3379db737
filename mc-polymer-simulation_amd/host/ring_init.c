/*
 * ring_init.c -- model initialisation of the ring drop-in.
 * Restates gpupol2/gpupol_init.c (file:line per function below), keeping the
 * host LFSR113 stream position exactly: the seed table draws of SimStateInit
 * come before the stored-length shuffle of RedistribSL, so the t=0 file for
 * a given (N, L, density, seed) is the reference's.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "gpupol_ring.h"
#include "elp_host_common.h"

/* gpupol2/gpupol_init.c:3-37 */
int SimStateInit(SimState* ss, SimProperties* sp) {
    uint nLane = (uint)(sp->nwg * sp->ws);
    uint* wg = (uint*)malloc(sizeof(uint) * sp->R);
    for (int d = 0; d < sp->nDevices; d++) {
        SimState* cs = ss + d;
        cs->lattice = (uint8_t*)calloc(sp->latSize, 1);
        cs->gpuLattice = (uint8_t*)calloc(sp->latSize, 1);
        cs->nPol = 0;
        cs->pol = NULL; cs->polNumTranslate = NULL; cs->label2Id = NULL; cs->id2Label = NULL;
        elp_lfsr113_seed_gpu(&cs->rngState, sp->seed, d);
        cs->labelHead = 0x10;
        cs->headBit = 5;
        cs->labelBit = 18;
        /* The device engine is counter-based and needs no seed table, but the
           draws are made (and kept) so the host stream position matches. */
        cs->seeds = (uint*)malloc(sizeof(uint) * (size_t)nLane * 2 * sp->R);
        for (size_t w = 0; w < sp->nwg; w++) {
            for (uint i = 0; i < sp->R; i++) wg[i] = elp_lfsr113(&cs->rngState);
            for (size_t l = 0; l < sp->ws; l++) {
                size_t g = 2 * (size_t)sp->R * (l + w * sp->ws);
                for (uint i = 0; i < sp->R; i++) {
                    cs->seeds[g + i] = elp_lfsr113(&cs->rngState);
                    cs->seeds[g + i + sp->R] = wg[i];
                }
            }
        }
    }
    free(wg);
    return 0;
}

void SimStateFree(SimState* ss) {
    for (int i = 0; i < ss->nPol; i++) { free(ss->pol[i].bonds); free(ss->pol[i].label); }
    free(ss->pol); free(ss->polNumTranslate); free(ss->label2Id); free(ss->id2Label);
    free(ss->lattice); free(ss->gpuLattice); free(ss->seeds);
    memset(ss, 0, sizeof *ss);
}

/* gpupol2/gpupol_init.c:39-41 */
void ClearLattice(SimState* ss, SimProperties* sp) { memset(ss->lattice, 0, sp->latSize); }

/* gpupol2/gpupol_init.c:43-62 -- box edges truncated to multiples of `unit`
   (the reference: 32, the tile edge; 16 allows boxes such as 48 for tiles of 16) */
void SetBoxDimensionUnit(SimProperties* sp, int LT, int LU, int LV, int unit) {
    const int E = unit > 0 ? unit : ELP_RING_TILE;
    if (LT % E || LU % E || LV % E) printf("Warning changing box dimensions to be a multiple of 24!\n");
    sp->nwt = LT / E; sp->nwu = LU / E; sp->nwv = LV / E;
    sp->nwg = (size_t)sp->nwt * sp->nwu * sp->nwv;
    sp->LT = E * sp->nwt; sp->LU = E * sp->nwu; sp->LV = E * sp->nwv;
    sp->L = sp->LT * sp->LU * sp->LV;
    sp->latSize = sp->L;
    sp->maxPolLength = sp->LT * sp->LU * 4;
}

void SetBoxDimension(SimProperties* sp, int LT, int LU, int LV) { SetBoxDimensionUnit(sp, LT, LU, LV, ELP_RING_TILE); }

/* gpupol2/gpupol_init.c:64-81 */
int SimPropDefaultInit(SimProperties* sp) {
    memset(sp, 0, sizeof *sp);
    sp->tMax = 50000;
    sp->R = 4;
    sp->writeInterval = 100;
    sp->polLength = 2000;
    sp->fastEq = 0;
    sp->density = 1.0;
    sp->curT = 0;
    sp->ws = ELP_RING_WS_EDGE * ELP_RING_WS_EDGE * ELP_RING_WS_EDGE;
    sp->seed = 239864012;
    sp->dirOut = NULL;
    sp->trans = (uint*)malloc(4 * sizeof(uint));
    sp->nSteps = (int)(0.8 * ELP_RING_LCELL * ELP_RING_LCELL * ELP_RING_LCELL);
    sp->equilibrated = 0;
    sp->double_step = 0;
    sp->nDevices = 1;
    SetGPUTrans(sp->trans);
    return 0;
}

/* gpupol2/gpupol_init.c:154-180 */
static void FindBestBoxSize(SimProperties* sp, int nPol, int polSize, int* dRes) {
    int LT = (int)sp->LT, LU = (int)sp->LU, LV = (int)sp->LV;
    const double idealSL = 0.3;
    double best = -9999999;
    for (int dt = 2; dt < LT; dt += 2) {
        int nt = LT / dt;
        for (int du = 1; du < LU; du++) {
            int nu = LU / du;
            for (int dv = 2; dv < LV; dv += 2) {
                int nv = LV / dv;
                int vol = dt * du * dv;
                if (vol * 2 > polSize && nt * nu * nv >= nPol) {
                    double sl = (polSize - vol) / (double)polSize;
                    double score = vol / (double)(dt + du + dv) - pow(vol, 2. / 3.) * pow(sl - idealSL, 2);
                    if (score > best) { dRes[0] = dt; dRes[1] = du; dRes[2] = dv; best = score; }
                }
            }
        }
    }
}

/* gpupol2/gpupol_init.c:183-248 */
int GetBestPolymerPartitioning(SimProperties* sp, SimState* ss, int polLength, double density) {
    double nDouble = density * sp->L / (double)polLength;
    double dPol = fabs(nDouble - (int)(nDouble + 0.5));
    int nPol = (int)(nDouble + 0.5);
    double dDensity = dPol * polLength / (double)sp->L;
    if (dDensity / density > 1e-5)
        printf("Warning: density is different from the wanted one:\n difference: %lf, density: %lf\n", dDensity, density);
    int dRes[3] = {0, 0, 0};
    FindBestBoxSize(sp, nPol, polLength, dRes);
    uint dt = (uint)dRes[0], du = (uint)dRes[1], dv = (uint)dRes[2];
    printf("dt,du,dv = (%i,%i,%i)\n", dt, du, dv);
    if (!dt || !du || !dv) { printf("Error: not enough space to put polymer into\n"); return 192; }
    uint nt = sp->LT / dt, nu = sp->LU / du, nv = sp->LV / dv;
    printf("N=(%u,%u,%u) = %u >= %i\n", nt, nu, nv, nt * nu * nv, nPol);
    if (nt * nu * nv < (uint)nPol) { printf("Error: not enough space to put polymer into\n"); return 192; }
    printf("Stored length density = %lf\n", (polLength - (dt * du * dv)) / (double)polLength);
    uint sl = (uint)polLength - dt * du * dv;
    if (sl >= (uint)polLength) { printf("Internal error: Failed to find polymer partition\n"); exit(1); }
    int nAdded = 0;
    /* the final maxPolLength already here: it sizes every chain's arrays (the box
       bound SetBoxDimension leaves, 4 LT LU, would give each chain megabytes) */
    sp->maxPolLength = polLength + 1;
    /* note: positions step by LT/nt (not dt), as in the reference */
    for (uint t = 0; t < nt * dt && nAdded < nPol; t += sp->LT / nt)
        for (uint u = 0; u < nu * du && nAdded < nPol; u += sp->LU / nu)
            for (uint v = 0; v < nv * dv && nAdded < nPol; v += sp->LV / nv) {
                for (int d = 0; d < sp->nDevices; d++) {
                    int rc = ConstructCubeRingPolymer(ss + d, sp, t, u, v, dt, du, dv, sl);
                    if (rc) return rc;
                }
                nAdded++;
            }
    sp->maxPolLength = polLength + 1;
    if (GetAllRingPolymers(sp, ss)) return 193;
    for (int d = 0; d < sp->nDevices; d++)
        if (UpdatePolymerWithLabels(ss + d)) return 123;
    printf("------------------------------------------\n");
    return 0;
}

/* gpupol2/gpupol_init.c:298-365: a compact ring filling a dt x du x dv
   block (boustrophedon in t/u, stacked and mirrored in v), then sl
   stored-length monomers inserted in front of the first sl bonds. */
int ConstructCubeRingPolymer(SimState* ss, SimProperties* sp, uint tStart, uint uStart, uint vStart,
                             uint dt, uint du, uint dv, uint sl) {
    uint halfT = dt / 2;
    uint tUnit = 0x1;
    const uint uUnit = 0x2, vUnit = 0x4;
    size_t cap = 2 * (size_t)dt * du * dv + 4;
    uint* b = (uint*)malloc(cap * sizeof(uint));
    uint* out = (uint*)malloc((cap + sl) * sizeof(uint));
    uint n = 0;
    /* one layer: serpentine over (t, u) in the first half of t */
    for (uint u = 0; u < du; u++) {
        for (uint t = 0; t + 1 < halfT; t++) b[n++] = tUnit;
        if (u + 1 < du) { b[n++] = uUnit; tUnit = (~tUnit) & 0xf; }
    }
    b[n++] = vUnit;
    /* come back on the next v layer: reversed, negated path */
    for (uint i = 0; i + 1 < n; i++) b[n + i] = (~b[n - i - 2]) & 0xf;
    n += n - 1;
    b[n++] = vUnit;
    /* repeat the two-layer block dv/2 times */
    for (uint k = 1; k < dv / 2; k++) memcpy(b + k * n, b, sizeof(uint) * n);
    n = n * (dv / 2) - 1;
    b[n++] = 0xe;
    /* second half in t: mirror in u and v, keep t */
    for (uint i = 0; i + 1 < n; i++) {
        uint src = b[n - i - 2];
        b[n + i] = (src == 0x1 || src == 0xe) ? src : ((~src) & 0xf);
    }
    n += n - 1;
    b[n++] = 0x1;   /* closes the ring */
    uint m = 0, i;
    for (i = 0; i < sl; i++) { out[m++] = 0x0; out[m++] = b[i]; }
    for (; i < n; i++) out[m++] = b[i];
    if (sl + dt * du * dv != m) {
        printf("Internal error: Failed to create ring polymer, length not equal to its desired length (%i vs %i vs %i vs %i)\n",
               dt * du * dv + sl, m, n, sl);
        exit(1);
    }
    int rc = AddPolymerToSim(ss, sp, tStart, uStart, vStart, (int)m, out);
    free(b); free(out);
    return rc;
}

/* gpupol2/gpupol_init.c:415-425: does the 18-bit label create a false head? */
int UniqueLabel(int startLabel, int nLabelBit, int head, int headBit) {
    int mask = (1 << headBit) - 1;
    long tot = (long)head | ((long)startLabel << headBit) | ((long)head << (headBit + nLabelBit));
    for (int i = 1; i < headBit + nLabelBit; i++)
        if (((tot >> i) & mask) == head) return 0;
    return 1;
}

/* gpupol2/gpupol_init.c:427-499: labels = 18 id bits (MSB first),
   alternating 1,0,... and the head 1,0,0,0,0 at the end. */
int AddPolymerToSim(SimState* ss, SimProperties* sp, uint t, uint u, uint v, int length, uint* bonds) {
    Polymer* pol;
    if (!ss->nPol) {
        ss->pol = (Polymer*)malloc(sizeof(Polymer));
        ss->polNumTranslate = (int*)malloc(sizeof(int));
        ss->label2Id = (int*)malloc((1 << ss->labelBit) * sizeof(int));
        ss->id2Label = (int*)malloc((1 << ss->labelBit) * sizeof(int));
        for (int i = 0; i < (1 << ss->labelBit); i++) { ss->label2Id[i] = -1; ss->id2Label[i] = -1; }
    } else {
        ss->pol = (Polymer*)realloc(ss->pol, sizeof(Polymer) * (ss->nPol + 1));
        ss->polNumTranslate = (int*)realloc(ss->polNumTranslate, sizeof(int) * (ss->nPol + 1));
    }
    pol = ss->pol + ss->nPol;
    ss->nPol++;
    /* room for the chain and for the head + id label bits written below */
    size_t cap = sp->maxPolLength > (uint)length + 1 ? sp->maxPolLength : (uint)length + 1;
    if (cap < (size_t)(ss->headBit + ss->labelBit + 1)) cap = (size_t)(ss->headBit + ss->labelBit + 1);
    pol->label = (uint*)calloc(cap, sizeof(uint));
    pol->bonds = (uint*)calloc(cap, sizeof(uint));
    pol->labelStart = 0;
    int lab = (ss->nPol == 1) ? 0 : ss->id2Label[ss->nPol - 2] + 1;
    while (!UniqueLabel(lab, ss->labelBit, ss->labelHead, ss->headBit)) lab++;
    if (lab >= (1 << ss->labelBit)) {
        printf("Ran out of labels, please make a smaller box/longer polymers!\n");
        ss->nPol--;
        free(pol->label); free(pol->bonds);
        return 2;
    }
    for (int i = 0; i < ss->labelBit; i++) pol->label[i] = (lab >> (ss->labelBit - 1 - i)) & 0x1;
    int alt = 1;
    for (int i = ss->labelBit; i < length - ss->headBit; i++, alt ^= 1) pol->label[i] = alt;
    pol->label[length - ss->headBit] = 1;
    for (int i = length - ss->headBit + 1; i < length; i++) pol->label[i] = 0;
    ss->id2Label[ss->nPol - 1] = lab;
    ss->label2Id[lab] = ss->nPol - 1;
    ss->polNumTranslate[ss->nPol - 1] = ss->nPol - 1;
    pol->length = 0;
    if (SetBondVecs(ss->lattice, t, u, v, bonds, (uint)length, pol->label, sp)) return 1;
    memcpy(pol->bonds, bonds, sizeof(uint) * length);
    pol->length = length;
    pol->startTUV.t = t; pol->startTUV.u = u; pol->startTUV.v = v;
    pol->startLeftSL = 0;
    return 0;
}

/* gpupol2/gpupol_init.c:501-550: uniform reshuffle of the stored-length
   units over the bonds of each ring (inside-out Fisher-Yates on LFSR113). */
void RedistribSL(SimState* ss, SimProperties* sp) {
    uint* nb = (uint*)malloc(sizeof(uint) * (sp->maxPolLength + 2));
    uint* dist = (uint*)calloc(sp->maxPolLength + 2, sizeof(uint));
    for (int d = 0; d < sp->nDevices; d++) {
        ClearLattice(ss + d, sp);
        for (int p = 0; p < ss[d].nPol; p++) {
            Polymer* pol = ss[d].pol + p;
            int nBond = 0, nSl = 0;
            for (int i = 0; i < pol->length; i++) {
                if (pol->bonds[i] == 0x0) nSl++;
                else if (pol->bonds[i] != 0xf) nBond++;
            }
            for (int i = 0; i < nBond; i++) {
                int j = (int)(elp_lfsr113(&ss->rngState) % (uint)(i + 1));   /* device 0's stream */
                dist[i] = dist[j];
                dist[j] = (i < nSl) ? 1 : 0;
            }
            for (int i = 0, o = 0, s = 0; i < nBond; i++, o++, s++) {
                if (pol->bonds[s] == 0x0) s++;
                if (dist[i] == 1) nb[o++] = 0x0;
                nb[o] = pol->bonds[s];
            }
            memcpy(pol->bonds, nb, sizeof(uint) * (nBond + nSl));
        }
        SetLatticeFromSS(ss + d, sp);
    }
    free(nb); free(dist);
}

/* gpupol2/gpupol_init.c:552-584: for each valid sum s, the first bond of
   each of its two unordered decompositions, nibble 2*(s%4)+k of word s/4. */
int SetGPUTrans(uint* trans) { return elp_set_gpu_trans(trans); }

int elp_set_gpu_trans(uint trans[4]) {
    uint nFound[16] = {0};
    uint sum;
    for (int i = 0; i < 4; i++) trans[i] = 0;
    for (uint a = 1; a < 15; a++) {
        if (!elp_bond_valid(a)) continue;
        for (uint b = a + 1; b < 15; b++) {
            if (!elp_bond_valid(b)) continue;
            if (!elp_bond_add_valid(a, b, &sum)) continue;
            trans[sum / 4] |= a << (8 * (sum % 4) + 4 * nFound[sum]);
            if (++nFound[sum] > 2) {
                printf("Internal error: setting trans reactions failed: nFound[0x%x] = %i\n", sum, nFound[sum]);
                exit(1);
            }
        }
    }
    for (uint i = 1; i < 15; i++)
        if (i != 0x3 && i != 0xc && nFound[i] != 2) { printf("Error %i (%i)\n", i, nFound[i]); exit(1); }
    return 0;
}

long elp_t_from_name(const char* file) {
    /* strip the directory, cut at the first '_' after it, parse "t=<T>" */
    const char* base = strrchr(file, '/');
    base = base ? base + 1 : file;
    const char* p = strstr(base, "t=");
    if (!p) return 0;
    return strtol(p + 2, NULL, 10);
}
