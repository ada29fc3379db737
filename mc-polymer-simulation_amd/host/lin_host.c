/*
 * lin_host.c -- model, chain recovery and IO of the linear-chain drop-in.
 * Restates gpupol/gpupol_init.c, gpupol/oclpol.c and gpupol/gpupol_io.c
 * (file:line per function), keeping the host LFSR113 stream position.
 */
#include <dirent.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include "gpupol_lin.h"
#include "elp_host_common.h"

uint SiteIndex(uint t, uint u, uint v, const SimProperties* sp) { return t + sp->LT * (u + sp->LU * v); }

/* gpupol/oclpol.c:106-124 */
void AddUnitTUV(uint unit, uint* t, uint* u, uint* v, const SimProperties* sp) {
    int dt, du, dv;
    elp_bond_delta(unit, &dt, &du, &dv);
    *t = (uint)(((int)*t + dt + (int)sp->LT) % (int)sp->LT);
    *u = (uint)(((int)*u + du + (int)sp->LU) % (int)sp->LU);
    *v = (uint)(((int)*v + dv + (int)sp->LV) % (int)sp->LV);
}

/* gpupol/gpupol_init.c:3-38 */
int SimStateInit(SimState* ss, SimProperties* sp) {
    uint* wg = (uint*)malloc(sizeof(uint) * sp->R);
    for (int d = 0; d < sp->nDevices; d++) {
        SimState* cs = ss + d;
        cs->lattice = (uint16_t*)calloc(sp->L, sizeof(uint16_t));
        cs->gpuLattice = (uint16_t*)calloc(sp->L, sizeof(uint16_t));
        cs->nPol = 0;
        cs->pol = NULL; cs->polNumTranslate = NULL; cs->label2Id = NULL; cs->id2Label = NULL;
        elp_lfsr113_seed_gpu(&cs->rngState, sp->seed, d);
        cs->labelHead = 0x10; cs->headBit = 5; cs->labelBit = 18;
        cs->seeds = (uint*)malloc(sizeof(uint) * sp->nwg * sp->ws * 2 * sp->R);
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

/* gpupol/gpupol_init.c:46-65: box edges truncated to multiples of `unit`
   (the reference: 32, the tile edge) */
void SetBoxDimensionUnit(SimProperties* sp, int LT, int LU, int LV, int unit) {
    const int E = unit > 0 ? unit : ELP_LIN_TILE;
    sp->nwt = LT / E; sp->nwu = LU / E; sp->nwv = LV / E;
    sp->nwg = (size_t)sp->nwt * sp->nwu * sp->nwv;
    sp->LT = E * sp->nwt; sp->LU = E * sp->nwu; sp->LV = E * sp->nwv;
    sp->L = sp->LT * sp->LU * sp->LV;
    sp->latSize = sp->L;
    sp->gpuLatSize = sp->L;
    sp->maxPolLength = sp->LT * sp->LU * 4;
}

void SetBoxDimension(SimProperties* sp, int LT, int LU, int LV) { SetBoxDimensionUnit(sp, LT, LU, LV, ELP_LIN_TILE); }

/* gpupol/gpupol_init.c:67-84 */
int SimPropDefaultInit(SimProperties* sp) {
    memset(sp, 0, sizeof *sp);
    sp->tMax = 50000;
    sp->R = 4;
    sp->writeInterval = 100;
    sp->polLength = 2000;
    sp->fastEq = 0;
    sp->density = 1.0;
    sp->curT = 0;
    sp->ws = 512;
    sp->seed = 239864012;
    sp->trans = (uint*)malloc(4 * sizeof(uint));
    sp->nSteps = 64;
    sp->equilibrated = 0;
    sp->double_step = 0;
    sp->nDevices = 1;
    elp_set_gpu_trans(sp->trans);
    return 0;
}

/* gpupol/gpupol_init.c:152-177 -- note the stored-length estimate uses the
   number of boxes nt*nu*nv, as in the reference */
static void FindBestBoxSize(SimProperties* sp, int nPol, int polSize, int* dRes) {
    int LT = (int)sp->LT, LU = (int)sp->LU, LV = (int)sp->LV;
    double best = -9999999;
    for (int dt = 2; dt < LT; dt += 2) {
        int nt = LT / dt;
        for (int du = 1; du < LU; du++) {
            int nu = LU / du;
            for (int dv = 2; dv < LV; dv += 2) {
                int nv = LV / dv;
                int vol = dt * du * dv;
                if (vol * 2 > polSize && nt * nu * nv >= nPol) {
                    double sl = (polSize - nt * nu * nv) / (double)polSize;
                    double score = vol / (double)(dt + du + dv) - pow(vol, 2. / 3.) * pow(sl - 0.3, 2);
                    if (score > best) { dRes[0] = dt; dRes[1] = du; dRes[2] = dv; best = score; }
                }
            }
        }
    }
}

/* gpupol/gpupol_init.c:245-291: two-row hairpin of 2*dt sites ending in 0xf */
static int ConstructBarPolymer(SimState* ss, SimProperties* sp, int polId, int length) {
    int nt = (int)sp->LT / ((length + 1) / 2 - 1) + 1;
    int nu = (int)sp->LU / 2, nv = (int)sp->LV;
    int dt = (int)sp->LT / nt, du = 2, dv = 1;
    if (polId >= nt * nu * nv) { printf("Error constructing polymers: Too little space!\n"); return 120; }
    int sl = length - dt * du * dv;
    int t = (polId % nt) * dt, u = ((polId / nt) % nu) * du, v = (polId / (nt * nu)) * dv;
    uint* b = (uint*)calloc((size_t)length + 2 * dt + 2, sizeof(uint));
    uint* out = (uint*)calloc((size_t)length + 1, sizeof(uint));
    int m = 0;
    for (int i = 0; i < dt - 1; i++) b[m++] = 0x1;
    b[m++] = 0x2;
    for (int i = 0; i < dt - 1; i++) b[m++] = 0xe;
    b[m++] = 0xf;
    for (int i = 0; i < sl; i++) { out[2 * i] = 0x0; out[2 * i + 1] = b[i]; }
    for (int i = 2 * sl; i < length; i++) out[i] = b[i - sl];
    int rc = AddPolymerToSim(ss, sp, (uint)t, (uint)u, (uint)v, length, out);
    free(b); free(out);
    return rc;
}

/* gpupol/gpupol_init.c:293-360: the compact cube path of the ring builder,
   opened at the end (last bond 0xf). */
static int ConstructCubePolymer(SimState* ss, SimProperties* sp, uint t0, uint u0, uint v0, uint dt, uint du,
                                uint dv, uint sl) {
    uint halfT = dt / 2, tUnit = 0x1;
    size_t cap = 2 * (size_t)dt * du * dv + 4;
    uint* b = (uint*)malloc(cap * sizeof(uint));
    uint* out = (uint*)malloc((cap + sl) * sizeof(uint));
    uint n = 0;
    for (uint u = 0; u < du; u++) {
        for (uint t = 0; t + 1 < halfT; t++) b[n++] = tUnit;
        if (u + 1 < du) { b[n++] = 0x2; tUnit = (~tUnit) & 0xf; }
    }
    b[n++] = 0x4;
    for (uint i = 0; i + 1 < n; i++) b[n + i] = (~b[n - i - 2]) & 0xf;
    n += n - 1;
    b[n++] = 0x4;
    for (uint k = 1; k < dv / 2; k++) memcpy(b + k * n, b, sizeof(uint) * n);
    n = n * (dv / 2) - 1;
    b[n++] = 0xe;
    for (uint i = 0; i + 1 < n; i++) {
        uint s = b[n - i - 2];
        b[n + i] = (s == 0x1 || s == 0xe) ? s : ((~s) & 0xf);
    }
    n += n - 1;
    b[n++] = 0xf;
    uint m = 0, i;
    for (i = 0; i < sl; i++) { out[m++] = 0x0; out[m++] = b[i]; }
    for (; i < n; i++) out[m++] = b[i];
    if (sl + dt * du * dv != m) { printf("Internal error: Failed to create polymer (%u vs %u)\n", dt * du * dv + sl, m); exit(1); }
    int rc = AddPolymerToSim(ss, sp, t0, u0, v0, (int)m, out);
    free(b); free(out);
    return rc;
}

/* gpupol/gpupol_init.c:180-243 */
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
    /* the final maxPolLength already here: it sizes every chain's arrays (the box
       bound SetBoxDimension leaves, 4 LT LU, would give each chain megabytes) */
    sp->maxPolLength = polLength + 1;
    if (polLength < 100) {
        for (int d = 0; d < sp->nDevices; d++)
            for (int i = 0; i < nPol; i++) {
                int rc = ConstructBarPolymer(ss + d, sp, i, polLength);
                if (rc) return rc;
            }
    } else {
        int nAdded = 0;
        for (uint t = 0; t < nt * dt && nAdded < nPol; t += dt)
            for (uint u = 0; u < nu * du && nAdded < nPol; u += du)
                for (uint v = 0; v < nv * dv && nAdded < nPol; v += dv) {
                    for (int d = 0; d < sp->nDevices; d++) {
                        int rc = ConstructCubePolymer(ss + d, sp, t, u, v, dt, du, dv, sl);
                        if (rc) return rc;
                    }
                    nAdded++;
                }
    }
    sp->maxPolLength = polLength + 1;
    if (GetAllRingPolymers(sp, ss)) return 193;
    for (int d = 0; d < sp->nDevices; d++)
        if (UpdatePolymerWithLabels(ss + d)) return 123;
    printf("------------------------------------------\n");
    return 0;
}

static int UniqueLabel(int lab, int nLabelBit, int head, int headBit) {
    int mask = (1 << headBit) - 1;
    long tot = (long)head | ((long)lab << headBit) | ((long)head << (headBit + nLabelBit));
    for (int i = 1; i < headBit + nLabelBit; i++)
        if (((tot >> i) & mask) == head) return 0;
    return 1;
}

/* gpupol/gpupol_init.c:422-482: labels = head 1,0,0,0,0, 18 id bits (MSB
   first), then alternating 1,0,... */
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
    pol->label[0] = 1;
    for (int i = 1; i < ss->headBit; i++) pol->label[i] = 0;
    for (int i = ss->headBit; i < ss->headBit + ss->labelBit; i++)
        pol->label[i] = (lab >> (ss->headBit + ss->labelBit - 1 - i)) & 0x1;
    ss->id2Label[ss->nPol - 1] = lab;
    ss->label2Id[lab] = ss->nPol - 1;
    int alt = 1;
    for (int i = ss->labelBit + ss->headBit; i < length; i++, alt ^= 1) pol->label[i] = alt;
    ss->polNumTranslate[ss->nPol - 1] = ss->nPol - 1;
    pol->length = 0;
    pol->polCoor = NULL;
    if (SetBondVecs(ss, sp, t, u, v, bonds, (uint)length, pol->label)) return 1;
    memcpy(pol->bonds, bonds, sizeof(uint) * length);
    pol->length = length;
    pol->startTUV.t = t; pol->startTUV.u = u; pol->startTUV.v = v;
    pol->startLeftSL = 0;
    return 0;
}

/* gpupol/oclpol.c:126-162 (the chain-end monomer carries no label bit) */
int SetBondVecs(SimState* ss, SimProperties* sp, uint t, uint u, uint v, uint* bonds, uint polSize, uint* labels) {
    uint16_t* lat = ss->lattice;
    uint s = SiteIndex(t, u, v, sp);
    for (uint i = 0; i < polSize; i++) {
        if (bonds[i] == 0x0) {
            lat[s] |= (uint16_t)(0x100 | ((labels[i] & 1) << 10));
        } else if (bonds[i] == 0xf) {
            break;
        } else {
            if (LW_NEXT(lat[s])) { printf("Error setting bond vec: site already occupied\n(%i %i %i)\n", t, u, v); return 1; }
            lat[s] |= (uint16_t)(bonds[i] | ((labels[i] & 1) << 9));
            AddUnitTUV(bonds[i], &t, &u, &v, sp);
            s = SiteIndex(t, u, v, sp);
            if (LW_PREV(lat[s])) { printf("Error setting bond vec: site already occupied[2]\n(%i %i %i)\n", t, u, v); return 1; }
            lat[s] |= (uint16_t)(bonds[i] << 4);
        }
    }
    return 0;
}

/* gpupol/gpupol_init.c:484-532 */
void RedistribSL(SimState* ss, SimProperties* sp) {
    printf("Redistributing stored length distribution\n");
    uint* nb = (uint*)malloc(sizeof(uint) * (sp->maxPolLength + 2));
    uint* dist = (uint*)calloc(sp->maxPolLength + 2, sizeof(uint));
    for (int d = 0; d < sp->nDevices; d++) {
        memset(ss[d].lattice, 0, sizeof(uint16_t) * sp->L);
        for (int p = 0; p < ss[d].nPol; p++) {
            Polymer* pol = ss[d].pol + p;
            int nBond = 0, nSl = 0;
            for (int i = 0; i < pol->length; i++) {
                if (pol->bonds[i] == 0x0) nSl++;
                else if (pol->bonds[i] != 0xf) nBond++;
            }
            for (int i = 0; i < nBond; i++) {
                int j = (int)(elp_lfsr113(&ss->rngState) % (uint)(i + 1));
                dist[i] = dist[j];
                dist[j] = (i < nSl) ? 1 : 0;
            }
            for (int i = 0, o = 0, s = 0; i < nBond; i++, o++, s++) {
                if (pol->bonds[s] == 0x0) s++;
                if (dist[i] == 1) nb[o++] = 0x0;
                nb[o] = pol->bonds[s];
            }
            memcpy(pol->bonds, nb, sizeof(uint) * (nBond + nSl));
            SetBondVecs(ss + d, sp, pol->startTUV.t, pol->startTUV.u, pol->startTUV.v, pol->bonds, (uint)pol->length,
                        pol->label);
        }
    }
    printf("Finished redistributing stored length\n");
    free(nb); free(dist);
}

/* gpupol/oclpol.c:380-417: walk backwards (towards monomer 0) shifting label
   bits in from the top until the head 1,0,0,0,0 or the chain start. */
static int FindPolymerStart(const SimProperties* sp, const SimState* ss, uint* t, uint* u, uint* v, Polymer* pol,
                            const uint16_t* lat) {
    int reg = 0, i = 0;
    const int top = ss->headBit - 1;
    while (i < 100000) {
        uint w = lat[SiteIndex(*t, *u, *v, sp)];
        uint lab = LW_LABELS(w);
        pol->startLeftSL = 0;
        if (LW_SL(w)) {
            reg = (reg >> 1) | (int)((lab & 1) << top);
            i++;
            if (reg == ss->labelHead && i >= ss->headBit) return 0;
            lab >>= 1;
            pol->startLeftSL = 1;
        }
        reg = (reg >> 1) | (int)((lab & 1) << top);
        i++;
        if (reg == ss->labelHead && i >= ss->headBit) return 0;
        uint prev = LW_PREV(w);
        if (!prev) return 0;           /* chain start */
        AddUnitTUV((~prev) & 0xf, t, u, v, sp);
    }
    printf("Error: Did not find label head\n");
    return -1;
}

/* gpupol/oclpol.c:233-287: read forward from the start to the chain end;
   clears visited sites in the scratch copy `lat` */
static int GetLinPolymer(const SimProperties* sp, const SimState* ss, uint t, uint u, uint v, Polymer* pol,
                         uint16_t* lat, const uint16_t* orig) {
    if (FindPolymerStart(sp, ss, &t, &u, &v, pol, orig)) return -1;
    pol->startTUV.t = t; pol->startTUV.u = u; pol->startTUV.v = v;
    memset(pol->label, 0, sp->maxPolLength * sizeof(uint));
    int n = 0, leftSL = pol->startLeftSL;
    for (;;) {
        uint s = SiteIndex(t, u, v, sp);
        uint lab = LW_LABELS(orig[s]);
        if (leftSL == 1) {
            if (LW_SL(orig[s])) { pol->label[n] = lab >> 1; pol->bonds[n] = 0; }
            else n--;
        } else {
            pol->label[n] = lab & 1;
            uint nx = LW_NEXT(lat[s]);
            lat[s] &= (uint16_t)~0xff;
            if (!nx) { pol->bonds[n] = 0xf; return n + 1; }
            pol->bonds[n] = nx;
            AddUnitTUV(nx, &t, &u, &v, sp);
        }
        n++;
        leftSL ^= 1;
        if ((uint)n == sp->maxPolLength) { printf("polymer too long\n"); return -1; }
    }
}

/* gpupol/oclpol.c:187-231: scan with v fastest */
int GetAllRingPolymers(SimProperties* sp, SimState* ss) {
    uint16_t* copy = (uint16_t*)malloc(sizeof(uint16_t) * sp->L);
    int rc = 0;
    for (int d = 0; d < sp->nDevices && !rc; d++) {
        SimState* cs = ss + d;
        memcpy(copy, cs->lattice, sizeof(uint16_t) * sp->L);
        int n = 0;
        for (uint t = 0; t < sp->LT && !rc; t++)
            for (uint u = 0; u < sp->LU && !rc; u++)
                for (uint v = 0; v < sp->LV && !rc; v++) {
                    if (!(copy[SiteIndex(t, u, v, sp)] & 0xff)) continue;
                    if (n == cs->nPol) { printf("Too many polymers\n"); rc = -1; break; }
                    Polymer* pol = cs->pol + n;
                    int len = GetLinPolymer(sp, cs, t, u, v, pol, copy, cs->lattice);
                    if (len < 0) { rc = -1; break; }
                    pol->length = len;
                    n++;
                }
        if (!rc && n != cs->nPol) { printf("Found %i polymers, expected %i\n", n, cs->nPol); rc = -1; }
    }
    free(copy);
    return rc;
}

/* gpupol/oclpol.c:700-730: id bits are labels 5..22 */
int UpdatePolymerWithLabels(SimState* ss) {
    int* seen = (int*)calloc(ss->nPol, sizeof(int));
    int fail = 0;
    for (int p = 0; p < ss->nPol; p++) {
        Polymer* pol = ss->pol + p;
        int lab = 0;
        for (int i = ss->headBit; i < ss->headBit + ss->labelBit; i++)
            lab |= (int)pol->label[i] << (ss->headBit + ss->labelBit - i - 1);
        int id = ss->label2Id[lab];
        if (id < 0) { printf("Unknown polymer label %i (length %i)\n", lab, pol->length); free(seen); return -1; }
        ss->polNumTranslate[p] = id;
        seen[id]++;
    }
    for (int i = 0; i < ss->nPol; i++)
        if (seen[i] != 1) { printf("Account failure: %i %i\n", i, seen[i]); fail++; }
    free(seen);
    return fail ? -1 : 0;
}

/* gpupol/gpupol_io.c:91-149; box edges must be multiples of `unit` (the
   reference: 32), else 202 is returned */
int ReadLatticeFileUnit(SimState* ss, SimProperties* sp, const char* file, int unit) {
    FILE* f = fopen(file, "r");
    if (!f) { printf("Error reading file %s\n", file); return 199; }
    int LT, LU, LV, np = 0, maxLen = 0;
    if (fscanf(f, "%*s %i", &LT) != 1 || fscanf(f, "%*s %i", &LU) != 1 || fscanf(f, "%*s %i", &LV) != 1) {
        fclose(f); return 199;
    }
    const int E = unit > 0 ? unit : ELP_LIN_TILE;
    if (LT <= 0 || LU <= 0 || LV <= 0 || LT % E || LU % E || LV % E) {
        printf("Error: Box dimensions not correct. Work group dimensions: (%i, %i, %i) vs file (%i, %i, %i)\n",
               E, E, E, LT, LU, LV);
        fclose(f);
        return 202;
    }
    SetBoxDimensionUnit(sp, LT, LU, LV, E);
    SimStateInit(ss, sp);
    if (fscanf(f, "%*s %i", &np) != 1 || fscanf(f, "%*s %i", &maxLen) != 1 || np < 0 || maxLen <= 0) { fclose(f); return 199; }
    sp->maxPolLength = (uint)maxLen;
    uint* bonds = (uint*)malloc(sizeof(uint) * (maxLen + 1));
    char* text = (char*)malloc((size_t)maxLen + 2);
    char fmt[32];
    snprintf(fmt, sizeof fmt, "%%%ds", maxLen + 1);
    int longest = 0;
    for (int p = 0; p < np; p++) {
        int len; uint t, u, v;
        if (fscanf(f, "%*s %i", &len) != 1 || fscanf(f, "%u %u %u", &t, &u, &v) != 3 || fscanf(f, fmt, text) != 1 ||
            len <= 0 || len > maxLen || (int)strlen(text) < len || t >= sp->LT || u >= sp->LU || v >= sp->LV) {
            printf("Error: malformed polymer %i in %s\n", p, file);
            fclose(f); free(bonds); free(text); return 199;
        }
        for (int i = 0; i < len; i++) {
            char c = text[i];
            bonds[i] = (c >= '0' && c <= '9') ? (uint)(c - '0') : (uint)(10 + (c | 0x20) - 'a');
            if (bonds[i] > 0xf) {
                printf("Error: bad bond character in %s\n", file);
                fclose(f); free(bonds); free(text); return 199;
            }
        }
        if (len > longest) longest = len;
        if (AddPolymerToSim(ss, sp, t, u, v, len, bonds)) {
            printf("Error: polymer %i of %s overlaps another one\n", p, file);
            fclose(f); free(bonds); free(text); return 199;
        }
    }
    fclose(f);
    free(bonds); free(text);
    /* files whose header gives the chain length itself as maxPolLength (efpol
       writes them so) would stop the chain walk one monomer short */
    if (longest >= (int)sp->maxPolLength) sp->maxPolLength = (uint)longest + 1;
    printf("Succesfully read file %s\n", file);
    return 0;
}

int ReadLatticeFile(SimState* ss, SimProperties* sp, const char* file) {
    return ReadLatticeFileUnit(ss, sp, file, ELP_LIN_TILE);
}

/* gpupol/gpupol_io.c:151-187 */
int WriteLatticeFile(SimProperties* sp, SimState* ss, const char* file) {
    FILE* f = fopen(file, "w");
    if (!f) { printf("Error opening file %s\n", file); return -1; }
    int* rev = (int*)malloc(sizeof(int) * (ss->nPol ? ss->nPol : 1));
    fprintf(f, "LT= %i\nLU= %i\nLV= %i\nnp= %i\nmaxPolLength= %i\n", sp->LT, sp->LU, sp->LV, ss->nPol, sp->maxPolLength);
    for (int p = 0; p < ss->nPol; p++) rev[ss->polNumTranslate[p]] = p;
    for (int p = 0; p < ss->nPol; p++) {
        Polymer* pol = ss->pol + rev[p];
        fprintf(f, "len= %i\n%u  %u  %u\t", pol->length, pol->startTUV.t, pol->startTUV.u, pol->startTUV.v);
        for (int i = (int)pol->labelStart; i < pol->length; i++) fprintf(f, "%x", pol->bonds[i]);
        for (int i = 0; i < (int)pol->labelStart; i++) fprintf(f, "%x", pol->bonds[i]);
        fprintf(f, "\n");
    }
    free(rev);
    fclose(f);
    return 0;
}

/* gpupol/gpupol_io.c:32-60 */
void WriteSimulationSettings(SimProperties* sp, SimState* ss) {
    char file[4096];
    snprintf(file, sizeof file, "%s/simulation_settings.txt", sp->dirOut);
    FILE* f = fopen(file, "w");
    if (!f) return;
    fprintf(f, "Start_seed = %u\nLength = %i\nPolytype = lin\nDensity = %lf\nLatsize = %i\n", sp->seed, sp->polLength,
            sp->density, sp->LT);
    fprintf(f, "Start_polysize = %i\nDouble_step = %i\nInterval = %li\nEquilibrated = %i\nNpol = %i\n", sp->polLength,
            sp->double_step, sp->writeInterval, sp->equilibrated, ss->nPol);
    fprintf(f, "Polymodel = sl_equal\nExecutable = gpupol\nFast_equilibration = %i\nFEQ_timestep = %li\n",
            sp->fastEq ? 1 : 0, sp->fastEq);
    fprintf(f, "Release = mi355x-elp\n");
    fclose(f);
}

static void mkdir_p(const char* path) {
    char buf[4096];
    snprintf(buf, sizeof buf, "%s", path);
    for (char* p = buf + 1; *p; p++)
        if (*p == '/') { *p = 0; mkdir(buf, 0777); *p = '/'; }
    mkdir(buf, 0777);
}

/* Fraction of the sites a fresh self-avoiding-walk start covers, from
   ELP_LIN_INIT ("saw" or "saw:FRACTION"); 0 = the reference's start. */
static double saw_fraction_from_env(void) {
    const char* init = getenv("ELP_LIN_INIT");
    if (!init || strncmp(init, "saw", 3)) return 0;
    return init[3] == ':' ? atof(init + 4) : ELP_SAW_SITE_FRACTION;
}

/* gpupol/gpupol_init.c:86-150.  The device context of GPULibInit stays the
   library's own: one device, index 0. */
void LoadPolymers(SimProperties* sp, SimState* ss) {
    char file[8192];
    GPULibContext ctx;
    GPULibInit(sp, NULL, &ctx);
    const double sawFraction = saw_fraction_from_env();
    mkdir_p(sp->dirOut);
    long highT = -1;
    int nHigh = -1;
    char* highFiles[ELP_N_MAX_GPU];
    DIR* dir = opendir(sp->dirOut);
    struct dirent* de;
    while (dir && (de = readdir(dir))) {
        if (strncmp(de->d_name, "t=", 2)) continue;
        long tf = elp_t_from_name(de->d_name);
        if (tf > highT) { for (int i = 0; i < nHigh; i++) free(highFiles[i]); nHigh = 0; }
        if (tf >= highT && nHigh < ELP_N_MAX_GPU) { highFiles[nHigh++] = strdup(de->d_name); highT = tf; }
    }
    if (dir) closedir(dir);
    if (nHigh == sp->nDevices) {
        printf("Starting from t=%li\n", highT);
        sp->curT = highT;
        sp->tMax += highT;
        for (int i = 0; i < sp->nDevices; i++) {
            snprintf(file, sizeof file, "%s/%s", sp->dirOut, highFiles[i]);
            int rc = ReadLatticeFile(ss + i, sp, file);
            if (rc) exit(rc);
            if (UpdatePolymerWithLabels(ss + i)) exit(123);
        }
    } else {
        SetBoxDimension(sp, (int)sp->LT, (int)sp->LU, (int)sp->LV);
        SimStateInit(ss, sp);
        if (sawFraction > 0) {
            if (BuildSAWPolymers(sp, ss, sp->polLength, sp->density, sawFraction)) exit(192);
        } else {
            int rc = GetBestPolymerPartitioning(sp, ss, sp->polLength, sp->density);
            if (rc) exit(rc);
        }
        RedistribSL(ss, sp);
        for (int i = 0; i < sp->nDevices; i++) {
            snprintf(file, sizeof file, "%s/t=0_dev=%i.res", sp->dirOut, i);
            WriteLatticeFile(sp, ss + i, file);
        }
    }
    for (int i = 0; i < nHigh; i++) free(highFiles[i]);
}

long elp_t_from_name(const char* file) {
    const char* base = strrchr(file, '/');
    base = base ? base + 1 : file;
    const char* p = strstr(base, "t=");
    return p ? strtol(p + 2, NULL, 10) : 0;
}

int elp_set_gpu_trans(uint trans[4]) {
    uint nFound[16] = {0}, sum;
    for (int i = 0; i < 4; i++) trans[i] = 0;
    for (uint a = 1; a < 15; a++) {
        if (!elp_bond_valid(a)) continue;
        for (uint b = a + 1; b < 15; b++) {
            if (!elp_bond_valid(b) || !elp_bond_add_valid(a, b, &sum)) continue;
            trans[sum / 4] |= a << (8 * (sum % 4) + 4 * nFound[sum]);
            if (++nFound[sum] > 2) { printf("Internal error: setting trans reactions failed\n"); exit(1); }
        }
    }
    return 0;
}
