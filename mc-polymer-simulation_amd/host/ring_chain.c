/*
 * ring_chain.c -- site-byte encoding and ring recovery from the lattice.
 * Restates gpupol2/oclpol.c (file:line per function).  Chains are recovered
 * from the per-monomer label bits that travel with the monomers: walk the
 * chain until the 5-bit head 1,0,0,0,0 has been read, start on the next
 * monomer, read the 18 id bits that follow.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "gpupol_ring.h"
#include "elp_host_common.h"

#define SITE_BOND(x) ((x) & 0xf)
#define SITE_SL(x) (((x) >> 6) & 0x1)
#define SITE_LABELS(x) (((x) >> 4) & 0x3)

uint SiteIndex(uint t, uint u, uint v, const SimProperties* sp) { return t + sp->LT * (u + sp->LU * v); }

/* gpupol2/oclpol.c:130-148 */
void AddUnitTUV(uint unit, uint* t, uint* u, uint* v, const SimProperties* sp) {
    int dt, du, dv;
    elp_bond_delta(unit, &dt, &du, &dv);
    *t = (uint)(((int)*t + dt + (int)sp->LT) % (int)sp->LT);
    *u = (uint)(((int)*u + du + (int)sp->LU) % (int)sp->LU);
    *v = (uint)(((int)*v + dv + (int)sp->LV) % (int)sp->LV);
}

/* gpupol2/oclpol.c:157-181 */
int SetBondVecs(uint8_t* lattice, uint t, uint u, uint v, uint* bonds, uint polSize, uint* labels,
                SimProperties* sp) {
    uint site = SiteIndex(t, u, v, sp);
    for (uint i = 0; i < polSize; i++) {
        if (bonds[i] == 0x0) {
            lattice[site] = (uint8_t)((lattice[site] & ~0x60) | 0x40 | ((labels[i] & 1) << 5));
        } else if (bonds[i] == 0xf) {
            printf("Error: Linear polymer detected!\n");
            break;
        } else {
            if (SITE_BOND(lattice[site])) {
                printf("Error setting bond vector %i: site already occupied\n", i);
                printf("(%i %i %i)\n", t, u, v);
                return 1;
            }
            lattice[site] = (uint8_t)((lattice[site] & ~0x1f) | bonds[i] | ((labels[i] & 1) << 4));
            AddUnitTUV(bonds[i], &t, &u, &v, sp);
            site = SiteIndex(t, u, v, sp);
        }
    }
    return 0;
}

/* gpupol2/oclpol.c:146-155 */
void SetLatticeFromSS(SimState* ss, SimProperties* sp) {
    memset(ss->lattice, 0, sp->latSize);
    for (int p = 0; p < ss->nPol; p++) {
        Polymer* pol = ss->pol + p;
        SetBondVecs(ss->lattice, pol->startTUV.t, pol->startTUV.u, pol->startTUV.v, pol->bonds,
                    (uint)pol->length, pol->label, sp);
    }
}

/* gpupol2/oclpol.c:278-342: walk forward from (t,u,v) in chain order,
   shifting label bits into a 5-bit register, until it reads the head; then
   step to the monomer after the head.  Returns 0, or -1 if no head. */
static int FindPolymerStart(const SimProperties* sp, const SimState* ss, uint* t, uint* u, uint* v,
                            Polymer* pol, const uint8_t* lat) {
    int mask = (1 << ss->headBit) - 1;
    int reg = 0, i = 0;
    uint site = SiteIndex(*t, *u, *v, sp);
    long limit = 2L * sp->maxPolLength;
    while (i < limit) {
        uint8_t x = lat[site];
        int sl = SITE_SL(x), lab = SITE_LABELS(x);
        pol->startLeftSL = sl;
        /* the stored-length monomer (bit 5) precedes the bond carrier (bit 4) */
        reg = ((reg << 1) & mask) | ((lab >> sl) & 1);
        i++;
        if (reg == ss->labelHead && i >= ss->headBit) goto found;
        if (sl) {
            reg = ((reg << 1) & mask) | (lab & 1);
            i++;
            pol->startLeftSL = 0;
            if (reg == ss->labelHead && i >= ss->headBit) goto found;
        }
        uint b = SITE_BOND(x);
        if (!b) { printf("Linear polymer detected: not supported\n"); return -1; }
        AddUnitTUV(b, t, u, v, sp);
        site = SiteIndex(*t, *u, *v, sp);
    }
    printf("Error: Did not find label head\n");
    return -1;
found:
    if (pol->startLeftSL) {
        pol->startLeftSL = 0;          /* head ended on the SL monomer: start on its bond carrier */
    } else {
        AddUnitTUV(SITE_BOND(lat[site]), t, u, v, sp);
        pol->startLeftSL = SITE_SL(lat[SiteIndex(*t, *u, *v, sp)]);
    }
    return 0;
}

/* gpupol2/oclpol.c:220-262: read the ring from its start; bonds get a 0 for
   each stored-length monomer.  Clears the bond bits it walks over in `lat`
   (a scratch copy) so the scan does not find the ring again. */
static int GetRingPolymer(const SimProperties* sp, const SimState* ss, uint t, uint u, uint v,
                          Polymer* pol, uint8_t* lat) {
    if (FindPolymerStart(sp, ss, &t, &u, &v, pol, lat)) return -1;
    pol->startTUV.t = t; pol->startTUV.u = u; pol->startTUV.v = v;
    int n = 0;
    memset(pol->label, 0, sp->maxPolLength * sizeof(uint));
    int leftSL = pol->startLeftSL;
    uint site = SiteIndex(t, u, v, sp);
    uint lab = SITE_LABELS(lat[site]);
    do {
        if (leftSL == 1) {
            if (SITE_SL(lat[site])) { pol->label[n] = lab >> 1; pol->bonds[n] = 0; }
            else n--;
        } else {
            pol->label[n] = lab & 1;
            uint b = SITE_BOND(lat[site]);
            if (!b) { pol->bonds[n] = 0xf; return n + 1; }
            lat[site] &= (uint8_t)~0x0f;
            pol->bonds[n] = b;
            AddUnitTUV(b, &t, &u, &v, sp);
            site = SiteIndex(t, u, v, sp);
            lab = SITE_LABELS(lat[site]);
        }
        n++;
        leftSL ^= 1;
        if ((uint)n == sp->maxPolLength) { printf("polymer too long\n"); return -1; }
    } while (!(pol->startTUV.t == t && pol->startTUV.u == u && pol->startTUV.v == v && leftSL == pol->startLeftSL));
    return n;
}

/* gpupol2/oclpol.c:183-218: scan sites with v fastest, then u, then t. */
int GetAllRingPolymers(SimProperties* sp, SimState* ss) {
    uint8_t* copy = (uint8_t*)malloc(sp->latSize);
    int rc = 0;
    for (int d = 0; d < sp->nDevices && !rc; d++) {
        SimState* cs = ss + d;
        memcpy(copy, cs->lattice, sp->latSize);
        int n = 0;
        for (uint t = 0; t < sp->LT && !rc; t++)
            for (uint u = 0; u < sp->LU && !rc; u++)
                for (uint v = 0; v < sp->LV && !rc; v++) {
                    if (!SITE_BOND(copy[SiteIndex(t, u, v, sp)])) continue;
                    if (n == cs->nPol) { printf("Too many polymers\n"); rc = -1; break; }
                    Polymer* pol = cs->pol + n;
                    int len = GetRingPolymer(sp, cs, t, u, v, pol, copy);
                    if (len < 0) { rc = -1; break; }
                    pol->length = len;
                    n++;
                }
        if (!rc && n != cs->nPol) { printf("Found %i polymers, expected %i\n", n, cs->nPol); rc = -1; }
    }
    free(copy);
    return rc;
}

/* gpupol2/oclpol.c:345-376: map the 18 id bits of each recovered ring to
   its polymer id; every id must be found exactly once. */
int UpdatePolymerWithLabels(SimState* ss) {
    int* seen = (int*)calloc(ss->nPol, sizeof(int));
    int fail = 0;
    for (int p = 0; p < ss->nPol; p++) {
        Polymer* pol = ss->pol + p;
        int lab = 0;
        for (int i = 0; i < ss->labelBit; i++) lab |= (int)pol->label[i] << (ss->labelBit - i - 1);
        int id = ss->label2Id[lab];
        if (id < 0) {
            printf("Unknown polymer label %i (length %i, start %u %u %u)\n", lab, pol->length,
                   pol->startTUV.t, pol->startTUV.u, pol->startTUV.v);
            free(seen);
            return -1;
        }
        ss->polNumTranslate[p] = id;
        seen[id]++;
    }
    for (int i = 0; i < ss->nPol; i++)
        if (seen[i] != 1) { printf("Account failure: %i %i\n", i, seen[i]); fail++; }
    free(seen);
    return fail ? -1 : 0;
}
