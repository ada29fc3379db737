/*
 * ring_io.c -- tuv polymer files, simulation_settings.txt and the fresh/
 * resume start-up of the ring drop-in.
 * Restates gpupol2/gpupol_io.c and LoadPolymers (gpupol2/gpupol_init.c:83-152).
 *
 * tuv format: "LT= %i\nLU= %i\nLV= %i\nnp= %i\nmaxPolLength= %i\n", then per
 * chain "len= %i\n%u  %u  %u\t<hex bond codes>\n" (0 = stored length).
 */
#include <dirent.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include "gpupol_ring.h"
#include "elp_host_common.h"

/* gpupol2/gpupol_io.c:105-165; box edges must be multiples of `unit` (the
   reference: 32, :131-134), else 202 is returned */
int ReadLatticeFileUnit(SimState* ss, SimProperties* sp, const char* file, int unit) {
    printf("Reading lattice file\n");
    FILE* f = fopen(file, "r");
    if (!f) { printf("Error reading file %s\n", file); return 199; }
    int LT, LU, LV, np = 0, maxLen = 0;
    if (fscanf(f, "%*s %i", &LT) != 1 || fscanf(f, "%*s %i", &LU) != 1 || fscanf(f, "%*s %i", &LV) != 1) {
        printf("Error: malformed header in %s\n", file); fclose(f); return 199;
    }
    const int E = unit > 0 ? unit : ELP_RING_TILE;
    if (LT <= 0 || LU <= 0 || LV <= 0 || LT % E || LU % E || LV % E) {
        printf("Error: Box dimensions not correct. Work group dimensions: (%i, %i, %i) vs file (%i, %i, %i)\n",
               E, E, E, LT, LU, LV);
        fclose(f);
        return 202;
    }
    SetBoxDimensionUnit(sp, LT, LU, LV, E);
    SimStateInit(ss, sp);
    if (fscanf(f, "%*s %i", &np) != 1 || fscanf(f, "%*s %i", &maxLen) != 1 || np < 0 || maxLen <= 0) {
        printf("Error: malformed header in %s\n", file); fclose(f); return 199;
    }
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
            if (bonds[i] > 0xf) { printf("Error: bad bond character in %s\n", file); fclose(f); free(bonds); free(text); return 199; }
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
    printf("Finished Reading lattice file\n");
    return 0;
}

int ReadLatticeFile(SimState* ss, SimProperties* sp, const char* file) {
    return ReadLatticeFileUnit(ss, sp, file, ELP_RING_TILE);
}

/* gpupol2/gpupol_io.c:191-203 */
void WriteRingPolymer(FILE* f, Polymer* pol) {
    fprintf(f, "len= %i\n", pol->length);
    fprintf(f, "%u  %u  %u\t", pol->startTUV.t, pol->startTUV.u, pol->startTUV.v);
    for (int i = (int)pol->labelStart; i < pol->length; i++) fprintf(f, "%x", pol->bonds[i]);
    for (int i = 0; i < (int)pol->labelStart; i++) fprintf(f, "%x", pol->bonds[i]);
    fprintf(f, "\n");
}

/* gpupol2/gpupol_io.c:167-189: chains in polymer-id order */
int WriteLatticeFile(SimProperties* sp, SimState* ss, const char* file) {
    FILE* f = fopen(file, "w");
    if (!f) { printf("Error opening file %s\n", file); return -1; }
    int* rev = (int*)malloc(sizeof(int) * (ss->nPol ? ss->nPol : 1));
    fprintf(f, "LT= %i\nLU= %i\nLV= %i\nnp= %i\nmaxPolLength= %i\n", sp->LT, sp->LU, sp->LV, ss->nPol, sp->maxPolLength);
    for (int p = 0; p < ss->nPol; p++) rev[ss->polNumTranslate[p]] = p;
    for (int p = 0; p < ss->nPol; p++) WriteRingPolymer(f, ss->pol + rev[p]);
    free(rev);
    fclose(f);
    return 0;
}

/* gpupol2/gpupol_io.c:31-77 */
void WriteSimulationSettings(SimProperties* sp, SimState* ss) {
    char file[4096];
    snprintf(file, sizeof file, "%s/simulation_settings.txt", sp->dirOut);
    FILE* f = fopen(file, "w");
    if (!f) return;
    fprintf(f, "Start_seed = %u\n", sp->seed);
    fprintf(f, "Length = %i\n", sp->polLength);
    fprintf(f, "Polytype = ring\n");
    fprintf(f, "Density = %lf\n", sp->density);
    fprintf(f, "Latsize = %i\n", sp->LT);
    fprintf(f, "Start_polysize = %i\n", sp->polLength);
    fprintf(f, "Double_step = %i\n", sp->double_step);
    fprintf(f, "Interval = %li\n", sp->writeInterval);
    fprintf(f, "Equilibrated = %i\n", sp->equilibrated);
    fprintf(f, "Npol = %i\n", ss->nPol);
    fprintf(f, "Polymodel = sl_equal\n");
    fprintf(f, "Executable = gpupol3\n");
    fprintf(f, "Fast_equilibration = %i\n", sp->fastEq ? 1 : 0);
    fprintf(f, "FEQ_timestep = %li\n", sp->fastEq);
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

/* gpupol2/gpupol_init.c:83-152: resume from the highest t=<T>_dev=<i>.res
   files if there is one per device (seed += T/interval), else build the
   compact rings, reshuffle stored length and write t=0.  The device context
   of GPULibInit stays the library's own: one device, index 0. */
void LoadPolymers(SimProperties* sp, SimState* ss) {
    char file[8192];
    GPULibContext ctx;
    GPULibInit(sp, NULL, &ctx);
    mkdir_p(sp->dirOut);
    long highT = -1;
    int nHigh = -1;
    char* highFiles[ELP_N_MAX_GPU];
    DIR* dir = opendir(sp->dirOut);
    struct dirent* de;
    while (dir && (de = readdir(dir))) {
        if (strncmp(de->d_name, "t=", 2)) continue;
        long tf = elp_t_from_name(de->d_name);
        if (tf > highT) {
            for (int i = 0; i < nHigh; i++) free(highFiles[i]);
            nHigh = 0;
        }
        if (tf >= highT && nHigh < ELP_N_MAX_GPU) {
            highFiles[nHigh++] = strdup(de->d_name);
            highT = tf;
        }
    }
    if (dir) closedir(dir);
    if (nHigh == sp->nDevices) {
        printf("Starting from t=%li\n", highT);
        sp->curT = highT;
        sp->tMax += highT;
        sp->seed += (uint)(highT / sp->writeInterval);
        for (int i = 0; i < sp->nDevices; i++) {
            snprintf(file, sizeof file, "%s/%s", sp->dirOut, highFiles[i]);
            printf("file = %s\n", file);
            int rc = ReadLatticeFile(ss + i, sp, file);
            if (rc) exit(rc);
            if (UpdatePolymerWithLabels(ss + i)) exit(123);
        }
    } else {
        SetBoxDimension(sp, (int)sp->LT, (int)sp->LU, (int)sp->LV);
        SimStateInit(ss, sp);
        int rc = GetBestPolymerPartitioning(sp, ss, sp->polLength, sp->density);
        if (rc) exit(rc);
        RedistribSL(ss, sp);
        for (int i = 0; i < sp->nDevices; i++) {
            snprintf(file, sizeof file, "%s/t=0_dev=%i.res", sp->dirOut, i);
            WriteLatticeFile(sp, ss + i, file);
        }
    }
    for (int i = 0; i < nHigh; i++) free(highFiles[i]);
    if (GetAllRingPolymers(sp, ss)) exit(193);
    sp->polLength = ss[0].pol[0].length;
}
