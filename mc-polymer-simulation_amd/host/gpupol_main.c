/*
 * gpupol_main.c -- drop-in for bin/gpupol_cuda_lin (do_run.sh:262,355).
 * Same positional argv, output files and progress line as gpupol/gpupol.c:14-70:
 *   NMONO TIME SEED DIR DENSITY FAST_EQ INTERVAL LT LU LV EQUILIBRATED DBL_STEP
 * As in the reference, FAST_EQ reshuffles the stored length of the host copy
 * before each interval (gpupol/gpupol.c:54); the device lattice is not
 * reloaded, so it only advances the host random stream.
 * tuv files are written from the device (chain recovery and text on the
 * GPU, byte-identical) unless ELP_GPU_TUV=0 or FAST_EQ is set.
 * ELP_LIN_INIT=saw[:FRACTION] makes a fresh start grow the chains as random
 * self-avoiding walks covering FRACTION of the sites (default
 * ELP_SAW_SITE_FRACTION) instead of the reference's folded chains; resumed
 * runs read their tuv file as before.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include "gpupol_lin.h"

static SimProperties sp;
static GPULibContext gpuContext;
static SimState ss[ELP_N_MAX_GPU];
static GPUDeviceState devices[ELP_N_MAX_GPU];

int main(int argc, char** argv) {
    struct timeval t0, t1;
    char fileOut[8192];
    SimPropDefaultInit(&sp);
    if (argc > 1) sp.polLength = atoi(argv[1]);
    if (argc > 2) sp.tMax = (long)atof(argv[2]);
    if (argc > 3) sp.seed = (uint)atoi(argv[3]);
    if (argc > 4) sp.dirOut = argv[4];
    if (argc > 5) sp.density = atof(argv[5]);
    if (argc > 6) sp.fastEq = atoi(argv[6]);
    if (argc > 7) sp.writeInterval = (long)atof(argv[7]);
    if (argc > 8) sp.LT = (uint)atoi(argv[8]);
    if (argc > 9) sp.LU = (uint)atoi(argv[9]);
    if (argc > 10) sp.LV = (uint)atoi(argv[10]);
    if (argc > 11) sp.equilibrated = atoi(argv[11]);
    if (argc > 12) sp.double_step = atoi(argv[12]);
    if (!sp.dirOut || sp.writeInterval <= 0) {
        fprintf(stderr, "usage: %s NMONO TIME SEED DIR [DENSITY FAST_EQ INTERVAL LT LU LV EQUILIBRATED DBL_STEP]\n", argv[0]);
        return 2;
    }
    if (!sp.LT) sp.LT = 32;
    if (!sp.LU) sp.LU = sp.LT;
    if (!sp.LV) sp.LV = sp.LT;
    printf("Box: %ix%ix%i\n", sp.LT, sp.LU, sp.LV);
    LoadPolymers(&sp, ss);
    GPULibLoadBuffers(&sp, ss, devices, &gpuContext);
    const char* clean = "\b \b\b \b\b \b\b \r";
    double totStep = (double)(sp.tMax - sp.curT);
    /* FAST_EQ reshuffles the host copy each interval, which must then be current */
    const int deviceTuv = ElpGpuTuvEnabled() && !sp.fastEq;
    double kernelMs = 0;
    long units = 0;
    for (; sp.curT < sp.tMax; sp.curT += sp.writeInterval) {
        gettimeofday(&t0, NULL);
        if (sp.fastEq) RedistribSL(ss, &sp);
        if (deviceTuv) GPULibRunDevice(&sp, ss, devices, &gpuContext, (int)sp.writeInterval);
        else GPULibRun(&sp, ss, devices, &gpuContext, (int)sp.writeInterval);
        gettimeofday(&t1, NULL);
        kernelMs += devices[0].lastKernelMs;
        units += sp.writeInterval;
        double el = (t1.tv_sec - t0.tv_sec) + (t1.tv_usec - t0.tv_usec) / 1e6;
        printf("%s[%.1lf%%] Time elapsed = %.2f s at %.2f MF/s", clean,
               100 * (1 - (sp.tMax - sp.curT - sp.writeInterval) / totStep), el,
               sp.writeInterval * (double)sp.L * sp.nDevices / el / 1e6);
        fflush(stdout);
        for (int i = 0; i < sp.nDevices; i++) {
            snprintf(fileOut, sizeof fileOut, "%s/t=%li_dev=%i.res", sp.dirOut, sp.curT + sp.writeInterval, i);
            if (deviceTuv) WriteLatticeFileDevice(&sp, ss + i, devices + i, fileOut);
            else WriteLatticeFile(&sp, ss + i, fileOut);
        }
    }
    printf("\nSimulation complete\n");
    if (units > 0 && kernelMs > 0)
        printf("kernel time %.3f ms for %ld time units (%.2f MF/s on the device)\n", kernelMs, units,
               units * (double)sp.L / (kernelMs * 1e-3) / 1e6);
    WriteSimulationSettings(&sp, ss);
    GPULibRelease(&sp, devices, &gpuContext);
    return 0;
}
