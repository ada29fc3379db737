// refkernel_lin.hip -- ORACLE / TEST INFRASTRUCTURE ONLY.
//
// Runs the reference linear-chain kernel gpupol/linpol.cu, compiled
// unmodified from the reference tree (oracle/Makefile.ref), with the launch
// sequence of gpupol/gpupol_cuda.cu:84-112 (one in-place launch per time
// unit at a fresh tile offset) on the global layout of gpupol/oclpol.c:3-16.
// The tests compare its result with orc_lin_run_ref.
//
// usage: refkernel_lin L NSTEP NTIME lattice_in seeds_in offsets_in lattice_out seeds_out
//   lattice_*: uint16[L^3] row-major site words (next | prev<<4 | SL<<8 | labels<<9)
//   seeds_*  : uint32[(L/32)^3 * 512 * 8] (gpupol/gpupol_init.c:23-34)
//   offsets_in: int32[3*NTIME] tile offset (t,u,v) of each time unit
#include <hip/hip_runtime.h>
#include "linpol.cu"
#include <stdio.h>
#include <stdlib.h>
#include <vector>

#define CHK(x)                                                                         \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) { fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_)); exit(3); } \
    } while (0)

static size_t gpu_site(unsigned t, unsigned u, unsigned v, unsigned nw) {
    size_t s = ((t & 0x3) << 9) | ((u & 0x3) << 11) | ((v & 0x3) << 13);
    s |= ((t & 0x1c) >> 2) | ((u & 0x1c) << 1) | ((v & 0x1c) << 4);
    s |= (size_t)((t >> 5) + (u >> 5) * nw + (v >> 5) * nw * nw) << 15;
    return s;
}

template <class T>
static std::vector<T> slurp(const char* f, size_t n) {
    std::vector<T> v(n);
    FILE* p = fopen(f, "rb");
    if (!p || fread(v.data(), sizeof(T), n, p) != n) { fprintf(stderr, "cannot read %s\n", f); exit(2); }
    fclose(p);
    return v;
}
template <class T>
static void spill(const char* f, const std::vector<T>& v) {
    FILE* p = fopen(f, "wb");
    if (!p || fwrite(v.data(), sizeof(T), v.size(), p) != v.size()) { fprintf(stderr, "cannot write %s\n", f); exit(2); }
    fclose(p);
}

int main(int argc, char** argv) {
    if (argc != 9) { fprintf(stderr, "usage: %s L NSTEP NTIME lat_in seeds_in offsets_in lat_out seeds_out\n", argv[0]); return 2; }
    const int L = atoi(argv[1]), nStep = atoi(argv[2]), nTime = atoi(argv[3]);
    if (L % 32) { fprintf(stderr, "L must be a multiple of 32\n"); return 2; }
    const unsigned nw = L / 32, nwg = nw * nw * nw;
    const size_t nsite = (size_t)L * L * L, nseed = (size_t)nwg * 512 * 8;
    auto lat = slurp<unsigned short>(argv[4], nsite);
    auto seeds = slurp<unsigned int>(argv[5], nseed);
    auto off = slurp<int>(argv[6], 3 * (size_t)nTime);
    for (int x : off) if (x < 0 || x >= 32) { fprintf(stderr, "offset out of range\n"); return 2; }
    std::vector<unsigned int> g(nsite);
    for (unsigned v = 0; v < (unsigned)L; v++)
        for (unsigned u = 0; u < (unsigned)L; u++)
            for (unsigned t = 0; t < (unsigned)L; t++) g[gpu_site(t, u, v, nw)] = lat[t + (size_t)L * (u + (size_t)L * v)];

    unsigned int trans[4] = {0, 0, 0, 0}, nf[16] = {0};
    auto ok = [](unsigned a) { return a != 0 && a != 3 && a != 0xc && a != 0xf; };
    for (unsigned a = 1; a < 15; a++)
        for (unsigned b = a + 1; b < 15; b++) {
            if (!ok(a) || !ok(b) || ((a | b) != 0xf && (a & b))) continue;
            unsigned s = ((a | b) == 0xf) ? (a & b) : (a | b);
            if (!ok(s)) continue;
            trans[s / 4] |= a << (8 * (s % 4) + 4 * nf[s]);
            nf[s]++;
        }

    unsigned int *dlat, *dtrans;
    uint4* dseeds;
    CHK(hipMalloc(&dlat, nsite * 4));
    CHK(hipMalloc(&dseeds, nseed * 4));
    CHK(hipMalloc(&dtrans, 16));
    CHK(hipMemcpy(dlat, g.data(), nsite * 4, hipMemcpyHostToDevice));
    CHK(hipMemcpy(dseeds, seeds.data(), nseed * 4, hipMemcpyHostToDevice));
    CHK(hipMemcpy(dtrans, trans, 16, hipMemcpyHostToDevice));
    for (int i = 0; i < nTime; i++) {
        const unsigned tuv = off[3 * i] + off[3 * i + 1] * 32 + off[3 * i + 2] * 1024;
        hipLaunchKernelGGL(polmove, dim3(nwg), dim3(512), 0, 0, (uint)nStep, dseeds, dlat, dtrans, tuv, nw, nw, nw);
        CHK(hipGetLastError());
    }
    CHK(hipDeviceSynchronize());
    CHK(hipMemcpy(g.data(), dlat, nsite * 4, hipMemcpyDeviceToHost));
    CHK(hipMemcpy(seeds.data(), dseeds, nseed * 4, hipMemcpyDeviceToHost));
    for (unsigned v = 0; v < (unsigned)L; v++)
        for (unsigned u = 0; u < (unsigned)L; u++)
            for (unsigned t = 0; t < (unsigned)L; t++)
                lat[t + (size_t)L * (u + (size_t)L * v)] = (unsigned short)g[gpu_site(t, u, v, nw)];
    spill(argv[7], lat);
    spill(argv[8], seeds);
    return 0;
}
