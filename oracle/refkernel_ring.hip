// refkernel_ring.hip -- ORACLE / TEST INFRASTRUCTURE ONLY.
//
// Runs the reference ring kernel gpupol2/ringpol3.cu, compiled unmodified
// from the reference tree (oracle/Makefile.ref), with the launch sequence of
// gpupol2/gpupol_cuda.cu:29-50 (load: relayout launch with nStep=0) and
// :114-152 (one launch per time unit, double buffer), then undoes the
// shifted tile layout as gpupol2/oclpol.c:18-47 does.  The tests compare
// its result with the oracle's reference-mode replay (orc_ring_run_ref).
//
// usage: refkernel_ring L NSTEP NTIME lattice_in seeds_in shifts_in lattice_out seeds_out
//   lattice_*: L^3 bytes, row-major (t fastest), site bytes of gpupol2/oclpol.c:68-124
//   seeds_*  : uint32[(L/32)^3 * 512 * 8], the table of gpupol2/gpupol_init.c:22-33
//   shifts_in: int32[3*(NTIME+1)], the load-time shift then one shift per time unit
#include <hip/hip_runtime.h>
#include "ringpol3.cu"
#include <stdio.h>
#include <stdlib.h>
#include <vector>

#define CHK(x)                                                                         \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) { fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_)); exit(3); } \
    } while (0)

static size_t tile_site(int t, int u, int v, int nw) {   // host layout GetGpuSite (gpupol2/oclpol.c:3-16)
    size_t wid = (size_t)(t / 32) + (size_t)(u / 32) * nw + (size_t)(v / 32) * nw * nw;
    return (size_t)(t % 32 + 32 * (u % 32) + 1024 * (v % 32)) + wid * 32768;
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
    if (argc != 9) { fprintf(stderr, "usage: %s L NSTEP NTIME lat_in seeds_in shifts_in lat_out seeds_out\n", argv[0]); return 2; }
    const int L = atoi(argv[1]), nStep = atoi(argv[2]), nTime = atoi(argv[3]);
    if (L % 32) { fprintf(stderr, "L must be a multiple of 32\n"); return 2; }
    const int nw = L / 32, nwg = nw * nw * nw;
    const size_t nsite = (size_t)L * L * L, nseed = (size_t)nwg * 512 * 8;
    auto lat = slurp<unsigned char>(argv[4], nsite);
    auto seeds = slurp<unsigned int>(argv[5], nseed);
    auto d = slurp<int>(argv[6], 3 * (size_t)(nTime + 1));
    for (int x : d) if (x < 0 || x >= 32) { fprintf(stderr, "shift out of range\n"); return 2; }

    std::vector<unsigned char> host(nsite);
    for (int v = 0; v < L; v++)
        for (int u = 0; u < L; u++)
            for (int t = 0; t < L; t++) host[tile_site(t, u, v, nw)] = lat[t + (size_t)L * (u + (size_t)L * v)];

    // trans table: first bond of each decomposition (gpupol2/gpupol_init.c:552-584)
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

    char *lat0, *lat1;
    uint4* dseeds;
    unsigned int* dtrans;
    CHK(hipMalloc(&lat0, nsite));
    CHK(hipMalloc(&lat1, nsite));
    CHK(hipMalloc(&dseeds, nseed * 4));
    CHK(hipMalloc(&dtrans, 16));
    CHK(hipMemcpy(lat0, host.data(), nsite, hipMemcpyHostToDevice));
    CHK(hipMemcpy(dseeds, seeds.data(), nseed * 4, hipMemcpyHostToDevice));
    CHK(hipMemcpy(dtrans, trans, 16, hipMemcpyHostToDevice));
    char* buf[2] = {lat0, lat1};
    int cur = 0;
    long cum[3] = {d[0], d[1], d[2]};
    int tuv = d[0] + d[1] * 32 + d[2] * 1024;
    hipLaunchKernelGGL(polmove, dim3(nwg), dim3(512), 0, 0, 0, dseeds, buf[cur], buf[cur ^ 1], dtrans, 0, tuv,
                       (uint)nw, (uint)nw, (uint)nw);
    CHK(hipGetLastError());
    cur ^= 1;
    int last[3] = {d[0], d[1], d[2]};
    for (int i = 1; i <= nTime; i++) {
        const int* s = &d[3 * i];
        for (int k = 0; k < 3; k++) { cum[k] += s[k]; last[k] = s[k]; }
        const int prev = tuv;
        tuv = s[0] + s[1] * 32 + s[2] * 1024;
        hipLaunchKernelGGL(polmove, dim3(nwg), dim3(512), 0, 0, nStep, dseeds, buf[cur], buf[cur ^ 1], dtrans, prev, tuv,
                           (uint)nw, (uint)nw, (uint)nw);
        CHK(hipGetLastError());
        cur ^= 1;
    }
    CHK(hipDeviceSynchronize());
    std::vector<unsigned char> gpu(nsite);
    CHK(hipMemcpy(gpu.data(), buf[cur], nsite, hipMemcpyDeviceToHost));
    CHK(hipMemcpy(seeds.data(), dseeds, nseed * 4, hipMemcpyDeviceToHost));

    // undo the 8-piece shifted tile layout (gpupol2/oclpol.c:18-47)
    for (int wt = 0; wt < nw; wt++)
        for (int wu = 0; wu < nw; wu++)
            for (int wv = 0; wv < nw; wv++) {
                const size_t wid = wt + (size_t)wu * nw + (size_t)wv * nw * nw;
                size_t off = 0;
                for (int p = 0; p < 8; p++) {
                    const int bt = (p & 1) ? last[0] : 32 - last[0];
                    const int bu = (p & 2) ? last[1] : 32 - last[1];
                    const int bv = (p & 4) ? last[2] : 32 - last[2];
                    for (int i = 0; i < bt * bu * bv; i++) {
                        long t = ((p & 1) ? 32 - last[0] : 0) + i % bt;
                        long u = ((p & 2) ? 32 - last[1] : 0) + (i / bt) % bu;
                        long v = ((p & 4) ? 32 - last[2] : 0) + i / (bt * bu);
                        t = (t + cum[0] + wt * 32) % L;
                        u = (u + cum[1] + wu * 32) % L;
                        v = (v + cum[2] + wv * 32) % L;
                        lat[t + (size_t)L * (u + (size_t)L * v)] = gpu[i + off + wid * 32768];
                    }
                    off += (size_t)bt * bu * bv;
                }
            }
    spill(argv[7], lat);
    spill(argv[8], seeds);
    return 0;
}
