// elp_engine.hip -- the C ABI of libelp.so (include/elp_gpu.h): engine
// objects owning the device lattice, a HIP stream and timing events, and the
// per-time-unit launch loop.  Replaces the buffer management and launch loop
// of gpupol2/gpupol_cuda.cu:78-152 and gpupol/gpupol_cuda.cu:29-112.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include "../../include/elp_gpu.h"
#include "elp_kernels.h"

namespace {

thread_local std::string g_err = "";

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(ELP_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIPCHK(call)                                    \
    do {                                                \
        hipError_t e_ = (call);                         \
        if (e_ != hipSuccess) return hip_fail(e_, #call); \
    } while (0)

// first bond of each of the two decompositions of every valid bond sum,
// nibble 2*(s%4)+k of word s/4 (gpupol2/gpupol_init.c:552-584)
void make_trans(uint32_t trans[4]) {
    auto valid = [](uint32_t a) { return a != 0 && a != 3 && a != 0xc && a != 0xf; };
    uint32_t nfound[16] = {0};
    for (int i = 0; i < 4; i++) trans[i] = 0;
    for (uint32_t a = 1; a < 15; a++) {
        if (!valid(a)) continue;
        for (uint32_t b = a + 1; b < 15; b++) {
            if (!valid(b)) continue;
            if ((a | b) != 0xf && (a & b)) continue;
            uint32_t s = ((a | b) == 0xf) ? (a & b) : (a | b);
            if (!valid(s)) continue;
            trans[s / 4] |= a << (8 * (s % 4) + 4 * nfound[s]);
            nfound[s]++;
        }
    }
}

}  // namespace

struct elp_engine {
    int kind;              // 0 ring, 1 linear
    int device;
    int LT, LU, LV;
    int tt, tu, tv;
    int sublattice;
    int n_iter;
    int threads;
    size_t bytes_per_site;
    size_t nbytes;
    void* d_lat;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    bool uploaded;
    bool timed;
    uint32_t trans[4];
};

static int launch_one(elp_engine* e, bool probe, const elp::RingLaunch& p) {
    const int nblocks = (e->LT / e->tt) * (e->LU / e->tu) * (e->LV / e->tv);
    if (e->kind == 0)
        return elp::ring_launch(e->tt, e->tu, e->tv, e->sublattice == ELP_SUBLAT_DENSE, e->threads, probe, nblocks,
                                e->stream, (uint8_t*)e->d_lat, p);
    return elp::lin_launch(e->tt, e->tu, e->tv, e->sublattice == ELP_SUBLAT_DENSE, e->threads, probe, nblocks,
                           e->stream, (uint16_t*)e->d_lat, p);
}

// threads per workgroup: ELP_THREADS env override, else the default of the schedule
static int default_threads(int kind, int tt, int tu, int tv, int sublattice) {
    const char* env = getenv("ELP_THREADS");
    if (env && atoi(env) > 0) return atoi(env);
    const int ncell = (tt / 4) * (tu / 4) * (tv / 4);
    const int nslot = ncell * (sublattice == ELP_SUBLAT_DENSE ? 4 : 1);
    (void)kind;
    int th = nslot < 512 ? nslot : 512;
    return th;
}

static int create(elp_engine** out, int kind, int device, int LT, int LU, int LV, int tt, int tu, int tv,
                  int sublattice, int n_iter) {
    if (!out) return fail(ELP_E_ARG, "out is NULL");
    *out = nullptr;
    if (sublattice != ELP_SUBLAT_CUBIC && sublattice != ELP_SUBLAT_DENSE) return fail(ELP_E_ARG, "bad sublattice %d", sublattice);
    if (tt % 4 || tu % 4 || tv % 4 || tt < 8 || tu < 8 || tv < 8 || tt > 32 || tu > 32 || tv > 32)
        return fail(ELP_E_ARG, "tile %dx%dx%d: edges must be multiples of 4 in [8,32]", tt, tu, tv);
    if (LT <= 0 || LU <= 0 || LV <= 0 || LT % tt || LU % tu || LV % tv)
        return fail(ELP_E_ARG, "box %dx%dx%d not divisible by tile %dx%dx%d", LT, LU, LV, tt, tu, tv);
    if (n_iter < 0) return fail(ELP_E_ARG, "n_iter < 0");
    elp_engine* e = new elp_engine();
    e->kind = kind; e->device = device;
    e->LT = LT; e->LU = LU; e->LV = LV;
    e->tt = tt; e->tu = tu; e->tv = tv;
    e->sublattice = sublattice; e->n_iter = n_iter;
    e->threads = default_threads(kind, tt, tu, tv, sublattice);
    e->bytes_per_site = kind == 0 ? 1 : 2;
    e->nbytes = (size_t)LT * LU * LV * e->bytes_per_site;
    make_trans(e->trans);
    elp::RingLaunch dummy{};
    if (launch_one(e, true, dummy) != 0) {
        int th = e->threads;
        delete e;
        return fail(ELP_E_ARG, "schedule %s tile %dx%dx%d sublattice %d threads %d is not compiled in",
                    kind == 0 ? "ring" : "linear", tt, tu, tv, sublattice, th);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { delete e; return fail(ELP_E_NODEV, "no HIP device"); }
    if (device < 0 || device >= ndev) { delete e; return fail(ELP_E_ARG, "device %d out of range (%d devices)", device, ndev); }
    hipError_t err;
    if ((err = hipSetDevice(device)) != hipSuccess) { delete e; return hip_fail(err, "hipSetDevice"); }
    if ((err = hipMalloc(&e->d_lat, e->nbytes)) != hipSuccess) { delete e; return hip_fail(err, "hipMalloc lattice"); }
    if ((err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking)) != hipSuccess) {
        (void)hipFree(e->d_lat); delete e; return hip_fail(err, "hipStreamCreate");
    }
    (void)hipEventCreate(&e->ev0);
    (void)hipEventCreate(&e->ev1);
    *out = e;
    return ELP_OK;
}

extern "C" {

const char* elp_last_error(void) { return g_err.c_str(); }

int elp_device_count(int* n) {
    if (!n) return fail(ELP_E_ARG, "n is NULL");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return ELP_OK;
}

int elp_ring_create(elp_engine** out, int device, int LT, int LU, int LV, int tt, int tu, int tv, int sublattice,
                    int n_iter) {
    return create(out, 0, device, LT, LU, LV, tt, tu, tv, sublattice, n_iter);
}

int elp_lin_create(elp_engine** out, int device, int LT, int LU, int LV, int tt, int tu, int tv, int sublattice,
                   int n_iter) {
    return create(out, 1, device, LT, LU, LV, tt, tu, tv, sublattice, n_iter);
}

int elp_destroy(elp_engine* e) {
    if (!e) return ELP_OK;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    (void)hipEventDestroy(e->ev0);
    (void)hipEventDestroy(e->ev1);
    (void)hipStreamDestroy(e->stream);
    (void)hipFree(e->d_lat);
    delete e;
    return ELP_OK;
}

int elp_upload(elp_engine* e, const void* lattice, uint64_t nbytes) {
    if (!e || !lattice) return fail(ELP_E_ARG, "NULL argument");
    if (nbytes != e->nbytes) return fail(ELP_E_ARG, "upload of %llu bytes, lattice is %zu", (unsigned long long)nbytes, e->nbytes);
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipMemcpyAsync(e->d_lat, lattice, e->nbytes, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->uploaded = true;
    return ELP_OK;
}

int elp_download(elp_engine* e, void* lattice, uint64_t nbytes) {
    if (!e || !lattice) return fail(ELP_E_ARG, "NULL argument");
    if (nbytes != e->nbytes) return fail(ELP_E_ARG, "download of %llu bytes, lattice is %zu", (unsigned long long)nbytes, e->nbytes);
    if (!e->uploaded) return fail(ELP_E_STATE, "download before upload");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipMemcpyAsync(lattice, e->d_lat, e->nbytes, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return ELP_OK;
}

int elp_run(elp_engine* e, int n_time, const int32_t* origins, uint64_t tau0, uint32_t seed, uint32_t stream) {
    if (!e) return fail(ELP_E_ARG, "NULL engine");
    if (n_time < 0 || (n_time > 0 && !origins)) return fail(ELP_E_ARG, "bad n_time/origins");
    if (!e->uploaded) return fail(ELP_E_STATE, "run before upload");
    for (int i = 0; i < n_time; i++) {
        const int32_t* o = origins + 3 * i;
        if (o[0] < 0 || o[0] >= e->tt || o[1] < 0 || o[1] >= e->tu || o[2] < 0 || o[2] >= e->tv)
            return fail(ELP_E_ARG, "origin (%d,%d,%d) of time unit %d outside the tile", o[0], o[1], o[2], i);
    }
    HIPCHK(hipSetDevice(e->device));
    elp::RingLaunch p{};
    p.LT = e->LT; p.LU = e->LU; p.LV = e->LV;
    p.ntile_t = e->LT / e->tt; p.ntile_u = e->LU / e->tu;
    p.seed = seed; p.stream = stream;
    p.n_iter = e->n_iter;
    for (int i = 0; i < 4; i++) p.trans[i] = e->trans[i];
    HIPCHK(hipEventRecord(e->ev0, e->stream));
    for (int i = 0; i < n_time; i++) {
        p.ot = origins[3 * i]; p.ou = origins[3 * i + 1]; p.ov = origins[3 * i + 2];
        p.tau = (uint32_t)(tau0 + (uint64_t)i);
        int rc = launch_one(e, false, p);
        if (rc) return fail(ELP_E_HIP, "kernel launch failed: %s", hipGetErrorString((hipError_t)rc));
    }
    HIPCHK(hipEventRecord(e->ev1, e->stream));
    e->timed = true;
    return ELP_OK;
}

int elp_sync(elp_engine* e) {
    if (!e) return fail(ELP_E_ARG, "NULL engine");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    return ELP_OK;
}

int elp_last_run_ms(elp_engine* e, float* ms) {
    if (!e || !ms) return fail(ELP_E_ARG, "NULL argument");
    if (!e->timed) return fail(ELP_E_STATE, "no run recorded");
    HIPCHK(hipEventElapsedTime(ms, e->ev0, e->ev1));
    return ELP_OK;
}

int elp_launch_info(elp_engine* e, int* threads_per_block, int* blocks_per_time_unit, int* lds_bytes) {
    if (!e) return fail(ELP_E_ARG, "NULL engine");
    if (threads_per_block) *threads_per_block = e->threads;
    if (blocks_per_time_unit) *blocks_per_time_unit = (e->LT / e->tt) * (e->LU / e->tu) * (e->LV / e->tv);
    if (lds_bytes) *lds_bytes = (int)(e->tt * e->tu * e->tv * (e->kind == 0 ? 1 : 2));
    return ELP_OK;
}

__global__ void elp_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int elp_bench_copy(int device, uint64_t nbytes, int reps, float* ms) {
    if (!ms || nbytes == 0 || nbytes % 16 || reps <= 0) return fail(ELP_E_ARG, "bad copy arguments");
    HIPCHK(hipSetDevice(device));
    void *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, nbytes));
    hipError_t err = hipMalloc(&b, nbytes);
    if (err != hipSuccess) { (void)hipFree(a); return hip_fail(err, "hipMalloc"); }
    (void)hipMemset(a, 1, nbytes);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const size_t n = nbytes / 16;
    hipLaunchKernelGGL(elp_copy_kernel, dim3(2048), dim3(256), 0, 0, (const uint4*)a, (uint4*)b, n);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(elp_copy_kernel, dim3(2048), dim3(256), 0, 0, (const uint4*)a, (uint4*)b, n);
    (void)hipEventRecord(e1, 0);
    err = hipEventSynchronize(e1);
    float t = 0;
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (err != hipSuccess) return hip_fail(err, "copy kernel");
    return ELP_OK;
}

}  // extern "C"
