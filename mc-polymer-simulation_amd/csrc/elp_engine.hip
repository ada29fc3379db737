// elp_engine.hip -- the C ABI of libelp.so (include/elp_gpu.h): engine
// objects owning the device lattice, a HIP stream and timing events, and the
// per-time-unit launch loop.  Replaces the buffer management and launch loop
// of gpupol2/gpupol_cuda.cu:78-152 and gpupol/gpupol_cuda.cu:29-112.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>
#include <math.h>
#include <mutex>
#include "../../include/elp_gpu.h"
#include "elp_chains.h"
#include "elp_kernels.h"
#include "elp_learned.h"

namespace {

// last error message of this thread (a fixed buffer: reporting an error allocates nothing)
thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
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
    uint32_t rules;        // ELP_RULES_*
    // move counting (elp_count_moves): device counters of attempts at empty
    // sites, leaving the tile and accepted; attempts counted on the host
    unsigned long long* d_counts = nullptr;
    bool counting = false;
    uint64_t attempts = 0;
};

// Move counters: elp::kCountSlots slots of elp::kCountStride words; the
// totals (empty, face, accepted) are the sums over the slots.
constexpr size_t kCountBytes = sizeof(unsigned long long) * elp::kCountSlots * elp::kCountStride;
static hipError_t sum_counts(const unsigned long long* d, hipStream_t st, unsigned long long out[3]) {
    std::vector<unsigned long long> h((size_t)elp::kCountSlots * elp::kCountStride);
    hipError_t e = hipMemcpyAsync(h.data(), d, kCountBytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    out[0] = out[1] = out[2] = 0;
    for (int i = 0; i < elp::kCountSlots; i++)
        for (int k = 0; k < 3; k++) out[k] += h[(size_t)i * elp::kCountStride + k];
    return e;
}

// move attempts of one time unit: every slot of every tile in every substep
static uint64_t attempts_per_unit(int kind, int LT, int LU, int LV, int tt, int tu, int tv, int sublattice,
                                  int n_iter) {
    const uint64_t ntile = (uint64_t)(LT / tt) * (LU / tu) * (LV / tv);
    const uint64_t nslot = (uint64_t)(tt / 4) * (tu / 4) * (tv / 4) * (sublattice == ELP_SUBLAT_DENSE ? 4 : 1);
    return ntile * nslot * (uint64_t)n_iter * (kind == 0 ? 4u : 2u);
}

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
    if ((uint64_t)LT * LU * LV >= ((uint64_t)1 << 31))
        return fail(ELP_E_ARG, "box %dx%dx%d: at most 2^31 - 1 sites (the kernels index sites in 32 bits)", LT, LU, LV);
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

const char* elp_last_error(void) { return g_err; }

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
    if (e->d_counts) (void)hipFree(e->d_counts);
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
    p.rules = e->rules;
    p.counts = e->counting ? e->d_counts : nullptr;
    for (int i = 0; i < 4; i++) p.trans[i] = e->trans[i];
    if (e->counting)
        e->attempts += (uint64_t)n_time * attempts_per_unit(e->kind, e->LT, e->LU, e->LV, e->tt, e->tu, e->tv,
                                                            e->sublattice, e->n_iter);
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

int elp_count_moves(elp_engine* e, int enable) {
    if (!e) return fail(ELP_E_ARG, "NULL engine");
    HIPCHK(hipSetDevice(e->device));
    if (enable) {
        if (!e->d_counts) HIPCHK(hipMalloc(&e->d_counts, kCountBytes));
        HIPCHK(hipMemsetAsync(e->d_counts, 0, kCountBytes, e->stream));
        e->attempts = 0;
    }
    e->counting = enable != 0;
    return ELP_OK;
}

int elp_move_counts(elp_engine* e, uint64_t counts[4]) {
    if (!e || !counts) return fail(ELP_E_ARG, "NULL argument");
    if (!e->d_counts) return fail(ELP_E_STATE, "move counting was never enabled");
    HIPCHK(hipSetDevice(e->device));
    unsigned long long c[3];
    HIPCHK(sum_counts(e->d_counts, e->stream, c));
    HIPCHK(hipStreamSynchronize(e->stream));
    counts[0] = e->attempts;
    counts[1] = c[0]; counts[2] = c[1]; counts[3] = c[2];
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

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched replicas
// ---------------------------------------------------------------------------

__global__ void elp_set_cursor(int* c, int v) { *c = v; }
__global__ void elp_advance_cursor(int* c, int v) { *c += v; }

struct elp_batch {
    int kind;
    int device;
    int n_rep;
    int tt, tu, tv;
    int sublattice;
    int n_iter;
    int threads;
    size_t bytes_per_site;
    std::vector<int32_t> box;              // 3 per replica
    std::vector<elp::BatchRep> reps;
    uint64_t n_sites;                      // sites of the buffer (incl. alignment gaps)
    uint32_t n_tiles;
    void* d_lat = nullptr;
    elp::BatchRep* d_reps = nullptr;
    uint32_t* d_rep_of_tile = nullptr;
    elp::BatchStep* d_steps = nullptr;
    size_t steps_rows = 0;                 // capacity of d_steps in time units
    int* d_cursor = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<char> uploaded;
    bool timed = false;
    std::vector<elp::BatchStep> h_steps;   // staging of the last run's schedule
    hipGraphExec_t gexec = nullptr;        // cached graph of g_units launches + cursor advance
    int g_units = 0;
    const void* g_steps = nullptr;
    uint32_t rules = 0;                    // ELP_RULES_*
    unsigned long long* d_counts = nullptr;   // move counting, as in elp_engine
    bool counting = false;
    uint64_t attempts = 0;
};

static int batch_launch_one(elp_batch* b, bool probe, int k) {
    elp::BatchLaunch bl{b->d_reps, b->d_rep_of_tile, b->d_steps, b->d_cursor, k, b->n_rep, b->n_iter, b->rules,
                        b->counting ? b->d_counts : nullptr};
    const bool dense = b->sublattice == ELP_SUBLAT_DENSE;
    if (b->kind == 0)
        return elp::ring_batch_launch(b->tt, b->tu, b->tv, dense, b->threads, probe, (int)b->n_tiles, b->stream,
                                      (uint8_t*)b->d_lat, bl);
    return elp::lin_batch_launch(b->tt, b->tu, b->tv, dense, b->threads, probe, (int)b->n_tiles, b->stream,
                                 (uint16_t*)b->d_lat, bl);
}

static void batch_free(elp_batch* b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    (void)hipFree(b->d_lat);
    (void)hipFree(b->d_reps);
    (void)hipFree(b->d_rep_of_tile);
    (void)hipFree(b->d_steps);
    (void)hipFree(b->d_cursor);
    if (b->d_counts) (void)hipFree(b->d_counts);
    delete b;
}

static int batch_create(elp_batch** out, int kind, int device, int n_rep, const int32_t* boxes, int tt, int tu, int tv,
                        int sublattice, int n_iter) {
    if (!out) return fail(ELP_E_ARG, "out is NULL");
    *out = nullptr;
    if (n_rep <= 0 || n_rep > 65535 || !boxes) return fail(ELP_E_ARG, "n_rep %d out of [1, 65535] or boxes NULL", n_rep);
    if (sublattice != ELP_SUBLAT_CUBIC && sublattice != ELP_SUBLAT_DENSE) return fail(ELP_E_ARG, "bad sublattice %d", sublattice);
    if (n_iter < 0) return fail(ELP_E_ARG, "n_iter < 0");
    elp_batch* b = new elp_batch();
    b->kind = kind; b->device = device; b->n_rep = n_rep;
    b->tt = tt; b->tu = tu; b->tv = tv;
    b->sublattice = sublattice; b->n_iter = n_iter;
    b->threads = default_threads(kind, tt, tu, tv, sublattice);
    b->bytes_per_site = kind == 0 ? 1 : 2;
    b->box.assign(boxes, boxes + 3 * n_rep);
    uint64_t site = 0, tile = 0;
    for (int r = 0; r < n_rep; r++) {
        const int LT = boxes[3 * r], LU = boxes[3 * r + 1], LV = boxes[3 * r + 2];
        if (LT <= 0 || LU <= 0 || LV <= 0 || LT % tt || LU % tu || LV % tv) {
            delete b;
            return fail(ELP_E_ARG, "replica %d: box %dx%dx%d not divisible by tile %dx%dx%d", r, LT, LU, LV, tt, tu, tv);
        }
        if ((uint64_t)LT * LU * LV >= ((uint64_t)1 << 31)) {
            delete b;
            return fail(ELP_E_ARG, "replica %d: box %dx%dx%d has more than 2^31 - 1 sites", r, LT, LU, LV);
        }
        elp::BatchRep rep{};
        rep.site_off = site;
        rep.LT = LT; rep.LU = LU; rep.LV = LV;
        rep.ntile_t = LT / tt; rep.ntile_u = LU / tu;
        rep.tile0 = (uint32_t)tile;
        b->reps.push_back(rep);
        site += ((uint64_t)LT * LU * LV + 255) & ~(uint64_t)255;    // replicas start on 256-site boundaries
        tile += (uint64_t)(LT / tt) * (LU / tu) * (LV / tv);
    }
    if (tile >= (1ull << 31)) { delete b; return fail(ELP_E_ARG, "too many tiles (%llu)", (unsigned long long)tile); }
    b->n_sites = site;
    b->n_tiles = (uint32_t)tile;
    b->uploaded.assign(n_rep, 0);
    // the default launch shape, else the widest compiled one
    for (int th : {1024, 512, 256, 128, 64}) {
        if (batch_launch_one(b, true, 0) == 0) break;
        b->threads = th;
    }
    if (batch_launch_one(b, true, 0) != 0) {
        int th = b->threads;
        delete b;
        return fail(ELP_E_ARG, "batched schedule %s tile %dx%dx%d sublattice %d threads %d is not compiled in",
                    kind == 0 ? "ring" : "linear", tt, tu, tv, sublattice, th);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { delete b; return fail(ELP_E_NODEV, "no HIP device"); }
    if (device < 0 || device >= ndev) { delete b; return fail(ELP_E_ARG, "device %d out of range (%d devices)", device, ndev); }
    hipError_t err;
    std::vector<uint32_t> rep_of_tile(b->n_tiles);
    for (int r = 0; r < n_rep; r++) {
        const uint32_t n = (uint32_t)((b->box[3 * r] / tt) * (b->box[3 * r + 1] / tu) * (b->box[3 * r + 2] / tv));
        for (uint32_t i = 0; i < n; i++) rep_of_tile[b->reps[r].tile0 + i] = (uint32_t)r;
    }
#define BCHK(call, what)                                       \
    if ((err = (call)) != hipSuccess) {                        \
        batch_free(b);                                         \
        return hip_fail(err, what);                            \
    }
    BCHK(hipSetDevice(device), "hipSetDevice");
    BCHK(hipMalloc(&b->d_lat, b->n_sites * b->bytes_per_site), "hipMalloc batch lattice");
    BCHK(hipMalloc(&b->d_reps, sizeof(elp::BatchRep) * n_rep), "hipMalloc replica table");
    BCHK(hipMalloc(&b->d_rep_of_tile, sizeof(uint32_t) * b->n_tiles), "hipMalloc tile map");
    BCHK(hipMalloc(&b->d_cursor, sizeof(int)), "hipMalloc cursor");
    BCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
    BCHK(hipEventCreate(&b->ev0), "hipEventCreate");
    BCHK(hipEventCreate(&b->ev1), "hipEventCreate");
    BCHK(hipMemcpy(b->d_rep_of_tile, rep_of_tile.data(), sizeof(uint32_t) * b->n_tiles, hipMemcpyHostToDevice),
         "hipMemcpy tile map");
#undef BCHK
    *out = b;
    return ELP_OK;
}

static int batch_check_rep(elp_batch* b, int rep, const void* p, uint64_t nbytes, uint64_t* off, uint64_t* len) {
    if (!b || !p) return fail(ELP_E_ARG, "NULL argument");
    if (rep < 0 || rep >= b->n_rep) return fail(ELP_E_ARG, "replica %d out of range (%d replicas)", rep, b->n_rep);
    const uint64_t n = (uint64_t)b->box[3 * rep] * b->box[3 * rep + 1] * b->box[3 * rep + 2] * b->bytes_per_site;
    if (nbytes != n) return fail(ELP_E_ARG, "replica %d: %llu bytes, lattice is %llu", rep, (unsigned long long)nbytes,
                                 (unsigned long long)n);
    *off = b->reps[rep].site_off * b->bytes_per_site;
    *len = n;
    return ELP_OK;
}

// Graph of `units` launches (k = 0..units-1) followed by cursor += units.
static int batch_graph(elp_batch* b, int units) {
    if (b->gexec && b->g_units == units && b->g_steps == (const void*)b->d_steps) return ELP_OK;
    if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int k = 0; k < units && !rc; k++) rc = batch_launch_one(b, false, k);
    if (!rc) {
        hipLaunchKernelGGL(elp_advance_cursor, dim3(1), dim3(1), 0, b->stream, b->d_cursor, units);
        rc = (int)hipGetLastError();
    }
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &graph);
    if (rc) { if (graph) (void)hipGraphDestroy(graph); return fail(ELP_E_HIP, "capture: launch failed: %s", hipGetErrorString((hipError_t)rc)); }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) { b->gexec = nullptr; return hip_fail(e, "hipGraphInstantiate"); }
    b->g_units = units;
    b->g_steps = b->d_steps;
    return ELP_OK;
}

extern "C" {

int elp_ring_batch_create(elp_batch** out, int device, int n_rep, const int32_t* boxes, int tt, int tu, int tv,
                          int sublattice, int n_iter) {
    return batch_create(out, 0, device, n_rep, boxes, tt, tu, tv, sublattice, n_iter);
}

int elp_lin_batch_create(elp_batch** out, int device, int n_rep, const int32_t* boxes, int tt, int tu, int tv,
                         int sublattice, int n_iter) {
    return batch_create(out, 1, device, n_rep, boxes, tt, tu, tv, sublattice, n_iter);
}

int elp_batch_destroy(elp_batch* b) {
    if (!b) return ELP_OK;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    batch_free(b);
    return ELP_OK;
}

int elp_batch_upload(elp_batch* b, int rep, const void* lattice, uint64_t nbytes) {
    uint64_t off, len;
    int rc = batch_check_rep(b, rep, lattice, nbytes, &off, &len);
    if (rc) return rc;
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipMemcpyAsync((char*)b->d_lat + off, lattice, len, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    b->uploaded[rep] = 1;
    return ELP_OK;
}

int elp_batch_download(elp_batch* b, int rep, void* lattice, uint64_t nbytes) {
    uint64_t off, len;
    int rc = batch_check_rep(b, rep, lattice, nbytes, &off, &len);
    if (rc) return rc;
    if (!b->uploaded[rep]) return fail(ELP_E_STATE, "download of replica %d before upload", rep);
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipMemcpyAsync(lattice, (char*)b->d_lat + off, len, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return ELP_OK;
}

int elp_batch_count_moves(elp_batch* b, int enable) {
    if (!b) return fail(ELP_E_ARG, "NULL batch");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    if (enable) {
        if (!b->d_counts) HIPCHK(hipMalloc(&b->d_counts, kCountBytes));
        HIPCHK(hipMemsetAsync(b->d_counts, 0, kCountBytes, b->stream));
        b->attempts = 0;
    }
    if ((enable != 0) != b->counting && b->gexec) {   // the cached graph holds the other kernel
        (void)hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
        b->g_units = 0;
    }
    b->counting = enable != 0;
    return ELP_OK;
}

int elp_batch_move_counts(elp_batch* b, uint64_t counts[4]) {
    if (!b || !counts) return fail(ELP_E_ARG, "NULL argument");
    if (!b->d_counts) return fail(ELP_E_STATE, "move counting was never enabled");
    HIPCHK(hipSetDevice(b->device));
    unsigned long long c[3];
    HIPCHK(sum_counts(b->d_counts, b->stream, c));
    HIPCHK(hipStreamSynchronize(b->stream));
    counts[0] = b->attempts;
    counts[1] = c[0]; counts[2] = c[1]; counts[3] = c[2];
    return ELP_OK;
}

int elp_batch_run(elp_batch* b, int n_time, const int32_t* origins, const uint64_t* tau0, const uint32_t* keys,
                  int units_per_graph) {
    if (!b) return fail(ELP_E_ARG, "NULL batch");
    if (n_time < 0 || (n_time > 0 && (!origins || !tau0 || !keys))) return fail(ELP_E_ARG, "bad n_time/origins/tau0/keys");
    if (units_per_graph < 0) return fail(ELP_E_ARG, "units_per_graph < 0");
    for (int r = 0; r < b->n_rep; r++)
        if (!b->uploaded[r]) return fail(ELP_E_STATE, "run before upload of replica %d", r);
    const int R = b->n_rep;
    for (int i = 0; i < n_time; i++)
        for (int r = 0; r < R; r++) {
            const int32_t* o = origins + 3 * ((size_t)i * R + r);
            if (o[0] < 0 || o[0] >= b->tt || o[1] < 0 || o[1] >= b->tu || o[2] < 0 || o[2] >= b->tv)
                return fail(ELP_E_ARG, "origin (%d,%d,%d) of replica %d, time unit %d outside the tile", o[0], o[1], o[2], r, i);
        }
    HIPCHK(hipSetDevice(b->device));
    // the previous run may still read the schedule and the replica table
    HIPCHK(hipStreamSynchronize(b->stream));
    if (n_time == 0) return ELP_OK;
    if ((size_t)n_time > b->steps_rows) {
        (void)hipFree(b->d_steps);
        b->d_steps = nullptr;
        b->steps_rows = 0;
        HIPCHK(hipMalloc(&b->d_steps, sizeof(elp::BatchStep) * (size_t)n_time * R));
        b->steps_rows = (size_t)n_time;
    }
    b->h_steps.resize((size_t)n_time * R);
    for (int i = 0; i < n_time; i++)
        for (int r = 0; r < R; r++) {
            const int32_t* o = origins + 3 * ((size_t)i * R + r);
            b->h_steps[(size_t)i * R + r] = elp::BatchStep{o[0], o[1], o[2], (uint32_t)(tau0[r] + (uint64_t)i)};
        }
    for (int r = 0; r < R; r++) {
        b->reps[r].seed = keys[2 * r];
        b->reps[r].stream = keys[2 * r + 1];
    }
    HIPCHK(hipMemcpyAsync(b->d_steps, b->h_steps.data(), sizeof(elp::BatchStep) * b->h_steps.size(),
                          hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->d_reps, b->reps.data(), sizeof(elp::BatchRep) * R, hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(elp_set_cursor, dim3(1), dim3(1), 0, b->stream, b->d_cursor, 0);
    HIPCHK(hipGetLastError());
    if (b->counting)
        b->attempts += (uint64_t)n_time * b->n_tiles *
                       attempts_per_unit(b->kind, b->tt, b->tu, b->tv, b->tt, b->tu, b->tv, b->sublattice, b->n_iter);
    const int G = units_per_graph;
    const int replays = G > 0 ? n_time / G : 0;
    if (replays > 0) {
        int rc = batch_graph(b, G);
        if (rc) return rc;
    }
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    for (int i = 0; i < replays; i++) HIPCHK(hipGraphLaunch(b->gexec, b->stream));
    // remainder: rows cursor + k, cursor = replays * G after the replays
    for (int k = 0; k < n_time - replays * G; k++) {
        int rc = batch_launch_one(b, false, k);
        if (rc) return fail(ELP_E_HIP, "kernel launch failed: %s", hipGetErrorString((hipError_t)rc));
    }
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    b->timed = true;
    return ELP_OK;
}

int elp_batch_sync(elp_batch* b) {
    if (!b) return fail(ELP_E_ARG, "NULL batch");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    return ELP_OK;
}

int elp_batch_last_run_ms(elp_batch* b, float* ms) {
    if (!b || !ms) return fail(ELP_E_ARG, "NULL argument");
    if (!b->timed) return fail(ELP_E_STATE, "no run recorded");
    HIPCHK(hipEventElapsedTime(ms, b->ev0, b->ev1));
    return ELP_OK;
}

int elp_batch_info(elp_batch* b, int* threads_per_block, int* blocks_per_time_unit, uint64_t* sites) {
    if (!b) return fail(ELP_E_ARG, "NULL batch");
    if (threads_per_block) *threads_per_block = b->threads;
    if (blocks_per_time_unit) *blocks_per_time_unit = (int)b->n_tiles;
    if (sites) {
        uint64_t n = 0;
        for (int r = 0; r < b->n_rep; r++) n += (uint64_t)b->box[3 * r] * b->box[3 * r + 1] * b->box[3 * r + 2];
        *sites = n;
    }
    return ELP_OK;
}

int elp_batch_set_threads(elp_batch* b, int threads) {
    if (!b) return fail(ELP_E_ARG, "NULL batch");
    const int old = b->threads;
    b->threads = threads;
    if (threads <= 0 || batch_launch_one(b, true, 0) != 0) {
        b->threads = old;
        return fail(ELP_E_ARG, "batched launch with %d threads is not compiled for this schedule", threads);
    }
    if (threads != old && b->gexec) {
        (void)hipStreamSynchronize(b->stream);
        (void)hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
    }
    return ELP_OK;
}

int elp_set_rules(elp_engine* e, int rules) {
    if (!e) return fail(ELP_E_ARG, "NULL engine");
    if (rules != ELP_RULES_REFERENCE && !(rules == ELP_RULES_BALANCED && e->kind == 0))
        return fail(ELP_E_ARG, "rules %d not available for the %s engine", rules, e->kind == 0 ? "ring" : "linear");
    e->rules = (uint32_t)rules;
    return ELP_OK;
}

int elp_batch_set_rules(elp_batch* b, int rules) {
    if (!b) return fail(ELP_E_ARG, "NULL batch");
    if (rules != ELP_RULES_REFERENCE && !(rules == ELP_RULES_BALANCED && b->kind == 0))
        return fail(ELP_E_ARG, "rules %d not available for the %s batch", rules, b->kind == 0 ? "ring" : "linear");
    if ((uint32_t)rules != b->rules && b->gexec) {      // captured launches carry the old rules
        (void)hipStreamSynchronize(b->stream);
        (void)hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
    }
    b->rules = (uint32_t)rules;
    return ELP_OK;
}

int elp_set_threads(elp_engine* e, int threads) {
    if (!e) return fail(ELP_E_ARG, "NULL engine");
    const int old = e->threads;
    e->threads = threads;
    elp::RingLaunch dummy{};
    if (threads <= 0 || launch_one(e, true, dummy) != 0) {
        e->threads = old;
        return fail(ELP_E_ARG, "launch with %d threads is not compiled for this schedule", threads);
    }
    return ELP_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Learned proposal network
// ---------------------------------------------------------------------------

struct elp_net {
    int device;
    hipStream_t stream = nullptr;
    float* d_prm = nullptr;                // fp32 master parameters
    float* d_grad = nullptr;
    float* d_m1 = nullptr;
    float* d_m2 = nullptr;
    float* d_loss = nullptr;
    unsigned long long* d_stats = nullptr; // proposals, accepted
    int8_t* d_x = nullptr;                 // training features [cap][144]
    int32_t* d_y = nullptr;                // training moves [cap]
    int n_samples = 0, cap = 0;
    int adam_step = 0;
};

static void net_free(elp_net* n) {
    if (n->stream) (void)hipStreamDestroy(n->stream);
    (void)hipFree(n->d_prm); (void)hipFree(n->d_grad); (void)hipFree(n->d_m1); (void)hipFree(n->d_m2);
    (void)hipFree(n->d_loss); (void)hipFree(n->d_stats); (void)hipFree(n->d_x); (void)hipFree(n->d_y);
    delete n;
}

// Glorot-uniform weights from a 64-bit LCG of the seed, zero biases.
static void net_init_params(std::vector<float>& p, uint32_t seed) {
    p.assign(elp::LP_NPARAM, 0.f);
    uint64_t x = 0x9E3779B97F4A7C15ull ^ seed;
    auto uni = [&x]() {
        x = x * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(x >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int fan_in, int fan_out) {
        const double a = sqrt(6.0 / (fan_in + fan_out));
        for (int i = 0; i < fan_in * fan_out; i++) p[off + i] = (float)((2.0 * uni() - 1.0) * a);
    };
    fill(elp::LP_OFF_W1, elp::LP_NIN, elp::LP_NH);
    fill(elp::LP_OFF_W2, elp::LP_NH, elp::LP_NH);
    fill(elp::LP_OFF_W3, elp::LP_NH, elp::LP_NOUT);
}

extern "C" {

int elp_net_create(elp_net** out, int device, uint32_t seed) {
    if (!out) return fail(ELP_E_ARG, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(ELP_E_NODEV, "no HIP device");
    if (device < 0 || device >= ndev) return fail(ELP_E_ARG, "device %d out of range", device);
    elp_net* n = new elp_net();
    n->device = device;
    hipError_t err;
#define NCHK(call, what)                        \
    if ((err = (call)) != hipSuccess) {         \
        net_free(n);                            \
        return hip_fail(err, what);             \
    }
    NCHK(hipSetDevice(device), "hipSetDevice");
    NCHK(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking), "hipStreamCreate");
    const size_t pb = sizeof(float) * elp::LP_NPARAM;
    NCHK(hipMalloc(&n->d_prm, pb), "hipMalloc params");
    NCHK(hipMalloc(&n->d_grad, pb), "hipMalloc grad");
    NCHK(hipMalloc(&n->d_m1, pb), "hipMalloc adam m");
    NCHK(hipMalloc(&n->d_m2, pb), "hipMalloc adam v");
    NCHK(hipMalloc(&n->d_loss, sizeof(float)), "hipMalloc loss");
    NCHK(hipMalloc(&n->d_stats, 2 * sizeof(unsigned long long)), "hipMalloc stats");
    std::vector<float> p;
    net_init_params(p, seed);
    NCHK(hipMemcpy(n->d_prm, p.data(), pb, hipMemcpyHostToDevice), "upload params");
#undef NCHK
    int rc = 0;
    rc |= elp::lp_launch_fill(n->d_grad, elp::LP_NPARAM, 0.f, n->stream);
    rc |= elp::lp_launch_fill(n->d_m1, elp::LP_NPARAM, 0.f, n->stream);
    rc |= elp::lp_launch_fill(n->d_m2, elp::LP_NPARAM, 0.f, n->stream);
    if (rc || hipStreamSynchronize(n->stream) != hipSuccess) { net_free(n); return fail(ELP_E_HIP, "net init kernels"); }
    *out = n;
    return ELP_OK;
}

int elp_net_destroy(elp_net* n) {
    if (!n) return ELP_OK;
    (void)hipSetDevice(n->device);
    (void)hipStreamSynchronize(n->stream);
    net_free(n);
    return ELP_OK;
}

int elp_net_params(elp_net* n, float* params, uint64_t count) {
    if (!n || !params || count != (uint64_t)elp::LP_NPARAM) return fail(ELP_E_ARG, "need %d parameters", elp::LP_NPARAM);
    HIPCHK(hipSetDevice(n->device));
    HIPCHK(hipMemcpyAsync(params, n->d_prm, sizeof(float) * count, hipMemcpyDeviceToHost, n->stream));
    HIPCHK(hipStreamSynchronize(n->stream));
    return ELP_OK;
}

int elp_net_set_params(elp_net* n, const float* params, uint64_t count) {
    if (!n || !params || count != (uint64_t)elp::LP_NPARAM) return fail(ELP_E_ARG, "need %d parameters", elp::LP_NPARAM);
    HIPCHK(hipSetDevice(n->device));
    HIPCHK(hipMemcpyAsync(n->d_prm, params, sizeof(float) * count, hipMemcpyHostToDevice, n->stream));
    HIPCHK(hipStreamSynchronize(n->stream));
    return ELP_OK;
}

// device scratch of a host call, freed on return
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { (void)hipFree(p); }
};

int elp_net_forward(elp_net* n, const int8_t* x, int count, float* prob) {
    if (!n || !x || !prob || count <= 0) return fail(ELP_E_ARG, "bad forward arguments");
    HIPCHK(hipSetDevice(n->device));
    DevBuf dx, dp;
    HIPCHK(hipMalloc(&dx.p, (size_t)count * elp::LP_NIN));
    HIPCHK(hipMalloc(&dp.p, sizeof(float) * (size_t)count * elp::LP_NMOVE));
    HIPCHK(hipMemcpyAsync(dx.p, x, (size_t)count * elp::LP_NIN, hipMemcpyHostToDevice, n->stream));
    int rc = elp::lp_launch_forward(n->d_prm, (const int8_t*)dx.p, count, (float*)dp.p, n->stream);
    if (rc) return fail(ELP_E_HIP, "forward kernel: %s", hipGetErrorString((hipError_t)rc));
    HIPCHK(hipMemcpyAsync(prob, dp.p, sizeof(float) * (size_t)count * elp::LP_NMOVE, hipMemcpyDeviceToHost, n->stream));
    HIPCHK(hipStreamSynchronize(n->stream));
    return ELP_OK;
}

int elp_net_gradient(elp_net* n, const int8_t* x, const int32_t* y, int count, float* grad, float* loss) {
    if (!n || !x || !y || !grad || count <= 0) return fail(ELP_E_ARG, "bad gradient arguments");
    for (int i = 0; i < count; i++)
        if (y[i] < 0 || y[i] >= elp::LP_NMOVE) return fail(ELP_E_ARG, "label %d out of range", y[i]);
    HIPCHK(hipSetDevice(n->device));
    DevBuf dx, dy, dg;
    HIPCHK(hipMalloc(&dx.p, (size_t)count * elp::LP_NIN));
    HIPCHK(hipMalloc(&dy.p, sizeof(int32_t) * count));
    HIPCHK(hipMalloc(&dg.p, sizeof(float) * elp::LP_NPARAM));
    HIPCHK(hipMemcpyAsync(dx.p, x, (size_t)count * elp::LP_NIN, hipMemcpyHostToDevice, n->stream));
    HIPCHK(hipMemcpyAsync(dy.p, y, sizeof(int32_t) * count, hipMemcpyHostToDevice, n->stream));
    int rc = elp::lp_launch_fill((float*)dg.p, elp::LP_NPARAM, 0.f, n->stream);
    rc = rc ? rc : elp::lp_launch_fill(n->d_loss, 1, 0.f, n->stream);
    rc = rc ? rc : elp::lp_launch_grad(n->d_prm, (const int8_t*)dx.p, (const int32_t*)dy.p, count, 1.f / count,
                                       (float*)dg.p, n->d_loss, n->stream);
    if (rc) return fail(ELP_E_HIP, "gradient kernel: %s", hipGetErrorString((hipError_t)rc));
    HIPCHK(hipMemcpyAsync(grad, dg.p, sizeof(float) * elp::LP_NPARAM, hipMemcpyDeviceToHost, n->stream));
    float l = 0;
    HIPCHK(hipMemcpyAsync(&l, n->d_loss, sizeof(float), hipMemcpyDeviceToHost, n->stream));
    HIPCHK(hipStreamSynchronize(n->stream));
    if (loss) *loss = l / count;
    return ELP_OK;
}

int elp_net_harvest(elp_net* n, elp_engine* e, int count, uint32_t seed) {
    if (!n || !e || count <= 0) return fail(ELP_E_ARG, "bad harvest arguments");
    if (e->kind != 0) return fail(ELP_E_ARG, "harvest needs a ring engine");
    if (!e->uploaded) return fail(ELP_E_STATE, "harvest before upload");
    if (e->device != n->device) return fail(ELP_E_ARG, "engine and network on different devices");
    HIPCHK(hipSetDevice(n->device));
    HIPCHK(hipStreamSynchronize(e->stream));           // the lattice of the engine's last run
    if (n->n_samples + count > n->cap) {
        int cap = n->cap ? n->cap : 1 << 16;
        while (cap < n->n_samples + count) cap *= 2;
        int8_t* x = nullptr;
        int32_t* y = nullptr;
        HIPCHK(hipMalloc(&x, (size_t)cap * elp::LP_NIN));
        hipError_t err = hipMalloc(&y, sizeof(int32_t) * (size_t)cap);
        if (err != hipSuccess) { (void)hipFree(x); return hip_fail(err, "hipMalloc samples"); }
        if (n->n_samples) {
            HIPCHK(hipMemcpyAsync(x, n->d_x, (size_t)n->n_samples * elp::LP_NIN, hipMemcpyDeviceToDevice, n->stream));
            HIPCHK(hipMemcpyAsync(y, n->d_y, sizeof(int32_t) * n->n_samples, hipMemcpyDeviceToDevice, n->stream));
            HIPCHK(hipStreamSynchronize(n->stream));
        }
        (void)hipFree(n->d_x);
        (void)hipFree(n->d_y);
        n->d_x = x;
        n->d_y = y;
        n->cap = cap;
    }
    int rc = elp::lp_launch_harvest((const uint8_t*)e->d_lat, e->LT, e->LU, e->LV, count, seed, (uint32_t)n->n_samples,
                                    n->d_x + (size_t)n->n_samples * elp::LP_NIN, n->d_y + n->n_samples, n->stream);
    if (rc) return fail(ELP_E_HIP, "harvest kernel: %s", hipGetErrorString((hipError_t)rc));
    HIPCHK(hipStreamSynchronize(n->stream));
    n->n_samples += count;
    return ELP_OK;
}

int elp_net_samples(elp_net* n, int* count, int8_t* x, int32_t* y) {
    if (!n || !count) return fail(ELP_E_ARG, "NULL argument");
    *count = n->n_samples;
    HIPCHK(hipSetDevice(n->device));
    if (x && n->n_samples)
        HIPCHK(hipMemcpyAsync(x, n->d_x, (size_t)n->n_samples * elp::LP_NIN, hipMemcpyDeviceToHost, n->stream));
    if (y && n->n_samples)
        HIPCHK(hipMemcpyAsync(y, n->d_y, sizeof(int32_t) * n->n_samples, hipMemcpyDeviceToHost, n->stream));
    HIPCHK(hipStreamSynchronize(n->stream));
    return ELP_OK;
}

int elp_net_train(elp_net* n, int epochs, int batch, float lr, float* last_loss) {
    if (!n || epochs <= 0 || batch <= 0 || !(lr > 0)) return fail(ELP_E_ARG, "bad training arguments");
    if (n->n_samples == 0) return fail(ELP_E_STATE, "no training samples (elp_net_harvest)");
    HIPCHK(hipSetDevice(n->device));
    float l = 0;
    for (int ep = 0; ep < epochs; ep++) {
        int rc = elp::lp_launch_fill(n->d_loss, 1, 0.f, n->stream);
        for (int b0 = 0; b0 < n->n_samples && !rc; b0 += batch) {
            const int nb = (n->n_samples - b0) < batch ? (n->n_samples - b0) : batch;
            rc = elp::lp_launch_grad(n->d_prm, n->d_x + (size_t)b0 * elp::LP_NIN, n->d_y + b0, nb, 1.f / nb, n->d_grad,
                                     n->d_loss, n->stream);
            if (!rc)
                rc = elp::lp_launch_adam(n->d_prm, n->d_grad, n->d_m1, n->d_m2, elp::LP_NPARAM, lr, 0.9f, 0.999f, 1e-8f,
                                         ++n->adam_step, n->stream);
        }
        if (rc) return fail(ELP_E_HIP, "training kernels: %s", hipGetErrorString((hipError_t)rc));
        HIPCHK(hipMemcpyAsync(&l, n->d_loss, sizeof(float), hipMemcpyDeviceToHost, n->stream));
        HIPCHK(hipStreamSynchronize(n->stream));
    }
    if (last_loss) *last_loss = l / n->n_samples;
    return ELP_OK;
}

int elp_run_learned(elp_engine* e, elp_net* n, int n_time, const int32_t* origins, uint64_t tau0, uint32_t seed,
                    uint32_t stream, uint64_t* proposals, uint64_t* accepted) {
    if (!e || !n) return fail(ELP_E_ARG, "NULL engine or network");
    if (e->kind != 0 || e->tt != 16 || e->tu != 16 || e->tv != 16 || e->sublattice != ELP_SUBLAT_CUBIC)
        return fail(ELP_E_ARG, "the learned proposal runs on ring engines with 16^3 tiles and the cubic sub-lattice");
    if (e->device != n->device) return fail(ELP_E_ARG, "engine and network on different devices");
    if (n_time < 0 || (n_time > 0 && !origins)) return fail(ELP_E_ARG, "bad n_time/origins");
    if (!e->uploaded) return fail(ELP_E_STATE, "run before upload");
    for (int i = 0; i < n_time; i++) {
        const int32_t* o = origins + 3 * i;
        if (o[0] < 0 || o[0] >= 16 || o[1] < 0 || o[1] >= 16 || o[2] < 0 || o[2] >= 16)
            return fail(ELP_E_ARG, "origin (%d,%d,%d) of time unit %d outside the tile", o[0], o[1], o[2], i);
    }
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(n->stream));           // parameters from the last training step
    unsigned long long zero[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(n->d_stats, zero, sizeof zero, hipMemcpyHostToDevice, e->stream));
    elp::RingLaunch p{};
    p.LT = e->LT; p.LU = e->LU; p.LV = e->LV;
    p.ntile_t = e->LT / 16; p.ntile_u = e->LU / 16;
    p.seed = seed; p.stream = stream;
    p.n_iter = e->n_iter;
    const int nblocks = (e->LT / 16) * (e->LU / 16) * (e->LV / 16);
    HIPCHK(hipEventRecord(e->ev0, e->stream));
    for (int i = 0; i < n_time; i++) {
        p.ot = origins[3 * i]; p.ou = origins[3 * i + 1]; p.ov = origins[3 * i + 2];
        p.tau = (uint32_t)(tau0 + (uint64_t)i);
        int rc = elp::lp_launch_mc((uint8_t*)e->d_lat, p, nblocks, n->d_prm, n->d_stats, e->stream);
        if (rc) return fail(ELP_E_HIP, "learned kernel launch failed: %s", hipGetErrorString((hipError_t)rc));
    }
    HIPCHK(hipEventRecord(e->ev1, e->stream));
    e->timed = true;
    unsigned long long st[2];
    HIPCHK(hipMemcpyAsync(st, n->d_stats, sizeof st, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (proposals) *proposals = st[0];
    if (accepted) *accepted = st[1];
    return ELP_OK;
}

// Chain recovery on the device: per chain (in id order) its bond string,
// length and start site; optionally the tuv text or a frame summary.  Works on
// a view of one lattice in device memory: an engine's lattice or one replica
// of a batch.
struct LatView {
    int kind;                 // 0 ring, 1 linear
    int device;
    const void* lat;
    int LT, LU, LV;
    hipStream_t stream;
};

static int engine_view(elp_engine* e, int kind, LatView* v) {
    if (!e) return fail(ELP_E_ARG, "NULL engine");
    if (e->kind != kind) return fail(ELP_E_ARG, "chain recovery needs a %s engine", kind ? "linear" : "ring");
    if (!e->uploaded) return fail(ELP_E_STATE, "chain recovery before upload");
    *v = LatView{e->kind, e->device, e->d_lat, e->LT, e->LU, e->LV, e->stream};
    return ELP_OK;
}

static int batch_view(elp_batch* b, int rep, LatView* v) {
    if (!b) return fail(ELP_E_ARG, "NULL batch");
    if (rep < 0 || rep >= b->n_rep) return fail(ELP_E_ARG, "replica %d out of range (%d replicas)", rep, b->n_rep);
    if (!b->uploaded[rep]) return fail(ELP_E_STATE, "chain recovery of replica %d before upload", rep);
    *v = LatView{b->kind, b->device, (const char*)b->d_lat + b->reps[rep].site_off * b->bytes_per_site,
                 b->box[3 * rep], b->box[3 * rep + 1], b->box[3 * rep + 2], b->stream};
    return ELP_OK;
}

struct ChainBufs {
    DevBuf l2i, starts, count, bonds, lens, start3, line_len, rg2, nsl, err, offset, text, sum3, mid3, contacts;
    int n_max = 0;   // longest chain
};

static int chains_run(const LatView& v, const int32_t* label2id, int n_pol, int max_len, ChainBufs& B) {
    if (!label2id || n_pol <= 0 || max_len <= 0) return fail(ELP_E_ARG, "bad chain-recovery arguments");
    const int kind = v.kind;
    hipStream_t st = v.stream;
    HIPCHK(hipSetDevice(v.device));
    const size_t nl = (size_t)1 << 18;
    HIPCHK(hipMalloc(&B.l2i.p, sizeof(int32_t) * nl));
    HIPCHK(hipMalloc(&B.starts.p, sizeof(uint32_t) * (size_t)n_pol));
    HIPCHK(hipMalloc(&B.count.p, sizeof(uint32_t) * 2));
    HIPCHK(hipMalloc(&B.bonds.p, (size_t)n_pol * max_len));
    HIPCHK(hipMalloc(&B.lens.p, sizeof(int32_t) * n_pol));
    HIPCHK(hipMalloc(&B.start3.p, sizeof(int32_t) * 3 * n_pol));
    HIPCHK(hipMalloc(&B.line_len.p, sizeof(int32_t) * n_pol));
    HIPCHK(hipMalloc(&B.rg2.p, sizeof(double) * n_pol));
    HIPCHK(hipMalloc(&B.nsl.p, sizeof(int32_t) * n_pol));
    HIPCHK(hipMemcpyAsync(B.l2i.p, label2id, sizeof(int32_t) * nl, hipMemcpyHostToDevice, st));
    uint32_t zero[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(B.count.p, zero, sizeof zero, hipMemcpyHostToDevice, st));
    std::vector<int32_t> neg(n_pol, 0);
    HIPCHK(hipMemcpyAsync(B.lens.p, neg.data(), sizeof(int32_t) * n_pol, hipMemcpyHostToDevice, st));
    uint32_t* cnt = (uint32_t*)B.count.p;
    int rc = kind ? elp::chains_launch_lin_heads((const uint16_t*)v.lat, v.LT, v.LU, v.LV, (uint32_t*)B.starts.p,
                                                 (uint32_t)n_pol, cnt, st)
                  : elp::chains_launch_heads((const uint8_t*)v.lat, v.LT, v.LU, v.LV, (uint32_t*)B.starts.p,
                                             (uint32_t)n_pol, cnt, st);
    if (rc) return fail(ELP_E_HIP, "head kernel: %s", hipGetErrorString((hipError_t)rc));
    uint32_t nfound = 0;
    HIPCHK(hipMemcpyAsync(&nfound, cnt, sizeof nfound, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (nfound != (uint32_t)n_pol)
        return fail(ELP_E_STATE, "found %u %s, expected %d", nfound, kind ? "chain starts" : "ring heads", n_pol);
    if (kind)
        rc = elp::chains_launch_lin_walk((const uint16_t*)v.lat, v.LT, v.LU, v.LV, (const uint32_t*)B.starts.p,
                                         nfound, (const int32_t*)B.l2i.p, n_pol, max_len, (uint8_t*)B.bonds.p,
                                         (int32_t*)B.lens.p, (int32_t*)B.start3.p, (int32_t*)B.line_len.p,
                                         (double*)B.rg2.p, (int32_t*)B.nsl.p, cnt + 1, st);
    else
        rc = elp::chains_launch_walk((const uint8_t*)v.lat, v.LT, v.LU, v.LV, (const uint32_t*)B.starts.p, nfound,
                                     (const int32_t*)B.l2i.p, n_pol, max_len, (uint8_t*)B.bonds.p, (int32_t*)B.lens.p,
                                     (int32_t*)B.start3.p, (int32_t*)B.line_len.p, (double*)B.rg2.p, (int32_t*)B.nsl.p,
                                     cnt + 1, st);
    if (rc) return fail(ELP_E_HIP, "walk kernel: %s", hipGetErrorString((hipError_t)rc));
    uint32_t err = 0;
    HIPCHK(hipMemcpyAsync(&err, cnt + 1, sizeof err, hipMemcpyDeviceToHost, st));
    std::vector<int32_t> lens(n_pol);
    HIPCHK(hipMemcpyAsync(lens.data(), B.lens.p, sizeof(int32_t) * n_pol, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (err & 1u) return fail(ELP_E_STATE, "a chain label is not a known polymer id");
    if (err & 2u) return fail(ELP_E_STATE, "a chain is longer than %d", max_len);
    if (err & 4u) return fail(ELP_E_STATE, "a chain is shorter than its 23 head and id labels");
    for (int i = 0; i < n_pol; i++) {
        if (lens[i] <= 0) return fail(ELP_E_STATE, "polymer id %d found %s", i, lens[i] ? "twice" : "nowhere");
        B.n_max = lens[i] > B.n_max ? lens[i] : B.n_max;
    }
    return ELP_OK;
}

static int chains_get(elp_engine* e, int kind, const int32_t* label2id, int n_pol, int max_len, uint8_t* bonds,
                      int32_t* lens, int32_t* starts) {
    LatView v;
    int rc = engine_view(e, kind, &v);
    if (rc) return rc;
    ChainBufs B;
    if ((rc = chains_run(v, label2id, n_pol, max_len, B))) return rc;
    if (bonds) HIPCHK(hipMemcpyAsync(bonds, B.bonds.p, (size_t)n_pol * max_len, hipMemcpyDeviceToHost, v.stream));
    if (lens) HIPCHK(hipMemcpyAsync(lens, B.lens.p, sizeof(int32_t) * n_pol, hipMemcpyDeviceToHost, v.stream));
    if (starts) HIPCHK(hipMemcpyAsync(starts, B.start3.p, sizeof(int32_t) * 3 * n_pol, hipMemcpyDeviceToHost, v.stream));
    HIPCHK(hipStreamSynchronize(v.stream));
    return ELP_OK;
}

int elp_ring_chains(elp_engine* e, const int32_t* label2id, int n_pol, int max_len, uint8_t* bonds, int32_t* lens,
                    int32_t* starts) {
    return chains_get(e, 0, label2id, n_pol, max_len, bonds, lens, starts);
}

int elp_lin_chains(elp_engine* e, const int32_t* label2id, int n_pol, int max_len, uint8_t* bonds, int32_t* lens,
                   int32_t* starts) {
    return chains_get(e, 1, label2id, n_pol, max_len, bonds, lens, starts);
}

int elp_ring_chain_stats(elp_engine* e, const int32_t* label2id, int n_pol, int max_len, double* rg2,
                         int32_t* n_sl) {
    LatView v;
    int rc = engine_view(e, 0, &v);
    if (rc) return rc;
    ChainBufs B;
    if ((rc = chains_run(v, label2id, n_pol, max_len, B))) return rc;
    if (rg2) HIPCHK(hipMemcpyAsync(rg2, B.rg2.p, sizeof(double) * n_pol, hipMemcpyDeviceToHost, v.stream));
    if (n_sl) HIPCHK(hipMemcpyAsync(n_sl, B.nsl.p, sizeof(int32_t) * n_pol, hipMemcpyDeviceToHost, v.stream));
    HIPCHK(hipStreamSynchronize(v.stream));
    return ELP_OK;
}

static int frame_summary(const LatView& v, const int32_t* label2id, int n_pol, int max_len, int32_t* start3,
                         int64_t* sum3, int32_t* mid3, double* rg2, int32_t* n_sl, uint64_t* contacts) {
    if (!start3 || !sum3 || !mid3) return fail(ELP_E_ARG, "frame summary needs start3, sum3 and mid3");
    ChainBufs B;
    int rc = chains_run(v, label2id, n_pol, max_len, B);
    if (rc) return rc;
    if (B.n_max > elp::kSummaryMaxLen)
        return fail(ELP_E_ARG, "frame summary supports chains of up to %d monomers (found %d)", elp::kSummaryMaxLen,
                    B.n_max);
    hipStream_t st = v.stream;
    HIPCHK(hipMalloc(&B.sum3.p, sizeof(int64_t) * 3 * (size_t)n_pol));
    HIPCHK(hipMalloc(&B.mid3.p, sizeof(int32_t) * 3 * (size_t)n_pol));
    HIPCHK(hipMalloc(&B.contacts.p, sizeof(uint64_t) * (size_t)max_len));
    HIPCHK(hipMemsetAsync(B.contacts.p, 0, sizeof(uint64_t) * (size_t)max_len, st));
    rc = elp::chains_launch_summary((const uint8_t*)B.bonds.p, (const int32_t*)B.lens.p, n_pol, max_len, B.n_max,
                                    v.kind, (int64_t*)B.sum3.p, (int32_t*)B.mid3.p, (unsigned long long*)B.contacts.p,
                                    st);
    if (rc) return fail(ELP_E_HIP, "summary kernel: %s", hipGetErrorString((hipError_t)rc));
    HIPCHK(hipMemcpyAsync(start3, B.start3.p, sizeof(int32_t) * 3 * n_pol, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(sum3, B.sum3.p, sizeof(int64_t) * 3 * n_pol, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(mid3, B.mid3.p, sizeof(int32_t) * 3 * n_pol, hipMemcpyDeviceToHost, st));
    if (rg2) HIPCHK(hipMemcpyAsync(rg2, B.rg2.p, sizeof(double) * n_pol, hipMemcpyDeviceToHost, st));
    if (n_sl) HIPCHK(hipMemcpyAsync(n_sl, B.nsl.p, sizeof(int32_t) * n_pol, hipMemcpyDeviceToHost, st));
    if (contacts)
        HIPCHK(hipMemcpyAsync(contacts, B.contacts.p, sizeof(uint64_t) * max_len, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return ELP_OK;
}

int elp_ring_frame_summary(elp_engine* e, const int32_t* label2id, int n_pol, int max_len, int32_t* start3,
                           int64_t* sum3, int32_t* mid3, double* rg2, int32_t* n_sl, uint64_t* contacts) {
    LatView v;
    int rc = engine_view(e, 0, &v);
    return rc ? rc : frame_summary(v, label2id, n_pol, max_len, start3, sum3, mid3, rg2, n_sl, contacts);
}

int elp_lin_frame_summary(elp_engine* e, const int32_t* label2id, int n_pol, int max_len, int32_t* start3,
                          int64_t* sum3, int32_t* mid3, double* rg2, int32_t* n_sl, uint64_t* contacts) {
    LatView v;
    int rc = engine_view(e, 1, &v);
    return rc ? rc : frame_summary(v, label2id, n_pol, max_len, start3, sum3, mid3, rg2, n_sl, contacts);
}

int elp_batch_frame_summary(elp_batch* b, int rep, const int32_t* label2id, int n_pol, int max_len, int32_t* start3,
                            int64_t* sum3, int32_t* mid3, double* rg2, int32_t* n_sl, uint64_t* contacts) {
    LatView v;
    int rc = batch_view(b, rep, &v);
    return rc ? rc : frame_summary(v, label2id, n_pol, max_len, start3, sum3, mid3, rg2, n_sl, contacts);
}

// ---- frame summary from a static scratch array (no allocation) ----
// Every device buffer of elp_ring_frame_summary_noalloc lives in `scratch`, one
// copy per GPU, held until the process ends.  Host code never passes the
// array's name to a kernel or a copy: it uses the device address that
// hipGetSymbolAddress returns for the current GPU, looked up once per device.
constexpr size_t kFrameScratchBytes = (size_t)32 << 20;
constexpr int kMaxScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kFrameScratchBytes];
static void* g_scratch_ptr[kMaxScratchDevices];
static std::mutex g_scratch_mu;

// device address of this GPU's scratch (the device becomes current)
static int scratch_on(int device, void** out) {
    if (device < 0 || device >= kMaxScratchDevices) return fail(ELP_E_ARG, "device %d out of range", device);
    HIPCHK(hipSetDevice(device));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !p) return hip_fail(err, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return ELP_OK;
}

struct FrameScratch {
    int32_t* l2i;            // label -> polymer id (2^18)
    uint32_t* starts;        // ring start cursors
    uint32_t* stat;          // [0] heads found, [1] walk error bits, [2] longest ring, [3] ids not found
    uint8_t* bonds;          // n_pol x max_len bond strings
    int32_t *lens, *start3, *line_len, *nsl, *mid3;
    double* rg2;
    int64_t* sum3;
    unsigned long long* contacts;
};

// Buffers of one frame summary carved from `base` in 256-byte aligned pieces;
// returns the bytes they take (base = 0 only sizes them).
static size_t frame_scratch_layout(uintptr_t base, int n_pol, int max_len, FrameScratch* S) {
    size_t off = 0;
    auto take = [&](size_t bytes) { uintptr_t p = base + off; off += (bytes + 255) & ~(size_t)255; return p; };
    const size_t n = (size_t)n_pol;
    S->l2i = (int32_t*)take(sizeof(int32_t) << 18);
    S->starts = (uint32_t*)take(sizeof(uint32_t) * n);
    S->stat = (uint32_t*)take(sizeof(uint32_t) * 4);
    S->bonds = (uint8_t*)take(n * (size_t)max_len);
    S->lens = (int32_t*)take(sizeof(int32_t) * n);
    S->start3 = (int32_t*)take(sizeof(int32_t) * 3 * n);
    S->line_len = (int32_t*)take(sizeof(int32_t) * n);
    S->nsl = (int32_t*)take(sizeof(int32_t) * n);
    S->mid3 = (int32_t*)take(sizeof(int32_t) * 3 * n);
    S->rg2 = (double*)take(sizeof(double) * n);
    S->sum3 = (int64_t*)take(sizeof(int64_t) * 3 * n);
    S->contacts = (unsigned long long*)take(sizeof(unsigned long long) * (size_t)max_len);
    return off;
}

static int tuv_text(elp_engine* e, int kind, const int32_t* label2id, int n_pol, int max_len,
                    int header_max_pol_length, char* text, uint64_t cap, uint64_t* len) {
    if (!len) return fail(ELP_E_ARG, "len is NULL");
    LatView v;
    int rc = engine_view(e, kind, &v);
    if (rc) return rc;
    ChainBufs B;
    if ((rc = chains_run(v, label2id, n_pol, max_len, B))) return rc;
    std::vector<int32_t> ll(n_pol);
    HIPCHK(hipMemcpyAsync(ll.data(), B.line_len.p, sizeof(int32_t) * n_pol, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    char head[256];
    const int hn = snprintf(head, sizeof head, "LT= %i\nLU= %i\nLV= %i\nnp= %i\nmaxPolLength= %i\n", e->LT, e->LU,
                            e->LV, n_pol, header_max_pol_length);
    std::vector<uint64_t> off(n_pol);
    uint64_t total = 0;
    for (int i = 0; i < n_pol; i++) { off[i] = total; total += (uint64_t)ll[i]; }
    *len = (uint64_t)hn + total;
    if (!text) return ELP_OK;                          // size query
    if (cap < *len) return fail(ELP_E_ARG, "text buffer of %llu bytes, need %llu", (unsigned long long)cap,
                                (unsigned long long)*len);
    HIPCHK(hipMalloc(&B.offset.p, sizeof(uint64_t) * n_pol));
    HIPCHK(hipMalloc(&B.text.p, total ? total : 1));
    HIPCHK(hipMemcpyAsync(B.offset.p, off.data(), sizeof(uint64_t) * n_pol, hipMemcpyHostToDevice, e->stream));
    rc = elp::chains_launch_text((const uint8_t*)B.bonds.p, (const int32_t*)B.lens.p, (const int32_t*)B.start3.p,
                                 (const uint64_t*)B.offset.p, n_pol, max_len, (char*)B.text.p, e->stream);
    if (rc) return fail(ELP_E_HIP, "text kernel: %s", hipGetErrorString((hipError_t)rc));
    memcpy(text, head, (size_t)hn);
    HIPCHK(hipMemcpyAsync(text + hn, B.text.p, total, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return ELP_OK;
}

int elp_ring_frame_summary_noalloc(elp_engine* e, const int32_t* label2id, int n_pol, int max_len,
                                   int32_t* start3, int64_t* sum3, int32_t* mid3, double* rg2, int32_t* n_sl,
                                   uint64_t* contacts) {
    if (!e || !label2id || n_pol <= 0 || max_len <= 0 || !start3 || !sum3 || !mid3)
        return fail(ELP_E_ARG, "frame summary needs an engine, label2id, n_pol > 0, max_len > 0, start3, sum3, mid3");
    if (e->kind != 0) return fail(ELP_E_ARG, "frame summary needs a ring engine");
    if (!e->uploaded) return fail(ELP_E_STATE, "frame summary before upload");
    FrameScratch S;
    const size_t need = frame_scratch_layout(0, n_pol, max_len, &S);
    if (need > kFrameScratchBytes)
        return fail(ELP_E_ARG, "frame summary of %d rings of up to %d monomers needs %zu bytes of scratch; "
                    "the static scratch holds %zu", n_pol, max_len, need, kFrameScratchBytes);
    void* base = nullptr;
    int rc = scratch_on(e->device, &base);
    if (rc) return rc;
    frame_scratch_layout((uintptr_t)base, n_pol, max_len, &S);
    hipStream_t st = e->stream;
    // the previous call's contents stay in the scratch: clear what is accumulated
    if ((rc = elp::chains_launch_fill32(S.stat, 0u, 4, st)) ||
        (rc = elp::chains_launch_fill32((uint32_t*)S.lens, 0u, (size_t)n_pol, st)) ||
        (rc = elp::chains_launch_fill32((uint32_t*)S.contacts, 0u, 2 * (size_t)max_len, st)))
        return fail(ELP_E_HIP, "fill kernel: %s", hipGetErrorString((hipError_t)rc));
    HIPCHK(hipMemcpyAsync(S.l2i, label2id, sizeof(int32_t) << 18, hipMemcpyHostToDevice, st));
    rc = elp::chains_launch_heads((const uint8_t*)e->d_lat, e->LT, e->LU, e->LV, S.starts, (uint32_t)n_pol, S.stat, st);
    if (rc) return fail(ELP_E_HIP, "head kernel: %s", hipGetErrorString((hipError_t)rc));
    uint32_t stat[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(stat, S.stat, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (stat[0] != (uint32_t)n_pol) return fail(ELP_E_STATE, "found %u ring heads, expected %d", stat[0], n_pol);
    rc = elp::chains_launch_walk((const uint8_t*)e->d_lat, e->LT, e->LU, e->LV, S.starts, stat[0], S.l2i, n_pol,
                                 max_len, S.bonds, S.lens, S.start3, S.line_len, S.rg2, S.nsl, S.stat + 1, st);
    if (rc) return fail(ELP_E_HIP, "walk kernel: %s", hipGetErrorString((hipError_t)rc));
    rc = elp::chains_launch_check_lens(S.lens, n_pol, S.stat + 2, st);
    if (rc) return fail(ELP_E_HIP, "length check kernel: %s", hipGetErrorString((hipError_t)rc));
    HIPCHK(hipMemcpyAsync(stat, S.stat, sizeof stat, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (stat[1] & 1u) return fail(ELP_E_STATE, "a chain label is not a known polymer id");
    if (stat[1] & 2u) return fail(ELP_E_STATE, "a chain is longer than %d", max_len);
    if (stat[3]) return fail(ELP_E_STATE, "%u polymer ids found nowhere", stat[3]);
    const int n_max = (int)stat[2];
    if (n_max > elp::kSummaryMaxLen)
        return fail(ELP_E_ARG, "frame summary supports rings of up to %d monomers (found %d)", elp::kSummaryMaxLen,
                    n_max);
    rc = elp::chains_launch_summary(S.bonds, S.lens, n_pol, max_len, n_max, 0, S.sum3, S.mid3, S.contacts, st);
    if (rc) return fail(ELP_E_HIP, "summary kernel: %s", hipGetErrorString((hipError_t)rc));
    HIPCHK(hipMemcpyAsync(start3, S.start3, sizeof(int32_t) * 3 * n_pol, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(sum3, S.sum3, sizeof(int64_t) * 3 * n_pol, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(mid3, S.mid3, sizeof(int32_t) * 3 * n_pol, hipMemcpyDeviceToHost, st));
    if (rg2) HIPCHK(hipMemcpyAsync(rg2, S.rg2, sizeof(double) * n_pol, hipMemcpyDeviceToHost, st));
    if (n_sl) HIPCHK(hipMemcpyAsync(n_sl, S.nsl, sizeof(int32_t) * n_pol, hipMemcpyDeviceToHost, st));
    if (contacts) HIPCHK(hipMemcpyAsync(contacts, S.contacts, sizeof(uint64_t) * max_len, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return ELP_OK;
}

int elp_ring_tuv(elp_engine* e, const int32_t* label2id, int n_pol, int max_len, int header_max_pol_length,
                 char* text, uint64_t cap, uint64_t* len) {
    return tuv_text(e, 0, label2id, n_pol, max_len, header_max_pol_length, text, cap, len);
}

int elp_lin_tuv(elp_engine* e, const int32_t* label2id, int n_pol, int max_len, int header_max_pol_length,
                char* text, uint64_t cap, uint64_t* len) {
    return tuv_text(e, 1, label2id, n_pol, max_len, header_max_pol_length, text, cap, len);
}

int elp_net_patch_moves(int device, const uint8_t* patches, const int32_t* moves, int count, uint8_t* out,
                        int32_t* ok, int32_t* rev) {
    if (!patches || !moves || !out || !ok || !rev || count <= 0) return fail(ELP_E_ARG, "bad patch-move arguments");
    HIPCHK(hipSetDevice(device));
    DevBuf dp, dm, dout, dok, drev;
    HIPCHK(hipMalloc(&dp.p, (size_t)count * 125));
    HIPCHK(hipMalloc(&dm.p, sizeof(int32_t) * count));
    HIPCHK(hipMalloc(&dout.p, (size_t)count * 125));
    HIPCHK(hipMalloc(&dok.p, sizeof(int32_t) * count));
    HIPCHK(hipMalloc(&drev.p, sizeof(int32_t) * count));
    HIPCHK(hipMemcpy(dp.p, patches, (size_t)count * 125, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dm.p, moves, sizeof(int32_t) * count, hipMemcpyHostToDevice));
    int rc = elp::lp_launch_patch_move((const uint8_t*)dp.p, (const int32_t*)dm.p, count, (uint8_t*)dout.p,
                                       (int32_t*)dok.p, (int32_t*)drev.p, 0);
    if (rc) return fail(ELP_E_HIP, "patch-move kernel: %s", hipGetErrorString((hipError_t)rc));
    HIPCHK(hipMemcpy(out, dout.p, (size_t)count * 125, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ok, dok.p, sizeof(int32_t) * count, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(rev, drev.p, sizeof(int32_t) * count, hipMemcpyDeviceToHost));
    return ELP_OK;
}

int elp_net_patch_features(int device, const uint8_t* patches, int count, int8_t* x) {
    if (!patches || !x || count <= 0) return fail(ELP_E_ARG, "bad patch-feature arguments");
    HIPCHK(hipSetDevice(device));
    DevBuf dp, dx;
    HIPCHK(hipMalloc(&dp.p, (size_t)count * 125));
    HIPCHK(hipMalloc(&dx.p, (size_t)count * elp::LP_NIN));
    HIPCHK(hipMemcpy(dp.p, patches, (size_t)count * 125, hipMemcpyHostToDevice));
    int rc = elp::lp_launch_patch_features((const uint8_t*)dp.p, count, (int8_t*)dx.p, 0);
    if (rc) return fail(ELP_E_HIP, "patch-feature kernel: %s", hipGetErrorString((hipError_t)rc));
    HIPCHK(hipMemcpy(x, dx.p, (size_t)count * elp::LP_NIN, hipMemcpyDeviceToHost));
    return ELP_OK;
}

int elp_net_tile_check(int device, const uint8_t* tiles, int n_tiles, const int32_t* cases, int count,
                       int32_t* result) {
    if (!tiles || !cases || !result || n_tiles <= 0 || count <= 0) return fail(ELP_E_ARG, "bad tile-check arguments");
    for (int i = 0; i < count; i++) {
        const int32_t* c = cases + 5 * i;
        if (c[0] < 0 || c[0] >= n_tiles || c[1] < 0 || c[1] >= 16 || c[2] < 0 || c[2] >= 16 || c[3] < 0 || c[3] >= 16 ||
            c[4] < 0 || c[4] >= elp::LP_NMOVE)
            return fail(ELP_E_ARG, "tile-check case %d out of range", i);
    }
    HIPCHK(hipSetDevice(device));
    DevBuf dt, dc, dr;
    HIPCHK(hipMalloc(&dt.p, (size_t)n_tiles * 4096));
    HIPCHK(hipMalloc(&dc.p, sizeof(int32_t) * 5 * (size_t)count));
    HIPCHK(hipMalloc(&dr.p, sizeof(int32_t) * 2 * (size_t)count));
    HIPCHK(hipMemcpy(dt.p, tiles, (size_t)n_tiles * 4096, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dc.p, cases, sizeof(int32_t) * 5 * (size_t)count, hipMemcpyHostToDevice));
    int rc = elp::lp_launch_tile_check((const uint8_t*)dt.p, (const int32_t*)dc.p, count, (int32_t*)dr.p, 0);
    if (rc) return fail(ELP_E_HIP, "tile-check kernel: %s", hipGetErrorString((hipError_t)rc));
    HIPCHK(hipMemcpy(result, dr.p, sizeof(int32_t) * 2 * (size_t)count, hipMemcpyDeviceToHost));
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
