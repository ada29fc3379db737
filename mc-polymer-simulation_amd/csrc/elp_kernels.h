// elp_kernels.h -- launch parameters and launchers of the gfx950 kernels
// (internal to libelp.so; the public C ABI is include/elp_gpu.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace elp {

// Kernel arguments of one time unit (passed by value in the kernarg segment).
struct RingLaunch {
    int LT, LU, LV;          // box
    int ot, ou, ov;          // tile-grid origin of this time unit, in [0, tile edge)
    int ntile_t, ntile_u;    // tiles per axis (t, u)
    uint32_t tau;            // Philox counter word 0: time unit
    uint32_t seed, stream;   // Philox key
    int n_iter;              // iterations of the 4-substep sequence
    uint32_t trans[4];       // first bond of each decomposition (SetGPUTrans)
    uint32_t rules;          // ring move rates: 0 reference, 1 balanced (ELP_RULES_*)
    unsigned long long* counts;   // counting launches: attempts at empty sites, leaving the tile, accepted
};

using LinLaunch = RingLaunch;

// One replica of a batched launch (device table, read once per workgroup).
struct BatchRep {
    uint64_t site_off;       // first site of the replica in the batch lattice buffer
    int LT, LU, LV;          // box
    int ntile_t, ntile_u;    // tiles per axis (t, u)
    uint32_t tile0;          // first global tile of the replica
    uint32_t seed, stream;   // Philox key
};

// Tile-grid origin and Philox time-unit counter of one replica for one time unit.
struct BatchStep {
    int ot, ou, ov;
    uint32_t tau;
};

// Kernel arguments of one batched time unit: every tile of every replica.
// The time unit read is row (*cursor + k) of steps[n_time][n_rep], so a
// captured graph of G launches (k = 0..G-1) followed by cursor += G replays
// consecutive time units.
struct BatchLaunch {
    const BatchRep* reps;
    const uint32_t* rep_of_tile;   // global tile -> replica
    const BatchStep* steps;
    const int* cursor;
    int k;
    int n_rep;
    int n_iter;
    uint32_t rules;
    unsigned long long* counts;   // counting launches (see RingLaunch)
};

// Outcome of one move attempt, returned by the move functions: the site is
// empty; the move would touch a site outside the tile (after its random
// gates); the move is rejected otherwise; the move is made.
enum : uint32_t { kMoveEmpty = 0, kMoveFace = 1, kMoveRejected = 2, kMoveAccepted = 3 };

// Move counters of counting launches: kCountSlots slots of kCountStride
// words (one 128-byte line each); workgroup b adds to slot b % kCountSlots,
// so that the atomics of a launch spread over many lines instead of
// serialising on one (a single line cost about 12 ns per wave atomic).
// Word k of a slot: 0 empty, 1 face, 2 accepted; the totals are the sums
// over the slots.
constexpr int kCountSlots = 512;
constexpr int kCountStride = 16;

// Launch one time unit (nblocks = number of tiles).  probe=true only checks
// that the schedule is compiled in.  A non-null p.counts (kCountSlots x
// kCountStride words) launches the counting instantiation of the kernel,
// which adds the outcomes of its attempts to its slot (empty, face, accepted)
// and moves exactly as the plain one.  Returns 0, a hipError_t, or -1 when
// the schedule is not available.
int ring_launch(int tt, int tu, int tv, bool dense, int threads, bool probe, int nblocks, hipStream_t stream,
                uint8_t* glat, const RingLaunch& p);
int lin_launch(int tt, int tu, int tv, bool dense, int threads, bool probe, int nblocks, hipStream_t stream,
               uint16_t* glat, const LinLaunch& p);
// Batched time unit over all replicas (nblocks = total tiles).
int ring_batch_launch(int tt, int tu, int tv, bool dense, int threads, bool probe, int nblocks, hipStream_t stream,
                      uint8_t* gbase, const BatchLaunch& b);
int lin_batch_launch(int tt, int tu, int tv, bool dense, int threads, bool probe, int nblocks, hipStream_t stream,
                     uint16_t* gbase, const BatchLaunch& b);

}  // namespace elp
