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
};

using LinLaunch = RingLaunch;

// Launch one time unit (nblocks = number of tiles).  probe=true only checks
// that the schedule is compiled in.  Returns 0, a hipError_t, or -1 when the
// schedule is not available.
int ring_launch(int tt, int tu, int tv, bool dense, int threads, bool probe, int nblocks, hipStream_t stream,
                uint8_t* glat, const RingLaunch& p);
int lin_launch(int tt, int tu, int tv, bool dense, int threads, bool probe, int nblocks, hipStream_t stream,
               uint16_t* glat, const LinLaunch& p);

}  // namespace elp
