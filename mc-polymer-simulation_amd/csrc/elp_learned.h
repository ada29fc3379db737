// elp_learned.h -- layout of the learned-proposal network and the launchers
// of its kernels (internal to libelp.so; the C ABI is include/elp_gpu.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "elp_kernels.h"

namespace elp {

constexpr int LP_NIN = 144;     // features per site
constexpr int LP_NH = 64;       // hidden width
constexpr int LP_NOUT = 16;     // logits computed (11 used)
constexpr int LP_NMOVE = 11;    // moves of a site
// fp32 parameter vector: W1[144][64] | W2[64][64] | b2[64] | W3[64][16] | b3[16]
constexpr int LP_OFF_W1 = 0;
constexpr int LP_OFF_W2 = LP_OFF_W1 + LP_NIN * LP_NH;
constexpr int LP_OFF_B2 = LP_OFF_W2 + LP_NH * LP_NH;
constexpr int LP_OFF_W3 = LP_OFF_B2 + LP_NH;
constexpr int LP_OFF_B3 = LP_OFF_W3 + LP_NH * LP_NOUT;
constexpr int LP_NPARAM = LP_OFF_B3 + LP_NOUT;     // 14416

int lp_launch_forward(const float* prm, const int8_t* x, int n, float* prob, hipStream_t st);
int lp_launch_grad(const float* prm, const int8_t* x, const int32_t* y, int n, float scale, float* grad, float* loss,
                   hipStream_t st);
int lp_launch_adam(float* prm, float* grad, float* m1, float* m2, int n, float lr, float b1, float b2, float eps,
                   int step, hipStream_t st);
int lp_launch_fill(float* p, int n, float v, hipStream_t st);
int lp_launch_harvest(const uint8_t* lat, int LT, int LU, int LV, int n, uint32_t seed, uint32_t stream, int8_t* x,
                      int32_t* y, hipStream_t st);
int lp_launch_patch_move(const uint8_t* in, const int32_t* mv, int n, uint8_t* out, int32_t* ok, int32_t* rev,
                         hipStream_t st);
int lp_launch_patch_features(const uint8_t* in, int n, int8_t* x, hipStream_t st);
int lp_launch_tile_check(const uint8_t* tiles, const int32_t* cases, int n, int32_t* res, hipStream_t st);
int lp_launch_mc(uint8_t* glat, const RingLaunch& p, int nblocks, const float* prm, unsigned long long* stats,
                 hipStream_t st);

}  // namespace elp
