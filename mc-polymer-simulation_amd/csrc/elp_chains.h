// elp_chains.h -- launchers of the ring recovery and tuv text kernels
// (internal to libelp.so; the C ABI is include/elp_gpu.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace elp {

int chains_launch_heads(const uint8_t* lat, int LT, int LU, int LV, uint32_t* starts, uint32_t cap, uint32_t* count,
                        hipStream_t st);
int chains_launch_walk(const uint8_t* lat, int LT, int LU, int LV, const uint32_t* starts, uint32_t n_start,
                       const int32_t* label2id, int n_pol, int max_len, uint8_t* bonds, int32_t* lens, int32_t* start3,
                       int32_t* line_len, double* rg2, int32_t* nsl, uint32_t* err, hipStream_t st);
int chains_launch_text(const uint8_t* bonds, const int32_t* lens, const int32_t* start3, const uint64_t* offset,
                       int n_pol, int max_len, char* text, hipStream_t st);

}  // namespace elp
