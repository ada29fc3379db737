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
int chains_launch_lin_heads(const uint16_t* lat, int LT, int LU, int LV, uint32_t* starts, uint32_t cap,
                            uint32_t* count, hipStream_t st);
int chains_launch_lin_walk(const uint16_t* lat, int LT, int LU, int LV, const uint32_t* starts, uint32_t n_start,
                           const int32_t* label2id, int n_pol, int max_len, uint8_t* bonds, int32_t* lens,
                           int32_t* start3, int32_t* line_len, double* rg2, int32_t* nsl, uint32_t* err,
                           hipStream_t st);
int chains_launch_text(const uint8_t* bonds, const int32_t* lens, const int32_t* start3, const uint64_t* offset,
                       int n_pol, int max_len, char* text, hipStream_t st);
// Per chain (ring or linear: a linear chain's last bond 0xf adds no step):
// sum of its monomer positions and its middle monomer (index n/2 of a ring,
// (n-1)/2 of a linear chain) relative to the start; contacts[g] += pairs g
// apart that touch.  n_max = longest chain (<= kSummaryMaxLen: the LDS image
// is summary_lds_bytes(n_max)).
constexpr int kSummaryMaxLen = 8192;
constexpr size_t summary_lds_bytes(int n_max) { return 32 + 4 * 3 * 256 + 16 * (size_t)n_max; }
int chains_launch_summary(const uint8_t* bonds, const int32_t* lens, int n_pol, int max_len, int n_max, int linear,
                          int64_t* sum3, int32_t* mid3, unsigned long long* contacts, hipStream_t st);
// p[0..n) = v (clears scratch memory without a memset node)
int chains_launch_fill32(uint32_t* p, uint32_t v, size_t n, hipStream_t st);
// stat[0] = longest chain (atomicMax), stat[1] += chains with no length (lens <= 0)
int chains_launch_check_lens(const int32_t* lens, int n_pol, uint32_t* stat, hipStream_t st);

}  // namespace elp
