// elp_device.h -- device helpers shared by the gfx950 move kernels:
// Philox4x32-10 counter-based generator and FCC bond-code algebra.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace elp {

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC11).  Round: two 32x32->64 products of counter words
// 0 and 2 with the multipliers; key bumped by the Weyl constants between
// rounds.  Ten rounds.
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        // one 32x32->64 product per multiplier (v_mad_u64_u32) instead of
        // separate high and low multiplies: half the quarter-rate multiplies
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = make_uint4((uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0);
    }
    return c;
}

// The same generator for a counter that differs between lanes: each round's
// two three-way XORs as one v_bitop3_b32 each (truth table 0x96) instead of
// two v_xor_b32.  A tile-uniform counter keeps the plain form, which the
// compiler runs on the scalar unit.
__device__ __forceinline__ uint4 philox4x32_10_lanes(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = make_uint4(__builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c.y, k0, 0x96), (uint32_t)p1,
                       __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c.w, k1, 0x96), (uint32_t)p0);
    }
    return c;
}

// Work without lattice dependence that a move runs right after issuing its
// last dependent LDS read, so that it fills the wait for that read
// (NoFill: none).
struct NoFill {
    __device__ __forceinline__ void operator()() const {}
};

__device__ __forceinline__ uint32_t word_of(const uint4& w, int j) {
    return j == 0 ? w.x : j == 1 ? w.y : j == 2 ? w.z : w.w;
}

// floor(n / d) for n < 2^22 and 1 <= d < 2^22 (tile indices and tile
// counts: a box holds fewer than 2^31 sites and a tile at least 2^11).  The
// float quotient (n and d exact in float, v_rcp_f32 within one ulp) is within
// 0.75 of n / d, so its truncation is off by at most one, and one
// correction step by the remainder makes it exact.  About ten instructions
// instead of the generic 32-bit division's thirty.
__device__ __forceinline__ uint32_t udiv_small(uint32_t n, uint32_t d) {
    const uint32_t q = (uint32_t)((float)n * __builtin_amdgcn_rcpf((float)d));
    const int32_t r = (int32_t)(n - q * d);
    return r < 0 ? q - 1u : (r >= (int32_t)d ? q + 1u : q);
}

// Counter word 2 of the tile-uniform site-offset stream; slot streams use
// the slot index (always < 2^31) there instead.
constexpr uint32_t kOffsetStream = 0xFFFFFFFFu;

// Logical tile of a workgroup.  Workgroups are dispatched round-robin over
// the 8 XCDs (block b -> XCD b % 8), each with its own L2.  Giving XCD x the
// contiguous tile range [x*n/8, (x+1)*n/8) (t fastest) keeps the tiles that
// share 128-byte lines of the row-major lattice on one L2, so each line is
// fetched and written back once per time unit.  Tiles are independent and
// the random streams are keyed by the logical tile, so the mapping changes
// no result.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n) {
    return (n & 7u) ? b : (b & 7u) * (n >> 3) + (b >> 3);
}

// 4-bit bond code (t,u,v,w) -> displacement (t-w, u-w, v-w)
__device__ __forceinline__ int bdt(uint32_t b) { return (int)(b & 1u) - (int)(b >> 3); }
__device__ __forceinline__ int bdu(uint32_t b) { return (int)((b >> 1) & 1u) - (int)(b >> 3); }
__device__ __forceinline__ int bdv(uint32_t b) { return (int)((b >> 2) & 1u) - (int)(b >> 3); }
__device__ __forceinline__ uint32_t bond_add(uint32_t a, uint32_t b) { return ((a | b) == 0xfu) ? (a & b) : (a | b); }
// a+b is a single bond (codes 0x3 and 0xc are not bonds)
__device__ __forceinline__ bool bond_add_valid(uint32_t a, uint32_t b, uint32_t& c) {
    if ((a | b) != 0xfu && (a & b)) return false;
    c = bond_add(a, b);
    return c != 0x3u && c != 0xcu;
}

// First bond of decomposition k of bond `next` (the SetGPUTrans table,
// gpupol2/gpupol_init.c:552-584, as a 128-bit constant: nibble 2*next+k).
constexpr uint64_t kTransLo = 0x2172716500767500ull;   // table words 0,1
constexpr uint64_t kTransHi = 0x0064540021b2b1a9ull;   // table words 2,3
__device__ __forceinline__ uint32_t trans_first(uint32_t next, uint32_t k) {
    const uint32_t e = 2u * next + k;
    const uint64_t w = (e & 16u) ? kTransHi : kTransLo;
    return (uint32_t)(w >> (4u * (e & 15u))) & 0xfu;
}

// The ordered bond pair (nb1, nb2) that replaces bond `next` for random bits
// r & 3: nb1 = trans_first(next, r & 1), nb2 = the bond with nb1 + nb2 = next,
// swapped when bit 1 is clear; returned as nb1 | nb2 << 4.  The move kernels
// read it from a 64-entry LDS table after the tile (entry 4 next + (r & 3)).
__device__ __forceinline__ uint32_t fold_pair(uint32_t next, uint32_t r) {
    uint32_t nb1 = trans_first(next, r & 1u);
    uint32_t nb2 = bond_add((~nb1) & 0xfu, next);
    if (!(r & 2u)) { const uint32_t x = nb1; nb1 = nb2; nb2 = x; }
    return nb1 | (nb2 << 4);
}
constexpr int kFoldPairs = 64;

constexpr int ilog2c(int x) { return x <= 1 ? 0 : 1 + ilog2c(x / 2); }

// Packed tile coordinates P = t | u<<SU | v<<SV with one guard bit per field
// (tile edges are powers of two): a neighbour is P + delta(bond), "outside
// the tile" is (P & GUARD) != 0, and the LDS index of a site is either P
// itself (PAD=1, sparse: 4x the tile in bytes) or the compact t-fastest
// index rebuilt from P (PAD=0).  addr() masks P into the tile so that a
// speculative read of a rejected neighbour stays inside the array.
template <int TT, int TU, int TV, int PAD>
struct PTile {
    static constexpr int LT = ilog2c(TT), LU = ilog2c(TU), LV = ilog2c(TV);
    static_assert((1 << LT) == TT && (1 << LU) == TU && (1 << LV) == TV, "tile edges must be powers of two");
    static constexpr uint32_t SU = LT + 1, SV = LT + LU + 2;
    static constexpr uint32_t ONES = 1u | (1u << SU) | (1u << SV);
    // The guard stops at bit 23, so bits 24-31 of P are never read: a move
    // may add an LDS table word that carries a table byte there (an offset
    // in bits 0-23) to P directly.  An underflow still sets bits SV+LV..23.
    static_assert(SV + LV < 24, "packed coordinates must fit in 24 bits");
    static constexpr uint32_t GUARD = ((uint32_t)TT | ((uint32_t)TU << SU) | (~0u << (SV + LV))) & 0xFFFFFFu;
    static constexpr uint32_t MASK = (uint32_t)(TT - 1) | ((uint32_t)(TU - 1) << SU) | ((uint32_t)(TV - 1) << SV);
    static constexpr int NSITE = PAD ? (1 << (SV + LV)) : TT * TU * TV;
    __device__ static __forceinline__ uint32_t pack(uint32_t t, uint32_t u, uint32_t v) { return t | (u << SU) | (v << SV); }
    // Bond code b -> offset of P: bits 0-2 of b spread to the field origins
    // 0, SU, SV by one 24-bit multiply (the three shifted copies of the 4-bit
    // code sit at bits [0,3], [SU-1,SU+2], [SV-2,SV+1], disjoint since
    // LT, LU >= 4, so no carries; bit 3 lands outside ONES), then (1,1,1)
    // subtracted when bit 3 is set.  Callers hold b = site & 15 already.
    static constexpr uint32_t SPREAD = 1u | (1u << (SU - 1)) | (1u << (SV - 2));
    static_assert(LT >= 4 && LU >= 4, "spread multiply needs tile edges >= 16 in t and u");
    __device__ static __forceinline__ uint32_t delta(uint32_t b) {
        const uint32_t d = __umul24(b & 15u, SPREAD) & ONES;
        return d - (ONES & (uint32_t)((int32_t)(b << 28) >> 31));
    }
    __device__ static __forceinline__ bool out(uint32_t P) { return (P & GUARD) != 0u; }
    __device__ static __forceinline__ uint32_t addr(uint32_t P) {
        P &= MASK;
        if (PAD) return P;
        return (P & (uint32_t)(TT - 1)) | ((P >> 1) & ((uint32_t)(TU - 1) << LT)) |
               ((P >> 2) & ((uint32_t)(TV - 1) << (LT + LU)));
    }
};

}  // namespace elp
