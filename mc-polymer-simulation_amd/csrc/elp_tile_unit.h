// elp_tile_unit.h -- one tile's time unit, shared by the ring and linear
// move kernels and by their solo and batched launches.
//
// A workgroup stages its tile of the periodic box from HBM into LDS (indexed
// by packed guarded coordinates, PTile), runs n_iter iterations of the move
// set's substeps with a workgroup barrier after each substep, and writes the
// tile back in place.  The move set M supplies the site type, the number of
// substeps per iteration and the move of substep s, which returns its
// outcome (kMoveEmpty ... kMoveAccepted):
//   M::site_t, M::NSUB, M::move<G>(lds, P, s, r, rules[, fill]),
//   M::HAS_RULES (whether move reads rules; the loop is then instantiated
//   per rule set), M::DRAW_AHEAD (whether the next group of iterations'
//   random words are drawn inside a move, by the callable fill that the
//   move runs after issuing its last dependent LDS read),
// and fills any LDS tables it keeps after the tile: M::init<G>(lds, tid).
//
// Randomness (Philox4x32-10, key = (seed, stream)):
//   site offsets : counter (tau, tile, 0xFFFFFFFF, iter/4), word iter%4,
//                  byte s = coset of substep s (tile-uniform);
//   moves        : counter (tau, tile, slot, iter/4), word iter%4, byte s;
//                  under balanced rates its bits 6-7 (the move drawn) are
//                  replaced by bits 6-7 of the site-offset byte, so every
//                  slot of a substep makes the same kind of move.
// The tile index is the replica-local logical tile, so a replica gives the
// same result alone or inside a batch, whatever the launch geometry.
#pragma once
#include "elp_device.h"
#include "elp_kernels.h"
#include <type_traits>

namespace elp {

// Geometry and random key of one tile's time unit.
struct UnitArgs {
    int LT, LU, LV;          // replica box
    int ntile_t, ntile_u;    // tiles per axis (t, u)
    int ot, ou, ov;          // tile-grid origin of this time unit
    uint32_t tau, seed, stream;
    int n_iter;
    uint32_t rules;          // move rates (ring: 0 reference, 1 balanced)
    unsigned long long* counts;   // COUNT: empty, face, accepted
};

__device__ __forceinline__ UnitArgs unit_args(const RingLaunch& p) {
    return UnitArgs{p.LT, p.LU, p.LV, p.ntile_t, p.ntile_u, p.ot, p.ou, p.ov, p.tau, p.seed, p.stream, p.n_iter,
                    p.rules, p.counts};
}

// Sum of v over the wave (butterfly over the 64 lanes).
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Copy one tile between the box (row-major, periodic) and LDS, in either
// direction.  Site indices are 32-bit (boxes hold fewer than 2^31 sites).
// When the workgroup covers whole (t,u) planes (THREADS a multiple of
// TT*TU), each thread keeps its (t,u) and steps v: the box index advances by
// one plane per step and wraps once, one add and one conditional subtract.
// When it covers whole t-rows of a plane (THREADS a multiple of TT dividing
// TT*TU), each thread keeps its t, holds the row offsets of its NU rows of a
// plane, and steps v the same way.
template <bool LOAD, int TT, int TU, int TV, int THREADS, int PAD, class T>
__device__ __forceinline__ void tile_copy(T* __restrict__ glat, T* tile, const UnitArgs& p, int gt0, int gu0, int gv0,
                                          int tid) {
    using G = PTile<TT, TU, TV, PAD>;
    const uint32_t plane = (uint32_t)p.LT * (uint32_t)p.LU, nsite = plane * (uint32_t)p.LV;
    if constexpr (THREADS % (TT * TU) == 0) {
        constexpr int VSTEP = THREADS / (TT * TU);
        const uint32_t lt = tid & (TT - 1), lu = (tid >> G::LT) & (TU - 1), lv0 = (uint32_t)tid >> (G::LT + G::LU);
        int gt = gt0 + (int)lt, gu = gu0 + (int)lu, gv = gv0 + (int)lv0;
        gt -= (gt >= p.LT) ? p.LT : 0;
        gu -= (gu >= p.LU) ? p.LU : 0;
        gv -= (gv >= p.LV) ? p.LV : 0;
        uint32_t idx = (uint32_t)gt + (uint32_t)p.LT * (uint32_t)gu + plane * (uint32_t)gv;
#pragma unroll
        for (int k = 0; k < TV / VSTEP; k++) {
            const uint32_t lv = lv0 + (uint32_t)(k * VSTEP);
            const uint32_t a = PAD ? G::pack(lt, lu, lv) : lt + TT * (lu + TU * lv);
            if (LOAD) tile[a] = glat[idx];
            else glat[idx] = tile[a];
            idx += plane * VSTEP;
            idx -= (idx >= nsite) ? nsite : 0u;
        }
    } else if constexpr (THREADS % TT == 0 && (TT * TU) % THREADS == 0) {
        constexpr int USTEP = THREADS / TT, NU = TU / USTEP;
        const uint32_t lt = tid & (TT - 1), lu0 = (uint32_t)tid >> G::LT;
        int gt = gt0 + (int)lt;
        gt -= (gt >= p.LT) ? p.LT : 0;
        uint32_t row[NU];
#pragma unroll
        for (int j = 0; j < NU; j++) {
            int gu = gu0 + (int)lu0 + j * USTEP;
            gu -= (gu >= p.LU) ? p.LU : 0;
            row[j] = (uint32_t)gt + (uint32_t)p.LT * (uint32_t)gu;
        }
        int gv = gv0;
        gv -= (gv >= p.LV) ? p.LV : 0;
        uint32_t base = plane * (uint32_t)gv;
#pragma unroll
        for (int lv = 0; lv < TV; lv++) {
#pragma unroll
            for (int j = 0; j < NU; j++) {
                const uint32_t lu = lu0 + (uint32_t)(j * USTEP);
                const uint32_t a = PAD ? G::pack(lt, lu, (uint32_t)lv) : lt + TT * (lu + TU * (uint32_t)lv);
                if (LOAD) tile[a] = glat[base + row[j]];
                else glat[base + row[j]] = tile[a];
            }
            base += plane;
            base -= (base >= nsite) ? nsite : 0u;
        }
    } else {
        for (int i = tid; i < TT * TU * TV; i += THREADS) {
            const uint32_t lt = i & (TT - 1), lu = (i >> G::LT) & (TU - 1), lv = i >> (G::LT + G::LU);
            int gt = gt0 + (int)lt, gu = gu0 + (int)lu, gv = gv0 + (int)lv;
            gt -= (gt >= p.LT) ? p.LT : 0;
            gu -= (gu >= p.LU) ? p.LU : 0;
            gv -= (gv >= p.LV) ? p.LV : 0;
            const uint32_t idx = (uint32_t)gt + (uint32_t)p.LT * ((uint32_t)gu + (uint32_t)p.LU * (uint32_t)gv);
            const uint32_t a = PAD ? G::pack(lt, lu, lv) : (uint32_t)i;
            if (LOAD) tile[a] = glat[idx];
            else glat[idx] = tile[a];
        }
    }
}

// HBM -> LDS with lattice-independent work between issuing the loads and
// writing LDS: in the plane-stepping layout (THREADS a multiple of TT*TU)
// the loads go to registers first, `between` runs while they are in flight,
// and the values are written to LDS after it; in the other layouts `between`
// runs after the copy.
template <int TT, int TU, int TV, int THREADS, int PAD, class T, class F>
__device__ __forceinline__ void tile_load(T* __restrict__ glat, T* tile, const UnitArgs& p, int gt0, int gu0, int gv0,
                                          int tid, const F& between) {
    using G = PTile<TT, TU, TV, PAD>;
    if constexpr (THREADS % (TT * TU) == 0) {
        constexpr int VSTEP = THREADS / (TT * TU), NV = TV / VSTEP;
        const uint32_t plane = (uint32_t)p.LT * (uint32_t)p.LU, nsite = plane * (uint32_t)p.LV;
        const uint32_t lt = tid & (TT - 1), lu = (tid >> G::LT) & (TU - 1), lv0 = (uint32_t)tid >> (G::LT + G::LU);
        int gt = gt0 + (int)lt, gu = gu0 + (int)lu, gv = gv0 + (int)lv0;
        gt -= (gt >= p.LT) ? p.LT : 0;
        gu -= (gu >= p.LU) ? p.LU : 0;
        gv -= (gv >= p.LV) ? p.LV : 0;
        uint32_t idx = (uint32_t)gt + (uint32_t)p.LT * (uint32_t)gu + plane * (uint32_t)gv;
        T v[NV];
#pragma unroll
        for (int k = 0; k < NV; k++) {
            v[k] = glat[idx];
            idx += plane * VSTEP;
            idx -= (idx >= nsite) ? nsite : 0u;
        }
        between();
#pragma unroll
        for (int k = 0; k < NV; k++) {
            const uint32_t lv = lv0 + (uint32_t)(k * VSTEP);
            tile[PAD ? G::pack(lt, lu, lv) : lt + TT * (lu + TU * lv)] = v[k];
        }
    } else {
        tile_copy<true, TT, TU, TV, THREADS, PAD>(glat, tile, p, gt0, gu0, gv0, tid);
        between();
    }
}

template <class M, int TT, int TU, int TV, bool DENSE, int THREADS, int PAD, bool COUNT = false>
__device__ __forceinline__ void tile_unit(typename M::site_t* __restrict__ glat, typename M::site_t* tile,
                                          const UnitArgs& p, uint32_t tileId) {
    using G = PTile<TT, TU, TV, PAD>;
    constexpr int NCT = TT / 4, NCU = TU / 4;
    constexpr int NCELL = (TT / 4) * (TU / 4) * (TV / 4);
    constexpr int SPC = DENSE ? 4 : 1;
    constexpr int NSLOT = NCELL * SPC;
    constexpr int SPT = NSLOT / THREADS;
    static_assert(NSLOT % THREADS == 0 && SPT >= 1, "slots must split evenly over threads");
    static_assert(THREADS >= 64, "M::init fills its tables with the first 64 threads");

    const int tid = threadIdx.x;
    const uint32_t q1 = udiv_small(tileId, (uint32_t)p.ntile_t);
    const uint32_t q2 = udiv_small(q1, (uint32_t)p.ntile_u);
    const int kt = (int)(tileId - q1 * (uint32_t)p.ntile_t);
    const int ku = (int)(q1 - q2 * (uint32_t)p.ntile_u);
    const int kv = (int)q2;
    const int gt0 = p.ot + kt * TT, gu0 = p.ou + ku * TU, gv0 = p.ov + kv * TV;

    // HBM -> LDS (periodic wrap; origins are < one tile edge, so one
    // subtraction suffices); while the loads are in flight, the random words
    // of the first group of four iterations (the empty statements keep the
    // compiler from sinking them past the barrier to their first use) and the
    // move set's LDS tables
    uint4 ow, sw[SPT];
    tile_load<TT, TU, TV, THREADS, PAD>(glat, tile, p, gt0, gu0, gv0, tid, [&] {
        ow = philox4x32_10(make_uint4(p.tau, tileId, kOffsetStream, 0u), p.seed, p.stream);
        asm volatile("" ::"s"(ow.x), "s"(ow.y), "s"(ow.z), "s"(ow.w));
#pragma unroll
        for (int k = 0; k < SPT; k++) {
            sw[k] = philox4x32_10_lanes(make_uint4(p.tau, tileId, (uint32_t)(tid + k * THREADS), 0u), p.seed,
                                        p.stream);
            asm volatile("" ::"v"(sw[k].x), "v"(sw[k].y), "v"(sw[k].z), "v"(sw[k].w));
        }
        M::template init<G>(tile, tid);
    });

    // slot k of this thread: cell corner (+ j along t for the dense cosets).
    // Dense coset j of substep offset c sits at corner + (j, (c&3) + 3j,
    // (c>>2) + 2j) mod 4: the u term is (cu + ju) masked to two bits, and
    // adding 2j mod 4 flips bit 1 for odd j, so the v term is an XOR of the
    // tile-uniform c>>2 into a corner that already holds that flip (vb).
    uint32_t base[SPT], ju[SPT], vb[SPT], slot[SPT];
#pragma unroll
    for (int k = 0; k < SPT; k++) {
        slot[k] = (uint32_t)(tid + k * THREADS);
        const uint32_t cell = slot[k] / SPC, j = slot[k] % SPC;
        base[k] = G::pack(4u * (cell % NCT) + (DENSE ? j : 0u), 4u * ((cell / NCT) % NCU), 4u * (cell / (NCT * NCU)));
        ju[k] = (3u * j) << G::SU;
        vb[k] = base[k] | ((2u * (j & 1u)) << G::SV);
    }
    __syncthreads();
    uint32_t nEmpty = 0, nFace = 0, nAcc = 0;   // COUNT: outcomes of this thread's attempts

    // The move rates are a compile-time constant of the loop, so no substep
    // branches on them.  With M::DRAW_AHEAD, each group of four iterations
    // draws the next group's words in the last substep of its first
    // iteration, after that substep's last dependent LDS read (the draws need
    // no lattice data, so they fill the wait; the last group draws one unused
    // set); otherwise each group draws its own words at its start.
    auto iterate = [&](auto rulesC) {
        constexpr uint32_t RULES = decltype(rulesC)::value;
        for (int it0 = 0; it0 < p.n_iter; it0 += 4) {
            if constexpr (!M::DRAW_AHEAD) {
                if (it0 > 0) {
                    ow = philox4x32_10(make_uint4(p.tau, tileId, kOffsetStream, (uint32_t)(it0 >> 2)), p.seed, p.stream);
#pragma unroll
                    for (int k = 0; k < SPT; k++)
                        sw[k] = philox4x32_10_lanes(make_uint4(p.tau, tileId, slot[k], (uint32_t)(it0 >> 2)), p.seed,
                                                    p.stream);
                }
            }
            uint4 owNext = make_uint4(0u, 0u, 0u, 0u), swNext[SPT];
#pragma unroll
            for (int k = 0; k < SPT; k++) swNext[k] = owNext;
            const int nj = (p.n_iter - it0) < 4 ? (p.n_iter - it0) : 4;
            // unrolled, so that word j of each Philox output is a fixed register
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (j >= nj) break;
                const uint32_t offw = word_of(ow, j);
#pragma unroll
                for (int s = 0; s < M::NSUB; s++) {
                    // the next group's words; the empty statements keep them
                    // in the move's read block (the compiler would otherwise
                    // sink them to the barrier, past the move's writes)
                    auto draw = [&] {
                        const uint32_t grp = (uint32_t)(it0 >> 2) + 1u;
                        owNext = philox4x32_10(make_uint4(p.tau, tileId, kOffsetStream, grp), p.seed, p.stream);
#pragma unroll
                        for (int k = 0; k < SPT; k++) {
                            swNext[k] = philox4x32_10_lanes(make_uint4(p.tau, tileId, slot[k], grp), p.seed, p.stream);
                            asm volatile("" ::"v"(swNext[k].x), "v"(swNext[k].y), "v"(swNext[k].z), "v"(swNext[k].w));
                        }
                        asm volatile("" ::"s"(owNext.x), "s"(owNext.y), "s"(owNext.z), "s"(owNext.w));
                    };
                    const uint32_t c = (offw >> (8 * s)) & (DENSE ? 0xfu : 0x3fu);
                    const uint32_t cu = (c & 3u) << G::SU, cv = (c >> 2) << G::SV;
#pragma unroll
                    for (int k = 0; k < SPT; k++) {
                        uint32_t P;
                        if (DENSE) P = ((cu + ju[k]) & (3u << G::SU)) | (vb[k] ^ cv);
                        else P = base[k] + G::pack(c & 3u, (c >> 2) & 3u, (c >> 4) & 3u);
                        // balanced rates: bits 6-7 (the move drawn) from the
                        // tile-uniform offset byte, the same for every slot
                        const uint32_t rb = (word_of(sw[k], j) >> (8 * s)) & 0xffu;
                        const uint32_t rk = RULES ? ((rb & 0x3fu) | ((offw >> (8 * s)) & 0xc0u)) : rb;
                        // the last slot of the first iteration's last substep draws
                        const uint32_t out = (M::DRAW_AHEAD && j == 0 && s == M::NSUB - 1 && k == SPT - 1)
                                                 ? M::template move<G>(tile, P, s, rk, RULES, draw)
                                                 : M::template move<G>(tile, P, s, rk, RULES);
                        if constexpr (COUNT) {
                            nEmpty += out == kMoveEmpty;
                            nFace += out == kMoveFace;
                            nAcc += out == kMoveAccepted;
                        }
                    }
                    __syncthreads();
                }
            }
            if constexpr (M::DRAW_AHEAD) {
                ow = owNext;
#pragma unroll
                for (int k = 0; k < SPT; k++) sw[k] = swNext[k];
            }
        }
    };
    if constexpr (M::HAS_RULES) {
        if (p.rules) iterate(std::integral_constant<uint32_t, 1u>{});
        else iterate(std::integral_constant<uint32_t, 0u>{});
    } else {
        iterate(std::integral_constant<uint32_t, 0u>{});
    }

    // LDS -> HBM.  The thread index passes through an empty statement, so the
    // compiler recomputes the box indices here instead of keeping the load's
    // (16 64-bit addresses) live in registers through the whole loop.
    int tidStore = tid;
    asm volatile("" : "+v"(tidStore));
    tile_copy<false, TT, TU, TV, THREADS, PAD>(glat, tile, p, gt0, gu0, gv0, tidStore);
    if constexpr (COUNT) {
        // one sum per wave, added by its first lane (vector atomics) to the
        // workgroup's counter slot
        nEmpty = wave_sum(nEmpty);
        nFace = wave_sum(nFace);
        nAcc = wave_sum(nAcc);
        if ((tid & 63) == 0) {
            unsigned long long* c = p.counts + (size_t)(blockIdx.x % kCountSlots) * kCountStride;
            atomicAdd(&c[0], (unsigned long long)nEmpty);
            atomicAdd(&c[1], (unsigned long long)nFace);
            atomicAdd(&c[2], (unsigned long long)nAcc);
        }
    }
}

// tile_unit with the occupied slots of each substep compacted to the front of
// the workgroup: each thread reads its slot's site, a ballot and a prefix sum
// over the waves place the occupied slots (packed coordinate and random byte)
// in list[0..count), and thread i < count makes the move of entry i.  An empty
// site makes no move, so the results equal tile_unit's (random bits and
// outcomes are keyed by the slot, not by the thread).  One slot per thread;
// list holds THREADS + THREADS/64 words.  No counting instantiation.
template <class M, int TT, int TU, int TV, bool DENSE, int THREADS, int PAD>
__device__ __forceinline__ void tile_unit_compact(typename M::site_t* __restrict__ glat, typename M::site_t* tile,
                                                  uint32_t* list, const UnitArgs& p, uint32_t tileId) {
    using G = PTile<TT, TU, TV, PAD>;
    constexpr int NCT = TT / 4, NCU = TU / 4;
    constexpr int NCELL = (TT / 4) * (TU / 4) * (TV / 4);
    constexpr int SPC = DENSE ? 4 : 1;
    constexpr int NW = THREADS / 64;
    static_assert(NCELL * SPC == THREADS, "one slot per thread");
    static_assert(THREADS % 64 == 0, "whole waves");
    uint32_t* cnt = list + THREADS;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int kt = (int)(tileId % (uint32_t)p.ntile_t);
    const int ku = (int)((tileId / (uint32_t)p.ntile_t) % (uint32_t)p.ntile_u);
    const int kv = (int)(tileId / (uint32_t)(p.ntile_t * p.ntile_u));
    const int gt0 = p.ot + kt * TT, gu0 = p.ou + ku * TU, gv0 = p.ov + kv * TV;
    tile_copy<true, TT, TU, TV, THREADS, PAD>(glat, tile, p, gt0, gu0, gv0, tid);
    const uint32_t slot = (uint32_t)tid;
    const uint32_t cell = slot / SPC, j = slot % SPC;
    const uint32_t base = G::pack(4u * (cell % NCT) + (DENSE ? j : 0u), 4u * ((cell / NCT) % NCU), 4u * (cell / (NCT * NCU)));
    const uint32_t ju = (3u * j) << G::SU, vb = base | ((2u * (j & 1u)) << G::SV);
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    M::template init<G>(tile, tid);
    __syncthreads();
    for (int it0 = 0; it0 < p.n_iter; it0 += 4) {
        const uint4 ow = philox4x32_10(make_uint4(p.tau, tileId, kOffsetStream, (uint32_t)(it0 >> 2)), p.seed, p.stream);
        const uint4 sw = philox4x32_10_lanes(make_uint4(p.tau, tileId, slot, (uint32_t)(it0 >> 2)), p.seed, p.stream);
        const int nj = (p.n_iter - it0) < 4 ? (p.n_iter - it0) : 4;
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            if (jj >= nj) break;
            const uint32_t offw = word_of(ow, jj);
#pragma unroll
            for (int s = 0; s < M::NSUB; s++) {
                const uint32_t c = (offw >> (8 * s)) & (DENSE ? 0xfu : 0x3fu);
                const uint32_t cu = (c & 3u) << G::SU, cv = (c >> 2) << G::SV;
                const uint32_t P = DENSE ? (((cu + ju) & (3u << G::SU)) | (vb ^ cv))
                                         : base + G::pack(c & 3u, (c >> 2) & 3u, (c >> 4) & 3u);
                const uint32_t rb = (word_of(sw, jj) >> (8 * s)) & 0xffu;
                // balanced rates: bits 6-7 from the tile-uniform offset byte (tile_unit)
                const uint32_t r = p.rules ? ((rb & 0x3fu) | ((offw >> (8 * s)) & 0xc0u)) : rb;
                const bool occ = tile[G::addr(P)] != 0;
                const uint64_t bal = __ballot(occ);
                if (lane == 0) cnt[w] = (uint32_t)__popcll(bal);
                __syncthreads();
                uint32_t before = (uint32_t)__popcll(bal & below), total = 0;
#pragma unroll
                for (int k = 0; k < NW; k++) {
                    const uint32_t ck = cnt[k];
                    before += (k < w) ? ck : 0u;
                    total += ck;
                }
                if (occ) list[before] = P | (r << 24);
                __syncthreads();
                if ((uint32_t)tid < total) {
                    const uint32_t e = list[tid];
                    M::template move<G>(tile, e & 0xFFFFFFu, s, e >> 24, p.rules);
                }
                __syncthreads();
            }
        }
    }
    int tidStore = tid;
    asm volatile("" : "+v"(tidStore));
    tile_copy<false, TT, TU, TV, THREADS, PAD>(glat, tile, p, gt0, gu0, gv0, tidStore);
}

// Workgroup of a batched launch: global tile g (XCD-aware order over all
// replicas' tiles) -> replica r = rep_of_tile[g], its local tile g - tile0,
// and the origin / time unit of replica r for time unit (*cursor + k).
template <class M, int TT, int TU, int TV, bool DENSE, int THREADS, int PAD, bool COUNT = false>
__device__ __forceinline__ void batch_unit(typename M::site_t* __restrict__ gbase, typename M::site_t* tile,
                                           const BatchLaunch& b) {
    const uint32_t g = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t r = b.rep_of_tile[g];
    const BatchRep rep = b.reps[r];
    const BatchStep st = b.steps[(size_t)(*b.cursor + b.k) * (size_t)b.n_rep + r];
    const UnitArgs p{rep.LT, rep.LU, rep.LV, rep.ntile_t, rep.ntile_u, st.ot, st.ou, st.ov,
                     st.tau, rep.seed, rep.stream, b.n_iter, b.rules, b.counts};
    tile_unit<M, TT, TU, TV, DENSE, THREADS, PAD, COUNT>(gbase + rep.site_off, tile, p, g - rep.tile0);
}

}  // namespace elp
