// elp_ring.hip -- gfx950 move kernel of the ring-polymer ELP engine.
//
// One workgroup owns one tile of the box for one time unit: it stages the
// tile (1 byte per site) from HBM into LDS, runs n_iter iterations of the
// substeps {TransForw, DiffuseSL, TransForw, TransBack} (the ringpol3 move
// set, gpupol2/ringpol3.cu:92-218 and :278-292) and writes the tile back in
// place.  Tiles of one time unit partition the box, so the in-place update
// is race-free; the grid origin moves between time units (host shift).
//
// Within a substep every slot of the tile attempts the same move type at a
// site of one coset of an independent sub-lattice, chosen per tile and
// substep from the tile-uniform offset stream:
//   cubic : one site per 4^3 cell, coset offset 6 bits (t,u,v mod 4)
//           -- the reference's workgroup-uniform cell offset;
//   dense : four sites per 4^3 cell, the coset of 4Z^3 + <(1,3,2)> with
//           representative (0, c&3, c>>2); slot j of a cell is the point
//           rep + j*(1,3,2) mod 4.
// Any two sites of one coset differ by a lattice vector of squared length
// >= 10 (FCC units, bond = 2), so the sites a move reads or writes (the site
// and its 12 neighbours) never overlap between slots: all slots of a
// substep commute, and a sequential replay in any order is exact.
// Moves that would touch a site outside the tile are rejected, as in the
// reference (OOB, gpupol2/ringpol3.cu:71-73).
//
// Randomness (Philox4x32-10, key = (seed, stream)):
//   offsets : counter (tau, tile, 0xFFFFFFFF, iter/4), word iter%4,
//             byte s = substep s;
//   moves   : counter (tau, tile, slot, iter/4), word iter%4, byte s.
// LDS layout: cell-major, addr = cell + NCELL * (t&3 | (u&3)<<2 | (v&3)<<4),
// so the site reads of one substep hit consecutive bytes.
#include "elp_device.h"
#include "elp_kernels.h"
#include "elp_tile_unit.h"

namespace elp {

template <int TT, int TU, int TV>
struct RingTile {
    static constexpr int NCT = TT / 4, NCU = TU / 4, NCV = TV / 4;
    static constexpr int NCELL = NCT * NCU * NCV;
    static constexpr int NBYTE = TT * TU * TV;
    __device__ static __forceinline__ int addr(int t, int u, int v) {
        const int cell = (t >> 2) + NCT * ((u >> 2) + NCU * (v >> 2));
        return cell + NCELL * ((t & 3) | ((u & 3) << 2) | ((v & 3) << 4));
    }
    __device__ static __forceinline__ bool inside(int t, int u, int v) {
        return (unsigned)t < (unsigned)TT && (unsigned)u < (unsigned)TU && (unsigned)v < (unsigned)TV;
    }
};

// Insert a monomer on an empty common neighbour of the site and its bond
// partner, taking it from whichever end holds stored length
// (gpupol2/ringpol3.cu:92-145).  r: bit0 decomposition, bit1 order,
// bit2 which monomer moves when both ends hold stored length.
template <class G>
__device__ __forceinline__ void trans_forw(uint8_t* lat, int t, int u, int v, uint32_t r, const uint32_t* trans) {
    const int a = G::addr(t, u, v);
    uint32_t s = lat[a];
    if (!s) return;
    const uint32_t next = s & 0xfu, label = s & 0x30u, sl = s & 0x40u;
    const int nt = t + bdt(next), nu = u + bdu(next), nv = v + bdv(next);
    if (!G::inside(nt, nu, nv)) return;
    const int na = G::addr(nt, nu, nv);
    uint32_t ns = lat[na];
    const uint32_t nsl = ns & 0x40u;
    if (!(sl | nsl)) return;
    uint32_t nb1 = (trans[next >> 2] >> (4 * (2 * (next & 3u) + (r & 1u)))) & 0xfu;
    uint32_t nb2 = bond_add((~nb1) & 0xfu, next);
    if (!(r & 2u)) { const uint32_t x = nb1; nb1 = nb2; nb2 = x; }
    const int dt = t + bdt(nb1), du = u + bdu(nb1), dv = v + bdv(nb1);
    if (!G::inside(dt, du, dv)) return;
    const int da = G::addr(dt, du, dv);
    if (lat[da]) return;
    const bool moveFirst = (sl && nsl) ? ((r >> 2) & 1u) : (sl != 0);
    uint32_t d = nb2;
    if (moveFirst) {
        s = nb1 | ((label >> 1) & 0x10u);
        d |= label & 0x10u;
    } else {
        s = nb1 | label | sl;
        d |= (ns & 0x20u) >> 1;
        ns &= 0x1fu;
    }
    lat[a] = (uint8_t)s;
    lat[da] = (uint8_t)d;
    if (!moveFirst) lat[na] = (uint8_t)ns;
}

// Retract a corner: site -> src -> third site forming a triangle; the src
// monomer folds onto one of its neighbours as stored length
// (gpupol2/ringpol3.cu:147-192).  r: bit0 choice when neither end has SL.
template <class G>
__device__ __forceinline__ void trans_back(uint8_t* lat, int t, int u, int v, uint32_t r) {
    const int a = G::addr(t, u, v);
    uint32_t s = lat[a];
    if (!s) return;
    const uint32_t next = s & 0xfu, label = s & 0x30u, sl = s & 0x40u;
    const int st = t + bdt(next), su = u + bdu(next), sv = v + bdv(next);
    if (!G::inside(st, su, sv)) return;
    const int sa = G::addr(st, su, sv);
    const uint32_t ss = lat[sa];
    if (ss & 0x40u) return;
    uint32_t newNext;
    if (!bond_add_valid(next, ss & 0xfu, newNext)) return;
    const int wt = t + bdt(newNext), wu = u + bdu(newNext), wv = v + bdv(newNext);
    if (!G::inside(wt, wu, wv)) return;
    const int wa = G::addr(wt, wu, wv);
    uint32_t ws = lat[wa];
    const uint32_t wsl = ws & 0x40u;
    if (sl && wsl) return;
    const bool moveFirst = (!sl && !wsl) ? (r & 1u) : (sl == 0);
    const uint32_t srcLabel = ss & 0x30u;
    if (moveFirst) {
        s = newNext | (label << 1) | srcLabel | 0x40u;
    } else {
        s = newNext | label | sl;
        ws = (ws & 0x3fu) | (srcLabel << 1) | 0x40u;
    }
    lat[sa] = 0;
    lat[a] = (uint8_t)s;
    lat[wa] = (uint8_t)ws;
}

// Move one stored-length unit across the bond when exactly one end holds it
// (gpupol2/ringpol3.cu:194-218).
template <class G>
__device__ __forceinline__ void diffuse_sl(uint8_t* lat, int t, int u, int v) {
    const int a = G::addr(t, u, v);
    uint32_t s = lat[a];
    if (!s) return;
    const uint32_t next = s & 0xfu, label = s & 0x30u, sl = s & 0x40u;
    const int nt = t + bdt(next), nu = u + bdu(next), nv = v + bdv(next);
    if (!G::inside(nt, nu, nv)) return;
    const int na = G::addr(nt, nu, nv);
    uint32_t ns = lat[na];
    if (((ns & 0x40u) + sl) != 0x40u) return;
    if (sl) {
        ns = ns | ((label & 0x10u) << 1) | 0x40u;
        s = next | ((label >> 1) & 0x10u);
    } else {
        s = next | (label << 1) | (((ns & 0x30u) >> 1) & 0x10u) | 0x40u;
        ns &= 0x1fu;
    }
    lat[a] = (uint8_t)s;
    lat[na] = (uint8_t)ns;
}

template <int TT, int TU, int TV, bool DENSE, int THREADS>
__global__ __launch_bounds__(THREADS) void ring_time_unit(uint8_t* __restrict__ glat, RingLaunch p) {
    using G = RingTile<TT, TU, TV>;
    constexpr int SPC = DENSE ? 4 : 1;
    constexpr int NSLOT = G::NCELL * SPC;
    constexpr int SPT = NSLOT / THREADS;
    static_assert(NSLOT % THREADS == 0 && SPT >= 1, "slots must split evenly over threads");
    __shared__ uint8_t tile[G::NBYTE];

    const int tid = threadIdx.x;
    const uint32_t tileId = xcd_tile(blockIdx.x, gridDim.x);
    const int kt = (int)(tileId % (uint32_t)p.ntile_t);
    const int ku = (int)((tileId / (uint32_t)p.ntile_t) % (uint32_t)p.ntile_u);
    const int kv = (int)(tileId / (uint32_t)(p.ntile_t * p.ntile_u));
    const int gt0 = p.ot + kt * TT, gu0 = p.ou + ku * TU, gv0 = p.ov + kv * TV;

    // HBM -> LDS (periodic wrap of the shifted tile)
    for (int i = tid; i < G::NBYTE; i += THREADS) {
        const int lt = i % TT, lu = (i / TT) % TU, lv = i / (TT * TU);
        int gt = gt0 + lt, gu = gu0 + lu, gv = gv0 + lv;
        gt -= (gt >= p.LT) ? p.LT : 0;
        gu -= (gu >= p.LU) ? p.LU : 0;
        gv -= (gv >= p.LV) ? p.LV : 0;
        tile[G::addr(lt, lu, lv)] = glat[(size_t)gt + (size_t)p.LT * ((size_t)gu + (size_t)p.LU * (size_t)gv)];
    }

    // per-slot cell origin
    int ct[SPT], cu[SPT], cv[SPT], jj[SPT];
    uint32_t slot[SPT];
#pragma unroll
    for (int k = 0; k < SPT; k++) {
        slot[k] = (uint32_t)(tid + k * THREADS);
        const int cell = (int)(slot[k] / SPC);
        jj[k] = (int)(slot[k] % SPC);
        ct[k] = 4 * (cell % G::NCT);
        cu[k] = 4 * ((cell / G::NCT) % G::NCU);
        cv[k] = 4 * (cell / (G::NCT * G::NCU));
    }
    __syncthreads();

    for (int it0 = 0; it0 < p.n_iter; it0 += 4) {
        const uint4 ow = philox4x32_10(make_uint4(p.tau, tileId, kOffsetStream, (uint32_t)(it0 >> 2)), p.seed, p.stream);
        uint4 sw[SPT];
#pragma unroll
        for (int k = 0; k < SPT; k++)
            sw[k] = philox4x32_10(make_uint4(p.tau, tileId, slot[k], (uint32_t)(it0 >> 2)), p.seed, p.stream);
        const int nj = (p.n_iter - it0) < 4 ? (p.n_iter - it0) : 4;
        for (int j = 0; j < nj; j++) {
            const uint32_t offw = word_of(ow, j);
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const uint32_t c = (offw >> (8 * s)) & (DENSE ? 0xfu : 0x3fu);
#pragma unroll
                for (int k = 0; k < SPT; k++) {
                    int t, u, v;
                    if (DENSE) {
                        t = ct[k] + jj[k];
                        u = cu[k] + (int)((c + 3u * (uint32_t)jj[k]) & 3u);
                        v = cv[k] + (int)(((c >> 2) + 2u * (uint32_t)jj[k]) & 3u);
                    } else {
                        t = ct[k] + (int)(c & 3u);
                        u = cu[k] + (int)((c >> 2) & 3u);
                        v = cv[k] + (int)((c >> 4) & 3u);
                    }
                    const uint32_t r = (word_of(sw[k], j) >> (8 * s)) & 0xffu;
                    if (s == 1) diffuse_sl<G>(tile, t, u, v);
                    else if (s == 3) trans_back<G>(tile, t, u, v, r);
                    else trans_forw<G>(tile, t, u, v, r, p.trans);
                }
                __syncthreads();
            }
        }
    }

    // LDS -> HBM
    for (int i = tid; i < G::NBYTE; i += THREADS) {
        const int lt = i % TT, lu = (i / TT) % TU, lv = i / (TT * TU);
        int gt = gt0 + lt, gu = gu0 + lu, gv = gv0 + lv;
        gt -= (gt >= p.LT) ? p.LT : 0;
        gu -= (gu >= p.LU) ? p.LU : 0;
        gv -= (gv >= p.LV) ? p.LV : 0;
        glat[(size_t)gt + (size_t)p.LT * ((size_t)gu + (size_t)p.LU * (size_t)gv)] = tile[G::addr(lt, lu, lv)];
    }
}

// =====================================================================
// Second formulation of the same kernel (selected by default): sites are
// addressed by packed tile coordinates P = t | u<<SU | v<<SV whose fields
// carry one guard bit each, so
//   * a neighbour is P + delta(bond) (bit arithmetic, no per-axis code),
//   * "outside the tile" is one AND with the guard bits,
//   * the LDS tile is plain t-fastest (PAD=0: index rebuilt from P in three
//     bit-field operations) or indexed by P itself (PAD=1, 16^3 tiles only),
//   * every LDS read a move needs is issued before the accept/reject test;
//     DiffuseSL and TransForw (site, bond-pair table word, target) wait for
//     two and three dependent LDS round trips, TransBack for three (site,
//     source, then the combine table beside the target): the small LDS
//     tables after the tile replace about 30 VALU instructions per
//     iteration.  Accept/reject rules and writes are those of the functions
//     above.
// =====================================================================
constexpr int kCombine = 256;
// LDS after the tile (ring, with tables): 64 entries of two words for
// (next, r & 3): word 0 the fold pair in bits 24-31 above the packed offset
// of its first bond in bits 0-23 (P + word is the target, bits 24-31 being
// outside every guard and address mask), word 1 the packed offset of next
// itself; then 256 bytes, the combine code of (next, srcNext).
constexpr int kRingTabTotal = 8 * kFoldPairs + kCombine;

template <class G>
__device__ __forceinline__ uint2* ring_tab(uint8_t* lat) {
    return reinterpret_cast<uint2*>(lat + G::NSITE);
}
template <class G>
__device__ __forceinline__ uint8_t* ring_combine_tab(uint8_t* lat) {
    return lat + G::NSITE + 8 * kFoldPairs;
}

// a where the mask m is set, b elsewhere (one v_bfi_b32): lets a move pick
// between its two outcomes without a divergent branch.
__device__ __forceinline__ uint32_t bitsel(uint32_t m, uint32_t a, uint32_t b) { return (a & m) | (b & ~m); }
// all ones when bit 6 (stored length) of x is set
__device__ __forceinline__ uint32_t bit6_mask(uint32_t x) { return (uint32_t)((int32_t)(x << 25) >> 31); }

// moveFirst of TransForw, (sl & nsl) ? r bit 2 : sl != 0, as one bit test
// (checked for all s, ns < 128 and r < 256).
__device__ __forceinline__ bool forw_move_first(uint32_t s, uint32_t ns, uint32_t r) {
    return ((s & (~ns | (r << 4))) & 0x40u) != 0u;
}

// PAIRS: read fold_pair, the packed offset of its first bond and that of
// next as one 64-bit entry of the LDS table (entry 4 next + (r & 3), filled
// by RingMovesT::init) instead of computing them: about 22 fewer VALU
// instructions; the bond partner is read beside the target, so the move
// still waits for three dependent LDS reads.  An empty site needs no test of
// its own: next = 0 puts the partner on the site itself, and neither end
// then holds stored length.
// Returns the outcome (kMoveEmpty ... kMoveAccepted) for counting launches.
template <class G, bool PAIRS = false, class F = NoFill>
__device__ __forceinline__ uint32_t trans_forw2(uint8_t* lat, uint32_t P, uint32_t r, const F& fill = F{}) {
    const uint32_t a = G::addr(P);
    uint32_t s = lat[a];
    const uint32_t next = s & 0xfu;
    const uint32_t pi = (next << 2) | (r & 3u);
    uint32_t pair, Pd, Pn;
    if (PAIRS) {
        const uint2 w = ring_tab<G>(lat)[pi];
        pair = w.x >> 24;
        Pd = P + w.x;
        Pn = P + w.y;
    } else {
        pair = fold_pair(next, r);
        Pd = P + G::delta(pair & 0xfu);
        Pn = P + G::delta(next);
    }
    const uint32_t nb1 = pair & 0xfu, nb2 = pair >> 4;
    const uint32_t na = G::addr(Pn), da = G::addr(Pd);
    uint32_t ns = lat[na];
    const uint32_t ds = lat[da];
    fill();
    // the parts of both outcomes that need neither read, computed while the
    // reads are in flight (the empty statement keeps them ahead of the wait)
    const uint32_t s1mf = nb1 | ((s >> 1) & 0x10u), s1no = nb1 | (s & 0x70u), d1mf = nb2 | (s & 0x10u);
    const bool inTile = !G::out(Pn | Pd);
    asm volatile("" ::"v"(s1mf), "v"(s1no), "v"(d1mf));
    // both outcomes as selects, one guarded write block (no early returns)
    const bool ok = inTile & (((s | ns) & 0x40u) != 0u) & (ds == 0u);
    const bool moveFirst = forw_move_first(s, ns, r);
    const uint32_t s1 = moveFirst ? s1mf : s1no;
    const uint32_t d1 = moveFirst ? d1mf : (nb2 | ((ns >> 1) & 0x10u));
    if (ok) {
        lat[a] = (uint8_t)s1;
        lat[da] = (uint8_t)d1;
        if (!moveFirst) lat[na] = (uint8_t)(ns & 0x1fu);
    }
    return s == 0u ? kMoveEmpty : !inTile ? kMoveFace : ok ? kMoveAccepted : kMoveRejected;
}

// rules = 1 (balanced): a fold forced by stored length at exactly one end
// (reference rate 1) is accepted with probability 1/4 (r bits 1-2 zero), the
// rate of the TransForw that undoes it, so every move and its reverse have
// equal probabilities (with the move choice of RingMoves::move).
// Bond next + srcNext as one bond: bits 0-3 the combined bond, bit 4 set when
// it is a bond (bond_add_valid).  TABLE: read it from the 256-byte LDS table
// (entry 16 next + srcNext).  When it is a bond its displacement is the sum
// of the two (checked for all 256 pairs), so the fold target P + combined
// bond is Ps + delta(srcNext): read beside the code, not after it.
__device__ __forceinline__ uint32_t combine_code(uint32_t next, uint32_t srcNext) {
    uint32_t c;
    const bool ok = bond_add_valid(next, srcNext, c);
    return bond_add(next, srcNext) | (ok ? 0x10u : 0u);
}

// TransBack's tests on bit 6 (stored length) of the site s, its source ss and
// the fold target ws, with the combine code cc (checked for every combination
// of those bits):
//   back_ok_bits: the combined bond is valid, the source holds no stored
//     length, and not both the site and the target hold it;
//   back_move_first: (sl | wsl) == 0 ? r bit 0 : sl == 0, as a mask.
__device__ __forceinline__ bool back_ok_bits(uint32_t s, uint32_t ss, uint32_t ws, uint32_t cc) {
    return (((cc << 2) & ~ss & ~(s & ws)) & 0x40u) != 0u;
}
__device__ __forceinline__ uint32_t back_move_first(uint32_t s, uint32_t ws, uint32_t r) {
    return bit6_mask(~s & (ws | (r << 6)));
}
// TransBack's new site and fold-target bytes: the source monomer joins the
// site (moveFirst) or the fold target.  BackParts holds what the site and
// source bytes fix before the target and the combine code are read.
struct BackParts {
    uint32_t siteMf, siteNo, tgtNo;
};
__device__ __forceinline__ BackParts back_parts(uint32_t s, uint32_t ss) {
    return BackParts{((s & 0x30u) << 1) | (ss & 0x30u) | 0x40u, s & 0x70u, ((ss & 0x30u) << 1) | 0x40u};
}
__device__ __forceinline__ uint32_t back_site(uint32_t mf, const BackParts& b, uint32_t newNext) {
    return newNext | bitsel(mf, b.siteMf, b.siteNo);
}
__device__ __forceinline__ uint32_t back_target(uint32_t mf, const BackParts& b, uint32_t ws) {
    return bitsel(mf, ws, (ws & 0x3fu) | b.tgtNo);
}
// DiffuseSL's new site and partner bytes, m = bit6_mask(s): the stored-length
// monomer crosses the bond towards the partner (m set) or onto the site.
// DiffuseParts holds what the site byte alone fixes.
struct DiffuseParts {
    uint32_t m, siteTo, siteFrom, partnerTo;
};
__device__ __forceinline__ DiffuseParts diffuse_parts(uint32_t s) {
    const uint32_t next = s & 0xfu;
    return DiffuseParts{bit6_mask(s), next | ((s >> 1) & 0x10u), next | ((s << 1) & 0x20u) | 0x40u,
                        ((s & 0x10u) << 1) | 0x40u};
}
__device__ __forceinline__ uint32_t diffuse_site(const DiffuseParts& d, uint32_t ns) {
    return bitsel(d.m, d.siteTo, d.siteFrom | ((ns >> 1) & 0x10u));
}
__device__ __forceinline__ uint32_t diffuse_partner(const DiffuseParts& d, uint32_t ns) {
    return bitsel(d.m, ns | d.partnerTo, ns & 0x1fu);
}

template <class G, bool TABLE = false, class F = NoFill>
__device__ __forceinline__ uint32_t trans_back2(uint8_t* lat, uint32_t P, uint32_t r, uint32_t rules,
                                                const F& fill = F{}) {
    const uint32_t a = G::addr(P);
    uint32_t s = lat[a];
    const uint32_t next = s & 0xfu;
    const uint32_t Ps = P + G::delta(next);
    const uint32_t sa = G::addr(Ps);
    const uint32_t ss = lat[sa];
    const uint32_t srcNext = ss & 0xfu;
    const uint32_t cc = TABLE ? (uint32_t)ring_combine_tab<G>(lat)[(next << 4) | srcNext] : combine_code(next, srcNext);
    const uint32_t Pw = Ps + G::delta(srcNext);
    const uint32_t newNext = cc & 0xfu;
    const uint32_t wa = G::addr(Pw);
    uint32_t ws = lat[wa];
    fill();
    // what the site and source fix, computed while the last reads are in
    // flight (the empty statement keeps it ahead of the wait)
    const BackParts bp = back_parts(s, ss);
    const bool pre = (s != 0u) & !G::out(Ps | Pw);
    asm volatile("" ::"v"(bp.siteMf), "v"(bp.siteNo), "v"(bp.tgtNo));
    const uint32_t sl = s & 0x40u, wsl = ws & 0x40u;
    const bool ok = pre & back_ok_bits(s, ss, ws, cc) & !((rules != 0u) & ((sl | wsl) != 0u) & ((r & 6u) != 0u));
    const uint32_t mf = back_move_first(s, ws, r);
    const uint32_t s1 = back_site(mf, bp, newNext);
    const uint32_t w1 = back_target(mf, bp, ws);
    if (ok) {
        lat[sa] = 0;
        lat[a] = (uint8_t)s1;
        lat[wa] = (uint8_t)w1;
    }
    return s == 0u ? kMoveEmpty : G::out(Ps | Pw) ? kMoveFace : ok ? kMoveAccepted : kMoveRejected;
}

template <class G, class F = NoFill>
__device__ __forceinline__ uint32_t diffuse_sl2(uint8_t* lat, uint32_t P, const F& fill = F{}) {
    const uint32_t a = G::addr(P);
    uint32_t s = lat[a];
    const uint32_t next = s & 0xfu;
    const uint32_t Pn = P + G::delta(next);
    const uint32_t na = G::addr(Pn);
    uint32_t ns = lat[na];
    fill();
    // what the site byte fixes, computed while the read is in flight (the
    // empty statement keeps it ahead of the wait)
    const DiffuseParts dp = diffuse_parts(s);
    const bool inTile = !G::out(Pn);
    asm volatile("" ::"v"(dp.m), "v"(dp.siteTo), "v"(dp.siteFrom), "v"(dp.partnerTo));
    // stored length at exactly one end; an empty site (next = 0) reads itself
    // as partner, so neither end holds it and nothing moves
    const bool ok = inTile & (((ns ^ s) & 0x40u) != 0u);
    const uint32_t s1 = diffuse_site(dp, ns);
    const uint32_t n1 = diffuse_partner(dp, ns);
    if (ok) {
        lat[a] = (uint8_t)s1;
        lat[na] = (uint8_t)n1;
    }
    return s == 0u ? kMoveEmpty : !inTile ? kMoveFace : ok ? kMoveAccepted : kMoveRejected;
}

// ringpol3's substep sequence {TransForw, DiffuseSL, TransForw, TransBack}
// (gpupol2/ringpol3.cu:278-292) as the move set of tile_unit.
template <bool PAIRS>
struct RingMovesT {
    using site_t = uint8_t;
    static constexpr int NSUB = 4;
    static constexpr bool HAS_RULES = true;
    static constexpr bool DRAW_AHEAD = true;
    static constexpr int EXTRA = PAIRS ? kRingTabTotal : 0;   // LDS bytes after the tile
    template <class G>
    __device__ static __forceinline__ void init(uint8_t* lat, int tid) {
        if (!PAIRS) return;
        if (tid < kFoldPairs) {
            const uint32_t pair = fold_pair((uint32_t)tid >> 2, (uint32_t)tid & 3u);
            ring_tab<G>(lat)[tid] = make_uint2((G::delta(pair & 0xfu) & 0xFFFFFFu) | (pair << 24),
                                               G::delta((uint32_t)tid >> 2));
            for (int i = tid; i < kCombine; i += kFoldPairs)
                ring_combine_tab<G>(lat)[i] = (uint8_t)combine_code((uint32_t)i >> 4, (uint32_t)i & 0xfu);
        }
    }
    template <class G, class F = NoFill>
    __device__ static __forceinline__ uint32_t move(uint8_t* lat, uint32_t P, int s, uint32_t r, uint32_t rules,
                                                    const F& fill = F{}) {
        if (rules) {
            // balanced rates: the move is drawn by bits 6-7 of the byte
            // (TransForw 1/2, DiffuseSL 1/4, TransBack 1/4 -- the proportions
            // of the reference sequence), so each substep is a symmetric
            // kernel and preserves the uniform measure by itself.  The tile
            // loop gives every slot of a substep the same two bits (from the
            // tile-uniform offset byte), so this is one scalar branch.
            const uint32_t ty = __builtin_amdgcn_readfirstlane(r >> 6);
            if (ty < 2u) return trans_forw2<G, PAIRS>(lat, P, r, fill);
            if (ty == 2u) return diffuse_sl2<G>(lat, P, fill);
            return trans_back2<G, PAIRS>(lat, P, r, rules, fill);
        }
        if (s == 1) return diffuse_sl2<G>(lat, P, fill);
        if (s == 3) return trans_back2<G, PAIRS>(lat, P, r, rules, fill);
        return trans_forw2<G, PAIRS>(lat, P, r, fill);
    }
};
using RingMoves = RingMovesT<false>;

template <int TT, int TU, int TV, bool DENSE, int THREADS, int PAD, bool PAIRS = false, bool COUNT = false>
__global__ __launch_bounds__(THREADS) void ring_time_unit2(uint8_t* __restrict__ glat, RingLaunch p) {
    __shared__ __align__(16) uint8_t tile[PTile<TT, TU, TV, PAD>::NSITE + RingMovesT<PAIRS>::EXTRA];
    tile_unit<RingMovesT<PAIRS>, TT, TU, TV, DENSE, THREADS, PAD, COUNT>(glat, tile, unit_args(p),
                                                                         xcd_tile(blockIdx.x, gridDim.x));
}

// ring_time_unit2 with the occupied slots of every substep compacted across
// the workgroup's waves (variant 5; measured against the default).
template <int TT, int TU, int TV, bool DENSE, int THREADS, int PAD>
__global__ __launch_bounds__(THREADS) void ring_time_unit_compact(uint8_t* __restrict__ glat, RingLaunch p) {
    __shared__ __align__(16) uint8_t tile[PTile<TT, TU, TV, PAD>::NSITE + RingMovesT<true>::EXTRA];
    __shared__ uint32_t list[THREADS + THREADS / 64];
    tile_unit_compact<RingMovesT<true>, TT, TU, TV, DENSE, THREADS, PAD>(glat, tile, list, unit_args(p),
                                                                         xcd_tile(blockIdx.x, gridDim.x));
}

// One time unit of every replica of a batch (one workgroup per tile).
template <int TT, int TU, int TV, bool DENSE, int THREADS, int PAD, bool PAIRS = false, bool COUNT = false>
__global__ __launch_bounds__(THREADS) void ring_batch_unit(uint8_t* __restrict__ gbase, BatchLaunch b) {
    __shared__ __align__(16) uint8_t tile[PTile<TT, TU, TV, PAD>::NSITE + RingMovesT<PAIRS>::EXTRA];
    batch_unit<RingMovesT<PAIRS>, TT, TU, TV, DENSE, THREADS, PAD, COUNT>(gbase, tile, b);
}

// ring_time_unit2 for variants 2-4: P-indexed (padded) LDS only for
// 16x16x{8,16} tiles (at most 16 KiB), compact otherwise; variant 4 adds the
// LDS bond-pair table of TransForw.
template <int TT, int TU, int TV, bool DENSE, int TH>
static void launch2(int variant, int nblocks, hipStream_t stream, uint8_t* glat, const RingLaunch& p) {
    constexpr int PADC = (TT == 16 && TU == 16 && TV <= 16) ? 1 : 0;
    if (p.counts) {   // counting instantiation of the default formulation (variant 4)
        hipLaunchKernelGGL((ring_time_unit2<TT, TU, TV, DENSE, TH, PADC, true, true>), dim3(nblocks), dim3(TH), 0,
                           stream, glat, p);
        return;
    }
    constexpr bool ONE_SLOT = (TT / 4) * (TU / 4) * (TV / 4) * (DENSE ? 4 : 1) == TH;
    if constexpr (ONE_SLOT) {
        if (variant == 5) {
            hipLaunchKernelGGL((ring_time_unit_compact<TT, TU, TV, DENSE, TH, PADC>), dim3(nblocks), dim3(TH), 0,
                               stream, glat, p);
            return;
        }
    }
    if constexpr (TT == 16 && TU == 16 && TV <= 16) {
        if (variant == 4)
            hipLaunchKernelGGL((ring_time_unit2<TT, TU, TV, DENSE, TH, 1, true>), dim3(nblocks), dim3(TH), 0, stream,
                               glat, p);
        else if (variant == 3)
            hipLaunchKernelGGL((ring_time_unit2<TT, TU, TV, DENSE, TH, 1>), dim3(nblocks), dim3(TH), 0, stream, glat, p);
        else
            hipLaunchKernelGGL((ring_time_unit2<TT, TU, TV, DENSE, TH, 0>), dim3(nblocks), dim3(TH), 0, stream, glat, p);
    } else {
        if (variant == 4)
            hipLaunchKernelGGL((ring_time_unit2<TT, TU, TV, DENSE, TH, 0, true>), dim3(nblocks), dim3(TH), 0, stream,
                               glat, p);
        else
            hipLaunchKernelGGL((ring_time_unit2<TT, TU, TV, DENSE, TH, 0>), dim3(nblocks), dim3(TH), 0, stream, glat, p);
    }
}

// ---- dispatch over the compiled schedules ----
// variant 1: ring_time_unit (cell-major LDS); 2: ring_time_unit2 compact;
// 3: ring_time_unit2 with P-indexed (padded) LDS where the tile allows it;
// 4 (default): as 3 with the LDS bond-pair table of TransForw; 5: as 4 with
// the occupied slots of each substep compacted across the waves (schedules
// with one slot per thread; the others run variant 4).
#define ELP_RING_CASE(TT, TU, TV, DENSE, TH)                                                              \
    if (tt == TT && tu == TU && tv == TV && dense == DENSE && threads == TH) {                             \
        if (probe) return 0;                                                                               \
        if (variant == 1 && (p.rules || p.counts))                                                         \
            return (int)hipErrorInvalidValue; /* first version: reference rates, no counting */           \
        if (variant == 1)                                                                                  \
            hipLaunchKernelGGL((ring_time_unit<TT, TU, TV, DENSE, TH>), dim3(nblocks), dim3(TH), 0, stream, \
                               glat, p);                                                                   \
        else                                                                                               \
            launch2<TT, TU, TV, DENSE, TH>(variant, nblocks, stream, glat, p);                             \
        return (int)hipGetLastError();                                                                     \
    }

int ring_launch(int tt, int tu, int tv, bool dense, int threads, bool probe, int nblocks, hipStream_t stream,
                uint8_t* glat, const RingLaunch& p) {
    // default: P-indexed LDS where the tile allows it, pair table (fastest measured)
    static const int variant = [] {
        const char* env = getenv("ELP_RING_VARIANT");
        return env ? atoi(env) : 4;
    }();
    ELP_RING_CASE(16, 16, 16, false, 64)
    ELP_RING_CASE(16, 16, 16, true, 64)
    ELP_RING_CASE(16, 16, 16, true, 128)
    ELP_RING_CASE(16, 16, 16, true, 256)
    ELP_RING_CASE(32, 32, 32, false, 256)
    ELP_RING_CASE(32, 32, 32, false, 512)
    ELP_RING_CASE(32, 32, 32, true, 512)
    ELP_RING_CASE(32, 32, 32, true, 1024)
    ELP_RING_CASE(32, 32, 16, false, 256)
    ELP_RING_CASE(32, 32, 16, true, 256)
    ELP_RING_CASE(32, 32, 16, true, 512)
    ELP_RING_CASE(16, 16, 32, false, 128)
    ELP_RING_CASE(16, 16, 32, true, 128)
    ELP_RING_CASE(16, 16, 32, true, 256)
    ELP_RING_CASE(16, 16, 8, true, 128)
    return -1;
}

// Batched launches use the second formulation (P-indexed LDS for 16^3 tiles).
#define ELP_RING_BATCH_CASE(TT, TU, TV, DENSE, TH)                                                         \
    if (tt == TT && tu == TU && tv == TV && dense == DENSE && threads == TH) {                             \
        if (probe) return 0;                                                                               \
        constexpr int PAD = (TT == 16 && TU == 16 && TV <= 16) ? 1 : 0;                                    \
        if (b.counts)                                                                                      \
            hipLaunchKernelGGL((ring_batch_unit<TT, TU, TV, DENSE, TH, PAD, true, true>), dim3(nblocks),    \
                               dim3(TH), 0, stream, gbase, b);                                             \
        else if (batch_pairs)                                                                              \
            hipLaunchKernelGGL((ring_batch_unit<TT, TU, TV, DENSE, TH, PAD, true>), dim3(nblocks), dim3(TH), 0, \
                               stream, gbase, b);                                                          \
        else                                                                                               \
            hipLaunchKernelGGL((ring_batch_unit<TT, TU, TV, DENSE, TH, PAD>), dim3(nblocks), dim3(TH), 0, stream, \
                               gbase, b);                                                                  \
        return (int)hipGetLastError();                                                                     \
    }

int ring_batch_launch(int tt, int tu, int tv, bool dense, int threads, bool probe, int nblocks, hipStream_t stream,
                      uint8_t* gbase, const BatchLaunch& b) {
    // LDS bond-pair table of TransForw unless ELP_RING_PAIRS=0 (measured +9 %)
    static const bool batch_pairs = [] {
        const char* env = getenv("ELP_RING_PAIRS");
        return !env || atoi(env) != 0;
    }();
    ELP_RING_BATCH_CASE(16, 16, 16, false, 64)
    ELP_RING_BATCH_CASE(16, 16, 16, true, 64)
    ELP_RING_BATCH_CASE(16, 16, 16, true, 128)
    ELP_RING_BATCH_CASE(16, 16, 16, true, 256)
    ELP_RING_BATCH_CASE(16, 16, 8, true, 128)
    ELP_RING_BATCH_CASE(32, 32, 32, false, 256)
    ELP_RING_BATCH_CASE(32, 32, 32, true, 512)
    return -1;
}

}  // namespace elp
