// elp_linear.hip -- gfx950 move kernel of the linear-chain ELP engine.
//
// Same tile/sub-lattice/RNG scheme as the ring kernel (elp_ring.hip), with
// one uint16 word per site in LDS and HBM:
//   bits 0-3 next bond, 4-7 prev bond (bond from the previous site into this
//   one), bit 8 stored length, bit 9 label of the monomer carrying the next
//   bond, bit 10 label of the stored-length monomer (which precedes it).
// Per iteration: {DiffuseSL, TransMove} (gpupol/linpol.cu:483-490); the
// reference runs 64 iterations per time unit (gpupol/gpupol_init.c:79).
// The reference packs these fields into two LDS arrays with shared words
// (gpupol/linpol.cu:97-105); here every field of a site lives in that site's
// own word, so per-site results are the same and no read-modify-write of a
// neighbour's word is needed.
// Randomness: Philox4x32-10 as in the ring kernel; TransMove of iteration i
// uses byte 1 of word i%4 of the slot stream, the coset of substep s is
// byte s of the tile-uniform offset word.
#include "elp_device.h"
#include "elp_kernels.h"
#include "elp_tile_unit.h"

namespace elp {

constexpr uint32_t kSL = 0x100u, kL0 = 0x200u, kL1 = 0x400u;

template <int TT, int TU, int TV>
struct LinTile {
    static constexpr int NCT = TT / 4, NCU = TU / 4, NCV = TV / 4;
    static constexpr int NCELL = NCT * NCU * NCV;
    static constexpr int NSITE = TT * TU * TV;
    __device__ static __forceinline__ int addr(int t, int u, int v) {
        const int cell = (t >> 2) + NCT * ((u >> 2) + NCU * (v >> 2));
        return cell + NCELL * ((t & 3) | ((u & 3) << 2) | ((v & 3) << 4));
    }
    __device__ static __forceinline__ bool inside(int t, int u, int v) {
        return (unsigned)t < (unsigned)TT && (unsigned)u < (unsigned)TU && (unsigned)v < (unsigned)TV;
    }
};

__device__ __forceinline__ bool bond_valid(uint32_t b) { return b != 0u && b != 3u && b != 0xcu && b != 0xfu; }

// Move one stored-length unit across the next bond when exactly one end
// holds it (gpupol/linpol.cu:314-363).
template <class G>
__device__ __forceinline__ void lin_diffuse_sl(uint16_t* lat, int t, int u, int v) {
    const int a = G::addr(t, u, v);
    uint32_t w = lat[a];
    const uint32_t next = w & 0xfu;
    if (!next) return;
    const int nt = t + bdt(next), nu = u + bdu(next), nv = v + bdv(next);
    if (!G::inside(nt, nu, nv)) return;
    const int na = G::addr(nt, nu, nv);
    uint32_t x = lat[na];
    if ((w & kSL) && !(x & kSL)) {
        const uint32_t l0 = (w >> 9) & 1u, l1 = (w >> 10) & 1u;
        w = (w & ~(kL0 | kL1 | kSL)) | (l1 << 9);
        x |= (l0 << 10) | kSL;
    } else if (!(w & kSL) && (x & kSL)) {
        const uint32_t q = (x >> 10) & 1u;
        x &= ~(kL1 | kSL);
        w = (w & ~kL0) | (((w >> 9) & 1u) << 10) | (q << 9) | kSL;
    } else {
        return;
    }
    lat[a] = (uint16_t)w;
    lat[na] = (uint16_t)x;
}

// Reptation / stored-length move of one monomer (gpupol/linpol.cu:106-311):
// with stored length at the site, split the next or prev bond (r bit 2)
// through an empty common neighbour, or grow the chain end into an empty
// neighbour (new bond r bits 3-6); without it, retract an end (gate r bits
// 4-7) or fold the monomer onto its prev/next neighbour (r bit 0) as stored
// length when the two bonds add up to one.
template <class G>
__device__ __forceinline__ void lin_trans_move(uint16_t* lat, int t, int u, int v, uint32_t r, const uint32_t* trans) {
    const int a = G::addr(t, u, v);
    uint32_t w = lat[a];
    if (!(w & 0xffu)) return;
    const uint32_t next = w & 0xfu, prev = (w >> 4) & 0xfu;
    if (w & kSL) {
        uint32_t pb = (r >> 2) & 1u;
        const uint32_t bond = pb ? prev : next;
        if (bond == 0) {  // grow the chain end out of the stored length
            pb ^= 1u;
            const uint32_t nb = (r >> 3) & 0xfu;
            if (!bond_valid(nb)) return;
            const uint32_t ins = pb ? nb : (nb ^ 0xfu);
            const int nt = t + bdt(nb), nu = u + bdu(nb), nv = v + bdv(nb);
            if (!G::inside(nt, nu, nv)) return;
            const int na = G::addr(nt, nu, nv);
            uint32_t x = lat[na];
            if (x & 0xffu) return;
            const uint32_t lab = (w >> 9) & 3u;
            w = (w & ~(kL0 | kL1 | kSL)) | (((lab >> pb) & 1u) << 9) | (ins << (4 * (pb ^ 1u)));
            x |= (((lab >> (pb ^ 1u)) & 1u) << 9) | (ins << (4 * pb));
            lat[a] = (uint16_t)w;
            lat[na] = (uint16_t)x;
            return;
        }
        const uint32_t dir = pb ? (bond ^ 0xfu) : bond;
        const int bt = t + bdt(dir), bu = u + bdu(dir), bv = v + bdv(dir);
        if (!G::inside(bt, bu, bv)) return;
        uint32_t nPrev = (trans[bond >> 2] >> (4 * (2 * (bond & 3u) + (r & 1u)))) & 0xfu;
        uint32_t nNext = bond_add((~nPrev) & 0xfu, bond);
        if (!(r & 2u)) { const uint32_t x = nPrev; nPrev = nNext; nNext = x; }
        const int pt = pb ? bt : t, pu = pb ? bu : u, pv = pb ? bv : v;   // old prev site
        const int nt = pt + bdt(nPrev), nu = pu + bdu(nPrev), nv = pv + bdv(nPrev);
        if (!G::inside(nt, nu, nv)) return;
        const int na = G::addr(nt, nu, nv);
        uint32_t x = lat[na];
        if (x & 0xffu) return;
        const int ba = G::addr(bt, bu, bv);
        uint32_t y = lat[ba];
        const uint32_t lab = (w >> (9 + pb)) & 1u;
        x |= nNext | (nPrev << 4) | (lab << 9);
        if (pb) {  // split prev bond: the stored-length monomer moves out
            w = (w & ~(0xf0u | kSL | kL1)) | (nNext << 4);
            y = (y & ~0xfu) | nPrev;
        } else {   // split next bond: the bond carrier moves out
            w = (w & ~(0xfu | kSL | kL0 | kL1)) | nPrev | (((w >> 10) & 1u) << 9);
            y = (y & ~0xf0u) | (nNext << 4);
        }
        lat[a] = (uint16_t)w;
        lat[na] = (uint16_t)x;
        lat[ba] = (uint16_t)y;
        return;
    }
    if ((r >> 4) & 1u) return;
    if (next == 0 || prev == 0) {  // retract a chain end onto its neighbour
        if ((r >> 5) & 7u) return;
        uint32_t bond, pb;
        if (next == 0) { bond = prev ^ 0xfu; pb = 1u; }
        else { bond = next; pb = 0u; }
        const int nt = t + bdt(bond), nu = u + bdu(bond), nv = v + bdv(bond);
        if (!G::inside(nt, nu, nv)) return;
        const int na = G::addr(nt, nu, nv);
        uint32_t x = lat[na];
        if (x & kSL) return;
        const uint32_t lab = (w >> 9) & 1u;
        const uint32_t nl = (lab << (pb ^ 1u)) | (((x >> 9) & 1u) << pb);
        w &= ~(kL0 | (0xfu << (4 * pb)));
        x = (x & ~(kL0 | kL1 | (0xfu << (4 * (pb ^ 1u))))) | (nl << 9) | kSL;
        lat[a] = (uint16_t)w;
        lat[na] = (uint16_t)x;
        return;
    }
    const uint32_t pb = r & 1u;
    const uint32_t back = (~prev) & 0xfu;
    const int pt = t + bdt(back), pu = u + bdu(back), pv = v + bdv(back);
    const int nt = t + bdt(next), nu = u + bdu(next), nv = v + bdv(next);
    if (!G::inside(pt, pu, pv) || !G::inside(nt, nu, nv)) return;
    uint32_t bond;
    if (!bond_add_valid(next, prev, bond)) return;
    const int pa = G::addr(pt, pu, pv), na = G::addr(nt, nu, nv);
    const int ba = pb ? pa : na;
    if (ba == a) return;
    uint32_t y = lat[ba];
    if (y & kSL) return;
    const uint32_t lab = (w >> 9) & 1u;
    y |= kSL;
    if (pb) y = (y & ~kL0) | (((y >> 9) & 1u) << 10) | (lab << 9);   // carrier becomes SL, monomer carries
    else y |= lab << 10;                                               // monomer is SL before the carrier
    lat[ba] = (uint16_t)y;
    lat[a] = (uint16_t)(w & ~(kL0 | 0xffu));
    lat[pa] = (uint16_t)((lat[pa] & ~0xfu) | bond);
    lat[na] = (uint16_t)((lat[na] & ~0xf0u) | (bond << 4));
}

template <int TT, int TU, int TV, bool DENSE, int THREADS>
__global__ __launch_bounds__(THREADS) void lin_time_unit(uint16_t* __restrict__ glat, LinLaunch p) {
    using G = LinTile<TT, TU, TV>;
    constexpr int SPC = DENSE ? 4 : 1;
    constexpr int NSLOT = G::NCELL * SPC;
    constexpr int SPT = NSLOT / THREADS;
    static_assert(NSLOT % THREADS == 0 && SPT >= 1, "slots must split evenly over threads");
    extern __shared__ uint16_t tile[];

    const int tid = threadIdx.x;
    const uint32_t tileId = xcd_tile(blockIdx.x, gridDim.x);
    const int kt = (int)(tileId % (uint32_t)p.ntile_t);
    const int ku = (int)((tileId / (uint32_t)p.ntile_t) % (uint32_t)p.ntile_u);
    const int kv = (int)(tileId / (uint32_t)(p.ntile_t * p.ntile_u));
    const int gt0 = p.ot + kt * TT, gu0 = p.ou + ku * TU, gv0 = p.ov + kv * TV;

    for (int i = tid; i < G::NSITE; i += THREADS) {
        const int lt = i % TT, lu = (i / TT) % TU, lv = i / (TT * TU);
        int gt = gt0 + lt, gu = gu0 + lu, gv = gv0 + lv;
        gt -= (gt >= p.LT) ? p.LT : 0;
        gu -= (gu >= p.LU) ? p.LU : 0;
        gv -= (gv >= p.LV) ? p.LV : 0;
        tile[G::addr(lt, lu, lv)] = glat[(size_t)gt + (size_t)p.LT * ((size_t)gu + (size_t)p.LU * (size_t)gv)];
    }

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
            for (int s = 0; s < 2; s++) {
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
                    if (s == 0) lin_diffuse_sl<G>(tile, t, u, v);
                    else lin_trans_move<G>(tile, t, u, v, (word_of(sw[k], j) >> 8) & 0xffu, p.trans);
                }
                __syncthreads();
            }
        }
    }

    for (int i = tid; i < G::NSITE; i += THREADS) {
        const int lt = i % TT, lu = (i / TT) % TU, lv = i / (TT * TU);
        int gt = gt0 + lt, gu = gu0 + lu, gv = gv0 + lv;
        gt -= (gt >= p.LT) ? p.LT : 0;
        gu -= (gu >= p.LU) ? p.LU : 0;
        gv -= (gv >= p.LV) ? p.LV : 0;
        glat[(size_t)gt + (size_t)p.LT * ((size_t)gu + (size_t)p.LU * (size_t)gv)] = tile[G::addr(lt, lu, lv)];
    }
}

// =====================================================================
// Second formulation (default): packed guarded coordinates (PTile, see
// elp_device.h), all neighbour words a move may need read right after the
// site word, then one accept/reject decision per case.  Same per-site
// results as the functions above.
// =====================================================================
// The move functions return their outcome (kMoveEmpty ... kMoveAccepted) for
// counting launches.
template <class G>
__device__ __forceinline__ uint32_t lin_diffuse_sl2(uint16_t* lat, uint32_t P) {
    const uint32_t a = G::addr(P);
    uint32_t w = lat[a];
    const uint32_t next = w & 0xfu;
    const uint32_t Pn = P + G::delta(next);
    const uint32_t na = G::addr(Pn);
    const uint32_t x = lat[na];
    // both directions as selects, one guarded write block
    const bool wsl = (w & kSL) != 0u, xsl = (x & kSL) != 0u;
    const bool toX = wsl & !xsl, toW = !wsl & xsl;
    const bool ok = (next != 0u) & !G::out(Pn) & (toX | toW);
    const uint32_t wX = (w & ~(kL0 | kL1 | kSL)) | (((w >> 10) & 1u) << 9);
    const uint32_t xX = x | (((w >> 9) & 1u) << 10) | kSL;
    const uint32_t wW = (w & ~kL0) | (((w >> 9) & 1u) << 10) | (((x >> 10) & 1u) << 9) | kSL;
    const uint32_t xW = x & ~(kL1 | kSL);
    if (ok) {
        lat[a] = (uint16_t)(toX ? wX : wW);
        lat[na] = (uint16_t)(toX ? xX : xW);
    }
    return (w & 0xffu) == 0u ? kMoveEmpty : ((next != 0u) & G::out(Pn)) ? kMoveFace : ok ? kMoveAccepted : kMoveRejected;
}

// The two sites a case reads are X2 = P + delta(c2) and X1 = base + delta(c1):
//   split   (stored length, bondS != 0): X2 = B (bond partner), X1 = N
//           (inserted site, from B when the bond is prev, else from P);
//   grow    (stored length, bondS == 0): X1 = X2 = Gr = P + delta(nb);
//   retract (no stored length, chain end): X1 = X2 = E;
//   fold    (no stored length, interior): X1 = Pp, X2 = Pn.
// Only the two offsets of the taken case are computed, and the bond pair of
// the split comes from the LDS table after the tile: word 4 bondS + (r & 3)
// holds fold_pair in bits 24-31 above the packed offset of its first bond
// (nPrev) in bits 0-23, so the split's X1 is one add of that word.
template <class G, class F = NoFill>
__device__ __forceinline__ uint32_t lin_trans_move_ends(uint16_t* lat, uint32_t P, uint32_t r, uint32_t a, uint32_t w,
                                                        uint32_t wp, const F& fill = F{}) {
    const uint32_t next = w & 0xfu, prev = (w >> 4) & 0xfu;
    const bool sl = (w & kSL) != 0u;
    // stored-length cases: split the chosen bond, or grow the end
    const uint32_t pbS = (r >> 2) & 1u;
    const uint32_t bondS = pbS ? prev : next;
    const uint32_t nb = (r >> 3) & 0xfu;
    const uint32_t pair = wp >> 24;
    const uint32_t nPrev = pair & 0xfu, nNext = pair >> 4;
    // cases without stored length: retract an end, or fold an interior monomer
    const bool end = (next == 0u) | (prev == 0u);
    const uint32_t bondE = (next == 0u) ? (prev ^ 0xfu) : next;
    // Every candidate is a named value, so each choice below is a select and
    // a wave whose lanes take different cases runs no divergent branch.
    const uint32_t c2S = pbS ? (bondS ^ 0xfu) : bondS;
    const uint32_t c2sl = bondS ? c2S : nb;
    const uint32_t c2no = end ? bondE : next;
    const uint32_t c1no = end ? bondE : (prev ^ 0xfu);
    const uint32_t c2 = sl ? c2sl : c2no;
    const uint32_t c1 = sl ? nb : c1no;
    const uint32_t X2 = P + G::delta(c2);
    // split: X1 = (pbS ? X2 : P) + delta(nPrev), the offset taken from the table word
    const uint32_t X1split = (pbS ? X2 : P) + wp;
    const uint32_t X1other = P + G::delta(c1);
    const uint32_t X1 = (sl & (bondS != 0u)) ? X1split : X1other;
    const uint32_t a1 = G::addr(X1), a2 = G::addr(X2);
    const uint32_t x1 = lat[a1];
    const uint32_t x2 = lat[a2];
    fill();
    // The four cases as selects and one guarded write block.
    const bool occ = (w & 0xffu) != 0u;
    const bool gate = ((r >> 4) & 1u) == 0u;            // cases without stored length
    // grow the chain end out of the stored length
    const bool nbBad = (nb == 0u) | (nb == 3u) | (nb == 0xcu) | (nb == 0xfu);
    const bool okGrow = sl & (bondS == 0u) & !nbBad & !G::out(X1) & ((x1 & 0xffu) == 0u);
    const uint32_t pbG = pbS ^ 1u, labG = (w >> 9) & 3u;
    const uint32_t ins = pbG ? nb : (nb ^ 0xfu);
    const uint32_t wG = (w & ~(kL0 | kL1 | kSL)) | (((labG >> pbG) & 1u) << 9) | (ins << (4 * (pbG ^ 1u)));
    const uint32_t x1G = x1 | (((labG >> (pbG ^ 1u)) & 1u) << 9) | (ins << (4 * pbG));
    // retract a chain end onto its neighbour as stored length
    const bool okEnd = !sl & gate & end & (((r >> 5) & 7u) == 0u) & !G::out(X1) & ((x1 & kSL) == 0u);
    const uint32_t pbE = (next == 0u) ? 1u : 0u;
    const uint32_t nl = (((w >> 9) & 1u) << (pbE ^ 1u)) | (((x1 >> 9) & 1u) << pbE);
    const uint32_t wE = w & ~(kL0 | (0xfu << (4 * pbE)));
    const uint32_t x1E = (x1 & ~(kL0 | kL1 | (0xfu << (4 * (pbE ^ 1u))))) | (nl << 9) | kSL;
    // split the chosen bond through an empty site
    const bool okSplit = sl & (bondS != 0u) & !G::out(X2 | X1) & ((x1 & 0xffu) == 0u);
    const uint32_t labS = (w >> (9 + pbS)) & 1u;
    const uint32_t x1S = x1 | nNext | (nPrev << 4) | (labS << 9);
    const uint32_t wSprev = (w & ~(0xf0u | kSL | kL1)) | (nNext << 4);
    const uint32_t wSnext = (w & ~(0xfu | kSL | kL0 | kL1)) | nPrev | (((w >> 10) & 1u) << 9);
    const uint32_t wS = pbS ? wSprev : wSnext;
    const uint32_t x2Sprev = (x2 & ~0xfu) | nPrev, x2Snext = (x2 & ~0xf0u) | (nNext << 4);
    const uint32_t x2S = pbS ? x2Sprev : x2Snext;
    // fold the monomer onto its prev (r bit 0) or next neighbour as stored
    // length; bond_add_valid(next, prev) written without an early return
    const uint32_t orb = next | prev, andb = next & prev;
    const uint32_t bond = (orb == 0xfu) ? andb : orb;
    const bool bondOk = ((orb == 0xfu) | (andb == 0u)) & (bond != 0x3u) & (bond != 0xcu);
    const uint32_t pbF = r & 1u, labF = (w >> 9) & 1u;
    const uint32_t xF = pbF ? x1 : x2;
    const bool okFold = !sl & gate & !end & !G::out(X1 | X2) & bondOk & ((xF & kSL) == 0u);
    const uint32_t x1fprev = ((x1 | kSL) & ~kL0) | (((x1 >> 9) & 1u) << 10) | (labF << 9);
    const uint32_t x2fnext = x2 | kSL | (labF << 10);
    const uint32_t x1f = pbF ? x1fprev : x1;
    const uint32_t x2f = pbF ? x2 : x2fnext;
    const uint32_t x1F = (x1f & ~0xfu) | bond, x2F = (x2f & ~0xf0u) | (bond << 4);
    const uint32_t wF = w & ~(kL0 | 0xffu);
    const bool acc = occ & (okGrow | okSplit | okEnd | okFold);
    if (acc) {
        lat[a] = (uint16_t)(okGrow ? wG : okSplit ? wS : okEnd ? wE : wF);
        lat[a1] = (uint16_t)(okGrow ? x1G : okSplit ? x1S : okEnd ? x1E : x1F);
        if (okSplit | okFold) lat[a2] = (uint16_t)(okSplit ? x2S : x2F);
    }
    // a face rejection: the random gates of the case pass (growth: a valid
    // new bond; without stored length: bit 4, and bits 5-7 for an end) and
    // the case's sites leave the tile
    const bool pass = sl ? !((bondS == 0u) & nbBad) : (gate & (!end | (((r >> 5) & 7u) == 0u)));
    return !occ ? kMoveEmpty : (pass & G::out(X1 | X2)) ? kMoveFace : acc ? kMoveAccepted : kMoveRejected;
}

// TransMove of a wave in which no occupied site is a chain end: every
// occupied lane has next != 0 and prev != 0, so only the split (stored
// length) and the fold (none) can happen, the split's bond is never 0 and
// the fold needs no end test.  Same outcome as lin_trans_move_ends for such a
// wave; empty lanes read in-tile sites and write nothing.
template <class G, class F = NoFill>
__device__ __forceinline__ uint32_t lin_trans_move_interior(uint16_t* lat, uint32_t P, uint32_t r, uint32_t a,
                                                            uint32_t w, uint32_t wp, const F& fill = F{}) {
    const uint32_t next = w & 0xfu, prev = (w >> 4) & 0xfu;
    const bool sl = (w & kSL) != 0u;
    const uint32_t pbS = (r >> 2) & 1u;
    const uint32_t bondS = pbS ? prev : next;
    const uint32_t pair = wp >> 24;
    const uint32_t nPrev = pair & 0xfu, nNext = pair >> 4;
    // split: X2 = bond partner, X1 = inserted site; fold: X2 = Pn, X1 = Pp.
    // X1 as a mask blend of its two candidates (a ternary over them compiles
    // to a divergent branch whose split arm waits for the table word before
    // the fold arm may start)
    const uint32_t c2 = sl ? (pbS ? (bondS ^ 0xfu) : bondS) : next;
    const uint32_t X2 = P + G::delta(c2);
    const uint32_t X1split = (pbS ? X2 : P) + wp, X1fold = P + G::delta(prev ^ 0xfu);
    const uint32_t slm = 0u - (uint32_t)sl;
    const uint32_t X1 = (X1split & slm) | (X1fold & ~slm);
    const uint32_t a1 = G::addr(X1), a2 = G::addr(X2);
    const uint32_t x1 = lat[a1];
    const uint32_t x2 = lat[a2];
    fill();
    const bool occ = (w & 0xffu) != 0u;
    const bool gate = ((r >> 4) & 1u) == 0u;
    const bool inTile = !G::out(X2 | X1);
    // split the chosen bond through an empty site
    const bool okSplit = sl & inTile & ((x1 & 0xffu) == 0u);
    const uint32_t labS = (w >> (9 + pbS)) & 1u;
    const uint32_t x1S = x1 | nNext | (nPrev << 4) | (labS << 9);
    const uint32_t wSprev = (w & ~(0xf0u | kSL | kL1)) | (nNext << 4);
    const uint32_t wSnext = (w & ~(0xfu | kSL | kL0 | kL1)) | nPrev | (((w >> 10) & 1u) << 9);
    const uint32_t wS = pbS ? wSprev : wSnext;
    const uint32_t x2Sprev = (x2 & ~0xfu) | nPrev, x2Snext = (x2 & ~0xf0u) | (nNext << 4);
    const uint32_t x2S = pbS ? x2Sprev : x2Snext;
    // fold the monomer onto its prev (r bit 0) or next neighbour
    const uint32_t orb = next | prev, andb = next & prev;
    const uint32_t bond = (orb == 0xfu) ? andb : orb;
    const bool bondOk = ((orb == 0xfu) | (andb == 0u)) & (bond != 0x3u) & (bond != 0xcu);
    const uint32_t pbF = r & 1u, labF = (w >> 9) & 1u;
    const uint32_t xF = pbF ? x1 : x2;
    const bool okFold = !sl & gate & inTile & bondOk & ((xF & kSL) == 0u);
    const uint32_t x1fprev = ((x1 | kSL) & ~kL0) | (((x1 >> 9) & 1u) << 10) | (labF << 9);
    const uint32_t x2fnext = x2 | kSL | (labF << 10);
    const uint32_t x1f = pbF ? x1fprev : x1;
    const uint32_t x2f = pbF ? x2 : x2fnext;
    const uint32_t x1F = (x1f & ~0xfu) | bond, x2F = (x2f & ~0xf0u) | (bond << 4);
    const uint32_t wF = w & ~(kL0 | 0xffu);
    // okSplit implies sl and okFold !sl, so sl picks the outcome
    const bool acc = occ & (okSplit | okFold);
    if (acc) {
        lat[a] = (uint16_t)(sl ? wS : wF);
        lat[a1] = (uint16_t)(sl ? x1S : x1F);
        lat[a2] = (uint16_t)(sl ? x2S : x2F);
    }
    return !occ ? kMoveEmpty : ((sl | gate) & !inTile) ? kMoveFace : acc ? kMoveAccepted : kMoveRejected;
}

// Growing and retracting need a chain end, 2 of every N monomers: a wave with
// no occupied chain-end site takes the two-case form on one scalar branch.
// Both forms read the table word of the chosen bond (bondS, r & 3), which is
// read before the branch so that the read and the test overlap.
template <class G, class F = NoFill>
__device__ __forceinline__ uint32_t lin_trans_move2(uint16_t* lat, uint32_t P, uint32_t r, const F& fill = F{}) {
    const uint32_t a = G::addr(P);
    const uint32_t w = lat[a];
    const uint32_t bondS = ((r >> 2) & 1u) ? (w >> 4) & 0xfu : w & 0xfu;
    const uint32_t wp = reinterpret_cast<const uint32_t*>(lat + G::NSITE)[(bondS << 2) | (r & 3u)];
    const bool occEnd = ((w & 0xffu) != 0u) & (((w & 0xfu) == 0u) | ((w & 0xf0u) == 0u));
    if (__ballot(occEnd) != 0ull) return lin_trans_move_ends<G>(lat, P, r, a, w, wp, fill);
    return lin_trans_move_interior<G>(lat, P, r, a, w, wp, fill);
}

// linpol's substep sequence {DiffuseSL, TransMove} (gpupol/linpol.cu:483-490)
// as the move set of tile_unit; TransMove takes byte 1 of the move word.
struct LinMoves {
    using site_t = uint16_t;
    static constexpr int NSUB = 2;
    static constexpr bool HAS_RULES = false;
    // the linear TransMove is closer to VALU-bound: the early draw (one unused
    // group per launch) measured 0.6 % slower than drawing at each group's start
    static constexpr bool DRAW_AHEAD = false;
    static constexpr int EXTRA = 4 * kFoldPairs;   // LDS bytes after the tile
    template <class G>
    __device__ static __forceinline__ void init(uint16_t* lat, int tid) {
        if (tid < kFoldPairs) {
            const uint32_t pair = fold_pair((uint32_t)tid >> 2, (uint32_t)tid & 3u);
            reinterpret_cast<uint32_t*>(lat + G::NSITE)[tid] = (G::delta(pair & 0xfu) & 0xFFFFFFu) | (pair << 24);
        }
    }
    template <class G, class F = NoFill>
    __device__ static __forceinline__ uint32_t move(uint16_t* lat, uint32_t P, int s, uint32_t r, uint32_t,
                                                    const F& fill = F{}) {
        if (s == 0) { fill(); return lin_diffuse_sl2<G>(lat, P); }
        return lin_trans_move2<G>(lat, P, r, fill);
    }
};

template <int TT, int TU, int TV, bool DENSE, int THREADS, int PAD, bool COUNT = false>
__global__ __launch_bounds__(THREADS) void lin_time_unit2(uint16_t* __restrict__ glat, LinLaunch p) {
    extern __shared__ uint16_t tile[];
    tile_unit<LinMoves, TT, TU, TV, DENSE, THREADS, PAD, COUNT>(glat, tile, unit_args(p),
                                                                xcd_tile(blockIdx.x, gridDim.x));
}

// One time unit of every replica of a batch (one workgroup per tile).
template <int TT, int TU, int TV, bool DENSE, int THREADS, int PAD, bool COUNT = false>
__global__ __launch_bounds__(THREADS) void lin_batch_unit(uint16_t* __restrict__ gbase, BatchLaunch b) {
    extern __shared__ uint16_t tile[];
    batch_unit<LinMoves, TT, TU, TV, DENSE, THREADS, PAD, COUNT>(gbase, tile, b);
}

// variant 1: lin_time_unit (cell-major LDS); 2: lin_time_unit2 compact;
// 3: lin_time_unit2 with P-indexed LDS for 16^3 tiles (default).
template <class K>
static int launch_lds(K kernel, size_t lds, int nblocks, int threads, hipStream_t stream, uint16_t* glat,
                      const LinLaunch& p) {
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return (int)e;
    }
    hipLaunchKernelGGL(kernel, dim3(nblocks), dim3(threads), lds, stream, glat, p);
    return (int)hipGetLastError();
}

#define ELP_LIN_CASE(TT, TU, TV, DENSE, TH)                                                                   \
    if (tt == TT && tu == TU && tv == TV && dense == DENSE && threads == TH) {                                 \
        if (probe) return 0;                                                                                   \
        if (p.counts) {                                                                                        \
            constexpr int PADC = (TT == 16 && TU == 16 && TV == 16) ? 1 : 0;                                   \
            return launch_lds(lin_time_unit2<TT, TU, TV, DENSE, TH, PADC, true>,                               \
                              sizeof(uint16_t) * PTile<TT, TU, TV, PADC>::NSITE + LinMoves::EXTRA, nblocks, TH, \
                              stream, glat, p);                                                                \
        }                                                                                                      \
        if (variant == 1)                                                                                      \
            return launch_lds(lin_time_unit<TT, TU, TV, DENSE, TH>, sizeof(uint16_t) * TT * TU * TV, nblocks, TH, \
                              stream, glat, p);                                                                \
        if constexpr (TT == 16 && TU == 16 && TV == 16)                                                        \
            if (variant == 3)                                                                                  \
                return launch_lds(lin_time_unit2<TT, TU, TV, DENSE, TH, 1>,                                    \
                                  sizeof(uint16_t) * PTile<TT, TU, TV, 1>::NSITE + LinMoves::EXTRA, nblocks, TH, \
                                  stream, glat, p);                                                            \
        return launch_lds(lin_time_unit2<TT, TU, TV, DENSE, TH, 0>,                                            \
                          sizeof(uint16_t) * TT * TU * TV + LinMoves::EXTRA, nblocks, TH, stream, glat, p);    \
    }

int lin_launch(int tt, int tu, int tv, bool dense, int threads, bool probe, int nblocks, hipStream_t stream,
               uint16_t* glat, const LinLaunch& p) {
    static const int variant = [] {
        const char* env = getenv("ELP_LIN_VARIANT");
        return env ? atoi(env) : 3;
    }();
    ELP_LIN_CASE(16, 16, 16, false, 64)
    ELP_LIN_CASE(16, 16, 16, true, 64)
    ELP_LIN_CASE(16, 16, 16, true, 128)
    ELP_LIN_CASE(16, 16, 16, true, 256)
    ELP_LIN_CASE(32, 32, 32, false, 256)
    ELP_LIN_CASE(32, 32, 32, false, 512)
    ELP_LIN_CASE(32, 32, 32, true, 512)
    ELP_LIN_CASE(32, 32, 32, true, 1024)
    ELP_LIN_CASE(32, 32, 16, false, 256)
    ELP_LIN_CASE(32, 32, 16, true, 512)
    ELP_LIN_CASE(16, 16, 32, false, 128)
    ELP_LIN_CASE(16, 16, 32, true, 512)
    return -1;
}

// Batched launches use the second formulation (P-indexed LDS for 16^3 tiles).
#define ELP_LIN_BATCH_CASE(TT, TU, TV, DENSE, TH)                                                     \
    if (tt == TT && tu == TU && tv == TV && dense == DENSE && threads == TH) {                         \
        if (probe) return 0;                                                                           \
        constexpr int PAD = (TT == 16 && TU == 16 && TV == 16) ? 1 : 0;                                \
        constexpr size_t lds = sizeof(uint16_t) * PTile<TT, TU, TV, PAD>::NSITE + LinMoves::EXTRA;     \
        auto kernel = b.counts ? lin_batch_unit<TT, TU, TV, DENSE, TH, PAD, true>                      \
                               : lin_batch_unit<TT, TU, TV, DENSE, TH, PAD>;                           \
        if (lds > 65536) {                                                                             \
            hipError_t e = hipFuncSetAttribute((const void*)kernel,                                    \
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);  \
            if (e != hipSuccess) return (int)e;                                                        \
        }                                                                                              \
        hipLaunchKernelGGL(kernel, dim3(nblocks), dim3(TH), lds, stream, gbase, b);                    \
        return (int)hipGetLastError();                                                                 \
    }

int lin_batch_launch(int tt, int tu, int tv, bool dense, int threads, bool probe, int nblocks, hipStream_t stream,
                     uint16_t* gbase, const BatchLaunch& b) {
    ELP_LIN_BATCH_CASE(16, 16, 16, false, 64)
    ELP_LIN_BATCH_CASE(16, 16, 16, true, 64)
    ELP_LIN_BATCH_CASE(16, 16, 16, true, 128)
    ELP_LIN_BATCH_CASE(16, 16, 16, true, 256)
    ELP_LIN_BATCH_CASE(32, 32, 32, false, 256)
    ELP_LIN_BATCH_CASE(32, 32, 32, true, 512)
    return -1;
}

}  // namespace elp
