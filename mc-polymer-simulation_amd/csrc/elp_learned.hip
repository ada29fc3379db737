// elp_learned.hip -- learned move proposal for ring equilibration (no
// reference counterpart; it biases which move the ring engine tries and is
// corrected by Metropolis-Hastings so that the uniform measure -- the
// equilibrium of the balanced ring rates -- stays stationary).
//
// Network (14416 parameters, fp32 master copy; every product runs on
// gfx950 MFMA with bf16 operands and fp32 accumulation):
//   x  = 144 features of a site, int8: 125 occupancy bits of the 5^3 patch
//        around the site, one-hot of its bond code (16), its stored-length
//        bit, the stored-length bit of its bond partner, and a constant 1
//        (the bias of layer 1);
//   h1 = relu(x W1)            W1: 144 x 64
//   h2 = relu(h1 W2 + b2)      W2: 64 x 64
//   z  = h2 W3 + b3            W3: 64 x 16, logits of the 11 moves of a site
//   q  = softmax(z[0..10])
// Moves of the site P (all change only P and sites one step away):
//   m 0..3   TransForw through common neighbour k = m (k bit 0: decomposition,
//            bit 1: order, as the r bits of the plain move); the monomer that
//            moves is the stored-length one of P if P holds stored length,
//            else that of the bond partner
//   m 4..7   TransForw k = m-4 moving the bond partner's stored length while
//            P holds stored length too (both ends)
//   m 8      TransBack folding the next monomer into P (P gains stored length)
//   m 9      TransBack folding it into the partner after it
//   m 10     DiffuseSL across the bond
// Each (state, move) gives a distinct new state and its reverse is a move
// of the same site, so q(m | S) and q(rev | S') give the exact MH ratio.
// Concurrent sites use the cubic sub-lattice (pitch 4): a patch (radius 2)
// never meets another site's moves (radius 1), so the sites of a substep
// are independent.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>
#include "elp_device.h"
#include "elp_learned.h"

namespace elp {

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);          // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ---------------------------------------------------------------- patch moves
// Patch: 5 x 5 x 5 site bytes, index t + 5u + 25v, the site at the centre.
// 0xFF marks a site outside the tile: a move touching it is invalid.
constexpr int PC = 62;
constexpr uint8_t WALL = 0xFF;

__device__ __forceinline__ int pstep(int i, uint32_t b) { return i + bdt(b) + 5 * bdu(b) + 25 * bdv(b); }

// Apply move m at the centre of the patch q.  Returns 1 and the reverse move
// index if the move is possible (q then holds the new state; ch[0..2] = the
// patch indices written), else 0 and q is unchanged.
__device__ int lp_move(uint8_t* q, int m, int* rev, int* ch) {
    const uint32_t s = q[PC];
    if (!s || s == WALL) return 0;
    const uint32_t next = s & 0xfu, label = s & 0x30u, sl = s & 0x40u;
    const int in = pstep(PC, next);
    uint32_t ns = q[in];
    if (ns == WALL) return 0;
    const uint32_t nsl = ns & 0x40u;
    if (m < 8) {                                       // TransForw
        const uint32_t k = (uint32_t)m & 3u;
        if (!(sl | nsl)) return 0;
        bool moveFirst;
        if (m < 4) moveFirst = sl != 0;                // P's stored length if any, else the partner's
        else { if (!(sl && nsl)) return 0; moveFirst = false; }
        uint32_t nb1 = trans_first(next, k & 1u);
        uint32_t nb2 = bond_add((~nb1) & 0xfu, next);
        if (!(k & 2u)) { const uint32_t x = nb1; nb1 = nb2; nb2 = x; }
        const int id = pstep(PC, nb1);
        if (q[id]) return 0;                           // occupied or outside
        uint32_t d = nb2, p = s;
        if (moveFirst) {
            p = nb1 | ((label >> 1) & 0x10u);
            d |= label & 0x10u;
        } else {
            p = nb1 | label | sl;
            d |= (ns & 0x20u) >> 1;
            ns &= 0x1fu;
        }
        q[PC] = (uint8_t)p;
        q[id] = (uint8_t)d;
        q[in] = (uint8_t)ns;
        *rev = moveFirst ? 8 : 9;
        ch[0] = PC; ch[1] = id; ch[2] = in;
        return 1;
    }
    if (m < 10) {                                      // TransBack
        if (nsl) return 0;                             // the folded monomer must be alone
        const uint32_t srcNext = ns & 0xfu;
        if ((next | srcNext) != 0xfu && (next & srcNext)) return 0;
        const uint32_t newNext = bond_add(next, srcNext);
        if (newNext == 0u || newNext == 0x3u || newNext == 0xcu) return 0;
        const int iw = pstep(PC, newNext);
        uint32_t ws = q[iw];
        if (ws == WALL) return 0;
        const uint32_t wsl = ws & 0x40u;
        const bool intoP = m == 8;
        if (intoP ? sl != 0 : wsl != 0) return 0;
        const uint32_t srcLabel = ns & 0x30u;
        uint32_t p;
        if (intoP) {
            p = newNext | (label << 1) | srcLabel | 0x40u;
        } else {
            p = newNext | label | sl;
            ws = (ws & 0x3fu) | (srcLabel << 1) | 0x40u;
        }
        // reverse: TransForw through the vacated site; k: its position
        // among the four common neighbours of P and the new partner
        int k = 0;
        for (; k < 4; k++) {
            uint32_t nb1 = trans_first(newNext, (uint32_t)k & 1u);
            const uint32_t nb2 = bond_add((~nb1) & 0xfu, newNext);
            if (!(k & 2)) nb1 = nb2;
            if (nb1 == next) break;
        }
        q[in] = 0;
        q[PC] = (uint8_t)p;
        q[iw] = (uint8_t)ws;
        // the partner's stored length moves back: forced if P has none
        *rev = intoP ? k : ((sl != 0) ? k + 4 : k);
        ch[0] = PC; ch[1] = in; ch[2] = iw;
        return 1;
    }
    if (m == 10) {                                     // DiffuseSL
        if ((nsl + sl) != 0x40u) return 0;
        uint32_t p;
        if (sl) {
            ns = ns | ((label & 0x10u) << 1) | 0x40u;
            p = next | ((label >> 1) & 0x10u);
        } else {
            p = next | (label << 1) | (((ns & 0x30u) >> 1) & 0x10u) | 0x40u;
            ns &= 0x1fu;
        }
        q[PC] = (uint8_t)p;
        q[in] = (uint8_t)ns;
        *rev = 10;
        ch[0] = PC; ch[1] = in; ch[2] = in;
        return 1;
    }
    return 0;
}

constexpr uint16_t BF_ONE = 0x3F80u;                  // bf16 1.0

// Features 125..143 of the centre site of a patch: bond one-hot, stored
// length of the site and of its bond partner, constant 1.
__device__ void lp_features_tail(const uint8_t* q, uint16_t* x) {
    const uint16_t one = BF_ONE;
    const uint32_t s = q[PC];
    const uint32_t b = (s == WALL) ? 0u : (s & 0xfu);
    for (int i = 0; i < 16; i++) x[125 + i] = (i == (int)b && s && s != WALL) ? one : 0;
    const int in = pstep(PC, b);
    x[141] = (s != WALL && (s & 0x40u)) ? one : 0;
    x[142] = (s && s != WALL && q[in] != WALL && (q[in] & 0x40u)) ? one : 0;
    x[143] = one;
}

// Features of the centre site of a patch (row of 144 bf16 values 0 / 1):
// occupancy of the 125 patch sites, then the tail above.
__device__ void lp_features(const uint8_t* q, uint16_t* x) {
    for (int i = 0; i < 125; i++) x[i] = q[i] ? BF_ONE : 0;
    lp_features_tail(q, x);
}

// ---------------------------------------------------------------- MFMA tiles
// One wave, 16 x 16 output tile: D(r, c) = sum_k A(r, k) B(k, c), K a
// multiple of 16; A(r, k) = a[r*ar + k*ak], B(k, c) = b[k*bk + c*bc] (bf16).
// v_mfma_f32_16x16x16_bf16: lane l holds A(l%16, 4(l/16)+j), B(4(l/16)+j, l%16),
// j < 4, and D(4(l/16)+i, l%16), i < 4.
__device__ __forceinline__ f32x4 mfma_tile(const uint16_t* a, int ar, int ak, const uint16_t* b, int bk, int bc, int K,
                                           f32x4 acc) {
    const int lane = threadIdx.x & 63;
    const int rc = lane & 15, kq = 4 * (lane >> 4);
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 av, bv;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            av[j] = (short)a[rc * ar + (k0 + kq + j) * ak];
            bv[j] = (short)b[(k0 + kq + j) * bk + rc * bc];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

// Shared-memory image of the network and one batch of 64 sites.
struct LpShared {
    uint16_t W1T[LP_NH][LP_NIN];     // [out][in]
    uint16_t W2T[LP_NH][LP_NH];
    uint16_t W3T[LP_NOUT][LP_NH];
    float b2[LP_NH], b3[LP_NOUT];
    uint16_t X[64][LP_NIN];
    uint16_t H1[64][LP_NH];
    uint16_t H2[64][LP_NH];
    float Z[64][LP_NOUT];
};

__device__ void lp_load_net(LpShared& S, const float* __restrict__ prm) {
    for (int i = threadIdx.x; i < LP_NIN * LP_NH; i += blockDim.x) {
        const int in = i / LP_NH, out = i % LP_NH;
        S.W1T[out][in] = to_bf16(prm[LP_OFF_W1 + i]);
    }
    for (int i = threadIdx.x; i < LP_NH * LP_NH; i += blockDim.x) {
        const int in = i / LP_NH, out = i % LP_NH;
        S.W2T[out][in] = to_bf16(prm[LP_OFF_W2 + i]);
    }
    for (int i = threadIdx.x; i < LP_NH * LP_NOUT; i += blockDim.x) {
        const int in = i / LP_NOUT, out = i % LP_NOUT;
        S.W3T[out][in] = to_bf16(prm[LP_OFF_W3 + i]);
    }
    for (int i = threadIdx.x; i < LP_NH; i += blockDim.x) S.b2[i] = prm[LP_OFF_B2 + i];
    for (int i = threadIdx.x; i < LP_NOUT; i += blockDim.x) S.b3[i] = prm[LP_OFF_B3 + i];
}

// Forward pass of the 64 rows of S.X: H1, H2 (bf16) and logits Z (fp32).
// Called by all threads of the workgroup (blockDim = 256, 4 waves).
__device__ void lp_forward(LpShared& S) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 15, r4 = 4 * (lane >> 4);
    for (int t = wave; t < 16; t += 4) {               // layer 1: 4 x 4 tiles
        const int rb = t >> 2, cb = t & 3;
        f32x4 acc = {0, 0, 0, 0};
        acc = mfma_tile(&S.X[rb * 16][0], LP_NIN, 1, &S.W1T[cb * 16][0], 1, LP_NIN, LP_NIN, acc);
#pragma unroll
        for (int i = 0; i < 4; i++) S.H1[rb * 16 + r4 + i][cb * 16 + col] = to_bf16(fmaxf(acc[i], 0.f));
    }
    __syncthreads();
    for (int t = wave; t < 16; t += 4) {               // layer 2
        const int rb = t >> 2, cb = t & 3;
        f32x4 acc = {0, 0, 0, 0};
        acc = mfma_tile(&S.H1[rb * 16][0], LP_NH, 1, &S.W2T[cb * 16][0], 1, LP_NH, LP_NH, acc);
#pragma unroll
        for (int i = 0; i < 4; i++)
            S.H2[rb * 16 + r4 + i][cb * 16 + col] = to_bf16(fmaxf(acc[i] + S.b2[cb * 16 + col], 0.f));
    }
    __syncthreads();
    {                                                  // layer 3: one tile per wave
        const int rb = wave;
        f32x4 acc = {0, 0, 0, 0};
        acc = mfma_tile(&S.H2[rb * 16][0], LP_NH, 1, &S.W3T[0][0], 1, LP_NH, LP_NH, acc);
#pragma unroll
        for (int i = 0; i < 4; i++) S.Z[rb * 16 + r4 + i][col] = acc[i] + S.b3[col];
    }
    __syncthreads();
}

// softmax over the 11 moves of row b (fp32)
__device__ void lp_softmax(const float* z, float* p) {
    float mx = z[0];
    for (int m = 1; m < LP_NMOVE; m++) mx = fmaxf(mx, z[m]);
    float sum = 0;
    for (int m = 0; m < LP_NMOVE; m++) { p[m] = __expf(z[m] - mx); sum += p[m]; }
    const float inv = 1.f / sum;
    for (int m = 0; m < LP_NMOVE; m++) p[m] *= inv;
}

// ------------------------------------------------ inference forward pass
// The inference kernels (lp_forward_kernel, lp_ring_time_unit) keep every
// weight operand in registers for the whole launch: wave w computes columns
// 16w..16w+15 of layers 1 and 2 for all 64 rows (the same B fragments for
// its four row blocks) and rows 16w..16w+15 of layer 3 (all 16 logits).
// Lane l holds B(4(l/16)+j, l%16), j < 4, of every k-step (17 bf16x4 + two
// biases: 36 VGPRs).  The A operands (activation rows) are read from LDS as
// one 8-byte read per lane and k-step; row strides of 148 and 68 bf16 (74 and
// 34 dwords) put the 16 rows of a tile on distinct even banks.
constexpr int LP_XS = LP_NIN + 4;
constexpr int LP_HS = LP_NH + 4;
struct LpAct {
    uint16_t X[64][LP_XS];
    uint16_t H1[64][LP_HS];
    uint16_t H2[64][LP_HS];
    float Z[64][LP_NOUT];
};
struct LpFrag {
    bf16x4 w1[LP_NIN / 16];
    bf16x4 w2[LP_NH / 16];
    bf16x4 w3[LP_NH / 16];
    float b2, b3;
};

__device__ __forceinline__ void lp_load_frag(LpFrag& F, const float* __restrict__ prm) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int c = lane & 15, kq = 4 * (lane >> 4);
#pragma unroll
    for (int kk = 0; kk < LP_NIN / 16; kk++)
#pragma unroll
        for (int j = 0; j < 4; j++) F.w1[kk][j] = (short)to_bf16(prm[LP_OFF_W1 + (16 * kk + kq + j) * LP_NH + 16 * w + c]);
#pragma unroll
    for (int kk = 0; kk < LP_NH / 16; kk++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            F.w2[kk][j] = (short)to_bf16(prm[LP_OFF_W2 + (16 * kk + kq + j) * LP_NH + 16 * w + c]);
            F.w3[kk][j] = (short)to_bf16(prm[LP_OFF_W3 + (16 * kk + kq + j) * LP_NOUT + c]);
        }
    F.b2 = prm[LP_OFF_B2 + 16 * w + c];
    F.b3 = prm[LP_OFF_B3 + c];
}

__device__ __forceinline__ bf16x4 lds_bf16x4(const uint16_t* p) { return *reinterpret_cast<const bf16x4*>(p); }

// Forward pass of the 64 rows of A.X into A.Z; all 4 waves of the workgroup
// (blockDim = 256) call it; ends with a barrier.
__device__ void lp_forward_reg(LpAct& A, const LpFrag& F) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int c = lane & 15, kq = 4 * (lane >> 4);
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {                   // layer 1 (bias through the constant feature)
        const uint16_t* a = &A.X[16 * rb + c][kq];
        f32x4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int kk = 0; kk < LP_NIN / 16; kk++)
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lds_bf16x4(a + 16 * kk), F.w1[kk], acc, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++) A.H1[16 * rb + kq + i][16 * w + c] = to_bf16(fmaxf(acc[i], 0.f));
    }
    __syncthreads();
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {                   // layer 2
        const uint16_t* a = &A.H1[16 * rb + c][kq];
        f32x4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int kk = 0; kk < LP_NH / 16; kk++)
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lds_bf16x4(a + 16 * kk), F.w2[kk], acc, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++) A.H2[16 * rb + kq + i][16 * w + c] = to_bf16(fmaxf(acc[i] + F.b2, 0.f));
    }
    __syncthreads();
    {                                                  // layer 3: row block w
        const uint16_t* a = &A.H2[16 * w + c][kq];
        f32x4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int kk = 0; kk < LP_NH / 16; kk++)
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lds_bf16x4(a + 16 * kk), F.w3[kk], acc, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++) A.Z[16 * w + kq + i][c] = acc[i] + F.b3;
    }
    __syncthreads();
}

// ---------------------------------------------------------------- kernels
// Probabilities of n feature rows (tests and diagnostics).
__global__ __launch_bounds__(256) void lp_forward_kernel(const float* __restrict__ prm, const int8_t* __restrict__ x,
                                                          int n, float* __restrict__ prob) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    LpAct& A = *reinterpret_cast<LpAct*>(lds_raw);
    LpFrag F;
    lp_load_frag(F, prm);
    for (int b0 = blockIdx.x * 64; b0 < n; b0 += gridDim.x * 64) {
        __syncthreads();
        for (int i = threadIdx.x; i < 64 * LP_NIN; i += blockDim.x) {
            const int r = i / LP_NIN, c = i % LP_NIN;
            A.X[r][c] = (b0 + r < n) ? to_bf16((float)x[(size_t)(b0 + r) * LP_NIN + c]) : 0;
        }
        __syncthreads();
        lp_forward_reg(A, F);
        if (threadIdx.x < 64 && b0 + threadIdx.x < n) {
            float p[LP_NMOVE];
            lp_softmax(A.Z[threadIdx.x], p);
            for (int m = 0; m < LP_NMOVE; m++) prob[(size_t)(b0 + threadIdx.x) * LP_NMOVE + m] = p[m];
        }
    }
}

// Gradient of scale * sum of cross-entropy losses over n samples, accumulated
// (atomically) into grad[LP_NPARAM]; loss sum into *loss.  Each wave keeps
// its weight-gradient tiles in registers over all batches of its workgroup.
struct LpTrainShared {
    LpShared F;
    uint16_t D3[64][LP_NOUT];
    uint16_t D2[64][LP_NH];
    uint16_t D1[64][LP_NH];
    float d3f[64][LP_NOUT];
    int y[64];
};

__global__ __launch_bounds__(256) void lp_grad_kernel(const float* __restrict__ prm, const int8_t* __restrict__ x,
                                                       const int32_t* __restrict__ y, int n, float scale,
                                                       float* __restrict__ grad, float* __restrict__ loss) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    LpTrainShared& T = *reinterpret_cast<LpTrainShared*>(lds_raw);
    LpShared& S = T.F;
    lp_load_net(S, prm);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 15, r4 = 4 * (lane >> 4);
    f32x4 gW1[9], gW2[4], gW3;
    for (int i = 0; i < 9; i++) gW1[i] = f32x4{0, 0, 0, 0};
    for (int i = 0; i < 4; i++) gW2[i] = f32x4{0, 0, 0, 0};
    gW3 = f32x4{0, 0, 0, 0};
    float gb2 = 0, gb3 = 0, lsum = 0;
    for (int b0 = blockIdx.x * 64; b0 < n; b0 += gridDim.x * 64) {
        __syncthreads();
        for (int i = threadIdx.x; i < 64 * LP_NIN; i += blockDim.x) {
            const int r = i / LP_NIN, c = i % LP_NIN;
            S.X[r][c] = (b0 + r < n) ? to_bf16((float)x[(size_t)(b0 + r) * LP_NIN + c]) : 0;
        }
        if (threadIdx.x < 64) T.y[threadIdx.x] = (b0 + threadIdx.x < n) ? y[b0 + threadIdx.x] : -1;
        __syncthreads();
        lp_forward(S);
        if (threadIdx.x < 64) {                        // dL/dz = scale * (softmax - onehot); padded rows 0
            const int b = threadIdx.x;
            float p[LP_NMOVE];
            lp_softmax(S.Z[b], p);
            const int yb = T.y[b];
            for (int m = 0; m < LP_NOUT; m++) {
                float d = 0;
                if (yb >= 0 && m < LP_NMOVE) d = scale * (p[m] - (m == yb ? 1.f : 0.f));
                T.d3f[b][m] = d;
                T.D3[b][m] = to_bf16(d);
            }
            if (yb >= 0) lsum += -__logf(fmaxf(p[yb], 1e-30f));
        }
        __syncthreads();
        // gW3 (64 x 16) += H2^T D3: row block = wave
        gW3 = mfma_tile(&S.H2[0][wave * 16], 1, LP_NH, &T.D3[0][0], LP_NOUT, 1, 64, gW3);
        // D2 = (D3 W3^T) * [H2 > 0]
        for (int t = wave; t < 16; t += 4) {
            const int rb = t >> 2, cb = t & 3;
            f32x4 acc = {0, 0, 0, 0};
            acc = mfma_tile(&T.D3[rb * 16][0], LP_NOUT, 1, &S.W3T[0][cb * 16], LP_NH, 1, LP_NOUT, acc);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int r = rb * 16 + r4 + i, c = cb * 16 + col;
                T.D2[r][c] = S.H2[r][c] ? to_bf16(acc[i]) : 0;
            }
        }
        if (threadIdx.x < LP_NOUT) {
            float s = 0;
            for (int b = 0; b < 64; b++) s += T.d3f[b][threadIdx.x];
            gb3 += s;
        }
        __syncthreads();
        // gW2 (64 x 64) += H1^T D2: tiles (rb, cb) with rb = wave
        for (int cb = 0; cb < 4; cb++)
            gW2[cb] = mfma_tile(&S.H1[0][wave * 16], 1, LP_NH, &T.D2[0][cb * 16], LP_NH, 1, 64, gW2[cb]);
        // D1 = (D2 W2^T) * [H1 > 0]
        for (int t = wave; t < 16; t += 4) {
            const int rb = t >> 2, cb = t & 3;
            f32x4 acc = {0, 0, 0, 0};
            acc = mfma_tile(&T.D2[rb * 16][0], LP_NH, 1, &S.W2T[0][cb * 16], LP_NH, 1, LP_NH, acc);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int r = rb * 16 + r4 + i, c = cb * 16 + col;
                T.D1[r][c] = S.H1[r][c] ? to_bf16(acc[i]) : 0;
            }
        }
        if (threadIdx.x < LP_NH) {
            float s = 0;
            for (int b = 0; b < 64; b++) s += from_bf16(T.D2[b][threadIdx.x]);
            gb2 += s;
        }
        __syncthreads();
        // gW1 (144 x 64) += X^T D1: 9 x 4 tiles, wave owns row blocks wave, wave+4, wave+8 (< 9)
        for (int j = 0; j < 9; j++) {
            const int t = wave + 4 * j;
            if (t >= 36) break;
            const int rb = t >> 2, cb = t & 3;
            gW1[j] = mfma_tile(&S.X[0][rb * 16], 1, LP_NIN, &T.D1[0][cb * 16], LP_NH, 1, 64, gW1[j]);
        }
    }
    // flush
    for (int j = 0; j < 9; j++) {
        const int t = wave + 4 * j;
        if (t >= 36) break;
        const int rb = t >> 2, cb = t & 3;
        for (int i = 0; i < 4; i++)
            atomicAdd(&grad[LP_OFF_W1 + (rb * 16 + r4 + i) * LP_NH + cb * 16 + col], gW1[j][i]);
    }
    for (int cb = 0; cb < 4; cb++)
        for (int i = 0; i < 4; i++)
            atomicAdd(&grad[LP_OFF_W2 + (wave * 16 + r4 + i) * LP_NH + cb * 16 + col], gW2[cb][i]);
    for (int i = 0; i < 4; i++) atomicAdd(&grad[LP_OFF_W3 + (wave * 16 + r4 + i) * LP_NOUT + col], gW3[i]);
    if (threadIdx.x < LP_NH) atomicAdd(&grad[LP_OFF_B2 + threadIdx.x], gb2);
    if (threadIdx.x < LP_NOUT) atomicAdd(&grad[LP_OFF_B3 + threadIdx.x], gb3);
    if (threadIdx.x < 64) atomicAdd(loss, lsum);
}

// Adam step on the fp32 master parameters; clears the gradient.
__global__ void lp_adam_kernel(float* __restrict__ prm, float* __restrict__ grad, float* __restrict__ m1,
                               float* __restrict__ m2, int n, float lr, float beta1, float beta2, float eps, float c1,
                               float c2) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float g = grad[i];
        const float a = beta1 * m1[i] + (1.f - beta1) * g;
        const float b = beta2 * m2[i] + (1.f - beta2) * g * g;
        m1[i] = a;
        m2[i] = b;
        prm[i] -= lr * (a * c1) / (sqrtf(b * c2) + eps);
        grad[i] = 0.f;
    }
}

__global__ void lp_fill_kernel(float* __restrict__ p, int n, float v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// Patch of site (t,u,v) of a periodic ring lattice (row-major, t fastest).
__device__ void gather_periodic(const uint8_t* __restrict__ lat, int LT, int LU, int LV, int t, int u, int v,
                                uint8_t* q) {
    for (int i = 0; i < 125; i++) {
        int a = t + i % 5 - 2, b = u + (i / 5) % 5 - 2, c = v + i / 25 - 2;
        a += (a < 0) ? LT : 0; a -= (a >= LT) ? LT : 0;
        b += (b < 0) ? LU : 0; b -= (b >= LU) ? LU : 0;
        c += (c < 0) ? LV : 0; c -= (c >= LV) ? LV : 0;
        q[i] = lat[(size_t)a + (size_t)LT * ((size_t)b + (size_t)LU * (size_t)c)];
    }
}

// Training samples from a ring lattice: per sample a uniformly drawn site
// with at least one possible move (up to 32 draws), and a move drawn from
// the probabilities with which the plain engine under balanced rates makes
// each move there (TransForw 1/2 over 4 neighbours x the stored-length
// choice, DiffuseSL 1/4, TransBack 1/4 with the forced fold at 1/4).
__global__ __launch_bounds__(64) void lp_harvest_kernel(const uint8_t* __restrict__ lat, int LT, int LU, int LV, int n,
                                                         uint32_t seed, uint32_t stream, int8_t* __restrict__ xo,
                                                         int32_t* __restrict__ yo) {
    __shared__ uint8_t qs[64][128];
    __shared__ uint8_t ws[64][128];
    __shared__ uint16_t xs[64][LP_NIN];
    const int tid = threadIdx.x;
    uint8_t* q = qs[tid];
    uint8_t* w = ws[tid];
    for (int i0 = blockIdx.x * 64; i0 < n; i0 += gridDim.x * 64) {
        const int i = i0 + tid;
        if (i >= n) continue;
        int label = -1;                                // no site with a possible move: skipped in training
        for (int tr = 0; tr < 32; tr++) {
            const uint4 r = philox4x32_10(make_uint4((uint32_t)i, (uint32_t)tr, 0xA5A5A5A5u, 0u), seed, stream);
            const uint64_t site = (((uint64_t)r.x << 32) | r.y) % ((uint64_t)LT * LU * LV);
            const int t = (int)(site % LT), u = (int)((site / LT) % LU), v = (int)(site / ((uint64_t)LT * LU));
            gather_periodic(lat, LT, LU, LV, t, u, v, q);
            float wt[LP_NMOVE], tot = 0;
            const uint32_t s = q[PC];
            bool bothSL = false, forcedTB = false;
            if (s && s != WALL) {
                const uint32_t nx = s & 0xfu, ns = q[pstep(PC, nx)];
                bothSL = (s & 0x40u) && ns != WALL && (ns & 0x40u);
                const uint32_t nn = bond_add(nx, ns & 0xfu);
                const uint32_t ws = q[pstep(PC, nn)];
                forcedTB = ((s & 0x40u) != 0) != (ws != WALL && (ws & 0x40u));
            }
            for (int m = 0; m < LP_NMOVE; m++) {
                for (int k = 0; k < 125; k++) w[k] = q[k];
                int rv, ch[3];
                wt[m] = 0;
                if (!lp_move(w, m, &rv, ch)) continue;
                if (m < 8) wt[m] = bothSL ? 1.f / 16 : 1.f / 8;
                else if (m < 10) wt[m] = forcedTB ? 1.f / 16 : 1.f / 8;
                else wt[m] = 0.25f;
                tot += wt[m];
            }
            if (tot > 0) {
                float u01 = (float)(r.z >> 8) * (1.f / 16777216.f) * tot;
                label = LP_NMOVE - 1;
                for (int m = 0; m < LP_NMOVE; m++) {
                    if (wt[m] > 0 && u01 < wt[m]) { label = m; break; }
                    u01 -= wt[m];
                }
                while (wt[label] == 0) label--;
                break;
            }
        }
        lp_features(q, xs[tid]);
        for (int c = 0; c < LP_NIN; c++) xo[(size_t)i * LP_NIN + c] = xs[tid][c] ? 1 : 0;
        yo[i] = label;
    }
}

// One move m at the centre of each of n patches (tests): out = patch after
// the move (unchanged if impossible), ok[i], rev[i].
__global__ void lp_patch_move_kernel(const uint8_t* __restrict__ in, const int32_t* __restrict__ mv, int n,
                                     uint8_t* __restrict__ out, int32_t* __restrict__ ok, int32_t* __restrict__ rev) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint8_t q[125];
    for (int k = 0; k < 125; k++) q[k] = in[(size_t)i * 125 + k];
    int r = -1, ch[3];
    ok[i] = lp_move(q, mv[i], &r, ch);
    rev[i] = r;
    for (int k = 0; k < 125; k++) out[(size_t)i * 125 + k] = q[k];
}

// Features of n patches (tests).
__global__ void lp_patch_features_kernel(const uint8_t* __restrict__ in, int n, int8_t* __restrict__ xo) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint8_t q[125];
    uint16_t x[LP_NIN];
    for (int k = 0; k < 125; k++) q[k] = in[(size_t)i * 125 + k];
    lp_features(q, x);
    for (int c = 0; c < LP_NIN; c++) xo[(size_t)i * LP_NIN + c] = x[c] ? 1 : 0;
}

// ---------------------------------------------------------------- learned MC
// One time unit of the learned-proposal engine on a 16^3 tile (cubic
// sub-lattice, 64 slots, 256 threads): n_iter iterations of 4 substeps; per
// substep every slot proposes m ~ q(. | S), evaluates q(. | S') after the
// move and accepts with min(1, q(rev | S') / q(m | S)).
// Randomness: Philox key (seed, stream); site offsets counter (tau, tile,
// 0xFFFFFFFF, iter/4) word iter%4 byte s (as the plain engine); per slot
// counter (tau, tile, slot | 0x40000000, iter) words 2s, 2s+1 = (u_move, u_accept)
// for s < 2 from the first call and s >= 2 from counter word 3 + 1.
struct LpMcShared {
    LpAct A;
    uint32_t poff[128];               // patch entry i -> (i%5) | (i/5%5) << 8 | (i/25) << 16
    uint8_t tile[4096];
    uint8_t qS[64][128];
    float pS[64][LP_NMOVE + 1];
    int mv[64], rv[64], ok[64];
    int ch[64][3];
};

__global__ __launch_bounds__(256) void lp_ring_time_unit(uint8_t* __restrict__ glat, RingLaunch p,
                                                          const float* __restrict__ prm, unsigned long long* __restrict__ stats) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    LpMcShared& M = *reinterpret_cast<LpMcShared*>(lds_raw);
    LpAct& S = M.A;
    constexpr int TT = 16;
    const int tid = threadIdx.x;
    const uint32_t tileId = xcd_tile(blockIdx.x, gridDim.x);
    const int kt = (int)(tileId % (uint32_t)p.ntile_t);
    const int ku = (int)((tileId / (uint32_t)p.ntile_t) % (uint32_t)p.ntile_u);
    const int kv = (int)(tileId / (uint32_t)(p.ntile_t * p.ntile_u));
    const int gt0 = p.ot + kt * TT, gu0 = p.ou + ku * TT, gv0 = p.ov + kv * TT;
    LpFrag F;
    lp_load_frag(F, prm);
    if (tid < 128) M.poff[tid] = (uint32_t)(tid % 5) | (uint32_t)(tid / 5 % 5) << 8 | (uint32_t)(tid / 25) << 16;
    for (int i = tid; i < 4096; i += 256) {
        int gt = gt0 + (i & 15), gu = gu0 + ((i >> 4) & 15), gv = gv0 + (i >> 8);
        gt -= (gt >= p.LT) ? p.LT : 0;
        gu -= (gu >= p.LU) ? p.LU : 0;
        gv -= (gv >= p.LV) ? p.LV : 0;
        M.tile[i] = glat[(size_t)gt + (size_t)p.LT * ((size_t)gu + (size_t)p.LU * (size_t)gv)];
    }
    unsigned long long nacc = 0, nprop = 0;
    __syncthreads();
    for (int it = 0; it < p.n_iter; it++) {
        const uint4 ow = philox4x32_10(make_uint4(p.tau, tileId, kOffsetStream, (uint32_t)(it >> 2)), p.seed, p.stream);
        const uint32_t offw = word_of(ow, it & 3);
        uint4 r0 = make_uint4(0, 0, 0, 0), r1 = r0;
        if (tid < 64) {
            r0 = philox4x32_10(make_uint4(p.tau, tileId, (uint32_t)tid | 0x40000000u, 2u * (uint32_t)it), p.seed, p.stream);
            r1 = philox4x32_10(make_uint4(p.tau, tileId, (uint32_t)tid | 0x40000000u, 2u * (uint32_t)it + 1u), p.seed,
                               p.stream);
        }
        for (int s = 0; s < 4; s++) {
            const uint32_t c = (offw >> (8 * s)) & 0x3fu;
            // patches of all 64 slots, by all threads: four threads per slot, entries
            // part, part+4, ... of its 5^3 patch (offsets from the LDS table)
            {
                const int sl = tid >> 2, part = tid & 3;
                const int a0 = 4 * (sl & 3) + (int)(c & 3u) - 2, b0 = 4 * ((sl >> 2) & 3) + (int)((c >> 2) & 3u) - 2,
                          d0 = 4 * (sl >> 4) + (int)((c >> 4) & 3u) - 2;
                for (int i = part; i < 125; i += 4) {
                    const uint32_t o = M.poff[i];
                    const int a = a0 + (int)(o & 0xffu), b = b0 + (int)((o >> 8) & 0xffu), d = d0 + (int)(o >> 16);
                    const uint8_t v = ((unsigned)a < 16u && (unsigned)b < 16u && (unsigned)d < 16u)
                                          ? M.tile[a + 16 * (b + 16 * d)] : WALL;
                    M.qS[sl][i] = v;
                    S.X[sl][i] = v ? BF_ONE : 0;
                }
            }
            __syncthreads();
            if (tid < 64) lp_features_tail(M.qS[tid], S.X[tid]);
            __syncthreads();
            lp_forward_reg(S, F);
            // propose m ~ q(. | S), apply it to the patch, features of S'
            if (tid < 64) {
                lp_softmax(S.Z[tid], M.pS[tid]);
                const uint32_t um = (s < 2) ? word_of(r0, 2 * s) : word_of(r1, 2 * (s - 2));
                // u in [0, sum of the probabilities as summed here): the loop
                // always stops on a move of nonzero probability
                float tot = 0;
                for (int k = 0; k < LP_NMOVE; k++) tot += M.pS[tid][k];
                float u = (float)(um >> 8) * (1.f / 16777216.f) * tot;
                int m = 0;
                for (int k = 0; k < LP_NMOVE; k++) {
                    if (M.pS[tid][k] > 0.f) m = k;
                    if (u < M.pS[tid][k]) break;
                    u -= M.pS[tid][k];
                }
                int rv = -1, ch[3] = {PC, PC, PC};
                const int ok = lp_move(M.qS[tid], m, &rv, ch);
                M.ok[tid] = ok;
                M.mv[tid] = m;
                M.rv[tid] = rv;
                for (int j = 0; j < 3; j++) {
                    M.ch[tid][j] = ch[j];
                    if (ok) S.X[tid][ch[j]] = M.qS[tid][ch[j]] ? BF_ONE : 0;
                }
                lp_features_tail(M.qS[tid], S.X[tid]);
            }
            __syncthreads();
            lp_forward_reg(S, F);
            // accept with min(1, q(rev | S') / q(m | S)); write the changed sites
            if (tid < 64) {
                nprop++;
                if (M.ok[tid]) {
                    float pT[LP_NMOVE];
                    lp_softmax(S.Z[tid], pT);
                    const uint32_t ua = (s < 2) ? word_of(r0, 2 * s + 1) : word_of(r1, 2 * (s - 2) + 1);
                    const float u = (float)(ua >> 8) * (1.f / 16777216.f);
                    if (u * M.pS[tid][M.mv[tid]] < pT[M.rv[tid]]) {
                        nacc++;
                        const int st = 4 * (tid & 3) + (int)(c & 3u), su = 4 * ((tid >> 2) & 3) + (int)((c >> 2) & 3u),
                                  sv = 4 * (tid >> 4) + (int)((c >> 4) & 3u);
                        for (int j = 0; j < 3; j++) {
                            const int i = M.ch[tid][j];
                            M.tile[(st + i % 5 - 2) + 16 * ((su + (i / 5) % 5 - 2) + 16 * (sv + i / 25 - 2))] = M.qS[tid][i];
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = tid; i < 4096; i += 256) {
        int gt = gt0 + (i & 15), gu = gu0 + ((i >> 4) & 15), gv = gv0 + (i >> 8);
        gt -= (gt >= p.LT) ? p.LT : 0;
        gu -= (gu >= p.LU) ? p.LU : 0;
        gv -= (gv >= p.LV) ? p.LV : 0;
        glat[(size_t)gt + (size_t)p.LT * ((size_t)gu + (size_t)p.LU * (size_t)gv)] = M.tile[i];
    }
    if (stats && tid < 64) {
        atomicAdd(&stats[0], nprop);
        atomicAdd(&stats[1], nacc);
    }
}

// ------------------------------------------- learned MC, wave-owned slots
// lp_ring_time_unit2: the chain of lp_ring_time_unit (same moves, proposal,
// acceptance rule and random streams) with every activation in registers.
// Wave w owns slots 16w..16w+15 of every substep.  Each layer is computed
// transposed, D = W^T X^T on v_mfma_f32_16x16x32_bf16 (gfx950): the weights
// are the A operand (rows = output units), the activations of the wave's 16
// slots the B operand (columns = slots).  Lane l = 16g + c holds
//   A(unit c of a 16-unit block, k-slot (g, kk, j)),  B(k-slot (g, kk, j), slot c),  j < 8,
//   D(unit 4g + i of the block, slot c),  i < 4,
// and the product sums over all k-slots, so which input unit a k-slot stands
// for is free as long as A and B agree.  Layer 1 gives k-slot (g, kk, j)
// feature j of patch row 4kk + g: rows of the five patch sites along t,
// padded to eight, then rows 25-27 with the bond one-hot, the two
// stored-length bits and the constant; a lane takes its row from one word of
// an occupancy bitmap of the tile and turns the bits into bf16 0/1 values
// with one read of a 256-entry LDS table.  Layers 2 and 3 give it hidden unit
// 16(2kk + j/4) + 4g + j%4, exactly the units the lane holds in D of the
// layer before, so a layer's output is the next layer's B operand without
// leaving the lane.  The weights are permuted to these maps at the start of
// a launch: layer 1's fragments stay in registers, those of layers 2 and 3
// in LDS (one 16-byte read per lane and product).
// The logits of a slot sit in its four lanes c + 16g (moves 4g..4g+3);
// softmax and sampling combine them with lane shuffles.  Lane c (g = 0)
// applies the proposed move to the tile at once and restores the old bytes
// when the MH test rejects it, so the second pass reads S' from the tile.
// A slot's writes (radius 1) never meet another slot's patch (radius 2,
// sites at pitch 4) and bitmap bits flip by LDS atomics, so the slots of a
// substep stay independent; DS operations of one wave complete in issue
// order, so hand-offs between its lanes need only a compiler barrier; one
// workgroup barrier per substep orders the tile writes of a substep before
// the patch reads of the next.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int LP_K1 = 7;           // layer-1 k-steps: 28 rows of 8 feature slots
constexpr int LP_OCCW = 20;        // bitmap rows per axis: u, v in [-2, 18)

struct LpMc2Shared {
    uint4 a2[4][2][64];            // layer-2 A fragments [unit block][k-step][lane]
    uint4 a3[2][64];               // layer-3 A fragments [k-step][lane]
    float b2[LP_NH], b3[LP_NOUT];
    // occupancy of tile row (u, v) at index (v + 2) * 20 + (u + 2): bit 8 + t
    // for t = 0..15, every other bit (and every row outside the tile) set:
    // sites outside the tile count as occupied, as in the patch features
    uint32_t occ[LP_OCCW * LP_OCCW];
    uint4 x8[256];                 // eight 0/1 features of a byte as a B fragment (lp_expand8)
    uint8_t tile[4096];
};

// input feature j of feature row r (rows of eight slots), or -1 for a zero slot
__device__ __forceinline__ int lp_feat_row(int r, int j) {
    if (r < 25) return j < 5 ? 5 * r + j : -1;     // patch site t = j of row r = u + 5v
    if (r < 27) return 125 + 8 * (r - 25) + j;     // bond one-hot
    return j < 3 ? 141 + j : -1;                   // stored length, partner's, constant
}
// input feature of layer-1 k-slot (g, kk, j): feature j of row 4kk + g
__device__ __forceinline__ int lp_feat1(int g, int kk, int j) { return lp_feat_row(4 * kk + g, j); }
// hidden unit of layer-2/3 k-slot (g, kk, j)
__device__ __forceinline__ int lp_hid(int g, int kk, int j) { return 16 * (2 * kk + (j >> 2)) + 4 * g + (j & 3); }

// Bitmap word i ((v + 2) * 20 + (u + 2)) of a 16^3 tile.
__device__ __forceinline__ uint32_t lp_occ_word(const uint8_t* tile, int i) {
    const int v = i / LP_OCCW - 2, u = i % LP_OCCW - 2;
    if ((unsigned)u >= 16u || (unsigned)v >= 16u) return 0xFFFFFFFFu;
    uint32_t o = 0;
    for (int t = 0; t < 16; t++) o |= (tile[t + 16 * (u + 16 * v)] != 0 ? 1u : 0u) << t;
    return 0xFF0000FFu | (o << 8);
}

// Bitmap offset of patch row r from the patch's row (u, v) = (-2, -2) (row
// 24 for the tail rows r > 24: read, then replaced by the tail features).
__device__ __forceinline__ int lp_row_offset_r(int r) {
    r = r < 25 ? r : 24;
    return LP_OCCW * (r / 5) + r % 5;
}
__device__ __forceinline__ int lp_row_offset(int g, int kk) { return lp_row_offset_r(4 * kk + g); }

// Eight features of a row as bits: a patch row from the bitmap (occ0: index of
// the slot's row (u, v) = (-2, -2), rowoff: the row's offset, tsh: bit of
// t = -2), or a tail row's bits.
__device__ __forceinline__ uint32_t lp_row_bits(const uint32_t* occ, int occ0, int rowoff, int tsh, uint32_t tail,
                                                bool isTail) {
    const uint32_t bits = (occ[occ0 + rowoff] >> tsh) & 0x1fu;
    return isTail ? tail : bits;
}

__device__ __forceinline__ uint32_t lp_pack2(float lo, float hi) {
    return (uint32_t)to_bf16(lo) | ((uint32_t)to_bf16(hi) << 16);
}

// Eight 0/1 features from the low byte of b: bf16 1.0 (0x3F80) where a bit is set.
__device__ __forceinline__ bf16x8 lp_expand8(uint32_t b) {
    uint4 d;
    d.x = (((b >> 0) & 1u) | (((b >> 1) & 1u) << 16)) * 0x3F80u;
    d.y = (((b >> 2) & 1u) | (((b >> 3) & 1u) << 16)) * 0x3F80u;
    d.z = (((b >> 4) & 1u) | (((b >> 5) & 1u) << 16)) * 0x3F80u;
    d.w = (((b >> 6) & 1u) | (((b >> 7) & 1u) << 16)) * 0x3F80u;
    return __builtin_bit_cast(bf16x8, d);
}

// relu of two D tiles (units 4g+i of blocks 2kk and 2kk+1) as a B operand
__device__ __forceinline__ bf16x8 lp_relu_pack(const f32x4& x, const f32x4& y) {
    bf16x8 r;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        r[i] = (__bf16)__builtin_amdgcn_fmed3f(x[i], 0.f, 3.0e38f);
        r[4 + i] = (__bf16)__builtin_amdgcn_fmed3f(y[i], 0.f, 3.0e38f);
    }
    return r;
}

__device__ __forceinline__ bf16x8 lds_frag(const uint4& v) { return __builtin_bit_cast(bf16x8, v); }

// max of two finite values in one instruction (median of a, b and +inf)
__device__ __forceinline__ float lp_max(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, __builtin_inff()); }

// LDS hand-off between lanes of one wave: its DS operations complete in
// issue order, so only the compiler must keep them in program order.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}

// Logits of moves 4g..4g+3 of this lane's slot.  occ0: bitmap index of the
// slot's patch row (u, v) = (-2, -2), rowoff[kk]: offset of patch row 4kk + g
// from it, tsh: bit of the patch's t = -2 in a bitmap word, tail: features of
// row 25 + (g - 1) (lanes g >= 1 at kk = 6).
__device__ __forceinline__ f32x4 lp_forward_t(const LpMc2Shared& M, const bf16x8 (&a1)[4][LP_K1],
                                              const int (&rowoff)[LP_K1], int occ0, int tsh, uint32_t tail, int g,
                                              int lane) {
    f32x4 d1[4];
#pragma unroll
    for (int rb = 0; rb < 4; rb++) d1[rb] = f32x4{0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < LP_K1; kk++) {
        const bf16x8 b = lds_frag(M.x8[lp_row_bits(M.occ, occ0, rowoff[kk], tsh, tail, kk == LP_K1 - 1 && g != 0)]);
#pragma unroll
        for (int rb = 0; rb < 4; rb++) d1[rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[rb][kk], b, d1[rb], 0, 0, 0);
    }
    bf16x8 h[2] = {lp_relu_pack(d1[0], d1[1]), lp_relu_pack(d1[2], d1[3])};
    f32x4 d2[4];
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
        d2[rb] = *reinterpret_cast<const f32x4*>(&M.b2[16 * rb + 4 * g]);
#pragma unroll
        for (int kk = 0; kk < 2; kk++)
            d2[rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lds_frag(M.a2[rb][kk][lane]), h[kk], d2[rb], 0, 0, 0);
    }
    h[0] = lp_relu_pack(d2[0], d2[1]);
    h[1] = lp_relu_pack(d2[2], d2[3]);
    f32x4 z = *reinterpret_cast<const f32x4*>(&M.b3[4 * g]);
#pragma unroll
    for (int kk = 0; kk < 2; kk++) z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lds_frag(M.a3[kk][lane]), h[kk], z, 0, 0, 0);
    return z;
}

// Tile byte at (t, u, v), WALL outside the tile.
__device__ __forceinline__ uint32_t lp_tsite(const uint8_t* tile, int t, int u, int v) {
    return ((unsigned)t < 16u && (unsigned)u < 16u && (unsigned)v < 16u) ? (uint32_t)tile[t + 16 * (u + 16 * v)] : WALL;
}

// Features of rows 25-27 of the slot at (t, u, v) (the lane's row for g >= 1).
__device__ __forceinline__ uint32_t lp_tail_bits(const uint8_t* tile, int t, int u, int v, int g) {
    const uint32_t s = tile[t + 16 * (u + 16 * v)];
    const uint32_t next = s & 0xfu;
    const uint32_t ps = lp_tsite(tile, t + bdt(next), u + bdu(next), v + bdv(next));
    const uint32_t hot = s ? (1u << next) : 0u;
    const uint32_t r27 = ((s >> 6) & 1u) | ((s && ps != WALL && (ps & 0x40u)) ? 2u : 0u) | 4u;
    return g == 1 ? (hot & 0xffu) : g == 2 ? (hot >> 8) : r27;
}

// Move m of lp_move at tile site (t, u, v) of a 16^3 tile, in select form:
// the three move types are evaluated side by side and one guarded block
// writes the outcome to the tile.  Returns 1 if the move is possible, with
// the reverse move in *rev, the sites it wrote in chi[0..*nch-1] (centre,
// bond partner, third site) and their old bytes in cho; else 0 and the tile
// is unchanged.  *flip: the one site whose occupancy the move changes (the
// TransForw target fills, the TransBack source empties; -1 for DiffuseSL and
// impossible moves).  The same moves as lp_move on the patch around the site.
__device__ __forceinline__ int lp_move_tile(uint8_t* tile, int t, int u, int v, int m, int* rev, int* chi,
                                            uint32_t* cho, int* nch, int* flip) {
    const int pc = t + 16 * (u + 16 * v);
    const uint32_t s = tile[pc];
    const uint32_t next = s & 0xfu, label = s & 0x30u, sl = s & 0x40u;
    const int nt = t + bdt(next), nu = u + bdu(next), nv = v + bdv(next);
    const bool nIn = ((unsigned)nt < 16u) & ((unsigned)nu < 16u) & ((unsigned)nv < 16u);
    const int in = nIn ? nt + 16 * (nu + 16 * nv) : pc;
    const uint32_t ns = tile[in];                      // the partner (meaningful when nIn)
    const uint32_t nsl = ns & 0x40u, srcNext = ns & 0xfu, srcLabel = ns & 0x30u;
    const bool isTF = m < 8, isTB = (m == 8) | (m == 9), intoP = m == 8;
    // TransForw: common neighbour k = m & 3 of the site and its partner
    const uint32_t k = (uint32_t)m & 3u;
    const uint32_t f = trans_first(next, k & 1u), fc = bond_add((~f) & 0xfu, next);
    const uint32_t nb1 = (k & 2u) ? f : fc, nb2 = (k & 2u) ? fc : f;
    // TransBack: fold the partner into the site; newNext = next + srcNext
    const uint32_t newNext = bond_add(next, srcNext);
    const bool sumOk = !(((next | srcNext) != 0xfu) && (next & srcNext)) & (newNext != 0u) & (newNext != 0x3u) &
                       (newNext != 0xcu);
    // third site: the TransForw target or the TransBack fold site
    const uint32_t xb = isTF ? nb1 : newNext;
    const int xt = t + bdt(xb), xu = u + bdu(xb), xv = v + bdv(xb);
    const bool xIn = ((unsigned)xt < 16u) & ((unsigned)xu < 16u) & ((unsigned)xv < 16u);
    const int ix = xIn ? xt + 16 * (xu + 16 * xv) : pc;
    const uint32_t x = tile[ix];
    const bool base = (s != 0u) & nIn;
    // TransForw outcome
    const bool mf = (m < 4) & (sl != 0u);
    const bool okF = isTF & base & ((sl | nsl) != 0u) & ((m < 4) | ((sl != 0u) & (nsl != 0u))) & xIn & (x == 0u);
    const uint32_t pF = mf ? (nb1 | ((label >> 1) & 0x10u)) : (nb1 | label | sl);
    const uint32_t dF = nb2 | (mf ? (label & 0x10u) : ((ns & 0x20u) >> 1));
    const uint32_t nF = mf ? ns : (ns & 0x1fu);
    // TransBack outcome
    const bool okB = isTB & base & (nsl == 0u) & sumOk & xIn & (intoP ? (sl == 0u) : ((x & 0x40u) == 0u));
    const uint32_t pB = intoP ? (newNext | (label << 1) | srcLabel | 0x40u) : (newNext | label | sl);
    const uint32_t wB = intoP ? x : ((x & 0x3fu) | (srcLabel << 1) | 0x40u);
    // reverse of a fold: TransForw through the vacated site, k = the position of
    // next among the four common neighbours of the site and its new partner
    const uint32_t g0 = trans_first(newNext, 0u), g1 = trans_first(newNext, 1u);
    const uint32_t kb = (bond_add((~g0) & 0xfu, newNext) == next) ? 0u
                        : (bond_add((~g1) & 0xfu, newNext) == next) ? 1u
                        : (g0 == next) ? 2u : (g1 == next) ? 3u : 4u;
    const int rB = (int)(intoP ? kb : (sl != 0u) ? kb + 4u : kb);
    // DiffuseSL outcome
    const bool okD = (m == 10) & base & ((nsl + sl) == 0x40u);
    const uint32_t pD = sl ? (next | ((label >> 1) & 0x10u)) : (next | (label << 1) | ((srcLabel >> 1) & 0x10u) | 0x40u);
    const uint32_t nD = sl ? (ns | ((label & 0x10u) << 1) | 0x40u) : (ns & 0x1fu);
    const bool ok = okF | okB | okD;
    chi[0] = pc; cho[0] = s;
    chi[1] = in; cho[1] = ns;
    chi[2] = ix; cho[2] = x;
    *nch = okD ? 2 : 3;
    *rev = okF ? (mf ? 8 : 9) : okB ? rB : 10;
    *flip = okF ? ix : okB ? in : -1;
    if (ok) {
        tile[pc] = (uint8_t)(okF ? pF : okB ? pB : pD);
        tile[in] = (uint8_t)(okF ? nF : okB ? 0u : nD);
        if (!okD) tile[ix] = (uint8_t)(okF ? dF : wB);
    }
    return ok ? 1 : 0;
}

// Flip the bitmap bit of tile site i (none for i < 0).
__device__ __forceinline__ void lp_occ_flip(uint32_t* occ, int i) {
    if (i >= 0) atomicXor(&occ[((i >> 8) + 2) * LP_OCCW + ((i >> 4) & 15) + 2], 1u << (8 + (i & 15)));
}

__global__ __launch_bounds__(256, 2) void lp_ring_time_unit2(uint8_t* __restrict__ glat, RingLaunch p,
                                                              const float* __restrict__ prm,
                                                              unsigned long long* __restrict__ stats) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    LpMc2Shared& M = *reinterpret_cast<LpMc2Shared*>(lds_raw);
    constexpr int TT = 16;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, g = lane >> 4, c = lane & 15;
    const uint32_t tileId = xcd_tile(blockIdx.x, gridDim.x);
    const int kt = (int)(tileId % (uint32_t)p.ntile_t);
    const int ku = (int)((tileId / (uint32_t)p.ntile_t) % (uint32_t)p.ntile_u);
    const int kv = (int)(tileId / (uint32_t)(p.ntile_t * p.ntile_u));
    const int gt0 = p.ot + kt * TT, gu0 = p.ou + ku * TT, gv0 = p.ov + kv * TT;

    // layer-1 A fragments: unit 16rb + c, k-slots (g, kk, j)
    bf16x8 a1[4][LP_K1];
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int kk = 0; kk < LP_K1; kk++) {
            uint4 d;
            uint32_t* dw = reinterpret_cast<uint32_t*>(&d);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int f0 = lp_feat1(g, kk, 2 * q), f1 = lp_feat1(g, kk, 2 * q + 1);
                dw[q] = lp_pack2(f0 >= 0 ? prm[LP_OFF_W1 + f0 * LP_NH + 16 * rb + c] : 0.f,
                                 f1 >= 0 ? prm[LP_OFF_W1 + f1 * LP_NH + 16 * rb + c] : 0.f);
            }
            a1[rb][kk] = __builtin_bit_cast(bf16x8, d);
        }
    // layers 2 and 3: fragments of lane l = 16gg + rr in LDS
    for (int e = tid; e < 4 * 2 * 64; e += 256) {
        const int rb = e >> 7, kk = (e >> 6) & 1, l = e & 63, gg = l >> 4, rr = l & 15;
        uint4 d;
        uint32_t* dw = reinterpret_cast<uint32_t*>(&d);
#pragma unroll
        for (int q = 0; q < 4; q++)
            dw[q] = lp_pack2(prm[LP_OFF_W2 + lp_hid(gg, kk, 2 * q) * LP_NH + 16 * rb + rr],
                             prm[LP_OFF_W2 + lp_hid(gg, kk, 2 * q + 1) * LP_NH + 16 * rb + rr]);
        M.a2[rb][kk][l] = d;
    }
    if (tid < 128) {
        const int kk = tid >> 6, l = tid & 63, gg = l >> 4, rr = l & 15;
        uint4 d;
        uint32_t* dw = reinterpret_cast<uint32_t*>(&d);
#pragma unroll
        for (int q = 0; q < 4; q++)
            dw[q] = lp_pack2(prm[LP_OFF_W3 + lp_hid(gg, kk, 2 * q) * LP_NOUT + rr],
                             prm[LP_OFF_W3 + lp_hid(gg, kk, 2 * q + 1) * LP_NOUT + rr]);
        M.a3[kk][l] = d;
    }
    if (tid < LP_NH) M.b2[tid] = prm[LP_OFF_B2 + tid];
    if (tid < LP_NOUT) M.b3[tid] = prm[LP_OFF_B3 + tid];
    M.x8[tid] = __builtin_bit_cast(uint4, lp_expand8((uint32_t)tid));
    for (int i = tid; i < 4096; i += 256) {
        int gt = gt0 + (i & 15), gu = gu0 + ((i >> 4) & 15), gv = gv0 + (i >> 8);
        gt -= (gt >= p.LT) ? p.LT : 0;
        gu -= (gu >= p.LU) ? p.LU : 0;
        gv -= (gv >= p.LV) ? p.LV : 0;
        M.tile[i] = glat[(size_t)gt + (size_t)p.LT * ((size_t)gu + (size_t)p.LU * (size_t)gv)];
    }
    __syncthreads();
    for (int i = tid; i < LP_OCCW * LP_OCCW; i += 256) M.occ[i] = lp_occ_word(M.tile, i);
    int rowoff[LP_K1];
#pragma unroll
    for (int kk = 0; kk < LP_K1; kk++) rowoff[kk] = lp_row_offset(g, kk);
    const int slot = 16 * w + c;
    uint32_t nacc = 0, nprop = 0;
    __syncthreads();
    for (int it = 0; it < p.n_iter; it++) {
        const uint4 ow = philox4x32_10(make_uint4(p.tau, tileId, kOffsetStream, (uint32_t)(it >> 2)), p.seed, p.stream);
        const uint32_t offw = word_of(ow, it & 3);
        // the slot's two Philox blocks of this iteration: counter 2it in lane c
        // (g = 0, substeps 0-1), 2it + 1 in lane c + 16 (g = 1, substeps 2-3)
        const uint4 rw = philox4x32_10_lanes(make_uint4(p.tau, tileId, (uint32_t)slot | 0x40000000u,
                                                        2u * (uint32_t)it + (uint32_t)(g & 1)), p.seed, p.stream);
#pragma unroll 1
        for (int s = 0; s < 4; s++) {
            const uint32_t c6 = (offw >> (8 * s)) & 0x3fu;
            const int st = 4 * (slot & 3) + (int)(c6 & 3u), su = 4 * ((slot >> 2) & 3) + (int)((c6 >> 2) & 3u),
                      sv = 4 * (slot >> 4) + (int)((c6 >> 4) & 3u);
            const int occ0 = LP_OCCW * sv + su, tsh = st + 6;
            const int src = c + ((s >> 1) << 4);
            const uint32_t um = __shfl(word_of(rw, 2 * (s & 1)), src, 64);
            const uint32_t ua = __shfl(word_of(rw, 2 * (s & 1) + 1), src, 64);
            // q(. | S): softmax over the slot's four lanes (moves 4g + i < 11)
            f32x4 z = lp_forward_t(M, a1, rowoff, occ0, tsh, lp_tail_bits(M.tile, st, su, sv, g), g, lane);
            float q[4];
            int m;
            float pm;
            {
                float mx = -3.0e38f;
#pragma unroll
                for (int i = 0; i < 4; i++) mx = (4 * g + i < LP_NMOVE) ? lp_max(mx, z[i]) : mx;
                mx = lp_max(mx, __shfl_xor(mx, 16, 64));
                mx = lp_max(mx, __shfl_xor(mx, 32, 64));
                float e[4], sg = 0.f;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    e[i] = (4 * g + i < LP_NMOVE) ? __expf(z[i] - mx) : 0.f;
                    sg += e[i];
                }
                const float S0 = __shfl(sg, c, 64), S1 = __shfl(sg, c + 16, 64), S2 = __shfl(sg, c + 32, 64),
                            S3 = __shfl(sg, c + 48, 64);
                const float inv = __builtin_amdgcn_rcpf(((S0 + S1) + S2) + S3);
                float Q = 0.f;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    q[i] = e[i] * inv;
                    Q += q[i];
                }
                // m ~ q: the first move whose interval of the running sum holds u
                // (the last of nonzero probability if rounding leaves u past the end)
                const float Q0 = __shfl(Q, c, 64), Q1 = __shfl(Q, c + 16, 64), Q2 = __shfl(Q, c + 32, 64),
                            Q3 = __shfl(Q, c + 48, 64);
                const float E1 = Q0, E2 = E1 + Q1, E3 = E2 + Q2, E4 = E3 + Q3;
                const float u = (float)(um >> 8) * (1.f / 16777216.f) * E4;
                const float Eg = g == 0 ? 0.f : g == 1 ? E1 : g == 2 ? E2 : E3;
                const float Eg1 = g == 0 ? E1 : g == 1 ? E2 : g == 2 ? E3 : E4;
                const int last = Q3 > 0.f ? 3 : Q2 > 0.f ? 2 : Q1 > 0.f ? 1 : 0;
                const bool win = (u >= Eg && u < Eg1) || (u >= E4 && g == last);
                float ur = u - Eg;
                int mi = -1, lastnz = -1;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    lastnz = (q[i] > 0.f) ? i : lastnz;
                    const bool open = mi < 0;
                    mi = (open & (ur < q[i])) ? i : mi;
                    ur = (mi < 0) ? ur - q[i] : ur;
                }
                mi = (mi < 0) ? lastnz : mi;
                const float pl = mi == 0 ? q[0] : mi == 1 ? q[1] : mi == 2 ? q[2] : q[3];
                int cand = win ? 4 * g + mi : -1;
                cand = max(cand, __shfl_xor(cand, 16, 64));
                m = max(cand, __shfl_xor(cand, 32, 64));
                float pc = win ? pl : 0.f;
                pc += __shfl_xor(pc, 16, 64);
                pm = pc + __shfl_xor(pc, 32, 64);
            }
            // lane c applies m to the tile (S -> S')
            int ok = 0, rv = -1, nch = 0, flip = -1, chi[3] = {0, 0, 0};
            uint32_t cho[3] = {0, 0, 0};
            // (m < 0 only if no interval held u, which non-finite logits would cause: no move)
            if (g == 0 && m >= 0) {
                ok = lp_move_tile(M.tile, st, su, sv, m, &rv, chi, cho, &nch, &flip);
                lp_occ_flip(M.occ, flip);
            }
            wave_lds_sync();
            // q(rev | S'), accept with min(1, q(rev | S') / q(m | S)) or restore S
            z = lp_forward_t(M, a1, rowoff, occ0, tsh, lp_tail_bits(M.tile, st, su, sv, g), g, lane);
            {
                float mx = -3.0e38f;
#pragma unroll
                for (int i = 0; i < 4; i++) mx = (4 * g + i < LP_NMOVE) ? lp_max(mx, z[i]) : mx;
                mx = lp_max(mx, __shfl_xor(mx, 16, 64));
                mx = lp_max(mx, __shfl_xor(mx, 32, 64));
                float e[4], sg = 0.f;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    e[i] = (4 * g + i < LP_NMOVE) ? __expf(z[i] - mx) : 0.f;
                    sg += e[i];
                }
                const float S0 = __shfl(sg, c, 64), S1 = __shfl(sg, c + 16, 64), S2 = __shfl(sg, c + 32, 64),
                            S3 = __shfl(sg, c + 48, 64);
                const float inv = __builtin_amdgcn_rcpf(((S0 + S1) + S2) + S3);
                const int rvb = __shfl(rv, c, 64);
                float qc = 0.f;
#pragma unroll
                for (int i = 0; i < 4; i++) qc = (rvb == 4 * g + i) ? e[i] * inv : qc;
                qc += __shfl_xor(qc, 16, 64);
                const float qr = qc + __shfl_xor(qc, 32, 64);
                if (g == 0) {
                    nprop++;
                    if (ok) {
                        const float uf = (float)(ua >> 8) * (1.f / 16777216.f);
                        if (uf * pm < qr) {
                            nacc++;
                        } else {
                            // the sites are distinct, so the order of restoring is free
                            lp_occ_flip(M.occ, flip);
#pragma unroll
                            for (int j = 0; j < 3; j++)
                                if (j < nch) M.tile[chi[j]] = (uint8_t)cho[j];
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = tid; i < 4096; i += 256) {
        int gt = gt0 + (i & 15), gu = gu0 + ((i >> 4) & 15), gv = gv0 + (i >> 8);
        gt -= (gt >= p.LT) ? p.LT : 0;
        gu -= (gu >= p.LU) ? p.LU : 0;
        gv -= (gv >= p.LV) ? p.LV : 0;
        glat[(size_t)gt + (size_t)p.LT * ((size_t)gu + (size_t)p.LU * (size_t)gv)] = M.tile[i];
    }
    if (stats) {
        // one sum per wave, added by its first lane
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            nprop += __shfl_xor(nprop, d, 64);
            nacc += __shfl_xor(nacc, d, 64);
        }
        if (lane == 0) {
            atomicAdd(&stats[0], (unsigned long long)nprop);
            atomicAdd(&stats[1], (unsigned long long)nacc);
        }
    }
}

// Check of the learned engine's tile path against the patch functions (tests):
// per case (tile, t, u, v, m), one workgroup compares the features built from
// the bitmap with lp_features of the patch around the site, lp_move_tile with
// lp_move on that patch (outcome, reverse move, written bytes; nothing outside
// the patch changes), the bitmap after the flips with the bitmap of the new
// tile, and the tile and bitmap after the revert with the originals.
// res[2i] = failed checks (bits 0-6 in that order), res[2i+1] = outcome.
__global__ __launch_bounds__(64) void lp_tile_check_kernel(const uint8_t* __restrict__ tiles,
                                                            const int32_t* __restrict__ cases, int32_t* __restrict__ res) {
    __shared__ uint8_t tile[4096], orig[4096];
    __shared__ uint32_t occ[LP_OCCW * LP_OCCW];
    __shared__ uint8_t q[128];
    __shared__ uint16_t x[LP_NIN];
    const int ci = blockIdx.x, tid = threadIdx.x;
    const int32_t* cs = cases + 5 * ci;
    const uint8_t* src = tiles + (size_t)cs[0] * 4096;
    for (int i = tid; i < 4096; i += 64) tile[i] = orig[i] = src[i];
    __syncthreads();
    for (int i = tid; i < LP_OCCW * LP_OCCW; i += 64) occ[i] = lp_occ_word(tile, i);
    __syncthreads();
    if (tid != 0) return;
    const int t = cs[1], u = cs[2], v = cs[3], m = cs[4];
    for (int i = 0; i < 125; i++) q[i] = (uint8_t)lp_tsite(tile, t - 2 + i % 5, u - 2 + i / 5 % 5, v - 2 + i / 25);
    lp_features(q, x);
    int flags = 0;
    const int occ0 = LP_OCCW * v + u, tsh = t + 6;
    for (int g = 0; g < 4; g++) {
        const uint32_t tail = lp_tail_bits(tile, t, u, v, g);
        for (int kk = 0; kk < LP_K1; kk++) {
            const uint32_t bits = lp_row_bits(occ, occ0, lp_row_offset(g, kk), tsh, tail, kk == LP_K1 - 1 && g != 0);
            for (int j = 0; j < 8; j++) {
                const int f = lp_feat1(g, kk, j);
                const uint32_t want = f >= 0 ? (x[f] != 0 ? 1u : 0u) : 0u;
                if (((bits >> j) & 1u) != want) flags |= 1;
            }
        }
    }
    int rvp = -1, chp[3];
    const int okp = lp_move(q, m, &rvp, chp);
    int rvt = -1, chi[3], nch = 0, flip = -1;
    uint32_t cho[3];
    const int okt = lp_move_tile(tile, t, u, v, m, &rvt, chi, cho, &nch, &flip);
    if (okp != okt) flags |= 2;
    if (okp && okt && rvp != rvt) flags |= 4;
    for (int i = 0; i < 125; i++) {
        const uint32_t w = lp_tsite(tile, t - 2 + i % 5, u - 2 + i / 5 % 5, v - 2 + i / 25);
        if (w != q[i]) flags |= 8;
    }
    for (int i = 0; i < 4096; i++) {
        const int dt = (i & 15) - t, du = ((i >> 4) & 15) - u, dv = (i >> 8) - v;
        const bool inPatch = dt >= -2 && dt <= 2 && du >= -2 && du <= 2 && dv >= -2 && dv <= 2;
        if (!inPatch && tile[i] != orig[i]) flags |= 16;
    }
    lp_occ_flip(occ, flip);
    for (int i = 0; i < LP_OCCW * LP_OCCW; i++)
        if (occ[i] != lp_occ_word(tile, i)) flags |= 32;
    if (okt) {
        lp_occ_flip(occ, flip);
        for (int j = 0; j < nch; j++) tile[chi[j]] = (uint8_t)cho[j];
    }
    for (int i = 0; i < 4096; i++)
        if (tile[i] != orig[i]) flags |= 64;
    for (int i = 0; i < LP_OCCW * LP_OCCW; i++)
        if (occ[i] != lp_occ_word(orig, i)) flags |= 64;
    res[2 * ci] = flags;
    res[2 * ci + 1] = okt;
}

// ---------------------------------------------------------------- launchers
template <class K>
static int set_lds(K kernel, size_t lds) {
    return (int)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

int lp_launch_forward(const float* prm, const int8_t* x, int n, float* prob, hipStream_t st) {
    const size_t lds = sizeof(LpAct);
    int rc = set_lds(lp_forward_kernel, lds);
    if (rc) return rc;
    const int blocks = n / 64 + 1 < 1024 ? n / 64 + 1 : 1024;
    hipLaunchKernelGGL(lp_forward_kernel, dim3(blocks), dim3(256), lds, st, prm, x, n, prob);
    return (int)hipGetLastError();
}

int lp_launch_grad(const float* prm, const int8_t* x, const int32_t* y, int n, float scale, float* grad, float* loss,
                   hipStream_t st) {
    const size_t lds = sizeof(LpTrainShared);
    int rc = set_lds(lp_grad_kernel, lds);
    if (rc) return rc;
    const int blocks = n / 64 + 1 < 512 ? n / 64 + 1 : 512;
    hipLaunchKernelGGL(lp_grad_kernel, dim3(blocks), dim3(256), lds, st, prm, x, y, n, scale, grad, loss);
    return (int)hipGetLastError();
}

int lp_launch_adam(float* prm, float* grad, float* m1, float* m2, int n, float lr, float b1, float b2, float eps,
                   int step, hipStream_t st) {
    const float c1 = 1.f / (1.f - powf(b1, (float)step)), c2 = 1.f / (1.f - powf(b2, (float)step));
    hipLaunchKernelGGL(lp_adam_kernel, dim3((n + 255) / 256), dim3(256), 0, st, prm, grad, m1, m2, n, lr, b1, b2, eps,
                       c1, c2);
    return (int)hipGetLastError();
}

int lp_launch_fill(float* p, int n, float v, hipStream_t st) {
    hipLaunchKernelGGL(lp_fill_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
    return (int)hipGetLastError();
}

int lp_launch_harvest(const uint8_t* lat, int LT, int LU, int LV, int n, uint32_t seed, uint32_t stream, int8_t* x,
                      int32_t* y, hipStream_t st) {
    const int blocks = (n + 63) / 64 < 4096 ? (n + 63) / 64 : 4096;
    hipLaunchKernelGGL(lp_harvest_kernel, dim3(blocks), dim3(64), 0, st, lat, LT, LU, LV, n, seed, stream, x, y);
    return (int)hipGetLastError();
}

int lp_launch_patch_move(const uint8_t* in, const int32_t* mv, int n, uint8_t* out, int32_t* ok, int32_t* rev,
                         hipStream_t st) {
    hipLaunchKernelGGL(lp_patch_move_kernel, dim3((n + 63) / 64), dim3(64), 0, st, in, mv, n, out, ok, rev);
    return (int)hipGetLastError();
}

int lp_launch_tile_check(const uint8_t* tiles, const int32_t* cases, int n, int32_t* res, hipStream_t st) {
    hipLaunchKernelGGL(lp_tile_check_kernel, dim3(n), dim3(64), 0, st, tiles, cases, res);
    return (int)hipGetLastError();
}

int lp_launch_patch_features(const uint8_t* in, int n, int8_t* x, hipStream_t st) {
    hipLaunchKernelGGL(lp_patch_features_kernel, dim3((n + 63) / 64), dim3(64), 0, st, in, n, x);
    return (int)hipGetLastError();
}

int lp_launch_mc(uint8_t* glat, const RingLaunch& p, int nblocks, const float* prm, unsigned long long* stats,
                 hipStream_t st) {
    // default: wave-owned slots (ELP_LP_VARIANT=1: the workgroup-wide first version)
    static const int variant = [] {
        const char* env = getenv("ELP_LP_VARIANT");
        return env ? atoi(env) : 2;
    }();
    if (variant == 1) {
        const size_t lds = sizeof(LpMcShared);
        int rc = set_lds(lp_ring_time_unit, lds);
        if (rc) return rc;
        hipLaunchKernelGGL(lp_ring_time_unit, dim3(nblocks), dim3(256), lds, st, glat, p, prm, stats);
        return (int)hipGetLastError();
    }
    const size_t lds = sizeof(LpMc2Shared);
    int rc = set_lds(lp_ring_time_unit2, lds);
    if (rc) return rc;
    hipLaunchKernelGGL(lp_ring_time_unit2, dim3(nblocks), dim3(256), lds, st, glat, p, prm, stats);
    return (int)hipGetLastError();
}

}  // namespace elp
