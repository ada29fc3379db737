// elp_chains.hip -- ring recovery and tuv text on the GPU.
//
// Replaces the host walk of GetAllRingPolymers / FindPolymerStart /
// GetRingPolymer / UpdatePolymerWithLabels (gpupol2/oclpol.c:183-376) and the
// per-character writer WriteLatticeFile (gpupol2/gpupol_io.c:167-203) for the
// ring engine, with the same result byte for byte:
//   * monomers are visited in chain order: on a site holding stored length
//     the stored-length monomer (label bit 5) comes first, then the monomer
//     carrying the site's bond (label bit 4), then the next site;
//   * a ring starts on the monomer after its label head 1,0,0,0,0; its id is
//     label2id[the 18 label bits that follow, most significant first];
//   * its bond string lists, from the start, 0 for a stored-length monomer
//     and the bond code otherwise; the tuv line is
//     "len= N\n" "t  u  v\t" hex string "\n", rings in id order.
// One thread per site finds heads (L^3 threads, five-monomer look-ahead),
// one thread per ring walks it (and sums the moments of its monomer
// positions: Rg^2 and the stored-length count, as ComputeGyration and
// ComputeSL of analysis/mode_compute.c:145-203), one thread per ring formats
// its line.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "elp_device.h"
#include "elp_chains.h"

namespace elp {

struct Cursor {
    uint32_t site;
    uint32_t sl;      // 1: the stored-length monomer of the site, 0: its bond carrier
};

struct Box {
    int LT, LU, LV;
    __device__ uint32_t step(uint32_t x, uint32_t b) const {
        int t = (int)(x % (uint32_t)LT), u = (int)((x / (uint32_t)LT) % (uint32_t)LU), v = (int)(x / ((uint32_t)LT * LU));
        t += bdt(b); u += bdu(b); v += bdv(b);
        t += (t < 0) ? LT : 0; t -= (t >= LT) ? LT : 0;
        u += (u < 0) ? LU : 0; u -= (u >= LU) ? LU : 0;
        v += (v < 0) ? LV : 0; v -= (v >= LV) ? LV : 0;
        return (uint32_t)t + (uint32_t)LT * ((uint32_t)u + (uint32_t)LU * (uint32_t)v);
    }
};

__device__ __forceinline__ uint32_t label_of(const uint8_t* lat, Cursor c) {
    const uint32_t s = lat[c.site];
    return c.sl ? (s >> 5) & 1u : (s >> 4) & 1u;
}

__device__ __forceinline__ Cursor next_of(const uint8_t* lat, const Box& B, Cursor c) {
    if (c.sl) return Cursor{c.site, 0u};
    const uint32_t n = B.step(c.site, lat[c.site] & 0xfu);
    return Cursor{n, (lat[n] >> 6) & 1u};
}

// Starts of all rings: the monomer after each head 1,0,0,0,0.
__global__ void ring_heads_kernel(const uint8_t* __restrict__ lat, Box B, uint32_t nsite, uint32_t* __restrict__ starts,
                                  uint32_t cap, uint32_t* __restrict__ count) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nsite) return;
    const uint32_t s = lat[x];
    if (!(s & 0xfu)) return;
    for (uint32_t k = 0; k < 2; k++) {
        Cursor c{x, k == 0 ? 1u : 0u};
        if (c.sl && !(s & 0x40u)) continue;
        if (!label_of(lat, c)) continue;
        bool head = true;
        for (int i = 0; i < 4 && head; i++) {
            c = next_of(lat, B, c);
            head = label_of(lat, c) == 0u;
        }
        if (!head) continue;
        c = next_of(lat, B, c);
        const uint32_t i = atomicAdd(count, 1u);
        if (i < cap) starts[i] = (c.site << 1) | c.sl;
    }
}

// Walk every ring from its start: id from the 18 label bits, bond string,
// length, start site and the length of its tuv line.
__global__ void ring_walk_kernel(const uint8_t* __restrict__ lat, Box B, const uint32_t* __restrict__ starts,
                                 uint32_t n_start, const int32_t* __restrict__ label2id, int n_pol, int max_len,
                                 uint8_t* __restrict__ bonds, int32_t* __restrict__ lens, int32_t* __restrict__ start3,
                                 int32_t* __restrict__ line_len, double* __restrict__ rg2, int32_t* __restrict__ nsl,
                                 uint32_t* __restrict__ err) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_start) return;
    const Cursor c0{starts[r] >> 1, starts[r] & 1u};
    Cursor c = c0;
    uint32_t lab = 0;
    for (int i = 0; i < 18; i++) {
        lab = (lab << 1) | label_of(lat, c);
        c = next_of(lat, B, c);
    }
    const int id = label2id[lab];
    if (id < 0 || id >= n_pol) { atomicOr(err, 1u); return; }
    c = c0;
    int n = 0, zeros = 0;
    // unwrapped (t,u,v) along the ring; moments of x = t+u-v, y = t-u, z = v
    int pt = 0, pu = 0, pv = 0;
    double sx = 0, sy = 0, sz = 0, sq = 0;
    do {
        if (n >= max_len) { atomicOr(err, 2u); return; }
        const uint32_t b = c.sl ? 0u : (lat[c.site] & 0xfu);
        bonds[(size_t)id * max_len + n] = (uint8_t)b;
        const double x = pt + pu - pv, y = pt - pu, z = pv;
        sx += x; sy += y; sz += z; sq += x * x + y * y + z * z;
        if (b) { pt += bdt(b); pu += bdu(b); pv += bdv(b); } else zeros++;
        n++;
        c = next_of(lat, B, c);
    } while (!(c.site == c0.site && c.sl == c0.sl));
    lens[id] = n;
    // squared radius of gyration in bond units (|bond|^2 = 2 in x, y, z)
    rg2[id] = 0.5 * (sq / n - (sx * sx + sy * sy + sz * sz) / ((double)n * n));
    nsl[id] = zeros;
    const int t = (int)(c0.site % (uint32_t)B.LT), u = (int)((c0.site / (uint32_t)B.LT) % (uint32_t)B.LU),
              v = (int)(c0.site / ((uint32_t)B.LT * B.LU));
    start3[3 * id] = t; start3[3 * id + 1] = u; start3[3 * id + 2] = v;
    auto digits = [](int a) { int d = 1; while (a >= 10) { a /= 10; d++; } return d; };
    // "len= " N "\n" t "  " u "  " v "\t" hex*N "\n"
    line_len[id] = 5 + digits(n) + 1 + digits(t) + 2 + digits(u) + 2 + digits(v) + 1 + n + 1;
}

__device__ char* put_int(char* p, int a) {
    char tmp[12];
    int k = 0;
    do { tmp[k++] = (char)('0' + a % 10); a /= 10; } while (a);
    while (k) *p++ = tmp[--k];
    return p;
}

__global__ void ring_text_kernel(const uint8_t* __restrict__ bonds, const int32_t* __restrict__ lens,
                                 const int32_t* __restrict__ start3, const uint64_t* __restrict__ offset, int n_pol,
                                 int max_len, char* __restrict__ text) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n_pol) return;
    char* p = text + offset[id];
    const int n = lens[id];
    *p++ = 'l'; *p++ = 'e'; *p++ = 'n'; *p++ = '='; *p++ = ' ';
    p = put_int(p, n);
    *p++ = '\n';
    p = put_int(p, start3[3 * id]);
    *p++ = ' '; *p++ = ' ';
    p = put_int(p, start3[3 * id + 1]);
    *p++ = ' '; *p++ = ' ';
    p = put_int(p, start3[3 * id + 2]);
    *p++ = '\t';
    const uint8_t* b = bonds + (size_t)id * max_len;
    for (int i = 0; i < n; i++) *p++ = "0123456789abcdef"[b[i] & 0xf];
    *p = '\n';
}

// ---- linear chains (uint16 site word: bits 0-3 next, 4-7 prev, 8 stored
// length, 9 carrier label, 10 stored-length label) ----
// Restates GetAllRingPolymers / FindPolymerStart / GetLinPolymer /
// UpdatePolymerWithLabels of gpupol/oclpol.c (host: lin_host.c) for the
// device: a chain starts on the site whose monomer has no prev bond (its
// label head 1,0,0,0,0 sits on monomers 0-4); walking forward, each site
// gives its stored-length monomer first (bond 0, label bit 10), then its
// carrier (the site's next bond, label bit 9; 0xf at the chain end); the id
// is the label bits of monomers 5-22, most significant first.

// Starts of all chains: occupied sites without a prev bond.
__global__ void lin_heads_kernel(const uint16_t* __restrict__ lat, uint32_t nsite, uint32_t* __restrict__ starts,
                                 uint32_t cap, uint32_t* __restrict__ count) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nsite) return;
    const uint32_t w = lat[x];
    if (!(w & 0xffu) || ((w >> 4) & 0xfu)) return;
    const uint32_t k = atomicAdd(count, 1u);
    if (k < cap) starts[k] = x;
}

__global__ void lin_walk_kernel(const uint16_t* __restrict__ lat, Box B, const uint32_t* __restrict__ starts,
                                uint32_t n_start, const int32_t* __restrict__ label2id, int n_pol, int max_len,
                                uint8_t* __restrict__ bonds, int32_t* __restrict__ lens, int32_t* __restrict__ start3,
                                int32_t* __restrict__ line_len, double* __restrict__ rg2, int32_t* __restrict__ nsl,
                                uint32_t* __restrict__ err) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_start) return;
    const uint32_t x0 = starts[r];
    // pass 1: the id from the label bits of monomers 5..22
    uint32_t x = x0, lab = 0;
    int n = 0;
    for (int guard = 0; n < 23 && guard < max_len; guard++) {
        const uint32_t w = lat[x];
        if (w & 0x100u) { if (n >= 5) lab = (lab << 1) | ((w >> 10) & 1u); n++; }
        if (n < 23) { if (n >= 5) lab = (lab << 1) | ((w >> 9) & 1u); n++; }
        const uint32_t nx = w & 0xfu;
        if (!nx) break;
        x = B.step(x, nx);
    }
    if (n < 23) { atomicOr(err, 4u); return; }   // shorter than head + id label
    const int id = label2id[lab];
    if (id < 0 || id >= n_pol) { atomicOr(err, 1u); return; }
    // pass 2: bond string and position moments (x = t+u-v, y = t-u, z = v)
    x = x0;
    n = 0;
    int pt = 0, pu = 0, pv = 0, zeros = 0;
    double sx = 0, sy = 0, sz = 0, sq = 0;
    uint8_t* out = bonds + (size_t)id * max_len;
    for (;;) {
        const uint32_t w = lat[x];
        const double cx = pt + pu - pv, cy = pt - pu, cz = pv;
        if (w & 0x100u) {
            if (n >= max_len) { atomicOr(err, 2u); return; }
            out[n++] = 0;
            zeros++;
            sx += cx; sy += cy; sz += cz; sq += cx * cx + cy * cy + cz * cz;
        }
        if (n >= max_len) { atomicOr(err, 2u); return; }
        const uint32_t nx = w & 0xfu;
        out[n++] = (uint8_t)(nx ? nx : 0xfu);
        sx += cx; sy += cy; sz += cz; sq += cx * cx + cy * cy + cz * cz;
        if (!nx) break;
        pt += bdt(nx); pu += bdu(nx); pv += bdv(nx);
        x = B.step(x, nx);
    }
    lens[id] = n;
    rg2[id] = 0.5 * (sq / n - (sx * sx + sy * sy + sz * sz) / ((double)n * n));
    nsl[id] = zeros;
    const int t = (int)(x0 % (uint32_t)B.LT), u = (int)((x0 / (uint32_t)B.LT) % (uint32_t)B.LU),
              v = (int)(x0 / ((uint32_t)B.LT * B.LU));
    start3[3 * id] = t; start3[3 * id + 1] = u; start3[3 * id + 2] = v;
    auto digits = [](int a) { int d = 1; while (a >= 10) { a /= 10; d++; } return d; };
    line_len[id] = 5 + digits(n) + 1 + digits(t) + 2 + digits(u) + 2 + digits(v) + 1 + n + 1;
}

// (dt,du,dv) in {-1,0,1}^3 -> bit (dt+1) + 3(du+1) + 9(dv+1): set for the
// same site and the 12 FCC neighbours (the displacements of the bond codes).
constexpr uint32_t contact_mask() {
    uint32_t m = 1u << 13;
    for (uint32_t b = 1; b < 15; b++) {
        if (b == 3u || b == 12u) continue;
        const int w = (int)(b >> 3);
        const int dt = (int)(b & 1u) - w, du = (int)((b >> 1) & 1u) - w, dv = (int)((b >> 2) & 1u) - w;
        m |= 1u << ((dt + 1) + 3 * (du + 1) + 9 * (dv + 1));
    }
    return m;
}
constexpr uint32_t kContactMask = contact_mask();
constexpr int kSumThreads = 256;

// One workgroup per chain (in id order): unwrapped (t,u,v) of every monomer
// relative to the chain start (a chunked prefix sum of the bond string in
// LDS), their sum, the middle monomer (index polSize/2: n/2 for a ring,
// (n-1)/2 for a linear chain, whose last bond 0xf adds no step) and the
// contact histogram: pairs i < j on one site or on nearest-neighbour sites in
// the chain's own unwrapped frame, counted at j - i
// (analysis/mode_compute.c:429-504).
// Dynamic LDS: 3 int64 sums, 3 x 256 chunk partials, then x, y, z, hist of
// n_max words each.
__global__ __launch_bounds__(kSumThreads) void chain_summary_kernel(const uint8_t* __restrict__ bonds,
                                                                   const int32_t* __restrict__ lens, int max_len,
                                                                   int n_max, int linear, int64_t* __restrict__ sum3,
                                                                   int32_t* __restrict__ mid3,
                                                                   unsigned long long* __restrict__ contacts) {
    extern __shared__ unsigned long long smem[];
    unsigned long long* ssum = smem;                                   // [4]
    int32_t* part = reinterpret_cast<int32_t*>(smem + 4);             // [3 * kSumThreads]
    int32_t* px = part + 3 * kSumThreads;                              // [n_max] each
    int32_t* py = px + n_max;
    int32_t* pz = py + n_max;
    uint32_t* hist = reinterpret_cast<uint32_t*>(pz + n_max);
    const int id = blockIdx.x, tid = threadIdx.x;
    const int n = lens[id];                                            // 1 <= n <= n_max
    const uint8_t* b = bonds + (size_t)id * max_len;
    const int chunk = (n + kSumThreads - 1) / kSumThreads;
    const int lo = tid * chunk < n ? tid * chunk : n;
    const int hi = lo + chunk < n ? lo + chunk : n;

    for (int i = tid; i < n; i += kSumThreads) hist[i] = 0u;
    if (tid < 3) ssum[tid] = 0ull;
    int st = 0, su = 0, sv = 0;
    for (int k = lo; k < hi; k++) { const uint32_t c = b[k] & 0xfu; if (c) { st += bdt(c); su += bdu(c); sv += bdv(c); } }
    part[3 * tid] = st; part[3 * tid + 1] = su; part[3 * tid + 2] = sv;
    __syncthreads();
    if (tid == 0) {                                                    // exclusive scan of the chunk sums
        int at = 0, au = 0, av = 0;
        for (int k = 0; k < kSumThreads; k++) {
            const int xt = part[3 * k], xu = part[3 * k + 1], xv = part[3 * k + 2];
            part[3 * k] = at; part[3 * k + 1] = au; part[3 * k + 2] = av;
            at += xt; au += xu; av += xv;
        }
    }
    __syncthreads();
    int t = part[3 * tid], u = part[3 * tid + 1], v = part[3 * tid + 2];
    long long mt = 0, mu = 0, mv = 0;
    for (int i = lo; i < hi; i++) {
        px[i] = t; py[i] = u; pz[i] = v;
        mt += t; mu += u; mv += v;
        const uint32_t c = b[i] & 0xfu;
        if (c) { t += bdt(c); u += bdu(c); v += bdv(c); }
    }
    if (hi > lo) {
        atomicAdd(&ssum[0], (unsigned long long)mt);
        atomicAdd(&ssum[1], (unsigned long long)mu);
        atomicAdd(&ssum[2], (unsigned long long)mv);
    }
    __syncthreads();
    for (int i = tid; i < n; i += kSumThreads) {
        const int ti = px[i] - 1, ui = py[i] - 1, vi = pz[i] - 1;
        for (int j = i + 1; j < n; j++) {
            // d = displacement + 1 per axis; a touching pair has every d in {0,1,2}
            const uint32_t d0 = (uint32_t)(px[j] - ti), d1 = (uint32_t)(py[j] - ui), d2 = (uint32_t)(pz[j] - vi);
            if (max(d0, max(d1, d2)) > 2u) continue;
            if ((kContactMask >> (d0 + 3u * d1 + 9u * d2)) & 1u) atomicAdd(&hist[j - i], 1u);
        }
    }
    __syncthreads();
    for (int g = tid; g < n; g += kSumThreads)
        if (hist[g]) atomicAdd(&contacts[g], (unsigned long long)hist[g]);
    if (tid == 0) {
        sum3[3 * id] = (int64_t)ssum[0]; sum3[3 * id + 1] = (int64_t)ssum[1]; sum3[3 * id + 2] = (int64_t)ssum[2];
        const int m = linear ? (n - 1) / 2 : n / 2;            // monomer polSize / 2
        mid3[3 * id] = px[m]; mid3[3 * id + 1] = py[m]; mid3[3 * id + 2] = pz[m];
    }
}

int chains_launch_summary(const uint8_t* bonds, const int32_t* lens, int n_pol, int max_len, int n_max, int linear,
                          int64_t* sum3, int32_t* mid3, unsigned long long* contacts, hipStream_t st) {
    const size_t lds = summary_lds_bytes(n_max);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)chain_summary_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return (int)e;
    }
    hipLaunchKernelGGL(chain_summary_kernel, dim3(n_pol), dim3(kSumThreads), lds, st, bonds, lens, max_len, n_max, linear,
                       sum3, mid3, contacts);
    return (int)hipGetLastError();
}

__global__ void fill32_kernel(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

int chains_launch_fill32(uint32_t* p, uint32_t v, size_t n, hipStream_t st) {
    if (!n) return 0;
    size_t nb = (n + 255) / 256;
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(fill32_kernel, dim3((unsigned)nb), dim3(256), 0, st, p, v, n);
    return (int)hipGetLastError();
}

__global__ void check_lens_kernel(const int32_t* __restrict__ lens, int n_pol, uint32_t* __restrict__ stat) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pol) return;
    const int n = lens[i];
    if (n <= 0) atomicAdd(&stat[1], 1u);
    else atomicMax(&stat[0], (uint32_t)n);
}

int chains_launch_check_lens(const int32_t* lens, int n_pol, uint32_t* stat, hipStream_t st) {
    hipLaunchKernelGGL(check_lens_kernel, dim3((n_pol + 255) / 256), dim3(256), 0, st, lens, n_pol, stat);
    return (int)hipGetLastError();
}

int chains_launch_heads(const uint8_t* lat, int LT, int LU, int LV, uint32_t* starts, uint32_t cap, uint32_t* count,
                        hipStream_t st) {
    const uint32_t n = (uint32_t)LT * LU * LV;
    hipLaunchKernelGGL(ring_heads_kernel, dim3((n + 255) / 256), dim3(256), 0, st, lat, Box{LT, LU, LV}, n, starts, cap,
                       count);
    return (int)hipGetLastError();
}

int chains_launch_walk(const uint8_t* lat, int LT, int LU, int LV, const uint32_t* starts, uint32_t n_start,
                       const int32_t* label2id, int n_pol, int max_len, uint8_t* bonds, int32_t* lens, int32_t* start3,
                       int32_t* line_len, double* rg2, int32_t* nsl, uint32_t* err, hipStream_t st) {
    hipLaunchKernelGGL(ring_walk_kernel, dim3((n_start + 63) / 64), dim3(64), 0, st, lat, Box{LT, LU, LV}, starts,
                       n_start, label2id, n_pol, max_len, bonds, lens, start3, line_len, rg2, nsl, err);
    return (int)hipGetLastError();
}

int chains_launch_lin_heads(const uint16_t* lat, int LT, int LU, int LV, uint32_t* starts, uint32_t cap,
                            uint32_t* count, hipStream_t st) {
    const uint32_t n = (uint32_t)LT * LU * LV;
    hipLaunchKernelGGL(lin_heads_kernel, dim3((n + 255) / 256), dim3(256), 0, st, lat, n, starts, cap, count);
    return (int)hipGetLastError();
}

int chains_launch_lin_walk(const uint16_t* lat, int LT, int LU, int LV, const uint32_t* starts, uint32_t n_start,
                           const int32_t* label2id, int n_pol, int max_len, uint8_t* bonds, int32_t* lens,
                           int32_t* start3, int32_t* line_len, double* rg2, int32_t* nsl, uint32_t* err,
                           hipStream_t st) {
    hipLaunchKernelGGL(lin_walk_kernel, dim3((n_start + 63) / 64), dim3(64), 0, st, lat, Box{LT, LU, LV}, starts,
                       n_start, label2id, n_pol, max_len, bonds, lens, start3, line_len, rg2, nsl, err);
    return (int)hipGetLastError();
}

int chains_launch_text(const uint8_t* bonds, const int32_t* lens, const int32_t* start3, const uint64_t* offset,
                       int n_pol, int max_len, char* text, hipStream_t st) {
    hipLaunchKernelGGL(ring_text_kernel, dim3((n_pol + 63) / 64), dim3(64), 0, st, bonds, lens, start3, offset, n_pol,
                       max_len, text);
    return (int)hipGetLastError();
}

}  // namespace elp
