// elp_autotune.cpp -- schedule autotuner of the batched engine (host code
// over the C ABI of libelp.so; no reference counterpart: the reference fixes
// 32^3 tiles, one cell per lane and 51 / 64 iterations at compile time,
// gpupol2/gpupol_init.c:76, gpupol/gpupol_init.c:79).
//
// Stage 1 times every compiled kernel configuration (tile shape, sub-lattice,
// threads per workgroup) whose tile divides every box, on scratch copies of
// the caller's lattices, by HIP events around the kernels, then runs it as
// long again with move counting.  A configuration is ranked by accepted
// moves per second (attempts per second times the accepted fraction), so
// attempts at empty sites and moves rejected at tile faces, which thinner
// tiles make more of, count for nothing.  Only configurations with the same
// statics compete: under the reference rates the stationary law depends on
// the sub-lattice (the move set is not reversible), so the caller fixes it;
// under balanced rates every sub-lattice samples the uniform measure.  The
// iterations per time unit follow the sub-lattice so that every candidate
// makes about the same attempts per site per time unit.  Stage 2 times the
// time units captured per HIP graph for the winner, by wall clock (host
// launch cost included).
// The tuned schedule is returned as a struct and as a one-line text record
// that elp_schedule_parse reads back, so a run can be replayed exactly.
#include <chrono>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>
#include "../../include/elp_gpu.h"

namespace {

struct Cand {
    int tt, tu, tv, sublattice, threads;
};

// every batched configuration compiled into libelp.so (elp_ring.hip / elp_linear.hip)
const Cand kCands[] = {
    {16, 16, 16, ELP_SUBLAT_CUBIC, 64},  {16, 16, 16, ELP_SUBLAT_DENSE, 64},
    {16, 16, 16, ELP_SUBLAT_DENSE, 128}, {16, 16, 16, ELP_SUBLAT_DENSE, 256},
    {16, 16, 8, ELP_SUBLAT_DENSE, 128},  {32, 32, 32, ELP_SUBLAT_CUBIC, 256}, {32, 32, 32, ELP_SUBLAT_DENSE, 512},
};

int default_iters(int kind, int sublattice) {
    if (kind == 0) return sublattice == ELP_SUBLAT_DENSE ? 13 : 51;
    return sublattice == ELP_SUBLAT_DENSE ? 16 : 64;
}

double attempts_per_unit(int kind, int sublattice, int n_iter, uint64_t sites) {
    const double per_cell = sublattice == ELP_SUBLAT_DENSE ? 4.0 : 1.0;
    return (double)n_iter * (kind == 0 ? 4 : 2) * per_cell * (double)(sites / 64);
}

struct Log {
    char* buf;
    uint64_t len, pos = 0;
    void add(const char* fmt, ...) __attribute__((format(printf, 2, 3)));
};

void Log::add(const char* fmt, ...) {
    if (!buf || len == 0 || pos + 1 >= len) return;
    va_list ap;
    va_start(ap, fmt);
    int n = vsnprintf(buf + pos, len - pos, fmt, ap);
    va_end(ap);
    if (n > 0) pos = (pos + (uint64_t)n < len) ? pos + (uint64_t)n : len - 1;
}

// scratch origins: any values in [0, tile) give the same kernel cost
void fill_origins(std::vector<int32_t>& o, int n_time, int n_rep, int tt, int tu, int tv) {
    o.resize((size_t)3 * n_time * n_rep);
    uint32_t x = 0x9E3779B9u;
    for (size_t i = 0; i < o.size(); i += 3) {
        x = x * 1664525u + 1013904223u; o[i] = (int32_t)((x >> 8) % (uint32_t)tt);
        x = x * 1664525u + 1013904223u; o[i + 1] = (int32_t)((x >> 8) % (uint32_t)tu);
        x = x * 1664525u + 1013904223u; o[i + 2] = (int32_t)((x >> 8) % (uint32_t)tv);
    }
}

struct BatchGuard {
    elp_batch* b = nullptr;
    ~BatchGuard() { elp_batch_destroy(b); }
};

}  // namespace

extern "C" {

int elp_schedule_format(const elp_schedule* s, char* buf, uint64_t len) {
    if (!s || !buf || len == 0) return ELP_E_ARG;
    int n = snprintf(buf, len,
                     "tile=%dx%dx%d sublattice=%s n_iter=%d threads=%d units_per_graph=%d attempts_per_s=%.6e "
                     "rules=%s accepted_per_s=%.6e",
                     s->tt, s->tu, s->tv, s->sublattice == ELP_SUBLAT_DENSE ? "dense" : "cubic", s->n_iter, s->threads,
                     s->units_per_graph, s->attempts_per_s, s->rules == ELP_RULES_BALANCED ? "balanced" : "reference",
                     s->accepted_per_s);
    return (n > 0 && (uint64_t)n < len) ? ELP_OK : ELP_E_ARG;
}

int elp_schedule_parse(const char* text, elp_schedule* s) {
    if (!text || !s) return ELP_E_ARG;
    char sub[16] = {0}, rules[16] = {0};
    elp_schedule r{};
    int used = 0;
    int n = sscanf(text, " tile=%dx%dx%d sublattice=%15s n_iter=%d threads=%d units_per_graph=%d%n", &r.tt, &r.tu,
                   &r.tv, sub, &r.n_iter, &r.threads, &r.units_per_graph, &used);
    if (n < 7) return ELP_E_ARG;
    if (!strcmp(sub, "dense")) r.sublattice = ELP_SUBLAT_DENSE;
    else if (!strcmp(sub, "cubic")) r.sublattice = ELP_SUBLAT_CUBIC;
    else return ELP_E_ARG;
    if (r.tt <= 0 || r.tu <= 0 || r.tv <= 0 || r.n_iter < 0 || r.threads <= 0 || r.units_per_graph < 0) return ELP_E_ARG;
    // optional fields, in order
    const char* rest = text + used;
    int k = 0;
    if (sscanf(rest, " attempts_per_s=%lf%n", &r.attempts_per_s, &k) == 1) rest += k;
    if (sscanf(rest, " rules=%15s%n", rules, &k) == 1) {
        rest += k;
        if (!strcmp(rules, "balanced")) r.rules = ELP_RULES_BALANCED;
        else if (strcmp(rules, "reference")) return ELP_E_ARG;
    }
    if (sscanf(rest, " accepted_per_s=%lf", &r.accepted_per_s) == 1) {}
    *s = r;
    return ELP_OK;
}

int elp_autotune(int device, int kind, int n_rep, const int32_t* boxes, const void* const* lattices, int time_units,
                 int rules, int sublattice, elp_schedule* best, char* log, uint64_t log_len) {
    if (!best || !boxes || !lattices || n_rep <= 0 || time_units <= 0 || (kind != 0 && kind != 1)) return ELP_E_ARG;
    if (rules != ELP_RULES_REFERENCE && rules != ELP_RULES_BALANCED) return ELP_E_ARG;
    if (sublattice != ELP_SUBLAT_CUBIC && sublattice != ELP_SUBLAT_DENSE && !(sublattice == -1 && rules == ELP_RULES_BALANCED))
        return ELP_E_ARG;   // reference rates: the sub-lattice fixes the statics
    if (kind == 1 && rules == ELP_RULES_BALANCED) return ELP_E_ARG;   // balanced rates exist for rings
    Log lg{log, log_len};
    if (log && log_len) log[0] = 0;
    const size_t bps = kind == 0 ? 1 : 2;
    uint64_t sites = 0;
    for (int r = 0; r < n_rep; r++) sites += (uint64_t)boxes[3 * r] * boxes[3 * r + 1] * boxes[3 * r + 2];
    std::vector<uint64_t> tau0(n_rep, 0);
    std::vector<uint32_t> keys(2 * (size_t)n_rep);
    for (int r = 0; r < n_rep; r++) { keys[2 * r] = 12345u + (uint32_t)r; keys[2 * r + 1] = (uint32_t)r; }
    std::vector<int32_t> org;
    elp_schedule win{};
    bool found = false;
    int last_err = ELP_E_ARG;

    auto make = [&](const Cand& c, int n_iter, BatchGuard& g) -> int {
        int rc = kind == 0 ? elp_ring_batch_create(&g.b, device, n_rep, boxes, c.tt, c.tu, c.tv, c.sublattice, n_iter)
                           : elp_lin_batch_create(&g.b, device, n_rep, boxes, c.tt, c.tu, c.tv, c.sublattice, n_iter);
        if (rc) return rc;
        if ((rc = elp_batch_set_threads(g.b, c.threads))) return rc;
        if (rules && (rc = elp_batch_set_rules(g.b, rules))) return rc;
        for (int r = 0; r < n_rep; r++) {
            const uint64_t nb = (uint64_t)boxes[3 * r] * boxes[3 * r + 1] * boxes[3 * r + 2] * bps;
            if ((rc = elp_batch_upload(g.b, r, lattices[r], nb))) return rc;
        }
        return ELP_OK;
    };

    // stage 1: kernel configuration
    double win_frac = 0;   // accepted fraction of the winner
    for (const Cand& c : kCands) {
        if (sublattice != -1 && c.sublattice != sublattice) continue;
        bool divides = true;
        for (int r = 0; r < n_rep && divides; r++)
            divides = boxes[3 * r] % c.tt == 0 && boxes[3 * r + 1] % c.tu == 0 && boxes[3 * r + 2] % c.tv == 0;
        if (!divides) continue;
        const int n_iter = default_iters(kind, c.sublattice);
        BatchGuard g;
        int rc = make(c, n_iter, g);
        if (rc == ELP_E_NODEV || rc == ELP_E_HIP) return rc;
        if (rc) { last_err = rc; continue; }     // not compiled for this kind
        fill_origins(org, 2, n_rep, c.tt, c.tu, c.tv);
        if ((rc = elp_batch_run(g.b, 2, org.data(), tau0.data(), keys.data(), 0)) || (rc = elp_batch_sync(g.b))) return rc;
        fill_origins(org, time_units, n_rep, c.tt, c.tu, c.tv);
        if ((rc = elp_batch_run(g.b, time_units, org.data(), tau0.data(), keys.data(), 0)) || (rc = elp_batch_sync(g.b)))
            return rc;
        float ms = 0;
        if ((rc = elp_batch_last_run_ms(g.b, &ms))) return rc;
        const double aps = attempts_per_unit(kind, c.sublattice, n_iter, sites) * time_units / (ms * 1e-3);
        // the same number of time units again, counting what the moves do
        uint64_t cnt[4] = {0, 0, 0, 0};
        if ((rc = elp_batch_count_moves(g.b, 1)) ||
            (rc = elp_batch_run(g.b, time_units, org.data(), tau0.data(), keys.data(), 0)) ||
            (rc = elp_batch_move_counts(g.b, cnt)) || (rc = elp_batch_count_moves(g.b, 0)))
            return rc;
        const double frac = cnt[0] ? (double)cnt[3] / (double)cnt[0] : 0.0;
        const double acc = aps * frac;
        lg.add("kernel tile=%dx%dx%d sublattice=%s threads=%d n_iter=%d: %.3f ms/unit %.4e attempts/s, accepted %.4f "
               "(empty %.4f, tile face %.4f): %.4e accepted/s\n",
               c.tt, c.tu, c.tv, c.sublattice == ELP_SUBLAT_DENSE ? "dense" : "cubic", c.threads, n_iter,
               ms / time_units, aps, frac, cnt[0] ? (double)cnt[1] / cnt[0] : 0.0,
               cnt[0] ? (double)cnt[2] / cnt[0] : 0.0, acc);
        if (!found || acc > win.accepted_per_s) {
            win = elp_schedule{c.tt, c.tu, c.tv, c.sublattice, n_iter, c.threads, 0, aps, rules, acc};
            win_frac = frac;
            found = true;
        }
    }
    if (!found) return last_err;

    // stage 2: time units per graph launch, by wall clock
    const Cand wc{win.tt, win.tu, win.tv, win.sublattice, win.threads};
    BatchGuard g;
    int rc = make(wc, win.n_iter, g);
    if (rc) return rc;
    double best_wall = 0;
    const int units = ((time_units + 63) / 64) * 64;
    fill_origins(org, units, n_rep, wc.tt, wc.tu, wc.tv);
    for (int G : {0, 4, 16, 64}) {
        // the first call with a new G also captures its graph: warm it up untimed
        if ((rc = elp_batch_run(g.b, G ? G : 1, org.data(), tau0.data(), keys.data(), G)) || (rc = elp_batch_sync(g.b)))
            return rc;
        const auto t0 = std::chrono::steady_clock::now();
        if ((rc = elp_batch_run(g.b, units, org.data(), tau0.data(), keys.data(), G)) || (rc = elp_batch_sync(g.b)))
            return rc;
        const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const double aps = attempts_per_unit(kind, win.sublattice, win.n_iter, sites) * units / wall;
        lg.add("units_per_graph=%d: %.3f ms/unit wall %.4e attempts/s %.4e accepted/s\n", G, wall * 1e3 / units, aps,
               aps * win_frac);
        if (best_wall == 0 || wall < best_wall) {
            best_wall = wall;
            win.units_per_graph = G;
            win.attempts_per_s = aps;
            win.accepted_per_s = aps * win_frac;
        }
    }
    *best = win;
    char rec[256];
    elp_schedule_format(&win, rec, sizeof rec);
    lg.add("chosen: %s\n", rec);
    return ELP_OK;
}

}  // extern "C"
