#!/usr/bin/env python3
"""Throughput of the MI355X ELP ring engine on one GPU.

A step is one time unit over the whole box: one launch of the move kernel
(every tile of the shifted tile grid: stage in LDS, n_iter iterations of
{TransForw, DiffuseSL, TransForw, TransBack}, write back).  The workload is
the dense ring melt the target is quoted on: L=128, N=1000, density 1.2
(2517 rings), built by the drop-in's own start-up (compact cube rings +
stored-length reshuffle) from seed 0 and relaxed by the warm-up steps.

Prints one JSON line.  value = move attempts/s (every slot of every substep
counts as one attempt, as in the reference kernel), by the wall clock around
exactly `--steps` time units with a synchronize of the engine's stream on
both sides (the only stream with work: the process does not load torch).
roofline.achieved = algorithmic HBM bytes per launch (the tile is read and
written once: 2 B/site x L^3) / average launch time from HIP events on the
engine's stream.  --rules balanced times the ring engine with balanced rates
(the benchmark line keeps the reference's).
cpu_baseline = the reference CPU engine efpol (cpupol/, built from its own
sources into oracle/_ref) timed on the host cores (the process's CPU affinity
mask), one independent process per core, on a bounded sample; the byte-exact
efpol restatement (oracle/bin/efpol_port_ring) when that build is absent.
accepted_moves_per_s and move_outcomes come from replaying the timed time
units with the counting kernels (elp_count_moves); "schedules" reports the
requested schedule and its companion (dense / the reference's cubic
sub-lattice on 16^3 tiles) side by side.
"""
import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "mc-polymer-simulation_amd")
sys.path.insert(0, PKG)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0      # MI355X dense bf16 MFMA


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline_reference(L, N, density, cores, warm, timed, ring=True):
    """The reference efpol itself (cpupol/, built from its sources into oracle/_ref by
    oracle/Makefile.ref): `cores` independent runs of warm + timed time units with a write
    interval of `timed`; each prints its DoMCStep rate per interval (cpupol/efpol.c:92-95),
    and the rate of its last interval (the timed one) is summed."""
    exe = os.path.join(ROOT, "oracle", "_ref", "efpol_ring" if ring else "efpol_lin")
    if not os.path.exists(exe) or warm % timed:
        return None
    work = tempfile.mkdtemp(prefix="efpol_ref_")
    try:
        procs = [subprocess.Popen([exe, str(2000 + 2 * i), os.path.join(work, str(i)), str(density), str(warm + timed),
                                   str(timed), str(N), "0", str(L), "sl_equal"],
                                  stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True) for i in range(cores)]
        rate = 0.0
        for p in procs:
            out, _ = p.communicate()
            rates = re.findall(r"at ([0-9.]+) M/s", out)
            if p.returncode != 0 or len(rates) != (warm + timed) // timed:
                return None
            rate += float(rates[-1]) * 1e6
    finally:
        shutil.rmtree(work, ignore_errors=True)
    return {
        "value": rate,
        "unit": "attempts/s",
        "cores": cores,
        "cpu_model": cpu_model(),
        # efpol's time unit is one attempt per monomer (cpupol/efpol.c:92-94)
        "monomer_time_units_per_s": rate,
        "kind": "reference",
        "sample": (f"reference efpol (cpupol, built from its sources), {'ring' if ring else 'linear'} L={L} N={N} "
                   f"density {density}, {cores} independent processes (seeds 2000, 2002, ...), {warm} untimed + "
                   f"{timed} timed time units from efpol's own start; its DoMCStep rate of the timed interval; "
                   f"1 attempt = 1 monomer move attempt"),
    }


def cpu_baseline(L, N, density, cores, warm, timed, ring=True):
    ref = cpu_baseline_reference(L, N, density, cores, warm, timed, ring)
    if ref:
        return ref
    exe = os.path.join(ROOT, "oracle", "bin", "efpol_port_ring" if ring else "efpol_port_lin")
    if not os.path.exists(exe):
        return None
    procs = [subprocess.Popen([exe, "bench", str(L), str(N), str(density), str(1000 + i), str(timed), str(warm)],
                              stdout=subprocess.PIPE, text=True) for i in range(cores)]
    rate = 0.0
    for p in procs:
        out, _ = p.communicate()
        kv = dict(x.split("=") for x in out.split())
        rate += float(kv["attempts"]) / float(kv["seconds"])
    return {
        "value": rate,
        "unit": "attempts/s",
        "cores": cores,
        "cpu_model": cpu_model(),
        "monomer_time_units_per_s": rate,
        "kind": "port",
        "sample": (f"efpol restatement, {'ring' if ring else 'linear'} L={L} N={N} density {density}, "
                   f"{cores} independent processes "
                   f"(seeds 1000..{1000 + cores - 1}), {warm} untimed + {timed} timed time units from efpol's "
                   f"two-site-loop start; DoMCStep region only; 1 attempt = 1 monomer move attempt"),
    }


def measured_traffic(kernel, chain, L, N, tile, sublattice):
    """HBM bytes per launch recorded from rocprofv3 FETCH_SIZE/WRITE_SIZE runs of this exact
    workload (profiles/hbm_traffic.txt), or None when the workload has not been profiled."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    key = [kernel, chain, str(L), str(N), "x".join(str(t) for t in tile), sublattice]
    with open(path) as f:
        for line in f:
            w = line.split()
            if len(w) == 7 and not line.startswith("#") and w[:6] == key:
                return int(w[6])
    return None


def bench_batch(args, tile):
    """64 independent ring melts (L = [32,48,64,96][r%4], N = [50,200,1000,2000][(r/4)%4],
    density 1.2, seed r): one launch per time unit over all 5040 tiles, replayed
    from captured HIP graphs.  value = aggregate move attempts/s."""
    import elp
    import numpy as np
    dense = args.sublattice == "dense"
    n_iter = args.n_iter if args.n_iter is not None else (13 if dense else 51)
    Ls, Ns = [32, 48, 64, 96], [50, 200, 1000, 2000]
    hosts = [elp.RingHost(N=Ns[(r // 4) % 4], L=Ls[r % 4], density=args.density, seed=r, box_unit=16)
             for r in range(64)]
    G = args.units_per_graph
    tuned = None
    if args.autotune:
        sched, log = elp.autotune("ring", [h.lattice() for h in hosts], [h.L for h in hosts], time_units=20)
        print(log, file=sys.stderr, end="")
        tuned = sched.record()
        tile, dense, n_iter, G = sched.tile, sched.sublattice == elp.SUBLAT_DENSE, sched.n_iter, sched.units_per_graph
        args.sublattice, args.threads = ("dense" if dense else "cubic"), sched.threads
    bat = elp.Batch("ring", [h.L for h in hosts], tile, elp.SUBLAT_DENSE if dense else elp.SUBLAT_CUBIC, n_iter)
    if args.threads:
        bat.set_threads(args.threads)
    for r, h in enumerate(hosts):
        bat.upload(r, h.lattice())
        h.load_shift(tile)
    seeds = [h.seed for h in hosts]
    streams = list(range(64))
    info = bat.launch_info()
    nsite = info["sites"]
    attempts_per_step = n_iter * 4 * (nsite // 64) * (4 if dense else 1)
    tau = 0
    if args.warmup:
        org = np.stack([h.origins(tile, args.warmup) for h in hosts], axis=1)
        bat.run(org, tau0=tau, seeds=seeds, streams=streams, units_per_graph=G)
        tau += args.warmup
    org = np.stack([h.origins(tile, args.steps) for h in hosts], axis=1)
    bat.sync()
    t0 = time.perf_counter()
    bat.run(org, tau0=tau, seeds=seeds, streams=streams, units_per_graph=G, sync=False)
    bat.sync()
    wall = time.perf_counter() - t0
    kernel_ms = bat.last_run_ms()
    for r, h in enumerate(hosts):          # every replica must still hold every labelled ring
        h.set_lattice(bat.download(r))
    value = attempts_per_step * args.steps / wall
    launch_s = kernel_ms * 1e-3 / args.steps
    bytes_per_launch = 2 * nsite
    achieved = bytes_per_launch / launch_s / 1e9
    rec = {
        "metric": "MC move attempts/s, %HBM roofline; equil. wall-time learned vs plain",
        "value": value,
        "unit": "attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": f"64 ring replicas L in {{32,48,64,96}}, N in {{50,200,1000,2000}}, density {args.density}, "
                        f"seeds 0..63 ({nsite} sites), one time unit of every replica per step",
            "tile": list(tile),
            "sublattice": args.sublattice,
            "iterations_per_time_unit": n_iter,
            "units_per_graph": G,
            "autotuned_schedule": tuned,
            "threads_per_block": info["threads"],
            "blocks_per_launch": info["blocks"],
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic("ring_batch_unit", "ring", "batch64", "mixed", tile, args.sublattice),
            "kernel": "ring_batch_unit",
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_us": launch_s * 1e6,
        },
        "cpu_baseline": None,
    }
    print(json.dumps(rec))
    return 0


def _rg2_now(host, eng):
    """Mean Rg^2 of the engine's rings from the device-side chain walk (no lattice download, no
    contact histogram)."""
    return float(np.mean(eng.ring_stats(host)[0]))


P_WINDOW = (500, 2000)          # contour separations averaged for the contact-probability probe


def _contact_now(host, eng):
    """Contact probability P(s) averaged over s in P_WINDOW, from one frame reduced on the device
    (elp_ring_frame_summary).  The compact start's rings meet far along their contour only at
    their turns, so it rises as they crumple and swell."""
    from elp.stats import Series
    ser = Series(True, host.L, host.n_pol, host.N, 1, summary=True)
    ser.add_summary(eng.ring_frame_summary(host))
    pc = ser.statics(0)["pc"]
    return float(np.mean(pc[P_WINDOW[0]:P_WINDOW[1] + 1]))


def bench_learned(args):
    """Equilibration A/B: ring melt L=128, N=4000, density 1.2 (629 rings) from
    the compact start, plain engine (balanced rates, 16^3 dense, 13
    iterations) vs the learned proposal (16^3 cubic, 51 iterations, network
    trained on 1e6 moves harvested from an L=96 N=1000 melt run by the plain
    engine), each for the same wall-clock budget; Rg^2 sampled along the way.
    value = wall-time ratio plain / learned to reach the Rg^2 both reach; the same ratio for the
    contact probability at large contour separations is reported beside it."""
    import elp
    budget = args.budget
    tile = (16, 16, 16)
    # training data: L=96, N=1000 melt run by the plain engine (balanced rates)
    h96 = elp.RingHost(N=1000, L=96, density=1.2, seed=0)
    e96 = elp.Engine("ring", h96.L, tile, elp.SUBLAT_DENSE, 13)
    e96.set_rules(elp.RULES_BALANCED)
    e96.upload(h96.lattice())
    h96.load_shift(tile)
    t0 = time.perf_counter()
    e96.run(h96.origins(tile, 20000), tau0=0, seed=h96.seed)
    net = elp.Net(seed=1)
    for k in range(4):
        e96.run(h96.origins(tile, 2000), tau0=20000 + 2000 * k, seed=h96.seed)
        net.harvest(e96, 250000, seed=100 + k)
    loss = net.train(epochs=4, batch=4096, lr=3e-3)
    t_train = time.perf_counter() - t0
    e96.close()
    curves = {}
    lp_kernel_ms, lp_launches, lp_blocks = 0.0, 0, 0
    for kind in ("plain", "learned"):
        h = elp.RingHost(N=4000, L=128, density=1.2, seed=0)
        sub, n_iter = (elp.SUBLAT_DENSE, 13) if kind == "plain" else (elp.SUBLAT_CUBIC, 51)
        eng = elp.Engine("ring", h.L, tile, sub, n_iter)
        eng.set_rules(elp.RULES_BALANCED)
        eng.upload(h.lattice())
        h.load_shift(tile)
        pts = [(0.0, 0, _rg2_now(h, eng), _contact_now(h, eng))]
        tau, gpu_s, chunk = 0, 0.0, (2000 if kind == "plain" else 20)
        while gpu_s < budget:
            org = h.origins(tile, chunk)
            t1 = time.perf_counter()
            if kind == "plain":
                eng.run(org, tau0=tau, seed=h.seed)
            else:
                net.run(eng, org, tau0=tau, seed=h.seed)
                lp_kernel_ms += eng.last_run_ms()        # HIP events around the launches of this chunk
                lp_launches += chunk
                lp_blocks = eng.launch_info()["blocks"]
            gpu_s += time.perf_counter() - t1
            tau += chunk
            pts.append((gpu_s, tau, _rg2_now(h, eng), _contact_now(h, eng)))
            chunk = min(chunk * 2, 64000 if kind == "plain" else 640)
        curves[kind] = pts
        eng.close()
    def t_reach(pts, col, target, rising):
        """GPU seconds and time units at which column col of the curve first reaches target
        (linear interpolation between samples)"""
        for a, b in zip(pts, pts[1:]):
            if (b[col] >= target) if rising else (b[col] <= target):
                f = (target - a[col]) / (b[col] - a[col]) if b[col] != a[col] else 1.0
                return a[0] + f * (b[0] - a[0]), a[1] + f * (b[1] - a[1])
        return pts[-1][0], pts[-1][1]

    def common(col):
        """the level both arms reach: the statistic's direction from the compact start is that of the
        plain arm (the longer run); the common level is the nearer of the two final values"""
        rising = curves["plain"][-1][col] >= curves["plain"][0][col]
        ends = (curves["plain"][-1][col], curves["learned"][-1][col])
        target = min(ends) if rising else max(ends)
        return target, t_reach(curves["plain"], col, target, rising), t_reach(curves["learned"], col, target, rising)

    # Rg^2 (rises as the compact rings swell) and the contact probability at large separations
    target, (tp, up), (tl, ul) = common(2)
    ptarget, (tpp, upp), (tlp, ulp) = common(3)
    # MFMA roofline of the learned kernel: per proposal two network evaluations of
    # 144*64 + 64*64 + 64*11 multiply-adds; one launch = every 16^3 tile x 64 slots
    # x 51 iterations x 4 substeps
    launch_s = lp_kernel_ms * 1e-3 / max(lp_launches, 1)
    flops_per_launch = 2 * 2 * (144 * 64 + 64 * 64 + 64 * 11) * lp_blocks * 64 * 51 * 4
    achieved_tf = flops_per_launch / launch_s / 1e12
    rec = {
        "metric": "MC move attempts/s, %HBM roofline; equil. wall-time learned vs plain",
        "value": tp / tl,
        "unit": "wall-time ratio plain/learned to reach the common Rg^2 (>1: learned faster)",
        "n_gpus": 1, "steps": None, "warmup": None, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8 lattice, bf16/fp32 network",
        "data": "synthetic",
        "config": {
            "workload": "ring melt L=128 N=4000 density 1.2 (629 rings) from the compact start; equilibration A/B "
                        f"with {budget:.0f} s of GPU wall time each",
            "plain": "balanced rates, 16^3 tiles, dense sub-lattice, 13 iterations",
            "learned": "balanced-rate MH chain with the learned proposal, 16^3 tiles, cubic sub-lattice, 51 iterations",
            "training": f"1e6 moves harvested from an L=96 N=1000 melt (plain engine, 28000 time units), 4 epochs "
                        f"Adam, final cross-entropy {loss:.3f}, {t_train:.1f} s including the L=96 run",
            "target_rg2": target,
            "plain_reach": {"seconds": tp, "time_units": up},
            "learned_reach": {"seconds": tl, "time_units": ul},
            "time_units_ratio_plain_over_learned": up / ul if ul else None,
            "contact_probability": {
                "statistic": f"mean P(s) over s = {P_WINDOW[0]}..{P_WINDOW[1]}", "target": ptarget,
                "plain_reach": {"seconds": tpp, "time_units": upp}, "learned_reach": {"seconds": tlp, "time_units": ulp},
                "wall_time_ratio_plain_over_learned": tpp / tlp if tlp else None,
                "time_units_ratio_plain_over_learned": upp / ulp if ulp else None},
            "curves": {k: [[round(a, 3), b, round(c, 4), float(f"{d:.4g}")] for a, b, c, d in v] for k, v in curves.items()},
            "curve_columns": ["GPU seconds", "time units", "mean Rg^2", "contact probability (statistic above)"],
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "mfma",
            "achieved": achieved_tf,
            "peak": BF16_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tf / BF16_PEAK_TFLOPS,
            "traffic": None,
            "kernel": "lp_ring_time_unit2",
            "flops_per_launch": flops_per_launch,
            "avg_launch_us": launch_s * 1e6,
            "launches": lp_launches,
        },
        "cpu_baseline": None,
    }
    print(json.dumps(rec))
    net.close()
    return 0


def bench_learned_units(args):
    """Throughput of the learned-proposal kernel alone: ring melt L=128,
    N=4000, density 1.2 from the compact start, 16^3 tiles, cubic sub-lattice,
    51 iterations, the network's own initial parameters; one time unit per
    step.  The MH chain's cost does not depend on the weights' values."""
    import elp
    tile = (16, 16, 16)
    h = elp.RingHost(N=4000, L=128, density=1.2, seed=0)
    eng = elp.Engine("ring", h.L, tile, elp.SUBLAT_CUBIC, 51)
    eng.set_rules(elp.RULES_BALANCED)
    eng.upload(h.lattice())
    h.load_shift(tile)
    net = elp.Net(seed=1)
    org = h.origins(tile, args.warmup + args.steps)
    if args.warmup:
        net.run(eng, org[:args.warmup], tau0=0, seed=h.seed)
    eng.sync()
    t0 = time.perf_counter()
    prop, acc = net.run(eng, org[args.warmup:], tau0=args.warmup, seed=h.seed)
    wall = time.perf_counter() - t0
    launch_s = eng.last_run_ms() * 1e-3 / args.steps
    blocks = eng.launch_info()["blocks"]
    flops_per_launch = 2 * 2 * (144 * 64 + 64 * 64 + 64 * 11) * blocks * 64 * 51 * 4
    achieved_tf = flops_per_launch / launch_s / 1e12
    rec = {
        "metric": "learned proposals/s (MH chain of the learned-proposal engine)",
        "value": prop / wall, "unit": "proposals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8 lattice, bf16/fp32 network",
        "data": "synthetic",
        "config": {"workload": "ring melt L=128 N=4000 density 1.2 (629 rings), learned-proposal time units, "
                               "16^3 tiles, cubic sub-lattice, 51 iterations, untrained network",
                   "accepted_fraction": acc / prop if prop else None, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": achieved_tf, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tf / BF16_PEAK_TFLOPS, "traffic": None, "kernel": "lp_ring_time_unit2",
                     "flops_per_launch": flops_per_launch, "avg_launch_us": launch_s * 1e6},
        "cpu_baseline": None,
    }
    print(json.dumps(rec))
    net.close()
    eng.close()
    return 0


STATICS = {
    "dense": "reference rates on the dense sub-lattice (four sites per 4^3 cell): the reference's moves; "
             "stored-length ratio about 1 % above the reference schedule's (the move set is not reversible)",
    "cubic": "reference rates on the reference's cubic sub-lattice (one site per 4^3 cell): the reference's statics",
}
STATICS_BALANCED = ("balanced rates (forced TransBack at 1/4, the move kind drawn per tile and substep): the "
                    "uniform measure over lattice states is stationary, the same statics on every schedule")


def binding_resource(kernel, chain, L, N, tile, sublattice, avg_launch_us):
    """VALU issue share of the kernel's wave time from recorded SQ counters of this exact workload
    (profiles/sq_counts.txt: waves, VALU and LDS instructions per wave, SQ_WAVE_CYCLES per wave in
    quad cycles), or None.  share = (waves per SIMD) x VALU per wave / wave quad-cycles: the fraction
    of a SIMD's cycles in which it issues VALU work (one wave64 VALU op takes 4 cycles = 1 quad)."""
    path = os.path.join(ROOT, "profiles", "sq_counts.txt")
    if not os.path.exists(path):
        return None
    key = [kernel, chain, str(L), str(N), "x".join(str(t) for t in tile), sublattice]
    with open(path) as f:
        for line in f:
            w = line.split()
            if len(w) == 11 and not line.startswith("#") and w[:6] == key:
                waves, valu, lds, quad, simds = (float(x) for x in w[6:])
                share = (waves / simds) * valu / quad
                return {"resource": "VALU issue", "achieved": share, "peak": 1.0, "frac": share,
                        "valu_per_wave": valu, "lds_per_wave": lds, "wave_quad_cycles": quad,
                        "profiled_us_per_launch": quad * 4 / 2.4e3, "current_avg_launch_us": avg_launch_us,
                        "source": "profiles/sq_counts.txt (rocprofv3 SQ_WAVES, SQ_INSTS_VALU, SQ_INSTS_LDS, "
                                  "SQ_WAVE_CYCLES)"}
    return None


def measure_schedule(elp, host, lat0, ring, tile, sublattice, n_iter, steps, warmup, rules=0):
    """Time `steps` time units of one schedule from lat0 (after `warmup` untimed ones), then replay the same
    time units with move counting from the same lattice (the counting kernels move identically, which is
    checked) for the accepted-move rate and the rejection shares."""
    dense = sublattice == "dense"
    n_iter = n_iter if n_iter is not None else ((13 if dense else 51) if ring else (16 if dense else 64))
    eng = elp.Engine("ring" if ring else "linear", host.L, tile, elp.SUBLAT_DENSE if dense else elp.SUBLAT_CUBIC,
                     n_iter)
    if rules:
        eng.set_rules(rules)
    eng.upload(lat0)
    host.set_lattice(lat0)
    host.load_shift(tile)
    nsite = host.L[0] * host.L[1] * host.L[2]
    attempts_per_step = n_iter * (4 if ring else 2) * (nsite // 64) * (4 if dense else 1)
    # the host's shift stream for warm-up and timed units, drawn up front: no host work between the
    # warm-up and the timed units (the GPU stays busy up to the timed region)
    org_w = np.ascontiguousarray(host.origins(tile, warmup), dtype=np.int32) if warmup else None
    org = np.ascontiguousarray(host.origins(tile, steps), dtype=np.int32)
    if warmup:
        eng.run(org_w, tau0=0, seed=host.seed, sync=False)
    eng.sync()
    t0 = time.perf_counter()
    eng.run(org, tau0=warmup, seed=host.seed, sync=False)
    eng.sync()
    wall = time.perf_counter() - t0
    kernel_ms = eng.last_run_ms()
    # host floor: the same bracket around a run of no time units (two event records and a sync)
    floor = []
    for _ in range(20):
        eng.sync()
        t1 = time.perf_counter()
        eng.run(org[:0], tau0=0, seed=host.seed, sync=False)
        eng.sync()
        floor.append(time.perf_counter() - t1)
    after = eng.download()
    # the lattice must still hold every labelled chain
    host.set_lattice(after)
    # the timed units again with move counting, from the same start (lat0 + the warm-up units)
    eng.upload(lat0)
    if warmup:
        eng.run(org_w, tau0=0, seed=host.seed)
    eng.count_moves(True)
    eng.run(org, tau0=warmup, seed=host.seed)
    c = eng.move_counts()
    eng.count_moves(False)
    if not np.array_equal(eng.download(), after):
        raise RuntimeError("the counting kernels moved differently from the plain ones")
    info = eng.launch_info()
    eng.close()
    a = max(c["attempts"], 1)   # a run of zero iterations (fixed-cost measurement) attempts nothing
    launch_s = kernel_ms * 1e-3 / steps
    return {
        "tile": list(tile), "sublattice": sublattice, "n_iter": n_iter, "rules": rules, "threads": info["threads"],
        "blocks": info["blocks"], "attempts_per_step": attempts_per_step, "wall_s": wall,
        "attempts_per_s": attempts_per_step * steps / wall,
        "accepted_moves_per_s": c["accepted"] / wall,
        "ms_per_step": wall * 1e3 / steps,
        "avg_launch_us": launch_s * 1e6,
        "host_floor_us": float(np.median(floor)) * 1e6,
        "achieved_GBs": 2 * (1 if ring else 2) * nsite / launch_s / 1e9,
        "move_outcomes": {"accepted": c["accepted"] / a, "empty_site": c["empty"] / a, "tile_face": c["face"] / a,
                          "rejected_other": (a - c["accepted"] - c["empty"] - c["face"]) / a},
    }


def schedule_entry(m, headline):
    return {
        "headline": headline,
        "tile": m["tile"], "sublattice": m["sublattice"], "iterations_per_time_unit": m["n_iter"],
        "statics": STATICS_BALANCED if m["rules"] else STATICS[m["sublattice"]],
        "attempts_per_s": m["attempts_per_s"], "accepted_moves_per_s": m["accepted_moves_per_s"],
        "ms_per_step": m["ms_per_step"], "avg_launch_us": m["avg_launch_us"],
        "roofline_frac": m["achieved_GBs"] / HBM_PEAK_GBS, "move_outcomes": m["move_outcomes"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--L", type=int, default=128)
    ap.add_argument("--N", type=int, default=None, help="monomers per chain (ring 1000, linear 500)")
    ap.add_argument("--chain", default="ring", choices=["ring", "linear"])
    ap.add_argument("--rules", default="reference", choices=["reference", "balanced"],
                    help="ring move rates (the benchmark line: the reference's)")
    ap.add_argument("--density", type=float, default=1.2)
    ap.add_argument("--tile", default="16")
    ap.add_argument("--sublattice", default="dense", choices=["cubic", "dense"])
    ap.add_argument("--n-iter", type=int, default=None)
    ap.add_argument("--cpu-cores", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-only", action="store_true",
                    help="only run the plain copy kernel (byte-counter calibration)")
    ap.add_argument("--workload", default="melt", choices=["melt", "batch", "learned", "learned-units"],
                    help="melt: one box (default); batch: 64 ring replicas, L in {32,48,64,96}, "
                         "N in {50,200,1000,2000}, seeds 0..63, in one launch per time unit; learned: "
                         "equilibration A/B of the learned proposal against the plain engine; learned-units: "
                         "time units of the learned-proposal kernel alone (ring L=128 N=4000, untrained network)")
    ap.add_argument("--budget", type=float, default=30.0, help="learned: GPU seconds per arm")
    ap.add_argument("--units-per-graph", type=int, default=16,
                    help="batch: time units captured per HIP graph (0 = direct launches)")
    ap.add_argument("--threads", type=int, default=0, help="threads per workgroup (0 = schedule default)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="batch: 1 = take tile, sub-lattice, threads and units per graph from elp_autotune")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(independent replicas run one engine per GPU with no exchange)", file=sys.stderr)
        return 2

    import elp

    if args.calib_only:
        out = {}
        for nb, reps in ((4 << 20, 200), (256 << 20, 20)):
            ms = elp.bench_copy(nb, reps)
            out[f"copy_{nb >> 20}MiB"] = {"ms": ms, "GBs_read_plus_write": 2 * nb / (ms * 1e-3) / 1e9, "reps": reps}
        print(json.dumps(out))
        return 0

    tile = tuple(int(x) for x in args.tile.split("x")) if "x" in args.tile else (int(args.tile),) * 3
    if args.workload == "batch":
        return bench_batch(args, tile)
    if args.workload == "learned-units":
        return bench_learned_units(args)
    if args.workload == "learned":
        return bench_learned(args)
    ring = args.chain == "ring"
    N = args.N or (1000 if ring else 500)
    host = (elp.RingHost if ring else elp.LinHost)(N=N, L=args.L, density=args.density, seed=0)
    lat0 = host.lattice()
    # the companion schedule: the reference's cubic sub-lattice on 16^3 tiles (its statics) beside the
    # dense one (fastest; stored-length ratio about 1 % higher under reference rates), or the reverse.
    # Measured first: the first timed run of a process pays a one-time runtime cost (about 20 us at the
    # first burst of launches), which then falls on the companion, not on the headline schedule.
    other_sub = "cubic" if args.sublattice == "dense" else "dense"
    rules = elp.RULES_BALANCED if args.rules == "balanced" else elp.RULES_REFERENCE
    if rules and not ring:
        print("bench.py: --rules balanced applies to the ring engine only", file=sys.stderr)
        return 2
    other = measure_schedule(elp, host, lat0, ring, (16, 16, 16), other_sub, None, args.steps, args.warmup, rules)
    head = measure_schedule(elp, host, lat0, ring, tile, args.sublattice, args.n_iter, args.steps, args.warmup,
                            rules)

    nsite = host.L[0] * host.L[1] * host.L[2]
    site_bytes = 1 if ring else 2
    bytes_per_launch = 2 * site_bytes * nsite
    copy_ms = elp.bench_copy(bytes_per_launch // 2, 200)
    copy_gbs = bytes_per_launch / (copy_ms * 1e-3) / 1e9
    kernel = "ring_time_unit2" if ring else "lin_time_unit2"
    traffic = measured_traffic(kernel, args.chain, host.L[0], host.N, tile, args.sublattice)
    cores = args.cpu_cores or len(os.sched_getaffinity(0))
    cpu = None if args.no_cpu else cpu_baseline(args.L, N, args.density, cores, warm=300, timed=100, ring=ring)

    wall = head["wall_s"]
    value = head["attempts_per_s"]
    # GPU attempts cover every site, empty ones included; normalised like efpol's
    # rate, a time unit of the whole melt counts once per monomer
    monomer_rate = host.n_pol * host.N * args.steps / wall
    rec = {
        "metric": "MC move attempts/s, %HBM roofline; equil. wall-time learned vs plain",
        "value": value,
        "unit": "attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "ms_per_step_over_kernel": wall * 1e3 / args.steps / (head["avg_launch_us"] * 1e-3),
        "host_floor_us": head["host_floor_us"],
        "time_units_per_s": args.steps / wall,
        "monomer_time_units_per_s": monomer_rate,
        "accepted_moves_per_s": head["accepted_moves_per_s"],
        "move_outcomes": head["move_outcomes"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8" if ring else "u16",
        "data": "synthetic",
        "config": {
            "workload": f"{'ring' if ring else 'linear-chain'} melt L={host.L[0]} N={host.N} density {args.density} "
                        f"({host.n_pol} chains), one time unit per step",
            "tile": list(tile),
            "sublattice": args.sublattice,
            "rules": args.rules,
            "iterations_per_time_unit": head["n_iter"],
            "attempts_per_site_per_time_unit": head["attempts_per_step"] / nsite,
            "threads_per_block": head["threads"],
            "blocks_per_launch": head["blocks"],
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": head["achieved_GBs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": head["achieved_GBs"] / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": "profiles/hbm_traffic.txt (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0)"
                              if traffic else None,
            "kernel": kernel,
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_us": head["avg_launch_us"],
            "copy_kernel_same_bytes_GBs": copy_gbs,
        },
        # the recorded SQ counters are those of the reference rates
        "binding": None if rules else binding_resource(kernel, args.chain, host.L[0], host.N, tile,
                                                       args.sublattice, head["avg_launch_us"]),
        "schedules": [schedule_entry(head, True), schedule_entry(other, False)],
        "cpu_baseline": cpu,
    }
    if cpu:
        rec["gpu_over_cpu"] = value / cpu["value"]
        rec["gpu_over_cpu_monomer_time_units"] = monomer_rate / cpu["monomer_time_units_per_s"]
    print(json.dumps(rec))
    return 0


if __name__ == "__main__":
    sys.exit(main())
