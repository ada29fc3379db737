"""Equilibrium statistics of the GPU ring engine (ring melt L=32, N=100,
density 1.2, 64 seeds run as one batch of 64 replicas, 200000 time units,
frames every 1000, the first 101 discarded), for the reference schedule (32^3 tiles, cubic sub-lattice, 51
iterations) and the dense schedule (16^3, dense, 13), against each other and
against efpol (tests/golden/efpol_stats/ring_L32_N100.json, written by the
script beside it).

With balanced rates (each substep a symmetric kernel, so the uniform measure
is stationary for any schedule) the two schedules must agree within 3 sigma.
With the reference rates the move set is not reversible and the fixed
substep sequence is a product of non-symmetric kernels, so the stored-length
ratio depends on the sub-lattice; the test pins that.
efpol samples a different model: it stacks any number of monomers
on a site (its equilibrium has about 10 % of bonds as stored length following
stored length), while a GPU site holds at most two, and the GPU move set is
not reversible (tests/test_ring_balance.py).  The test therefore pins the
structural difference and reports the efpol comparison instead of asserting
agreement.  efpol restricted to at most two monomers per site
(tests/golden/efpol_stats/ring2_L32_N100.json, from the GPU engine's own
starts) samples the uniform measure over the GPU state space (efpol's
sl_equal table is symmetric: tests/test_efpol_table.py), which is the
equilibrium of the balanced rates: there the statics must agree within
3 sigma, for both schedules, and, on the dense schedule, at N = 50 and 200 as
well (Rg^2(N); ring2_L32_N50.json, ring2_L32_N200.json).
"""
import json
import os
import sys

import numpy as np
import pytest

import elp
from elp.stats import Series

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REF = json.load(open(os.path.join(HERE, "golden", "efpol_stats", "ring_L32_N100.json")))
REF2 = json.load(open(os.path.join(HERE, "golden", "efpol_stats", "ring2_L32_N100.json")))
# the GPU runs follow the efpol-2 protocol (100000 time units, frames every 1000, the first 41 discarded:
# ring Rg^2 from the compact start still rises until about 40000 units)
L, N, T, INTERVAL, N_THERM = REF2["L"], REF2["N"], REF2["time_units"], REF2["interval"], REF2["n_therm_frames"]
LAGS = REF2["lags"]
assert LAGS == REF["lags"] and REF["interval"] == INTERVAL
S_POINTS = [2, 3, 5, 10, 20, 50]
# seeds per schedule: the stored-length ratio of one seed's window scatters by about 2e-4, so 64 seeds
# resolve 2.5e-5 (8 seeds give an unreliable standard error); the replicas run as one batch
NSEED = 64
SCHEDULES = {"reference 32^3 cubic 51": ((32, 32, 32), elp.SUBLAT_CUBIC, 51),
             "dense 16^3 13": ((16, 16, 16), elp.SUBLAT_DENSE, 13)}
# the reference's cubic sub-lattice on smaller tiles (8x the workgroups of 32^3)
OTHER = {"cubic 16^3 51": ((16, 16, 16), elp.SUBLAT_CUBIC, 51)}
_cache = {}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(name, rules=elp.RULES_REFERENCE, ref=REF2):
    """NSEED seeds of schedule `name` under the protocol of the efpol-2 fixture
    `ref` (chain length, time units, frame interval, frames discarded)."""
    n, t_units, interval, n_therm = ref["N"], ref["time_units"], ref["interval"], ref["n_therm_frames"]
    if (name, rules, n) in _cache:
        return _cache[(name, rules, n)]
    tile, sub, n_iter = {**SCHEDULES, **OTHER}[name]
    # the seeds as one batch of NSEED replicas (each bit-exact with its solo run)
    _say(f"[statistics] {name} rules {rules} N={n}, seeds 0..{NSEED - 1} as one batch")
    hosts = [elp.RingHost(N=n, L=L, density=REF["density"], seed=s) for s in range(NSEED)]
    bat = elp.Batch("ring", [h.L for h in hosts], tile, sub, n_iter)
    bat.set_rules(rules)
    series = []
    for r, h in enumerate(hosts):
        bat.upload(r, h.lattice())
        h.load_shift(tile)
        # frames reduced on the device, replica by replica in place: no lattice download per frame
        series.append(Series(True, h.L, h.n_pol, h.N, t_units // interval + 1, summary=True))
        series[r].add_summary(bat.frame_summary(r, h))
    tau = 0
    for _ in range(t_units // interval):
        org = np.stack([h.origins(tile, interval) for h in hosts], axis=1)
        bat.run(org, tau0=tau, seeds=[h.seed for h in hosts], streams=[0] * NSEED, units_per_graph=16)
        tau += interval
        for r, h in enumerate(hosts):
            series[r].add_summary(bat.frame_summary(r, h))
    res = []
    for r, (h, ser) in enumerate(zip(hosts, series)):
        h.set_lattice(bat.download(r))             # the last frame, for the stacking count
        stack3 = 0
        for c in range(h.n_pol):
            b = h.polymer(c)[1]
            stack3 += int(np.sum((b == 0) & (np.roll(b, 1) == 0)))
        st = ser.statics(n_therm)
        g1, g3 = ser.msd(n_therm, max(LAGS) + 1)
        res.append({"rg2": st["rg2"], "slrat": st["slrat"], "pc": st["pc"], "g1": g1[LAGS], "g3": g3[LAGS],
                    "stack3": stack3})
    bat.close()
    _cache[(name, rules, n)] = res
    return res


def _mean_se(vals):
    v = np.asarray(vals, dtype=float)
    return v.mean(axis=0), v.std(axis=0, ddof=1) / np.sqrt(len(v))


def _z(a, b):
    (ma, sa), (mb, sb) = _mean_se(a), _mean_se(b)
    return (ma - mb) / np.sqrt(sa ** 2 + sb ** 2)


def _compare_schedules(rules):
    a, b = _run("reference 32^3 cubic 51", rules), _run("dense 16^3 13", rules)
    zs = {}
    for key in ("rg2", "slrat"):
        zs[key] = _z([r[key] for r in a], [r[key] for r in b])
        _say(f"rules {rules} {key}: {_mean_se([r[key] for r in a])} vs {_mean_se([r[key] for r in b])}, "
             f"z = {zs[key]:.2f}")
    zs["pc"] = _z([r["pc"][S_POINTS] for r in a], [r["pc"][S_POINTS] for r in b])
    _say(f"rules {rules} P(s) at s={S_POINTS}: z = {np.round(zs['pc'], 2)}")
    return zs


def test_balanced_rules_schedules_agree_within_3_sigma():
    """Balanced rates: the uniform measure is stationary for every schedule, so
    the cubic and dense sub-lattices must give the same statics."""
    zs = _compare_schedules(elp.RULES_BALANCED)
    assert abs(zs["rg2"]) < 3 and abs(zs["slrat"]) < 3 and np.all(np.abs(zs["pc"]) < 3), zs


def test_reference_rules_schedule_dependence():
    """Reference rates: the move set is not reversible, and its equilibrium
    depends on the schedule.  Rg^2 agrees between the sub-lattices; the
    stored-length ratio does not (dense about 1 % higher: measured z ~ 10)."""
    zs = _compare_schedules(elp.RULES_REFERENCE)
    assert abs(zs["rg2"]) < 3, zs
    assert zs["slrat"] < -3, zs


def test_reference_rules_cubic_tile_size_independent():
    """Reference rates: the cubic sub-lattice on 16^3 tiles samples the same
    statics as on the reference's 32^3 tiles (only the sub-lattice matters)."""
    a, b = _run("reference 32^3 cubic 51"), _run("cubic 16^3 51")
    for key in ("rg2", "slrat"):
        z = _z([r[key] for r in a], [r[key] for r in b])
        _say(f"cubic 32^3 vs 16^3 {key}: z = {z:.2f}")
        assert abs(z) < 3, (key, z)
    z = _z([r["pc"][S_POINTS] for r in a], [r["pc"][S_POINTS] for r in b])
    _say(f"cubic 32^3 vs 16^3 P(s): z = {np.round(z, 2)}")
    assert np.all(np.abs(z) < 3), z


@pytest.mark.parametrize("rules", [elp.RULES_REFERENCE, elp.RULES_BALANCED])
def test_efpol_comparison_model_difference(rules):
    ef = REF["seeds"]
    for name in SCHEDULES:
        g = _run(name, rules)
        assert all(r["stack3"] == 0 for r in g)               # at most two monomers per site
        rows = []
        for key in ("rg2", "slrat"):
            (mg, sg), (me, se) = _mean_se([r[key] for r in g]), _mean_se([r[key] for r in ef])
            rows.append(f"{key} GPU {mg:.4f}+-{sg:.4f} efpol {me:.4f}+-{se:.4f} z {(mg - me) / np.hypot(sg, se):.1f}")
        pg, pe = _mean_se([r["pc"][S_POINTS] for r in g]), _mean_se([np.asarray(r["pc"])[S_POINTS] for r in ef])
        rows.append(f"P(s) s={S_POINTS} GPU {np.round(pg[0], 5)} efpol {np.round(pe[0], 5)}")
        g3g, g3e = _mean_se([r["g3"] for r in g])[0], _mean_se([r["g3"] for r in ef])[0]
        rows.append(f"g3 at lags {LAGS} x {INTERVAL}: GPU {np.round(g3g, 3)} efpol {np.round(g3e, 3)}")
        _say(f"[efpol comparison] {name} rules {rules}\n  " + "\n  ".join(rows))
    assert np.mean([r["stack3"] for r in ef]) > 0.05


@pytest.mark.parametrize("name", list(SCHEDULES))
def test_balanced_rules_equal_efpol_with_at_most_two_per_site(name):
    """Balanced rates sample the uniform measure over GPU lattice states; so
    does efpol restricted to at most two monomers per site.  Rg^2, the
    stored-length ratio and P(s) agree within 3 sigma.  Reported beside it:
    the stacking effect (efpol vs efpol-2) and the rates effect (GPU
    reference vs balanced rates) that make up the efpol comparison."""
    g, e2, ef = _run(name, elp.RULES_BALANCED), REF2["seeds"], REF["seeds"]
    gr = _run(name, elp.RULES_REFERENCE)
    assert all(r["stack3"] == 0 for r in e2)
    zs = {}
    for key in ("rg2", "slrat"):
        zs[key] = _z([r[key] for r in g], [r[key] for r in e2])
        (m_g, s_g), (m_2, s_2) = _mean_se([r[key] for r in g]), _mean_se([r[key] for r in e2])
        m_e, m_r = _mean_se([r[key] for r in ef])[0], _mean_se([r[key] for r in gr])[0]
        _say(f"[efpol-2] {name} {key}: GPU balanced {m_g:.5f}+-{s_g:.5f} efpol-2 {m_2:.5f}+-{s_2:.5f} "
             f"z {zs[key]:.2f}; stacking (efpol - efpol-2) {m_e - m_2:+.4f}, rates (GPU reference - balanced) "
             f"{m_r - m_g:+.4f}")
    zs["pc"] = _z([r["pc"][S_POINTS] for r in g], [np.asarray(r["pc"])[S_POINTS] for r in e2])
    _say(f"[efpol-2] {name} P(s) at s={S_POINTS}: z = {np.round(zs['pc'], 2)}")
    g3g, g3e = _mean_se([r["g3"] for r in g])[0], _mean_se([r["g3"] for r in e2])[0]
    _say(f"[efpol-2] {name} g3 at lags {LAGS} x {INTERVAL}: GPU {np.round(g3g, 3)} efpol-2 {np.round(g3e, 3)}")
    assert abs(zs["rg2"]) < 3 and abs(zs["slrat"]) < 3 and np.all(np.abs(zs["pc"]) < 3), zs


def _time_factor(tau, g_gpu, g_ref):
    """f with g_gpu(t) = g_ref(f t): least squares in log g over the lags whose f t stays inside the
    reference's lag range, the reference curve interpolated linearly in log-log."""
    lt, lr, lg = np.log(tau), np.log(g_ref), np.log(g_gpu)

    def cost(lf):
        x = lt + lf
        inside = (x >= lt[0]) & (x <= lt[-1])
        return np.mean((lg[inside] - np.interp(x[inside], lt, lr)) ** 2) if inside.sum() >= 3 else np.inf
    grid = np.linspace(-1.0, 1.0, 4001)
    return float(np.exp(grid[np.argmin([cost(x) for x in grid])]))


def _predict(tau, f, g_ref):
    x = np.log(tau) + np.log(f)
    inside = (x >= np.log(tau[0])) & (x <= np.log(tau[-1]))
    return inside, np.exp(np.interp(x, np.log(tau), np.log(g_ref)))


@pytest.mark.parametrize("case", ["reference rates vs efpol", "balanced rates vs efpol-2"])
def test_time_unit_factor_from_g3_predicts_g1(case):
    """One time-unit factor f (a GPU time unit = f efpol time units) fitted on
    the centre-of-mass displacement g3 predicts the middle-monomer
    displacement g1 within 15 % at every lag it reaches.  For the reference
    rates on the reference schedule (gpupol3) against efpol, f agrees within
    15 % with the ratio of the reference's own time factors,
    scripts/ne_list.dat:11,20 (efpol 0.85, gpupol3 1.1 at density 1.2:
    relaxation times scale with the factor, so f = 0.85 / 1.1)."""
    rules, ref = (elp.RULES_REFERENCE, REF) if case.startswith("reference") else (elp.RULES_BALANCED, REF2)
    g = _run("reference 32^3 cubic 51", rules)
    tau = np.asarray(LAGS, float)
    g3g, g1g = _mean_se([r["g3"] for r in g])[0], _mean_se([r["g1"] for r in g])[0]
    g3e, g1e = _mean_se([r["g3"] for r in ref["seeds"]])[0], _mean_se([r["g1"] for r in ref["seeds"]])[0]
    f = _time_factor(tau, g3g, g3e)
    inside, pred = _predict(tau, f, g1e)
    dev = g1g[inside] / pred[inside] - 1.0
    _say(f"[time unit] {case}: f = {f:.3f} efpol units per GPU unit (reference ne_list ratio {0.85 / 1.1:.3f}); "
         f"g1 GPU {np.round(g1g, 3)} predicted {np.round(pred, 3)} at lags {LAGS} x {INTERVAL}, "
         f"deviation {np.round(dev, 3)}")
    assert inside.sum() >= 3
    assert np.all(np.abs(dev) < 0.15), dev
    if rules == elp.RULES_REFERENCE:
        assert abs(f / (0.85 / 1.1) - 1.0) < 0.15, f


CHAIN_LENGTHS = [50, 100, 200]


def _ref2(n):
    """the efpol-2 fixture of chain length n (ring2_L32_N{n}.json)"""
    return REF2 if n == REF2["N"] else json.load(open(os.path.join(HERE, "golden", "efpol_stats", f"ring2_L32_N{n}.json")))


def test_rg2_of_n_equals_efpol_with_at_most_two_per_site():
    """Rg^2(N), the stored-length ratio and P(s): the dense schedule with
    balanced rates against efpol restricted to at most two monomers per site,
    each from the GPU engine's own starts, at N = 50, 100 and 200 (L=32,
    density 1.2; N = 200 runs 300000 time units and drops the first 151
    frames, its rings relaxing about four times slower).  Every statistic
    agrees within 3 sigma at every N, and Rg^2 grows with N on both sides
    with the same exponent within its error.  Dynamics: at every N one
    time-unit factor fitted on g3 predicts g1 within 15 %, and the factor
    (a ratio of local move rates) is the same at every N within 15 %."""
    rg, fs = {}, {}
    for n in CHAIN_LENGTHS:
        ref = _ref2(n)
        g, e2 = _run("dense 16^3 13", elp.RULES_BALANCED, ref), ref["seeds"]
        assert all(r["stack3"] == 0 for r in g) and all(r["stack3"] == 0 for r in e2)
        pts = [s for s in S_POINTS if s <= n // 2]
        zs = {key: _z([r[key] for r in g], [r[key] for r in e2]) for key in ("rg2", "slrat")}
        zs["pc"] = _z([r["pc"][pts] for r in g], [np.asarray(r["pc"])[pts] for r in e2])
        rg[n] = (_mean_se([r["rg2"] for r in g]), _mean_se([r["rg2"] for r in e2]))
        _say(f"[Rg^2(N)] N={n}: GPU {rg[n][0][0]:.4f}+-{rg[n][0][1]:.4f} efpol-2 {rg[n][1][0]:.4f}+-{rg[n][1][1]:.4f}; "
             f"z Rg^2 {zs['rg2']:.2f}, stored length {zs['slrat']:.2f}, P(s) at {pts} {np.round(zs['pc'], 2)}")
        assert abs(zs["rg2"]) < 3 and abs(zs["slrat"]) < 3 and np.all(np.abs(zs["pc"]) < 3), (n, zs)
        tau = np.asarray(LAGS, float)
        g3g, g1g = _mean_se([r["g3"] for r in g])[0], _mean_se([r["g1"] for r in g])[0]
        g3e, g1e = _mean_se([r["g3"] for r in e2])[0], _mean_se([r["g1"] for r in e2])[0]
        fs[n] = _time_factor(tau, g3g, g3e)
        inside, pred = _predict(tau, fs[n], g1e)
        dev = g1g[inside] / pred[inside] - 1.0
        _say(f"[Rg^2(N)] N={n}: time-unit factor {fs[n]:.3f} from g3, g1 deviation {np.round(dev, 3)}")
        assert inside.sum() >= 3 and np.all(np.abs(dev) < 0.15), (n, dev)
    assert max(fs.values()) / min(fs.values()) < 1.15, fs
    # exponent of Rg^2 ~ N^a between the shortest and longest chains, with its propagated error
    n0, n1 = min(CHAIN_LENGTHS), max(CHAIN_LENGTHS)
    for side, k in (("GPU", 0), ("efpol-2", 1)):
        (m1, s1), (m2, s2) = rg[n0][k], rg[n1][k]
        a = np.log(m2 / m1) / np.log(n1 / n0)
        rg[side] = (a, np.hypot(s1 / m1, s2 / m2) / np.log(n1 / n0))
        _say(f"[Rg^2(N)] {side}: Rg^2 ~ N^{rg[side][0]:.4f} +- {rg[side][1]:.4f}")
    assert rg["GPU"][0] > 0.5 and rg["efpol-2"][0] > 0.5
    assert abs(rg["GPU"][0] - rg["efpol-2"][0]) < 3 * np.hypot(rg["GPU"][1], rg["efpol-2"][1])
