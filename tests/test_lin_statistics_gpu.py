"""Equilibrium statistics of the GPU linear-chain engine (melt L=32, N=100,
density 1.2, 8 seeds run as one batch of 8 replicas, 400000 time units from
the compact start, frames every 10000, the first 11 discarded: the protocol
of the efpol reference tests/golden/efpol_stats/lin_L32_N100.json, written
by the script beside it), for the cubic (16^3, 64 iterations) and dense
(16^3, 16 iterations) sub-lattices, against each other and against efpol.

Linear chains relax slowly (the chain centre of mass moves by one radius of
gyration in ~30000 time units, and Rg^2 is still rising after 60000), so the
window starts at 110000.  The
two schedules must agree within 3 sigma in Rg^2, stored-length ratio and
P(s).  efpol stacks any number of monomers on a site and so stores more
length (the test pins that difference); its chains' Rg^2 per bond that is
not stored length must agree with the GPU's within 3 sigma.  efpol's
sl_double model restricted to at most two monomers per site, from the GPU
engine's own starts (tests/golden/efpol_stats/lin2_L32_N100.json), samples the
GPU engine's own stationary law: there every static agrees within 3 sigma.
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
REF = json.load(open(os.path.join(HERE, "golden", "efpol_stats", "lin_L32_N100.json")))
L, N = REF["L"], REF["N"]
T, INTERVAL, N_THERM = REF["time_units"], REF["interval"], REF["n_therm_frames"]
S_POINTS = [2, 3, 5, 10, 20, 50]
LAGS = [1, 2, 4, 8, 16, 29]          # frames (x INTERVAL time units), as the efpol fixtures
SCHEDULES = {"cubic 16^3 64": ((16, 16, 16), elp.SUBLAT_CUBIC, 64),
             "dense 16^3 16": ((16, 16, 16), elp.SUBLAT_DENSE, 16)}
_cache = {}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(name):
    if name in _cache:
        return _cache[name]
    tile, sub, n_iter = SCHEDULES[name]
    _say(f"[linear statistics] {name}, seeds 0..7 as one batch")
    hosts = [elp.LinHost(N=N, L=L, density=REF["density"], seed=s) for s in range(8)]
    bat = elp.Batch("linear", [h.L for h in hosts], tile, sub, n_iter)
    series = []
    for r, h in enumerate(hosts):
        bat.upload(r, h.lattice())
        # frames reduced on the device, replica by replica in place: no lattice download per frame
        series.append(Series(False, h.L, h.n_pol, h.N, T // INTERVAL + 1, summary=True))
        series[r].add_summary(bat.frame_summary(r, h))
    tau = 0
    for k in range(T // INTERVAL):
        org = np.stack([h.origins(tile, INTERVAL) for h in hosts], axis=1)
        bat.run(org, tau0=tau, seeds=[h.seed for h in hosts], streams=[0] * 8, units_per_graph=16)
        tau += INTERVAL
        for r, h in enumerate(hosts):
            series[r].add_summary(bat.frame_summary(r, h))
        if k % 10 == 9:
            _say(f"  {tau} time units")
    res = []
    for ser in series:
        st = ser.statics(N_THERM)
        g1, g3 = ser.msd(N_THERM, max(LAGS) + 1)
        res.append({"rg2": st["rg2"], "slrat": st["slrat"], "pc": st["pc"][S_POINTS], "g1": g1[LAGS], "g3": g3[LAGS]})
    bat.close()
    _cache[name] = res
    return res


def _mean_se(vals):
    v = np.asarray(vals, dtype=float)
    return v.mean(axis=0), v.std(axis=0, ddof=1) / np.sqrt(len(v))


def _z(a, b):
    (ma, sa), (mb, sb) = _mean_se(a), _mean_se(b)
    return (ma - mb) / np.sqrt(sa ** 2 + sb ** 2)


def _per_bond(rows):
    # Rg^2 per bond that is not stored length (N - 1 bonds per chain)
    return [r["rg2"] / ((N - 1) * (1.0 - r["slrat"])) for r in rows]


def test_linear_schedules_agree_within_3_sigma():
    a, b = _run("cubic 16^3 64"), _run("dense 16^3 16")
    for key in ("rg2", "slrat"):
        z = _z([r[key] for r in a], [r[key] for r in b])
        _say(f"linear {key}: {_mean_se([r[key] for r in a])} vs {_mean_se([r[key] for r in b])}, z = {z:.2f}")
        assert abs(z) < 3, (key, z)
    z = _z([r["pc"] for r in a], [r["pc"] for r in b])
    _say(f"linear P(s) at s={S_POINTS}: z = {np.round(z, 2)}")
    assert np.all(np.abs(z) < 3), z


def test_linear_efpol_comparison():
    ef = REF["seeds"]
    for name in SCHEDULES:
        g = _run(name)
        z_sl = _z([r["slrat"] for r in g], [s["slrat"] for s in ef])
        z_pb = _z(_per_bond(g), _per_bond(ef))
        _say(f"linear {name} vs efpol: Rg^2 {_mean_se([r['rg2'] for r in g])[0]:.3f} vs "
             f"{_mean_se([s['rg2'] for s in ef])[0]:.3f}; stored-length ratio z = {z_sl:.1f}; "
             f"Rg^2 per non-stored-length bond {_mean_se(_per_bond(g))[0]:.4f} vs {_mean_se(_per_bond(ef))[0]:.4f}, "
             f"z = {z_pb:.2f}")
        assert z_sl < -10                  # efpol stores more length (stacked stored length)
        assert abs(z_pb) < 3, z_pb         # the same conformations per bond that is not stored length


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


@pytest.mark.parametrize("name", list(SCHEDULES))
def test_linear_statics_equal_efpol_sl_double_with_at_most_two_per_site(name):
    """The linear move set is in detailed balance with weight 2 per stored-length
    bond (tests/test_lin_balance.py), and so is efpol's sl_double table
    (tests/test_efpol_table.py); efpol restricted to at most two monomers per
    site, from the GPU engine's own starts (tests/golden/efpol_stats/
    lin2_L32_N100.json, same protocol), samples that law over the states a GPU
    lattice holds.  Rg^2, the stored-length ratio and P(s) agree within 3
    sigma; one time-unit factor fitted on g3 predicts g1 within 15 %."""
    ref = json.load(open(os.path.join(HERE, "golden", "efpol_stats", "lin2_L32_N100.json")))
    assert (ref["time_units"], ref["interval"], ref["n_therm_frames"], ref["lags"]) == (T, INTERVAL, N_THERM, LAGS)
    e2, g = ref["seeds"], _run(name)
    assert all(r["stack3"] == 0 for r in e2)
    zs = {key: _z([r[key] for r in g], [r[key] for r in e2]) for key in ("rg2", "slrat")}
    zs["pc"] = _z([r["pc"] for r in g], [np.asarray(r["pc"])[S_POINTS] for r in e2])
    _say(f"[linear efpol-2] {name}: Rg^2 {_mean_se([r['rg2'] for r in g])} vs {_mean_se([r['rg2'] for r in e2])}, "
         f"stored length {_mean_se([r['slrat'] for r in g])} vs {_mean_se([r['slrat'] for r in e2])}; "
         f"z Rg^2 {zs['rg2']:.2f}, stored length {zs['slrat']:.2f}, P(s) at {S_POINTS} {np.round(zs['pc'], 2)}")
    tau = np.asarray(LAGS, float)
    g3g, g1g = _mean_se([r["g3"] for r in g])[0], _mean_se([r["g1"] for r in g])[0]
    g3e, g1e = _mean_se([r["g3"] for r in e2])[0], _mean_se([r["g1"] for r in e2])[0]
    f = _time_factor(tau, g3g, g3e)
    x = np.log(tau) + np.log(f)
    inside = (x >= np.log(tau[0])) & (x <= np.log(tau[-1]))
    dev = g1g[inside] / np.exp(np.interp(x, np.log(tau), np.log(g1e)))[inside] - 1.0
    _say(f"[linear efpol-2] {name}: time-unit factor {f:.3f} from g3, g1 deviation {np.round(dev, 3)}")
    assert abs(zs["rg2"]) < 3 and abs(zs["slrat"]) < 3 and np.all(np.abs(zs["pc"]) < 3), zs
    assert inside.sum() >= 3 and np.all(np.abs(dev) < 0.15), dev
