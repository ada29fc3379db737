"""Transition-rate ratios of the linear-chain move set, by exhaustive
enumeration (CPU, oracle).

For sites of a randomised linear melt the oracle's moves (orc_lin_moves:
DiffuseSL, and TransMove with every value of its random byte) are applied,
each resulting state weighted by its probability per iteration (one
DiffuseSL and one TransMove substep, every site equally likely to be
attempted).  The probability of the return to the start is enumerated the
same way over every occupied site next to the changed ones.

Reading gpupol/linpol.cu the rates are:
  * growth of a chain end out of stored length (:124-172): side 1/2 (bit 2)
    x direction 1/16 (bits 3-6); its reverse, the retraction of a one-monomer
    end onto its neighbour (:229-270): 1/2 (bit 4) x 1/8 (bits 5-7);
  * stored length split off to one of the four common neighbours of a bond
    (:175-225): side 1/2 x neighbour 1/4 (bits 0-1); its reverse, an interior
    monomer folded onto one of its two neighbours (:271-309): 1/2 (bit 4) x
    1/2 (bit 0);
  * DiffuseSL (:314-363): deterministic at the site, and its own reverse.
So every forward/reverse ratio is 2^(n_SL' - n_SL): the move set is in
detailed balance with the weight 2^(stored-length monomers), which is what
the test asserts for every transition found.  Planting any one wrong gate in
the oracle (orc_lin_plant) breaks it, which the second test shows.
"""
import numpy as np
import pytest

import elp
from lincheck import lin_reachable, lin_states, n_stored, planted, site_moves
from ringcheck import replay_lin

T = 16
CLASSES = ("diffuse", "grow", "insert", "retract_end", "retract")


def _melt():
    h = elp.LinHost(N=30, L=T, density=1.2, seed=5, box_unit=16)
    lat0 = h.lattice()
    h.load_shift((T, T, T))
    lat = replay_lin(lat0, h.L, (T, T, T), elp.SUBLAT_DENSE, 16, h.origins((T, T, T), 30), 0, h.seed)
    h.close()
    return lat, (T, T, T)


def _weight(k):
    return 1.0 if k == 0 else 1.0 / 256


def _forward(lat, L, sites):
    """{changes: [probability per iteration, [(k, site), ...]]}"""
    out = {}
    for x in sites:
        for k, ch in site_moves(lat, L, int(x)):
            e = out.setdefault(ch, [0.0, []])
            e[0] += _weight(k)
            e[1].append((k, int(x)))
    return out


def _near(idx, L):
    LT, LU, LV = L
    s = set()
    for x in idx:
        t, u, v = x % LT, (x // LT) % LU, x // (LT * LU)
        for dt in (-1, 0, 1):
            for du in (-1, 0, 1):
                for dv in (-1, 0, 1):
                    s.add((t + dt) % LT + LT * ((u + du) % LU + LU * ((v + dv) % LV)))
    return sorted(s)


def _back(lat, L, ch):
    b = lat.copy()
    for g, w in ch:
        b[g] = w
    want = tuple(sorted((g, int(lat[g])) for g, _ in ch))
    p = 0.0
    for y in _near([g for g, _ in ch], L):
        if b[y] & 0xFF:
            p += sum(_weight(k) for k, c2 in site_moves(b, L, y) if c2 == want)
    return p


def _classify(lat, ch):
    old = np.array([int(lat[g]) for g, _ in ch])
    new = np.array([w for _, w in ch])
    d = n_stored(new) - n_stored(old)
    if d == 0:
        return "diffuse", d
    is_end = lambda w: (w & 0xF) == 0 or ((w >> 4) & 0xF) == 0
    if d < 0:
        w = int(new[(old & 0xFF) == 0][0])             # the newly occupied site
        return ("grow" if is_end(w) else "insert"), d
    w = int(old[(new & 0xFF) == 0][0])                 # the vacated site
    return ("retract_end" if is_end(w) else "retract"), d


def _violations(lat, L, sites, stop_at_first=False):
    fwd = _forward(lat, L, sites)
    seen = dict.fromkeys(CLASSES, 0)
    bad = []
    for ch, (pf, how) in fwd.items():
        cls, d = _classify(lat, ch)
        seen[cls] += 1
        pb = _back(lat, L, ch)
        want = 2.0 ** d
        if pb == 0 or not np.isclose(pf / pb, want):
            bad.append((cls, pf, pb, want, how[:2]))
            if stop_at_first:
                break
    return fwd, seen, bad


def test_lin_move_set_rate_ratios():
    lat, L = _melt()
    rng = np.random.default_rng(1)
    sites = rng.choice(np.flatnonzero(lat & 0xFF), 500, replace=False)
    fwd, seen, bad = _violations(lat, L, sites)
    assert len(fwd) > 300
    assert not bad, bad[:5]
    assert all(seen[c] > 0 for c in CLASSES), seen


@pytest.mark.parametrize("which", [1, 2, 3])
def test_planted_wrong_gate_breaks_rate_ratios(which):
    """The enumeration can tell: with one gate of the oracle's TransMove
    changed, some transition's forward/reverse ratio leaves 2^(change in
    stored length)."""
    lat, L = _melt()
    rng = np.random.default_rng(1)
    sites = rng.choice(np.flatnonzero(lat & 0xFF), 500, replace=False)
    with planted(which):
        _, _, bad = _violations(lat, L, sites, stop_at_first=True)
    assert bad, f"planted mistake {which} not detected"


def test_four_monomer_chain_moves_connect_every_state():
    """The moves connect all 1812 conformations a GPU lattice can hold for one
    chain of four monomers (the state space of the stationary GPU test)."""
    states = lin_states(4)
    assert len(states) == 1812
    got = lin_reachable(4, tuple(int(b) for b in states[0]))
    assert got == {tuple(int(b) for b in s) for s in states}
