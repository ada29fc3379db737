"""Transition-rate ratios of the ring move set, by exhaustive enumeration.

For sites of a randomised ring melt the oracle's single-move entry point
(orc_ring_move) is applied with every value of the random bits a move reads
(TransForw: 3 bits, TransBack: 3 -- bit 0 picks the fold, bits 1-2 gate the
forced fold under balanced rates --, DiffuseSL: none).  Each resulting state S'
is weighted by its probability per iteration (two TransForw substeps, one
DiffuseSL, one TransBack; every site equally likely to be attempted), and the
probability of the return S' -> S is enumerated the same way over every site
next to the changed ones.

From the reference rules (gpupol2/ringpol3.cu:92-218):
  * every move has a reverse move;
  * DiffuseSL, and TransForw / TransBack where at most one end of the bond
    holds stored length: forward rate = reverse rate;
  * TransForw splitting stored length off a bond whose two ends both hold
    stored length: rate 2 * 1/4 * 1/2 = 1/4 (site choice, which monomer
    moves), undone by a forced TransBack at rate 1: ratio 1/4; the forced
    TransBack itself: ratio 4.
The test pins these ratios for every transition found, so a wrong neighbour,
bit or rate in the oracle's moves (which the GPU kernels match bit for bit)
shows up as a missing reverse move or a ratio off its class.

Consequence (documented with the statistics): the ratios are not those of any
stationary weight.  From A = ...P(SL)->Pn(SL)->R..., TransForw at P gives
B = P->D->Pn(SL) (ratio 1/4); DiffuseSL Pn->R gives C = P->D->Pn->R(SL)
(ratio 1); the same C is reached from A by DiffuseSL Pn->R (1) and TransForw
at P with P alone holding stored length (1).  The product of ratios around
that cycle is 1/4, not 1, so the move set violates Kolmogorov's criterion:
the chain is not reversible and its equilibrium is not the uniform measure
over its states (stored length per site <= 2) that efpol's table targets.

With balanced rates (ELP_RULES_BALANCED: the forced TransBack accepted with
probability 1/4, and each substep of each tile drawing TransForw / DiffuseSL /
TransBack with probabilities 1/2, 1/4, 1/4 -- the proportions used as
weights here, which are each slot's marginal probabilities) every
ratio must be exactly 1: each substep is then a symmetric kernel, so the
uniform measure is stationary for every schedule.
"""
import ctypes as C

import numpy as np
import pytest

import elp
from ringcheck import oracle, replay_ring

T = 16
WEIGHT = {0: 2.0, 1: 1.0, 2: 1.0}                 # substeps per iteration: TF, DSL, TF, TB
RBITS = {0: 8, 1: 1, 2: 8}                        # random values each move distinguishes


def _lib():
    lib = oracle()
    lib.orc_ring_move.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_uint32, C.c_int]
    return lib


def _tuv(x):
    return x % T, (x // T) % T, x // (T * T)


def _step(x, b):
    t, u, v = _tuv(x)
    w = (b >> 3) & 1
    t, u, v = t + (b & 1) - w, u + ((b >> 1) & 1) - w, v + ((b >> 2) & 1) - w
    return t + T * (u + T * v) if 0 <= t < T and 0 <= u < T and 0 <= v < T else None


def _badd(a, b):
    return (a & b) if (a | b) == 0xF else (a | b)


def _moves_from(lib, lat, sites, rules):
    """{state bytes: [probability per iteration, [(type, site), ...]]} of the moves at the sites."""
    out = {}
    base = lat.tobytes()
    for x in sites:
        t, u, v = _tuv(x)
        for typ in (0, 1, 2):
            for r in range(RBITS[typ]):
                a = lat.copy()
                lib.orc_ring_move(a.ctypes.data, T, T, T, typ, t, u, v, r, rules)
                b = a.tobytes()
                if b != base:
                    e = out.setdefault(b, [0.0, []])
                    e[0] += WEIGHT[typ] / RBITS[typ]
                    e[1].append((typ, int(x)))
    return out


def _near(idx):
    s = set()
    for x in idx:
        t, u, v = _tuv(x)
        for dt in (-1, 0, 1):
            for du in (-1, 0, 1):
                for dv in (-1, 0, 1):
                    a, b, c = t + dt, u + du, v + dv
                    if 0 <= a < T and 0 <= b < T and 0 <= c < T:
                        s.add(a + T * (b + T * c))
    return sorted(s)


def _expected_ratio(lat, typ, x, rules):
    if rules:
        return 1.0
    s = int(lat[x])
    n = _step(x, s & 0xF)
    if typ == 0:                                   # TransForw: both ends of the bond hold stored length
        return 0.25 if (s & 0x40) and (lat[n] & 0x40) else 1.0
    if typ == 2:                                   # TransBack: exactly one end of the new bond holds it
        w = _step(x, _badd(s & 0xF, int(lat[n]) & 0xF))
        return 4.0 if bool(s & 0x40) != bool(lat[w] & 0x40) else 1.0
    return 1.0


@pytest.mark.parametrize("rules", [elp.RULES_REFERENCE, elp.RULES_BALANCED])
def test_ring_move_set_rate_ratios(rules):
    h = elp.RingHost(N=30, L=T, density=1.2, seed=5, box_unit=16)
    lat0 = h.lattice()
    h.load_shift((T, T, T))
    lat = replay_ring(lat0, h.L, (T, T, T), elp.SUBLAT_DENSE, 13, h.origins((T, T, T), 30), 0, h.seed)
    lib = _lib()
    rng = np.random.default_rng(1)
    sites = rng.choice(np.flatnonzero(lat), 800, replace=False)
    fwd = _moves_from(lib, lat, sites, rules)
    assert len(fwd) > 400
    base = lat.tobytes()
    seen = {0.25: 0, 1.0: 0, 4.0: 0}
    for sb, (pf, how) in fwd.items():
        assert len(how) <= RBITS[how[0][0]]
        s2 = np.frombuffer(sb, dtype=np.uint8).copy()
        changed = np.flatnonzero(s2 != lat)
        assert 2 <= len(changed) <= 3
        back = _moves_from(lib, s2, _near(changed), rules).get(base, [0.0])[0]
        assert back > 0, f"no reverse move for {how}"
        want = _expected_ratio(lat, *how[0], rules)
        assert np.isclose(pf / back, want), (pf / back, want, how)
        seen[want] += 1
    if rules == elp.RULES_BALANCED:
        assert seen[1.0] == len(fwd)
    else:
        assert all(n > 0 for n in seen.values()), seen
