"""Test helpers for the linear-chain move set (test infrastructure): every
move of a site through the oracle (orc_lin_moves), the exact state space of
one short chain, its lattice encoding and decoding, and the set of states
the moves connect.

Site word of a linear lattice (uint16): bits 0-3 next bond, 4-7 prev bond
(the code of the bond entering the site, in chain direction), bit 8 stored
length, bit 9 label of the monomer carrying the next bond, bit 10 label of the
stored-length monomer (the first of the two on the site).
"""
import contextlib
import ctypes as C
import itertools

import numpy as np

from ringcheck import oracle

VALID = [b for b in range(1, 15) if b not in (3, 12)]
W = 4                                             # window edge: every touched site lies within +-1 of the move's site


def delta(b):
    w = (b >> 3) & 1
    return ((b & 1) - w, ((b >> 1) & 1) - w, ((b >> 2) & 1) - w)


_typed = False


def lib():
    global _typed
    o = oracle()
    if not _typed:
        o.orc_lin_moves.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
        o.orc_lin_move.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_uint32]
        o.orc_lin_plant.argtypes = [C.c_int]
        _typed = True
    return o


@contextlib.contextmanager
def planted(which):
    """Run the oracle's moves with a planted wrong gate (1: chain-end retraction
    gate 1/4 instead of 1/8; 2: interior retraction without its 1/2 gate;
    3: stored length split off to only two of the four common neighbours)."""
    o = lib()
    assert o.orc_lin_plant(which) == 0
    try:
        yield
    finally:
        o.orc_lin_plant(0)


def window_index(x, L):
    """Global indices (t fastest) of the W^3 periodic window whose (1,1,1) is site x."""
    LT, LU, LV = L
    t, u, v = x % LT, (x // LT) % LU, x // (LT * LU)
    d = np.arange(-1, W - 1)
    tt, uu, vv = np.meshgrid((t + d) % LT, (u + d) % LU, (v + d) % LV, indexing="ij")
    return (tt + LT * (uu + LU * vv)).transpose(2, 1, 0).reshape(-1)


def site_moves(lat, L, x):
    """All moves at site x of a periodic lattice: list of (k, changes), where
    k = 0 is DiffuseSL and k = 1 + r TransMove with random byte r, and changes
    is a tuple of (global index, new word) pairs (empty if nothing changed)."""
    idx = window_index(x, L)
    win = np.ascontiguousarray(lat[idx], dtype=np.uint16)
    out = np.empty((257, W ** 3), np.uint16)
    assert lib().orc_lin_moves(win.ctypes.data, W, W, W, 1, 1, 1, out.ctypes.data) == 0
    diff = out != win
    res = []
    for k in np.flatnonzero(diff.any(axis=1)):
        j = np.flatnonzero(diff[k])
        res.append((int(k), tuple(sorted(zip(idx[j].tolist(), out[k, j].tolist())))))
    return res


def n_stored(words):
    return int(np.count_nonzero((np.asarray(words, dtype=np.uint16) >> 8) & 1))


# ---------------- one short chain ----------------
def lin_states(n):
    """All conformations of one linear chain of n monomers that a GPU lattice
    can hold, as bond sequences b_0..b_{n-2} (b_i from monomer i to i+1,
    0 = stored length): no two stored-length bonds in a row (at most two
    monomers per site), distinct sites otherwise.  Returns an int array."""
    codes = np.array([0] + VALID)
    d = np.array([(0, 0, 0)] + [delta(b) for b in VALID])
    idx = np.array(list(itertools.product(range(13), repeat=n - 1)), dtype=np.int8)
    z = idx == 0
    ok = ~np.any(z[:, 1:] & z[:, :-1], axis=1) if n > 2 else np.ones(len(idx), bool)
    idx = idx[ok]
    pos = np.concatenate([np.zeros((len(idx), 1, 3), int), np.cumsum(d[idx], axis=1)], axis=1)
    good = np.ones(len(idx), bool)
    for i in range(n):
        for j in range(i + 1, n):
            same = np.all(pos[:, i] == pos[:, j], axis=1)
            if j == i + 1:
                same &= idx[:, i] != 0
            good &= ~same
    return codes[idx[good]]


def lin_lattice(bonds, L=16, start=(8, 8, 8), labels=None):
    """Lattice (uint16, L^3, t fastest) holding one chain with the given bond
    sequence from monomer 0 at `start`; labels[i] is monomer i's label bit
    (default: 1 for monomer 0, else 0)."""
    n = len(bonds) + 1
    lab = [1 if i == 0 else 0 for i in range(n)] if labels is None else list(labels)
    lat = np.zeros(L ** 3, np.uint16)
    p = np.array(start)
    pos = [p.copy()]
    for b in bonds:
        if b:
            p = p + np.array(delta(b))
        pos.append(p.copy())
    i = 0
    while i < n:
        t, u, v = pos[i] % L
        x = t + L * (u + L * v)
        prev = bonds[i - 1] if i > 0 else 0
        if i + 1 < n and bonds[i] == 0:               # monomers i, i+1 share the site
            nxt = bonds[i + 1] if i + 1 < n - 1 else 0
            lat[x] = nxt | prev << 4 | 1 << 8 | lab[i + 1] << 9 | lab[i] << 10
            i += 2
        else:
            nxt = bonds[i] if i < n - 1 else 0
            lat[x] = nxt | prev << 4 | lab[i] << 9
            i += 1
    return lat


def lin_decode(lat, n, L=16, labels=None):
    """(bond sequence from monomer 0, site of monomer 0) of the single chain
    in a lattice, or None if the lattice does not hold one chain of n monomers
    with the given label bits."""
    lab = [1 if i == 0 else 0 for i in range(n)] if labels is None else list(labels)
    lat = np.asarray(lat, dtype=np.uint16)
    occ = np.flatnonzero((lat & 0xFF) != 0)
    starts = [int(x) for x in occ if ((int(lat[x]) >> 4) & 0xF) == 0]
    if len(starts) != 1:
        return None
    x0 = x = starts[0]
    bonds, got = [], []
    seen = 0
    while True:
        w = int(lat[x])
        seen += 1
        if w & 0x100:
            got.append((w >> 10) & 1)
            bonds.append(0)
        got.append((w >> 9) & 1)
        nx = w & 0xF
        if nx == 0:
            break
        bonds.append(nx)
        dt, du, dv = delta(nx)
        t, u, v = x % L, (x // L) % L, x // (L * L)
        x = (t + dt) % L + L * ((u + du) % L + L * ((v + dv) % L))
        if seen > n:
            return None
    if len(got) != n or seen != len(occ) or got != lab:
        return None
    return tuple(bonds), x0


def lin_observables(bonds):
    """(stored-length bonds, Rg^2 in bond units x 36 rounded) of bond sequences."""
    bonds = np.atleast_2d(bonds)
    d = np.array([(0, 0, 0) if b == 0 else delta(b) for b in range(16)])
    steps = d[bonds]
    pos = np.concatenate([np.zeros((len(bonds), 1, 3)), np.cumsum(steps, axis=1)], axis=1)
    t, u, v = pos[..., 0], pos[..., 1], pos[..., 2]
    xyz = np.stack([t + u - v, t - u, v], axis=-1) / np.sqrt(2.0)
    c = xyz.mean(axis=1, keepdims=True)
    rg2 = ((xyz - c) ** 2).sum(axis=2).mean(axis=1)
    return (bonds == 0).sum(axis=1), np.rint(rg2 * 36).astype(int)


def lin_reachable(n, start, L=12):
    """Conformations connected to `start` by the oracle's moves (breadth-first,
    every site of the chain, DiffuseSL and every TransMove random byte), in a
    box large enough that no move meets a face or a periodic image."""
    c = (L // 2,) * 3
    todo, seen = [tuple(start)], {tuple(start)}
    while todo:
        cur = todo.pop()
        lat = lin_lattice(cur, L, c)
        for x in np.flatnonzero((lat & 0xFF) != 0):
            for _, ch in site_moves(lat, (L, L, L), int(x)):
                if not ch:
                    continue
                b = lat.copy()
                for g, w in ch:
                    b[g] = w
                dec = lin_decode(b, n, L)
                assert dec is not None, (cur, ch)
                if dec[0] not in seen:
                    seen.add(dec[0])
                    todo.append(dec[0])
    return seen


def chi2_weighted(samples, states, weight):
    """chi^2 of the joint histogram of (stored-length bonds, Rg^2) of sampled
    conformations against the law proportional to weight(n_SL) over `states`;
    bins with fewer than 10 expected samples are left out.  (chi2, dof, p)"""
    from scipy import stats
    nsl, rg = lin_observables(np.array(samples))
    ensl, erg = lin_observables(states)
    w = weight(ensl).astype(float)
    keys = sorted(set(zip(ensl.tolist(), erg.tolist())))
    expect = np.array([w[(ensl == a) & (erg == b)].sum() for a, b in keys]) / w.sum()
    got = np.array([np.sum((nsl == a) & (rg == b)) for a, b in keys], float)
    e = expect * len(samples)
    big = e >= 10
    chi2 = float(np.sum((got[big] - e[big]) ** 2 / e[big]))
    dof = int(big.sum()) - 1
    return chi2, dof, float(stats.chi2.sf(chi2, dof))
