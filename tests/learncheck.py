"""Test helpers for the learned proposal: the FP64 network oracle, patch
construction, and the exact state space of one short ring."""
import ctypes as C
import itertools

import numpy as np

from ringcheck import oracle

NIN, NMOVE, NPARAM = 144, 11, 14416
VALID = [b for b in range(1, 15) if b not in (3, 12)]


def delta(b):
    w = (b >> 3) & 1
    return ((b & 1) - w, ((b >> 1) & 1) - w, ((b >> 2) & 1) - w)


def lib():
    o = oracle()
    o.orc_mlp_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    o.orc_mlp_grad.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    o.orc_mlp_grad_d.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    o.orc_ring_move.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                C.c_uint32, C.c_int]
    return o


def mlp_forward(prm, x):
    prm = np.ascontiguousarray(prm, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.int8).reshape(-1, NIN)
    prob = np.empty((len(x), NMOVE))
    lib().orc_mlp_forward(prm.ctypes.data, x.ctypes.data, len(x), prob.ctypes.data)
    return prob


def mlp_grad(prm, x, y, double=False):
    x = np.ascontiguousarray(x, dtype=np.int8).reshape(-1, NIN)
    y = np.ascontiguousarray(y, dtype=np.int32)
    g = np.empty(NPARAM)
    loss = C.c_double()
    if double:
        p = np.ascontiguousarray(prm, dtype=np.float64)
        lib().orc_mlp_grad_d(p.ctypes.data, x.ctypes.data, y.ctypes.data, len(x), g.ctypes.data, C.byref(loss))
    else:
        p = np.ascontiguousarray(prm, dtype=np.float32)
        lib().orc_mlp_grad(p.ctypes.data, x.ctypes.data, y.ctypes.data, len(x), g.ctypes.data, C.byref(loss))
    return g, loss.value


def bf16_round(a):
    """round fp32 values to the nearest bf16 (ties to even), as fp32"""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def random_params(rng, scale=1.0):
    p = np.zeros(NPARAM, np.float32)
    a1, a2, a3 = np.sqrt(6 / (144 + 64)), np.sqrt(6 / 128), np.sqrt(6 / 80)
    p[0:9216] = rng.uniform(-a1, a1, 9216) * scale
    p[9216:13312] = rng.uniform(-a2, a2, 4096) * scale
    p[13312:13376] = rng.uniform(-0.1, 0.1, 64)
    p[13376:14400] = rng.uniform(-a3, a3, 1024) * scale
    p[14400:14416] = rng.uniform(-0.1, 0.1, 16)
    return p


def random_features(rng, n):
    x = np.zeros((n, NIN), np.int8)
    x[:, :125] = rng.random((n, 125)) < 0.8
    x[np.arange(n), 125 + rng.choice(VALID, n)] = 1
    x[:, 141] = rng.random(n) < 0.3
    x[:, 142] = rng.random(n) < 0.3
    x[:, 143] = 1
    return x


def patches_of(lat, L, sites):
    """5^3 patches (index t + 5u + 25v, site at the centre) of a periodic ring lattice"""
    LT, LU, LV = L
    lat3 = np.asarray(lat, np.uint8).reshape(LV, LU, LT)
    out = np.empty((len(sites), 125), np.uint8)
    d = np.arange(-2, 3)
    for n, x in enumerate(sites):
        t, u, v = x % LT, (x // LT) % LU, x // (LT * LU)
        blk = lat3[np.ix_((v + d) % LV, (u + d) % LU, (t + d) % LT)]
        out[n] = blk.reshape(-1)
    return out


def features_of(patch):
    """The feature row of a patch, restated from the definition in include/elp_gpu.h"""
    x = np.zeros(NIN, np.int8)
    x[:125] = patch != 0
    s = int(patch[62])
    if s and s != 0xFF:
        b = s & 0xF
        x[125 + b] = 1
        x[141] = 1 if s & 0x40 else 0
        dt, du, dv = delta(b)
        n = int(patch[62 + dt + 5 * du + 25 * dv])
        x[142] = 1 if (n != 0xFF and n & 0x40) else 0
    elif s == 0xFF:
        x[141] = 0
    x[143] = 1
    return x


def ring_states(n):
    """All labelled ring conformations of n monomers (bond sequences b_0..b_{n-1}
    from monomer 0, 0 = stored length) that a GPU lattice can hold: closed,
    no two stored-length bonds in a row (at most two monomers per site), and
    distinct sites otherwise.  Returns an int array (states, n)."""
    codes = np.array([0] + VALID)
    d = np.array([(0, 0, 0)] + [delta(b) for b in VALID])
    idx = np.array(list(itertools.product(range(13), repeat=n)), dtype=np.int8)
    steps = d[idx]                                         # (S, n, 3)
    closed = np.all(steps.sum(axis=1) == 0, axis=1)
    idx, steps = idx[closed], steps[closed]
    z = idx == 0
    ok = ~np.any(z & np.roll(z, -1, axis=1), axis=1)
    idx, steps = idx[ok], steps[ok]
    pos = np.concatenate([np.zeros((len(idx), 1, 3), int), np.cumsum(steps, axis=1)[:, :-1]], axis=1)
    good = np.ones(len(idx), bool)
    for i in range(n):
        for j in range(i + 1, n):
            same = np.all(pos[:, i] == pos[:, j], axis=1)
            adjacent_sl = ((j == i + 1) & (idx[:, i] == 0)) | ((i == 0) & (j == n - 1) & (idx[:, n - 1] == 0))
            good &= ~same | adjacent_sl
    return codes[idx[good]]


def ring_observables(bonds):
    """(number of stored-length bonds, Rg^2 in bond units x 36 rounded) of bond sequences"""
    bonds = np.atleast_2d(bonds)
    d = np.array([(0, 0, 0) if b == 0 else delta(b) for b in range(16)])
    steps = d[bonds]
    pos = np.concatenate([np.zeros((len(bonds), 1, 3)), np.cumsum(steps, axis=1)[:, :-1]], axis=1)
    t, u, v = pos[..., 0], pos[..., 1], pos[..., 2]
    xyz = np.stack([t + u - v, t - u, v], axis=-1) / np.sqrt(2.0)
    c = xyz.mean(axis=1, keepdims=True)
    rg2 = ((xyz - c) ** 2).sum(axis=2).mean(axis=1)
    return (bonds == 0).sum(axis=1), np.rint(rg2 * 36).astype(int)


def ring_lattice(bonds, L=16, centre=(8, 8, 8)):
    """Ring lattice (uint8, L^3) holding one ring with the given bond sequence
    from monomer 0, whose label bit is 1 (all others 0): a site carrying a
    bond stores its code, the label of its carrying monomer (bit 4) and, with a
    stored-length monomer before it on the same site, bit 6 and that monomer's
    label (bit 5)."""
    n = len(bonds)
    lat = np.zeros(L ** 3, np.uint8)
    p = np.array(centre)
    pos = []
    for i in range(n):
        pos.append(p.copy())
        if bonds[i]:
            p = p + np.array(delta(bonds[i]))
    for i in range(n):
        if bonds[i] == 0:
            continue                                    # stored-length monomer: recorded with its carrier
        t, u, v = pos[i] % L
        x = t + L * (u + L * v)
        byte = bonds[i] | (0x10 if i == 0 else 0)
        j = (i - 1) % n
        if bonds[j] == 0:                               # monomer j shares this site, before i
            byte |= 0x40 | (0x20 if j == 0 else 0)
        lat[x] = byte
    return lat


def ring_decode(lat, n, L=16):
    """Bond sequence from monomer 0 (the monomer with label 1) of the single
    ring in a lattice, or None if the lattice does not hold one valid ring."""
    occ = np.flatnonzero(lat)
    start = None
    for x in occ:
        s = int(lat[x])
        if s & 0x20:
            start, first_sl = x, True
        elif s & 0x10:
            start, first_sl = x, False
    if start is None:
        return None
    out = []
    x, on_sl = start, first_sl
    while len(out) < n:
        s = int(lat[x])
        if not s:
            return None
        if on_sl:
            out.append(0)
            on_sl = False
            continue
        b = s & 0xF
        out.append(b)
        dt, du, dv = delta(b)
        t, u, v = x % L, (x // L) % L, x // (L * L)
        x = (t + dt) % L + L * ((u + du) % L + L * ((v + dv) % L))
        on_sl = bool(int(lat[x]) & 0x40)
    if x != start or on_sl != first_sl:
        return None
    return tuple(out)


def has_corner(bonds):
    """some two consecutive bonds add up to one bond (a TransBack is possible)"""
    n = len(bonds)
    for i in range(n):
        a, b = int(bonds[i]), int(bonds[(i + 1) % n])
        if not a or not b or ((a | b) != 0xF and (a & b)):
            continue
        c = (a & b) if (a | b) == 0xF else (a | b)
        if c not in (0, 3, 12, 15):
            return True
    return False


def ring_reachable(n):
    """States of ring_states(n) the move set connects: all but those without
    stored length and without a corner, where no move applies (and which no
    move produces, since TransForw always leaves a corner)."""
    s = ring_states(n)
    keep = [(r == 0).any() or has_corner(r) for r in s]
    return s[np.array(keep)]
