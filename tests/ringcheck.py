"""Independent checks of ring lattices (test infrastructure).

Works on the row-major site bytes (t fastest) directly, without the host
library: bond displacement of the FCC code, in-degree counting, ring walks
and the cyclic label sequence of every ring.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DT = np.array([(b & 1) - (b >> 3) for b in range(16)])
DU = np.array([((b >> 1) & 1) - (b >> 3) for b in range(16)])
DV = np.array([((b >> 2) & 1) - (b >> 3) for b in range(16)])
VALID = np.array([b not in (0, 3, 12, 15) for b in range(16)])


def targets(lat, L):
    LT, LU, LV = L
    idx = np.arange(lat.size)
    t, u, v = idx % LT, (idx // LT) % LU, idx // (LT * LU)
    b = lat & 0xF
    return ((t + DT[b]) % LT) + LT * (((u + DU[b]) % LU) + LU * ((v + DV[b]) % LV))


def site_invariants(lat, L):
    """Per-site encoding rules and the in-degree rule (every occupied site
    is the bond target of exactly one occupied site).  Returns monomer count."""
    lat = np.asarray(lat, dtype=np.uint8)
    b = lat & 0xF
    occ = b != 0
    sl = (lat >> 6) & 1
    assert not np.any(lat & 0x80), "bit 7 set"
    assert np.all(VALID[b[occ]]), "invalid bond code"
    assert not np.any((sl == 1) & ~occ), "stored length on an empty site"
    assert not np.any(((lat >> 5) & 1) & (sl == 0)), "SL label bit without stored length"
    assert not np.any(((lat >> 4) & 1) & ~occ), "label bit on an empty site"
    tg = targets(lat, L)[occ]
    indeg = np.bincount(tg, minlength=lat.size)
    assert np.array_equal(indeg > 0, occ), "bond into an empty site / occupied site without incoming bond"
    assert indeg.max(initial=0) <= 1, "two bonds into one site"
    return int(occ.sum() + sl.sum())


def ring_label_sequences(lat, L):
    """Walk every ring; return sorted list of canonical cyclic label strings
    (rotated to start right after the head 1,0,0,0,0) and ring lengths."""
    lat = np.asarray(lat, dtype=np.uint8)
    tg = targets(lat, L)
    occ = (lat & 0xF) != 0
    seen = np.zeros(lat.size, dtype=bool)
    out = []
    for s0 in np.flatnonzero(occ):
        if seen[s0]:
            continue
        seq = []
        s = s0
        while not seen[s]:
            seen[s] = True
            x = int(lat[s])
            if x & 0x40:
                seq.append((x >> 5) & 1)
            seq.append((x >> 4) & 1)
            s = int(tg[s])
        assert s == s0, "chain does not close"
        n = len(seq)
        heads = [i for i in range(n) if [seq[(i + k) % n] for k in range(5)] == [1, 0, 0, 0, 0]]
        assert len(heads) == 1, f"ring with {len(heads)} label heads"
        h = (heads[0] + 5) % n
        out.append("".join(str(seq[(h + k) % n]) for k in range(n)))
    return sorted(out)


_orc = None


def oracle():
    global _orc
    if _orc is None:
        lib = C.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
        lib.orc_ring_run.argtypes = [C.c_void_p] + [C.c_int] * 9 + [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int]
        lib.orc_philox4x32_10.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _orc = lib
    return _orc


def replay_ring(lat, L, tile, sublat, n_iter, origins, tau0, seed, stream=0, order=0, rules=0, counts=False):
    """The oracle's replay; with counts=True also (empty, leaving the tile, accepted) of its attempts."""
    lat = np.ascontiguousarray(lat, dtype=np.uint8).copy()
    o = np.ascontiguousarray(origins, dtype=np.int32).reshape(-1, 3)
    lib = oracle()
    lib.orc_ring_run_rules.argtypes = [C.c_void_p] + [C.c_int] * 9 + [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32,
                                                                       C.c_int, C.c_int]
    lib.orc_ring_counts.argtypes = [C.c_int, C.c_void_p]
    if counts:
        lib.orc_ring_counts(1, None)
    rc = lib.orc_ring_run_rules(lat.ctypes.data, *L, *tile, int(sublat), int(n_iter), len(o), o.ctypes.data,
                                int(tau0), int(seed), int(stream), int(order), int(rules))
    assert rc == 0
    if counts:
        c = np.zeros(3, np.uint64)
        lib.orc_ring_counts(0, c.ctypes.data)
        return lat, tuple(int(x) for x in c)
    return lat


# ---------------- linear chains (uint16 site words) ----------------
def lin_site_invariants(lat, L):
    """Encoding rules and bond consistency of a linear-chain lattice:
    next(x) = b  <=>  prev(x+b) = b.  Returns (monomers, chains)."""
    lat = np.asarray(lat, dtype=np.uint16)
    LT, LU, LV = L
    nxt = (lat & 0xF).astype(np.int64)
    prv = ((lat >> 4) & 0xF).astype(np.int64)
    sl = (lat >> 8) & 1
    occ = (lat & 0xFF) != 0
    assert not np.any(lat & 0xF800), "unused bits set"
    assert np.all(VALID[nxt[nxt != 0]]) and np.all(VALID[prv[prv != 0]]), "invalid bond code"
    assert not np.any((sl == 1) & ~occ), "stored length on an empty site"
    assert not np.any(((lat >> 10) & 1) & (sl == 0)), "SL label without stored length"
    idx = np.arange(lat.size)
    t, u, v = idx % LT, (idx // LT) % LU, idx // (LT * LU)
    has = nxt != 0
    tg = ((t + DT[nxt]) % LT) + LT * (((u + DU[nxt]) % LU) + LU * ((v + DV[nxt]) % LV))
    assert np.array_equal(prv[tg[has]], nxt[has]), "next/prev mismatch"
    assert int(np.count_nonzero(prv != 0)) == int(np.count_nonzero(has)), "dangling prev bonds"
    starts = int(np.count_nonzero(occ & (prv == 0)))
    ends = int(np.count_nonzero(occ & (nxt == 0)))
    assert starts == ends
    return int(occ.sum() + sl.sum()), starts


def lin_label_sequences(lat, L):
    """Label sequence of every chain read from its start (monomer 0)."""
    lat = np.asarray(lat, dtype=np.uint16)
    LT, LU, LV = L
    out = []
    for s0 in np.flatnonzero(((lat & 0xFF) != 0) & (((lat >> 4) & 0xF) == 0)):
        seq = []
        s = int(s0)
        for _ in range(lat.size + 1):
            w = int(lat[s])
            if (w >> 8) & 1:
                seq.append((w >> 10) & 1)
            nx = w & 0xF
            if nx == 0:
                seq.append(-1)  # end monomer: its label bit is not stored at start-up
                break
            seq.append((w >> 9) & 1)
            t, u, v = s % LT, (s // LT) % LU, s // (LT * LU)
            s = ((t + DT[nx]) % LT) + LT * (((u + DU[nx]) % LU) + LU * ((v + DV[nx]) % LV))
        out.append("".join("x" if b < 0 else str(b) for b in seq))
    return sorted(out)


_lin_typed = False


def replay_lin(lat, L, tile, sublat, n_iter, origins, tau0, seed, stream=0, order=0, counts=False):
    """The oracle's replay; with counts=True also (empty, accepted) of its attempts."""
    global _lin_typed
    lib = oracle()
    if not _lin_typed:
        lib.orc_lin_run.argtypes = [C.c_void_p] + [C.c_int] * 9 + [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int]
        lib.orc_lin_counts.argtypes = [C.c_int, C.c_void_p]
        _lin_typed = True
    lat = np.ascontiguousarray(lat, dtype=np.uint16).copy()
    o = np.ascontiguousarray(origins, dtype=np.int32).reshape(-1, 3)
    if counts:
        lib.orc_lin_counts(1, None)
    rc = lib.orc_lin_run(lat.ctypes.data, *L, *tile, int(sublat), int(n_iter), len(o), o.ctypes.data, int(tau0),
                         int(seed), int(stream), int(order))
    assert rc == 0
    if counts:
        c = np.zeros(2, np.uint64)
        lib.orc_lin_counts(0, c.ctypes.data)
        return lat, tuple(int(x) for x in c)
    return lat
