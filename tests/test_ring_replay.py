"""CPU tests of the ring oracle (C replay) and the ring host model."""
import ctypes as C

import numpy as np
import pytest

import elp
from ringcheck import oracle, replay_ring, ring_label_sequences, site_invariants

KAT = [  # Random123 kat_vectors, philox4x32 10 rounds: ctr, key, expected
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(ctr, key, want):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    oracle().orc_philox4x32_10(c, k, o)
    assert list(o) == want


def _state(N, L, seed):
    h = elp.RingHost(N=N, L=L, density=1.2, seed=seed)
    return h, h.lattice()


def test_host_fresh_start_invariants():
    h, lat = _state(100, 32, 1)
    assert h.L == (32, 32, 32) and h.n_pol == 393 and h.N == 100
    assert site_invariants(lat, h.L) == 393 * 100
    seqs = ring_label_sequences(lat, h.L)
    assert len(seqs) == 393 and all(len(s) == 100 for s in seqs)
    # ids: first 18 bits after the head are distinct labels
    ids = {int(s[:18], 2) for s in seqs}
    assert len(ids) == 393


def test_host_stored_length_at_most_one_per_site_and_reshuffled():
    h, lat = _state(100, 32, 5)
    sl = (lat >> 6) & 1
    occ = (lat & 0xF) != 0
    # cube rings of 4x3x6=72 sites + 28 stored-length units, spread by the shuffle
    assert int(sl.sum()) == 393 * 28
    assert int(occ.sum()) == 393 * 72


@pytest.mark.parametrize("tile,sublat,n_iter", [((16, 16, 16), 0, 51), ((16, 16, 16), 1, 13), ((32, 32, 32), 0, 51),
                                                ((32, 32, 16), 1, 7)])
def test_replay_slot_order_independent(tile, sublat, n_iter):
    """Slots of one substep commute: ascending, descending and shuffled
    visiting orders must give the same lattice (brute-force check of the
    sub-lattice independence the kernel relies on)."""
    h, lat = _state(100, 32, 3)
    org = np.array([[3, 7, 1], [0, 0, 0], [tile[0] - 1, 2, tile[2] - 1]], dtype=np.int32)
    a = replay_ring(lat, h.L, tile, sublat, n_iter, org, 10, 77, 0, order=0)
    b = replay_ring(lat, h.L, tile, sublat, n_iter, org, 10, 77, 0, order=1)
    c = replay_ring(lat, h.L, tile, sublat, n_iter, org, 10, 77, 0, order=2)
    assert not np.array_equal(a, lat)
    assert np.array_equal(a, b) and np.array_equal(a, c)


def test_dense_sublattice_neighbourhoods_disjoint():
    """Closed form behind the dense schedule: every nonzero vector of
    4Z^3 + <(1,3,2)> has squared FCC length >= 10, i.e. is no sum of two
    bonds, so the 13-site move neighbourhoods of one coset never overlap."""
    def norm(t, u, v):
        return (t + u - v) ** 2 + (t - u) ** 2 + v ** 2
    bonds = [(b & 1) - (b >> 3) for b in range(16)]
    nb = {((b & 1) - (b >> 3), ((b >> 1) & 1) - (b >> 3), ((b >> 2) & 1) - (b >> 3))
          for b in range(16) if b not in (0, 3, 12, 15)}
    assert len(nb) == 12 and all(norm(*x) == 2 for x in nb)
    two = {(a[0] + b[0], a[1] + b[1], a[2] + b[2]) for a in nb | {(0, 0, 0)} for b in nb | {(0, 0, 0)}}
    for j in range(4):
        for a in range(-2, 3):
            for b in range(-2, 3):
                for c in range(-2, 3):
                    vec = (j + 4 * a, (3 * j) % 4 + 4 * b, (2 * j) % 4 + 4 * c)
                    if vec == (0, 0, 0):
                        continue
                    assert vec not in two
                    assert norm(*vec) >= 10
    del bonds


def test_no_denser_sublattice_than_dense():
    """The dense schedule's sub-lattice is the densest one made of cosets of
    4Z^3 whose move neighbourhoods stay disjoint: among all subgroups of
    (Z/4)^3, the ones whose nonzero elements (every lift into [-4, 7]^3) avoid
    the differences of two neighbourhood offsets have 1, 2 or 4 elements, and
    the one with 4 is <(1,3,2)>."""
    nb = {((b & 1) - (b >> 3), ((b >> 1) & 1) - (b >> 3), ((b >> 2) & 1) - (b >> 3))
          for b in range(16) if b not in (0, 3, 12, 15)} | {(0, 0, 0)}
    diff = {(a[0] - b[0], a[1] - b[1], a[2] - b[2]) for a in nb for b in nb}
    assert len(diff) == 55
    pts = [(t, u, v) for t in range(4) for u in range(4) for v in range(4)]

    def join(group, g):
        # the subgroup generated by group and g (abelian: group + <g>)
        return frozenset(((x[0] + k * g[0]) % 4, (x[1] + k * g[1]) % 4, (x[2] + k * g[2]) % 4)
                         for x in group for k in range(4))

    def span(gens):
        group = frozenset({(0, 0, 0)})
        for g in gens:
            group = join(group, g)
        return group

    # every subgroup, built up one generator at a time from the trivial one
    groups, frontier = set(), {span([])}
    while frontier:
        groups |= frontier
        frontier = {join(g, x) for g in frontier for x in pts if x not in g} - groups

    def independent(group):
        return all((x[0] + 4 * i, x[1] + 4 * j, x[2] + 4 * k) not in diff
                   for x in group if x != (0, 0, 0) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1))
    good = [g for g in groups if independent(g)]
    assert sorted(len(g) for g in good) == [1, 2, 4]
    assert max(good, key=len) == span([(1, 3, 2)])


@pytest.mark.parametrize("tile,sublat,n_iter", [((16, 16, 16), 0, 51), ((16, 16, 16), 1, 13)])
def test_replay_preserves_rings_and_labels(tile, sublat, n_iter):
    h, lat = _state(100, 32, 2)
    before = ring_label_sequences(lat, h.L)
    rng = np.random.default_rng(0)
    org = rng.integers(0, tile[0], size=(20, 3)).astype(np.int32)
    out = replay_ring(lat, h.L, tile, sublat, n_iter, org, 0, 1234, 0)
    assert site_invariants(out, h.L) == 393 * 100
    assert ring_label_sequences(out, h.L) == before
    moved = np.count_nonzero(out != lat)
    assert moved > lat.size // 10
    # host recovery agrees (labels -> ids, tuv writer)
    h.set_lattice(out)


def test_replay_counter_based_split_runs():
    """Counter-based RNG: one run of 6 time units equals 2+4 with tau0 carried."""
    h, lat = _state(100, 32, 4)
    org = np.array([[i % 16, (3 * i) % 16, (5 * i) % 16] for i in range(6)], dtype=np.int32)
    a = replay_ring(lat, h.L, (16, 16, 16), 0, 51, org, 100, 9, 0)
    b = replay_ring(lat, h.L, (16, 16, 16), 0, 51, org[:2], 100, 9, 0)
    b = replay_ring(b, h.L, (16, 16, 16), 0, 51, org[2:], 102, 9, 0)
    assert np.array_equal(a, b)
