"""Self-avoiding-walk start of the linear-chain melt (host/lin_saw.c): every
chain is a walk over distinct free sites, stored length is at most one unit
per site and never on the last site, the lattice encoding is consistent, the
start is fixed by the seed, and the move oracle runs from it."""
import numpy as np
import pytest

import elp
from ringcheck import DT, DU, DV, lin_label_sequences, lin_site_invariants, replay_lin

DIRS = (0x1, 0x2, 0x4, 0x5, 0x6, 0x7, 0x8, 0x9, 0xA, 0xB, 0xD, 0xE)


def _sites_per_chain(N, npol, nsite, frac):
    m = int(frac * nsite / npol + 0.5)
    return max(min(m, N), (N + 2) // 2)


def _check_chains(h, frac):
    LT, LU, LV = h.L
    nsite = LT * LU * LV
    m = _sites_per_chain(h.N, h.n_pol, nsite, frac)
    owner = np.full(nsite, -1, np.int64)
    steps = np.zeros(16, np.int64)
    for pid in range(h.n_pol):
        (t, u, v), bonds = h.polymer(pid)
        assert len(bonds) == h.N and bonds[-1] == 0xF
        b = bonds[:-1]
        assert np.count_nonzero(b) == m - 1, "sites per chain"
        assert not np.any((b[:-1] == 0) & (b[1:] == 0)), "two stored-length units on one site"
        assert b[-1] != 0, "stored length on the last site"
        for bond in b:
            if bond == 0:
                continue
            s = t + LT * (u + LU * v)
            assert owner[s] == -1, "site visited twice"
            owner[s] = pid
            steps[bond] += 1
            t, u, v = (t + DT[bond]) % LT, (u + DU[bond]) % LU, (v + DV[bond]) % LV
        s = t + LT * (u + LU * v)
        assert owner[s] == -1, "site visited twice"
        owner[s] = pid
    assert np.count_nonzero(owner >= 0) == h.n_pol * m
    return owner, steps


@pytest.mark.parametrize("N,L,frac", [(100, 32, 0.7), (50, 32, 0.6), (200, (32, 64, 32), 0.8), (30, 32, 0.5)])
def test_saw_start_is_self_avoiding(N, L, frac):
    h = elp.LinHost(N=N, L=L, density=1.2, seed=5, init="saw", site_fraction=frac)
    nsite = h.nsite
    npol = int(1.2 * nsite / N + 0.5)
    assert h.n_pol == npol and h.N == N
    lat = h.lattice()
    assert lin_site_invariants(lat, h.L) == (npol * N, npol)
    seqs = lin_label_sequences(lat, h.L)
    assert all(s.startswith("10000") for s in seqs)
    assert len({s[5:23] for s in seqs}) == npol
    owner, steps = _check_chains(h, frac)
    assert np.array_equal(owner >= 0, (lat & 0xFF) != 0), "lattice occupancy is the union of the walks"
    # steps are drawn uniformly among free neighbours: every FCC direction is used about equally
    used = steps[list(DIRS)] / steps.sum()
    assert used.min() > 0.5 / 12 and used.max() < 1.5 / 12, used


def test_saw_start_seeded():
    a = elp.LinHost(N=100, L=32, density=1.2, seed=9, init="saw").lattice()
    b = elp.LinHost(N=100, L=32, density=1.2, seed=9, init="saw").lattice()
    c = elp.LinHost(N=100, L=32, density=1.2, seed=10, init="saw").lattice()
    ref = elp.LinHost(N=100, L=32, density=1.2, seed=9).lattice()
    assert np.array_equal(a, b)
    assert np.count_nonzero(a != c) > a.size // 4
    assert np.count_nonzero(a != ref) > a.size // 4


def test_saw_start_rejects_overfull_box():
    with pytest.raises(elp.ElpError):
        elp.LinHost(N=100, L=32, density=1.2, seed=1, init="saw", site_fraction=0.97)
    with pytest.raises(elp.ElpError):
        elp.RingHost(N=100, L=32, density=1.2, seed=1, init="saw")


def test_saw_start_full_size():
    """Linear target size: L=128, N=500, density 1.2 -> 5033 chains."""
    h = elp.LinHost(N=500, L=128, density=1.2, seed=1, init="saw")
    assert h.n_pol == 5033
    assert lin_site_invariants(h.lattice(), h.L) == (5033 * 500, 5033)
    _check_chains(h, 0.7)


@pytest.mark.parametrize("tile,sublat,n_iter", [((16, 16, 16), 1, 16), ((32, 32, 32), 0, 64)])
def test_replay_from_saw_start(tile, sublat, n_iter):
    h = elp.LinHost(N=100, L=32, density=1.2, seed=4, init="saw")
    lat = h.lattice()
    seq0 = lin_label_sequences(lat, h.L)
    org = h.origins(tile, 8)
    a = replay_lin(lat, h.L, tile, sublat, n_iter, org, 0, h.seed)
    assert np.count_nonzero(a != lat) > lat.size // 8
    assert lin_site_invariants(a, h.L) == (h.n_pol * 100, h.n_pol)
    assert lin_label_sequences(a, h.L) == seq0
    h.set_lattice(a)
