"""CPU tests of the linear-chain host model and oracle replay."""
import numpy as np
import pytest

import elp
from ringcheck import lin_label_sequences, lin_site_invariants, replay_lin


@pytest.mark.parametrize("N,L,npol", [(200, 32, 197), (50, 32, 786), (500, 32, 79)])
def test_lin_host_fresh_start(N, L, npol):
    """Cube start (N >= 100) and two-row bar start (N < 100) of
    gpupol/gpupol_init.c:221-238; one start and one end per chain."""
    h = elp.LinHost(N=N, L=L, density=1.2, seed=3)
    assert h.n_pol == npol and h.N == N
    lat = h.lattice()
    assert lin_site_invariants(lat, h.L) == (npol * N, npol)
    seqs = lin_label_sequences(lat, h.L)
    assert all(s.startswith("10000") for s in seqs)   # head at monomer 0
    assert len({s[5:23] for s in seqs}) == npol         # distinct 18-bit ids


@pytest.mark.parametrize("tile,sublat,n_iter", [((16, 16, 16), 0, 64), ((16, 16, 16), 1, 16), ((32, 32, 32), 0, 64),
                                                ((32, 32, 16), 1, 16)])
def test_lin_replay_order_independent_and_invariant(tile, sublat, n_iter):
    h = elp.LinHost(N=200, L=32, density=1.2, seed=4)
    lat = h.lattice()
    seq0 = lin_label_sequences(lat, h.L)
    org = np.random.default_rng(2).integers(0, min(tile), size=(8, 3)).astype(np.int32)
    a = replay_lin(lat, h.L, tile, sublat, n_iter, org, 3, 11, 0, order=0)
    b = replay_lin(lat, h.L, tile, sublat, n_iter, org, 3, 11, 0, order=1)
    c = replay_lin(lat, h.L, tile, sublat, n_iter, org, 3, 11, 0, order=2)
    assert np.array_equal(a, b) and np.array_equal(a, c)
    assert np.count_nonzero(a != lat) > lat.size // 4
    assert lin_site_invariants(a, h.L) == (197 * 200, 197)
    assert lin_label_sequences(a, h.L) == seq0
    h.set_lattice(a)


def test_lin_replay_end_moves_happen():
    """Chain ends grow out of and retract into stored length: after a run
    the set of end sites has changed for most chains."""
    h = elp.LinHost(N=50, L=32, density=1.2, seed=1)
    lat = h.lattice()
    org = np.random.default_rng(3).integers(0, 16, size=(40, 3)).astype(np.int32)
    a = replay_lin(lat, h.L, (16, 16, 16), 1, 16, org, 0, 2, 0)
    ends0 = set(np.flatnonzero(((lat & 0xFF) != 0) & ((lat & 0xF) == 0)))
    ends1 = set(np.flatnonzero(((a & 0xFF) != 0) & ((a & 0xF) == 0)))
    assert len(ends0 - ends1) > len(ends0) // 4
