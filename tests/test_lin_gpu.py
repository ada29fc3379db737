"""GPU parity of the linear-chain engine against the packed-layout C replay."""
import sys

import numpy as np
import pytest

import elp
from ringcheck import lin_label_sequences, lin_site_invariants, replay_lin

pytestmark = pytest.mark.gpu

SCHEDULES = [
    ((32, 32, 32), elp.SUBLAT_CUBIC, 64),
    ((16, 16, 16), elp.SUBLAT_CUBIC, 64),
    ((16, 16, 16), elp.SUBLAT_DENSE, 16),
    ((32, 32, 16), elp.SUBLAT_DENSE, 16),
]


@pytest.mark.parametrize("tile,sublat,n_iter", SCHEDULES)
# Chain length sets how many waves hold an occupied chain end (the kernel's
# TransMove has a full form for those and a split/fold form for the rest):
# N=50 most waves, N=200 about a third, N=2000 (20 chains, 40 ends) few.
@pytest.mark.parametrize("N,seed", [(200, 0), (50, 1), (2000, 2)])
def test_lin_bit_exact_vs_replay(tile, sublat, n_iter, N, seed):
    print(f"[lin parity] tile={tile} sublat={sublat} n_iter={n_iter} N={N} seed={seed} L=32", flush=True)
    h = elp.LinHost(N=N, L=32, density=1.2, seed=seed)
    lat0 = h.lattice()
    org = h.origins(tile, 10)
    eng = elp.Engine("linear", h.L, tile, sublat, n_iter)
    eng.upload(lat0)
    eng.run(org[:4], tau0=0, seed=h.seed)
    eng.run(org[4:], tau0=4, seed=h.seed)
    got = eng.download()
    want = replay_lin(lat0, h.L, tile, sublat, n_iter, org, 0, h.seed)
    assert np.count_nonzero(want != lat0) > 1000
    assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} sites differ"


def test_lin_odd_tile_counts_bit_exact():
    """An 80 x 112 x 48 box: 5, 7 and 3 tiles of 16 per axis (the tile
    coordinates come from a float reciprocal with one correction step)."""
    print("[lin parity] box 80x112x48 dense 16^3", flush=True)
    h = elp.LinHost(N=200, L=(80, 112, 48), density=1.2, seed=7, box_unit=16)
    assert h.L == (80, 112, 48)
    lat0 = h.lattice()
    tile = (16, 16, 16)
    org = h.origins(tile, 6)
    eng = elp.Engine("linear", h.L, tile, elp.SUBLAT_DENSE, 16)
    eng.upload(lat0)
    eng.run(org, tau0=0, seed=h.seed)
    want = replay_lin(lat0, h.L, tile, elp.SUBLAT_DENSE, 16, org, 0, h.seed)
    assert np.count_nonzero(want != lat0) > 1000
    assert np.array_equal(eng.download(), want)
    eng.close()


def test_lin_full_size_invariants():
    """Config of the linear target: L=128, N=500 (5033 chains), 100 time units."""
    print("[lin full size] L=128 N=500", flush=True)
    sys.stdout.flush()
    h = elp.LinHost(N=500, L=128, density=1.2, seed=5)
    lat0 = h.lattice()
    tile = (16, 16, 16)
    org = h.origins(tile, 100)
    eng = elp.Engine("linear", h.L, tile, elp.SUBLAT_DENSE, 16)
    eng.upload(lat0)
    eng.run(org, tau0=0, seed=h.seed)
    got = eng.download()
    assert lin_site_invariants(got, h.L) == (5033 * 500, 5033)
    assert np.count_nonzero(got != lat0) > lat0.size // 4
    h.set_lattice(got)


@pytest.mark.parametrize("tile,sublat,n_iter", [SCHEDULES[0], SCHEDULES[2]])
def test_lin_saw_start_bit_exact_vs_replay(tile, sublat, n_iter):
    print(f"[lin parity, self-avoiding-walk start] tile={tile} sublat={sublat} L=32", flush=True)
    h = elp.LinHost(N=100, L=32, density=1.2, seed=2, init="saw")
    lat0 = h.lattice()
    org = h.origins(tile, 10)
    eng = elp.Engine("linear", h.L, tile, sublat, n_iter)
    eng.upload(lat0)
    eng.run(org, tau0=0, seed=h.seed)
    got = eng.download()
    want = replay_lin(lat0, h.L, tile, sublat, n_iter, org, 0, h.seed)
    assert np.count_nonzero(want != lat0) > 1000
    assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} sites differ"


def test_lin_saw_start_full_size():
    """Linear target configuration: L=128, N=500 (5033 chains) grown as random
    self-avoiding walks, 200 time units on the dense 16^3 schedule."""
    print("[lin full size, self-avoiding-walk start] L=128 N=500", flush=True)
    h = elp.LinHost(N=500, L=128, density=1.2, seed=3, init="saw")
    lat0 = h.lattice()
    seq0 = lin_label_sequences(lat0, h.L)
    tile = (16, 16, 16)
    org = h.origins(tile, 200)
    eng = elp.Engine("linear", h.L, tile, elp.SUBLAT_DENSE, 16)
    eng.upload(lat0)
    eng.run(org, tau0=0, seed=h.seed)
    got = eng.download()
    assert lin_site_invariants(got, h.L) == (5033 * 500, 5033)
    assert lin_label_sequences(got, h.L) == seq0
    assert np.count_nonzero(got != lat0) > lat0.size // 4
    h.set_lattice(got)


def test_lin_largest_box_bit_exact():
    """L=256 (16.7 M sites, 40 265 chains of N=500): two time units bit-exact
    against the C replay."""
    print("[lin parity] L=256 N=500 dense 16^3, 2 time units", flush=True)
    h = elp.LinHost(N=500, L=256, density=1.2, seed=12)
    lat0 = h.lattice()
    tile = (16, 16, 16)
    org = h.origins(tile, 2)
    eng = elp.Engine("linear", h.L, tile, elp.SUBLAT_DENSE, 16)
    eng.upload(lat0)
    eng.run(org, tau0=0, seed=h.seed)
    got = eng.download()
    want = replay_lin(lat0, h.L, tile, elp.SUBLAT_DENSE, 16, org, 0, h.seed)
    assert np.count_nonzero(want != lat0) > lat0.size // 4
    assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} sites differ"


def test_lin_c3_full_length():
    """C3 at its full length: linear L=128, N=500 (5033 chains), 10^5 time units
    on the dense 16^3 schedule.  Size-independent checks at that length: every
    chain intact (one start, one end, consistent bonds), the monomer count and
    every chain's label sequence unchanged; the same units run as two halves
    (the second resumed at time unit 50000) give the identical lattice; and the
    run as replica 0 of a batch gives the identical lattice."""
    print("[lin C3] L=128 N=500, 100000 time units, dense 16^3", flush=True)
    tile = (16, 16, 16)
    h = elp.LinHost(N=500, L=128, density=1.2, seed=5)
    lat0 = h.lattice()
    before = lin_label_sequences(lat0, h.L)
    org = h.origins(tile, 100000)
    eng = elp.Engine("linear", h.L, tile, elp.SUBLAT_DENSE, 16)
    eng.upload(lat0)
    eng.run(org, tau0=0, seed=h.seed)
    got = eng.download()
    assert lin_site_invariants(got, h.L) == (5033 * 500, 5033)
    assert lin_label_sequences(got, h.L) == before
    print("  one run: invariants and label sequences kept", flush=True)
    eng.upload(lat0)
    eng.run(org[:50000], tau0=0, seed=h.seed)
    eng.run(org[50000:], tau0=50000, seed=h.seed)
    assert np.array_equal(eng.download(), got)
    eng.close()
    print("  two halves: identical", flush=True)
    bat = elp.Batch("linear", [h.L], tile, elp.SUBLAT_DENSE, 16)
    bat.upload(0, lat0)
    bat.run(org[:, None, :], tau0=[0], seeds=[h.seed], streams=[0], units_per_graph=64)
    assert np.array_equal(bat.download(0), got)
    bat.close()
    print("  as replica 0 of a batch: identical", flush=True)
