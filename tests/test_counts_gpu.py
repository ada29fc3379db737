"""Move counting on the GPU (elp_count_moves / elp_batch_count_moves) against
the oracle's replay of the same schedule: the counting kernels move exactly
as the plain ones (bit-exact lattices), and their counts of attempts at
empty sites, of attempts leaving the tile (rings) and of accepted moves equal
the replay's."""
import sys

import numpy as np
import pytest

import elp
from ringcheck import replay_lin, replay_ring

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _attempts(L, tile, sublat, n_iter, nsub, n_time):
    ntile = (L[0] // tile[0]) * (L[1] // tile[1]) * (L[2] // tile[2])
    nslot = (tile[0] // 4) * (tile[1] // 4) * (tile[2] // 4) * (4 if sublat == elp.SUBLAT_DENSE else 1)
    return ntile * nslot * n_iter * nsub * n_time


@pytest.mark.parametrize("tile,sublat,n_iter,rules", [((16, 16, 16), elp.SUBLAT_DENSE, 13, elp.RULES_REFERENCE),
                                                      ((16, 16, 16), elp.SUBLAT_CUBIC, 51, elp.RULES_REFERENCE),
                                                      ((32, 32, 32), elp.SUBLAT_CUBIC, 51, elp.RULES_REFERENCE),
                                                      ((16, 16, 16), elp.SUBLAT_DENSE, 13, elp.RULES_BALANCED)])
def test_ring_counts_match_replay(tile, sublat, n_iter, rules):
    _say(f"[counts] ring L=64 N=200 tile={tile} sublat={sublat} rules={rules}")
    h = elp.RingHost(N=200, L=64, density=1.2, seed=6)
    lat0 = h.lattice()
    h.load_shift(tile)
    org = h.origins(tile, 6)
    eng = elp.Engine("ring", h.L, tile, sublat, n_iter)
    eng.set_rules(rules)
    eng.upload(lat0)
    eng.run(org[:2], tau0=0, seed=h.seed)
    plain = eng.download()
    eng.count_moves(True)
    eng.run(org[2:], tau0=2, seed=h.seed)
    c = eng.move_counts()
    got = eng.download()
    eng.close()
    mid = replay_ring(lat0, h.L, tile, sublat, n_iter, org[:2], 0, h.seed, rules=rules)
    want, (e, f, a) = replay_ring(mid, h.L, tile, sublat, n_iter, org[2:], 2, h.seed, rules=rules, counts=True)
    assert np.array_equal(plain, mid)
    assert np.array_equal(got, want)
    assert c["attempts"] == _attempts(h.L, tile, sublat, n_iter, 4, 4)
    assert (c["empty"], c["face"], c["accepted"]) == (e, f, a)
    assert 0 < a and 0 < f and e + f + a < c["attempts"]
    _say(f"  accepted {a / c['attempts']:.4f}, empty {e / c['attempts']:.4f}, tile face {f / c['attempts']:.4f}")


@pytest.mark.parametrize("tile,sublat,n_iter", [((16, 16, 16), elp.SUBLAT_DENSE, 16),
                                                ((32, 32, 32), elp.SUBLAT_CUBIC, 64)])
def test_lin_counts_match_replay(tile, sublat, n_iter):
    _say(f"[counts] linear L=64 N=200 tile={tile} sublat={sublat}")
    h = elp.LinHost(N=200, L=64, density=1.2, seed=6)
    lat0 = h.lattice()
    org = h.origins(tile, 4)
    eng = elp.Engine("linear", h.L, tile, sublat, n_iter)
    eng.upload(lat0)
    eng.count_moves(True)
    eng.run(org, tau0=0, seed=h.seed)
    c = eng.move_counts()
    got = eng.download()
    eng.close()
    want, (e, a) = replay_lin(lat0, h.L, tile, sublat, n_iter, org, 0, h.seed, counts=True)
    assert np.array_equal(got, want)
    assert c["attempts"] == _attempts(h.L, tile, sublat, n_iter, 2, 4)
    assert (c["empty"], c["accepted"]) == (e, a)
    assert 0 < c["face"] and e + c["face"] + a < c["attempts"]
    _say(f"  accepted {a / c['attempts']:.4f}, empty {e / c['attempts']:.4f}, "
         f"tile face {c['face'] / c['attempts']:.4f}")


def test_batch_counts_are_the_sum_of_solo_replays():
    """Four ring replicas (L 32 and 48, N 100 and 200) counted in one batch,
    from graphs of 2 time units and direct launches: the counts equal the sum
    of the replays of each replica."""
    _say("[counts] batch of 4 ring replicas")
    tile = (16, 16, 16)
    specs = [(32, 100, 1), (48, 200, 2), (32, 200, 3), (48, 100, 4)]
    hosts = [elp.RingHost(N=N, L=L, density=1.2, seed=s, box_unit=16) for L, N, s in specs]
    bat = elp.Batch("ring", [h.L for h in hosts], tile, elp.SUBLAT_DENSE, 13)
    for r, h in enumerate(hosts):
        bat.upload(r, h.lattice())
    rng = np.random.default_rng(3)
    org = rng.integers(0, 16, (5, 4, 3))
    seeds = np.array([11, 12, 13, 14])
    bat.count_moves(True)
    bat.run(org, tau0=np.zeros(4, np.int64), seeds=seeds, units_per_graph=2)
    c = bat.move_counts()
    tot = np.zeros(3, np.int64)
    for r, h in enumerate(hosts):
        want, cnt = replay_ring(h.lattice(), h.L, tile, elp.SUBLAT_DENSE, 13, org[:, r], 0, int(seeds[r]), counts=True)
        assert np.array_equal(bat.download(r), want)
        tot += np.array(cnt)
    assert c["attempts"] == sum(_attempts(h.L, tile, elp.SUBLAT_DENSE, 13, 4, 5) for h in hosts)
    assert (c["empty"], c["face"], c["accepted"]) == tuple(int(x) for x in tot)
    bat.count_moves(False)
    bat.close()
