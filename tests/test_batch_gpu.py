"""Batched replicas (elp_batch_* through the C ABI): every replica of a batch,
run through captured HIP graphs plus direct launches, must equal the same
replica run alone by the solo engine, and the solo engine equals the C replay
(checked here again for one replica per batch)."""
import sys

import numpy as np
import pytest

import elp
from ringcheck import lin_site_invariants, replay_lin, replay_ring, site_invariants

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _replicas(host_cls, spec, tile):
    hosts, lats = [], []
    for N, L, seed in spec:
        h = host_cls(N=N, L=L, density=1.2, seed=seed, box_unit=16)
        assert h.L == (tuple(L) if hasattr(L, "__len__") else (L, L, L))
        hosts.append(h)
        lats.append(h.lattice())
        h.load_shift(tile)
    return hosts, lats


def _run_batch_vs_solo(kind, host_cls, spec, tile, sublat, n_iter, runs, replay):
    hosts, lats = _replicas(host_cls, spec, tile)
    R = len(hosts)
    bat = elp.Batch(kind, [h.L for h in hosts], tile, sublat, n_iter)
    for r in range(R):
        bat.upload(r, lats[r])
    seeds = [h.seed for h in hosts]
    streams = list(range(R))
    all_org = []
    tau = 0
    for n_time, G in runs:
        org = np.stack([h.origins(tile, n_time) for h in hosts], axis=1)     # (n_time, R, 3)
        bat.run(org, tau0=tau, seeds=seeds, streams=streams, units_per_graph=G)
        all_org.append(org)
        tau += n_time
    org = np.concatenate(all_org, axis=0)
    for r in range(R):
        got = bat.download(r)
        eng = elp.Engine(kind, hosts[r].L, tile, sublat, n_iter)
        eng.upload(lats[r])
        eng.run(org[:, r], tau0=0, seed=seeds[r], stream=streams[r])
        solo = eng.download()
        eng.close()
        assert np.count_nonzero(solo != lats[r]) > solo.size // 50
        assert np.array_equal(got, solo), f"replica {r}: {np.count_nonzero(got != solo)} sites differ from solo"
    want = replay(lats[0], hosts[0].L, tile, sublat, n_iter, org[:, 0], 0, seeds[0], streams[0])
    assert np.array_equal(bat.download(0), want)
    bat.close()
    return hosts


@pytest.mark.parametrize("tile,sublat,n_iter", [((16, 16, 16), elp.SUBLAT_DENSE, 13),
                                                ((16, 16, 16), elp.SUBLAT_CUBIC, 51),
                                                ((16, 16, 8), elp.SUBLAT_DENSE, 13)])
def test_ring_batch_bit_exact_vs_solo(tile, sublat, n_iter):
    _say(f"[ring batch] tile={tile} sublat={sublat}")
    spec = [(50, 32, 0), (200, 48, 1), (100, (64, 32, 48), 2), (1000, 32, 3)]
    # 10 units: 2 graph replays of 4 + 2 direct; then 7 units: 1 replay + 3 direct; then 3 direct
    _run_batch_vs_solo("ring", elp.RingHost, spec, tile, sublat, n_iter, [(10, 4), (7, 4), (3, 0)], replay_ring)


@pytest.mark.parametrize("tile,sublat,n_iter", [((16, 16, 16), elp.SUBLAT_DENSE, 16),
                                                ((16, 16, 16), elp.SUBLAT_CUBIC, 64)])
def test_lin_batch_bit_exact_vs_solo(tile, sublat, n_iter):
    _say(f"[linear batch] tile={tile} sublat={sublat}")
    spec = [(50, 32, 0), (200, 48, 1), (100, (32, 64, 48), 2)]
    _run_batch_vs_solo("linear", elp.LinHost, spec, tile, sublat, n_iter, [(9, 4), (5, 0)], replay_lin)


def test_ring_batch_64_replicas_full_size():
    """64 ring melts, L in {32,48,64,96}, N in {50,200,1000,2000}, seeds 0..63
    (20 643 840 sites): 24 time units through 4 replays of a 5-unit graph and 4
    direct launches.  Every replica keeps every ring (monomer count, host
    recovery of all labelled rings); one replica per box size equals its solo run."""
    tile, sublat, n_iter = (16, 16, 16), elp.SUBLAT_DENSE, 13
    Ls, Ns = [32, 48, 64, 96], [50, 200, 1000, 2000]
    spec = [(Ns[(r // 4) % 4], Ls[r % 4], r) for r in range(64)]
    _say("[ring batch 64] building hosts")
    hosts, lats = _replicas(elp.RingHost, spec, tile)
    bat = elp.Batch("ring", [h.L for h in hosts], tile, sublat, n_iter)
    info = bat.launch_info()
    assert info["sites"] == 20643840 and info["blocks"] == 5040
    for r in range(64):
        bat.upload(r, lats[r])
    org = np.stack([h.origins(tile, 24) for h in hosts], axis=1)
    seeds = [h.seed for h in hosts]
    _say("[ring batch 64] running")
    bat.run(org, tau0=0, seeds=seeds, streams=list(range(64)), units_per_graph=5)
    for r in range(64):
        got = bat.download(r)
        assert site_invariants(got, hosts[r].L) == hosts[r].n_pol * hosts[r].N, f"replica {r}"
        hosts[r].set_lattice(got)
        if r < 4:
            eng = elp.Engine("ring", hosts[r].L, tile, sublat, n_iter)
            eng.upload(lats[r])
            eng.run(org[:, r], tau0=0, seed=seeds[r], stream=r)
            assert np.array_equal(got, eng.download()), f"replica {r} differs from solo"
            eng.close()
    bat.close()


@pytest.mark.parametrize("kind,rules", [("ring", elp.RULES_REFERENCE), ("linear", elp.RULES_REFERENCE),
                                        ("ring", elp.RULES_BALANCED)])
def test_autotuned_schedule_replays_bit_exact(kind, rules):
    """The tuner ranks by accepted moves per second among schedules with the
    statics of the request (reference rates: the dense sub-lattice only;
    balanced rates: both); its record, parsed back, drives a batch whose
    replicas equal the C replay of that recorded schedule."""
    _say(f"[autotune] {kind} rules={rules}")
    host_cls, replay = (elp.RingHost, replay_ring) if kind == "ring" else (elp.LinHost, replay_lin)
    spec = [(100, 32, 11), (200, 64, 12)]
    hosts = [host_cls(N=N, L=L, density=1.2, seed=s) for N, L, s in spec]
    lats = [h.lattice() for h in hosts]
    best, log = elp.autotune(kind, lats, [h.L for h in hosts], time_units=8, rules=rules)
    _say(log)
    assert all(np.array_equal(h.lattice(), l) for h, l in zip(hosts, lats))     # inputs untouched
    assert best.attempts_per_s > 0 and 0 < best.accepted_per_s < best.attempts_per_s and "chosen:" in log
    kernels = [ln for ln in log.splitlines() if ln.startswith("kernel ")]
    subs = {ln.split("sublattice=")[1].split()[0] for ln in kernels}
    assert subs == ({"dense"} if rules == elp.RULES_REFERENCE else {"dense", "cubic"}), subs
    assert all("accepted/s" in ln for ln in kernels)
    s = elp.Schedule.parse(best.record())
    assert s.rules == rules and s.accepted_per_s > 0
    bat = elp.Batch(kind, [h.L for h in hosts], s.tile, s.sublattice, s.n_iter)
    bat.set_threads(s.threads)
    bat.set_rules(rules)
    for r, h in enumerate(hosts):
        bat.upload(r, lats[r])
        h.load_shift(s.tile)
    n_time = 2 * max(s.units_per_graph, 1) + 1
    org = np.stack([h.origins(s.tile, n_time) for h in hosts], axis=1)
    bat.run(org, tau0=5, seeds=[h.seed for h in hosts], streams=[0, 0], units_per_graph=s.units_per_graph)
    for r, h in enumerate(hosts):
        kw = {"rules": rules} if kind == "ring" else {}
        want = replay(lats[r], h.L, s.tile, s.sublattice, s.n_iter, org[:, r], 5, h.seed, 0, **kw)
        assert np.array_equal(bat.download(r), want), f"replica {r}"
    bat.close()


def test_batch_rejects_bad_input():
    with pytest.raises(elp.ElpError, match="not divisible"):
        elp.Batch("ring", [(32, 32, 32), (40, 32, 32)], (16, 16, 16), elp.SUBLAT_DENSE, 13)
    bat = elp.Batch("ring", [(32, 32, 32)], (16, 16, 16), elp.SUBLAT_DENSE, 13)
    with pytest.raises(elp.ElpError, match="before upload"):
        bat.run(np.zeros((1, 1, 3)), 0, [1])
    bat.upload(0, np.zeros(32 ** 3, np.uint8))
    with pytest.raises(elp.ElpError, match="outside the tile"):
        bat.run(np.full((1, 1, 3), 16), 0, [1])
    with pytest.raises(elp.ElpError, match="bytes"):
        bat.upload(0, np.zeros(100, np.uint8))
    bat.close()
