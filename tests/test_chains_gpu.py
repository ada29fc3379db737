"""Ring recovery and tuv text on the device (elp_ring_chains / elp_ring_tuv)
against the host recovery and writer (GetAllRingPolymers + WriteLatticeFile
restated from gpupol2/oclpol.c and gpupol2/gpupol_io.c): the same bytes."""
import sys
import time

import numpy as np
import pytest

import elp

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(N, L, seed, n_time, tile=(16, 16, 16), sub=elp.SUBLAT_DENSE, n_iter=13):
    h = elp.RingHost(N=N, L=L, density=1.2, seed=seed)
    eng = elp.Engine("ring", h.L, tile, sub, n_iter)
    eng.upload(h.lattice())
    h.load_shift(tile)
    eng.run(h.origins(tile, n_time), tau0=0, seed=h.seed)
    return h, eng


@pytest.mark.parametrize("N,L,seed,n_time", [(100, 32, 3, 50), (200, 64, 4, 20), (1000, 128, 5, 30)])
def test_device_tuv_equals_host_writer(tmp_path, N, L, seed, n_time):
    _say(f"[chains] N={N} L={L}")
    h, eng = _run(N, L, seed, n_time)
    t0 = time.perf_counter()
    text = eng.ring_tuv(h)
    t1 = time.perf_counter()
    h.set_lattice(eng.download())
    path = tmp_path / "t.res"
    h.write_tuv(str(path))
    t2 = time.perf_counter()
    want = path.read_bytes()
    _say(f"  device recovery + text {1e3 * (t1 - t0):.1f} ms, host download + recovery + writer "
         f"{1e3 * (t2 - t1):.1f} ms, {len(want)} bytes")
    assert text == want
    bonds, lens, starts = eng.ring_chains(h)
    for pid in range(0, h.n_pol, max(1, h.n_pol // 50)):
        st, b = h.polymer(pid)
        assert lens[pid] == len(b) and tuple(starts[pid]) == st and np.array_equal(bonds[pid, :len(b)], b)
    eng.close()


def test_device_chain_stats_match_statistics_library():
    from elp.stats import Series
    h, eng = _run(1000, 96, 8, 40)
    rg2, nsl = eng.ring_stats(h)
    h.set_lattice(eng.download())
    s = Series(True, h.L, h.n_pol, h.N, 1)
    s.add_host(h)
    assert np.allclose(rg2, s.rg2(0), rtol=1e-12, atol=1e-9)
    assert np.isclose(nsl.sum() / (h.n_pol * h.N), s.statics(0)["slrat"])
    eng.close()


def test_device_recovery_rejects_a_damaged_lattice():
    h, eng = _run(100, 32, 6, 5)
    lat = eng.download()
    heads = np.flatnonzero((lat & 0x10) != 0)
    lat[heads] &= 0xEF                          # every carrier label 1 -> 0: no ring head left
    eng.upload(lat)
    with pytest.raises(elp.ElpError, match="ring heads"):
        eng.ring_tuv(h)
    eng.close()


@pytest.mark.parametrize("N,L,seed", [(100, 32, 7), (1000, 128, 8)])
def test_device_frame_summary_equals_host_reduction(N, L, seed):
    """elp_ring_frame_summary (positions, sums, middle monomer, contact histogram
    reduced on the device) equals the host reduction of the same frame."""
    from elp.stats import Series
    _say(f"[frame summary] N={N} L={L}")
    h, eng = _run(N, L, seed, 20)
    t0 = time.perf_counter()
    dev = eng.ring_frame_summary(h)
    t1 = time.perf_counter()
    h.set_lattice(eng.download())
    full = Series(True, h.L, h.n_pol, h.N, 1)
    full.add_host(h)
    want = full.summary(0)
    t2 = time.perf_counter()
    _say(f"  device summary {1e3 * (t1 - t0):.1f} ms, download + host recovery + reduction {1e3 * (t2 - t1):.1f} ms")
    for k in ("start3", "sum3", "mid3", "nsl", "contacts"):
        assert np.array_equal(dev[k], want[k]), k
    assert np.allclose(dev["rg2"], want["rg2"], rtol=1e-12, atol=0)
    # consecutive monomers always touch: N - 1 pairs one apart per ring, and the
    # closing pair (N - 1, 0) counts at separation N - 1
    assert dev["contacts"][1] == h.n_pol * (N - 1) and dev["contacts"][N - 1] >= h.n_pol


def test_device_summary_series_equals_full_series():
    """Six frames of a melt: statics and mean-square displacements from the
    device summaries equal those of the full host series (Rg^2 to rounding)."""
    from elp.stats import Series
    _say("[summary series] N=200 L=48")
    tile = (16, 16, 16)
    h = elp.RingHost(N=200, L=48, density=1.2, seed=9, box_unit=16)
    eng = elp.Engine("ring", h.L, tile, elp.SUBLAT_DENSE, 13)
    eng.upload(h.lattice())
    h.load_shift(tile)
    full = Series(True, h.L, h.n_pol, h.N, 6)
    summ = Series(True, h.L, h.n_pol, h.N, 6, summary=True)
    tau = 0
    for f in range(6):
        if f:
            eng.run(h.origins(tile, 200), tau0=tau, seed=h.seed)
            tau += 200
        summ.add_summary(eng.ring_frame_summary(h))
        h.set_lattice(eng.download())
        full.add_host(h)
    a, b = full.statics(1), summ.statics(1)
    assert np.isclose(a["rg2"], b["rg2"], rtol=1e-12) and a["slrat"] == b["slrat"]
    assert np.array_equal(a["pc"], b["pc"])
    ga, gb = full.msd(0, 6), summ.msd(0, 6)
    assert np.array_equal(ga[0], gb[0]) and np.array_equal(ga[1], gb[1])
    assert ga[1][5] > 0


def _run_lin(N, L, seed, n_time, init="reference"):
    tile = (16, 16, 16)
    h = elp.LinHost(N=N, L=L, density=1.2, seed=seed, init=init)
    eng = elp.Engine("linear", h.L, tile, elp.SUBLAT_DENSE, 16)
    eng.upload(h.lattice())
    eng.run(h.origins(tile, n_time), tau0=0, seed=h.seed)
    return h, eng


@pytest.mark.parametrize("N,L,seed,n_time,init", [(100, 32, 3, 50, "reference"), (200, 64, 4, 20, "reference"),
                                                  (500, 128, 5, 30, "reference"), (500, 128, 6, 30, "saw")])
def test_device_lin_tuv_equals_host_writer(tmp_path, N, L, seed, n_time, init):
    """Linear chains: device-side recovery and text equal the host recovery and writer byte for byte."""
    _say(f"[linear chains] N={N} L={L} start={init}")
    h, eng = _run_lin(N, L, seed, n_time, init)
    t0 = time.perf_counter()
    text = eng.lin_tuv(h)
    t1 = time.perf_counter()
    h.set_lattice(eng.download())
    path = tmp_path / "t.res"
    h.write_tuv(str(path))
    t2 = time.perf_counter()
    want = path.read_bytes()
    _say(f"  device recovery + text {1e3 * (t1 - t0):.1f} ms, host download + recovery + writer "
         f"{1e3 * (t2 - t1):.1f} ms, {len(want)} bytes")
    assert text == want
    bonds, lens, starts = eng.lin_chains(h)
    for pid in range(0, h.n_pol, max(1, h.n_pol // 50)):
        st, b = h.polymer(pid)
        assert lens[pid] == len(b) and tuple(starts[pid]) == st and np.array_equal(bonds[pid, :len(b)], b)
    eng.close()


def test_device_lin_recovery_rejects_a_damaged_lattice():
    h, eng = _run_lin(100, 32, 6, 5)
    lat = eng.download()
    occ = (lat & 0xFF) != 0
    lat[occ] ^= 0x200                           # every carrier label flipped: no known chain id left
    eng.upload(lat)
    with pytest.raises(elp.ElpError, match="known polymer id"):
        eng.lin_tuv(h)
    eng.close()


def _host_summary(ring, h, lat):
    from elp.stats import Series
    h.set_lattice(lat)
    full = Series(ring, h.L, h.n_pol, h.N, 1)
    full.add_host(h)
    return full.summary(0)


def _same_summary(dev, want):
    for k in ("start3", "sum3", "mid3", "nsl", "contacts"):
        assert np.array_equal(dev[k], want[k]), k
    assert np.allclose(dev["rg2"], want["rg2"], rtol=1e-12, atol=0)


@pytest.mark.parametrize("N,L,seed,n_time,init", [(100, 32, 3, 50, "reference"), (500, 128, 6, 30, "saw")])
def test_device_lin_frame_summary_equals_host_reduction(N, L, seed, n_time, init):
    """elp_lin_frame_summary equals the host reduction of the same frame (middle
    monomer (N-1)/2, N-1 stored-length candidates, contacts over all pairs)."""
    _say(f"[linear frame summary] N={N} L={L} start={init}")
    h, eng = _run_lin(N, L, seed, n_time, init)
    dev = eng.lin_frame_summary(h)
    _same_summary(dev, _host_summary(False, h, eng.download()))
    # consecutive monomers always touch: N - 1 pairs one apart per chain
    assert dev["contacts"][1] == h.n_pol * (N - 1)
    eng.close()


@pytest.mark.parametrize("kind", ["ring", "linear"])
def test_batch_frame_summary_equals_host_reduction(kind):
    """elp_batch_frame_summary reads each replica in place: equal to the host
    reduction of that replica's download, for replicas of different boxes."""
    _say(f"[batch frame summary] {kind}")
    ring = kind == "ring"
    cls = elp.RingHost if ring else elp.LinHost
    tile = (16, 16, 16)
    hosts = [cls(N=n, L=l, density=1.2, seed=s, box_unit=16) for n, l, s in ((100, 32, 1), (60, 48, 2), (100, 32, 3))]
    bat = elp.Batch(kind, [h.L for h in hosts], tile, elp.SUBLAT_DENSE, 13 if ring else 16)
    for r, h in enumerate(hosts):
        bat.upload(r, h.lattice())
        if ring:
            h.load_shift(tile)
    org = np.stack([h.origins(tile, 30) for h in hosts], axis=1)
    bat.run(org, tau0=0, seeds=[h.seed for h in hosts], streams=[0] * len(hosts))
    for r, h in enumerate(hosts):
        dev = bat.frame_summary(r, h)
        _same_summary(dev, _host_summary(ring, h, bat.download(r)))
    with pytest.raises(elp.ElpError, match="out of range"):
        bat.frame_summary(len(hosts), hosts[0])
    bat.close()
