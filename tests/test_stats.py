"""Statistics library (libelp_stats.so) pinned by closed forms and brute force."""
import os

import numpy as np
import pytest

import elp
from elp.stats import Series

VALID = [b for b in range(1, 15) if b not in (3, 12)]


def delta(b):
    w = (b >> 3) & 1
    return np.array([(b & 1) - w, ((b >> 1) & 1) - w, ((b >> 2) & 1) - w])


def xyz(tuv):
    t, u, v = tuv[..., 0], tuv[..., 1], tuv[..., 2]
    return np.stack([t + u - v, t - u, v], axis=-1) / np.sqrt(2.0)


def test_bond_codes_have_unit_length():
    for b in VALID:
        assert np.isclose(np.sum(xyz(delta(b)) ** 2), 1.0)
        assert np.array_equal(delta(b ^ 0xF), -delta(b))


def test_rg2_straight_linear_chain():
    n = 11
    s = Series(False, (64, 64, 64), 1, n, 1)
    s.add([[3, 5, 7]], [[1] * (n - 1) + [0xF]])
    assert np.isclose(s.rg2(0)[0], (n * n - 1) / 12.0)
    st = s.statics(0)
    assert np.isclose(st["rg2"], (n * n - 1) / 12.0) and st["slrat"] == 0.0
    # a straight chain: only consecutive monomers are in contact
    assert np.isclose(st["pc"][1], 1.0) and np.all(st["pc"][2:] == 0)


def test_rg2_regular_hexagon_ring():
    # the six bonds in a close-packed plane (normal (1,1,1) in x,y,z), taken in
    # angular order: a regular hexagon of side 1
    nrm = np.ones(3) / np.sqrt(3.0)
    dirs = [bb for bb in VALID if np.isclose(xyz(delta(bb)) @ nrm, 0)]
    assert len(dirs) == 6
    e1 = xyz(delta(dirs[0]))
    e2 = np.cross(nrm, e1)
    ang = {bb: np.arctan2(xyz(delta(bb)) @ e2, xyz(delta(bb)) @ e1) for bb in dirs}
    ring = sorted(dirs, key=lambda bb: ang[bb])
    assert np.array_equal(sum(delta(bb) for bb in ring), [0, 0, 0])
    s = Series(True, (32, 32, 32), 1, 6, 1)
    s.add([[0, 0, 0]], [ring])
    assert np.isclose(s.rg2(0)[0], 1.0)


def test_stored_length_two_site_ring():
    # N-2 stored-length bonds and one bond out and back: monomers on two sites
    n, k = 10, 3                              # monomers 0..k-1 on the first site, the rest on the second
    bonds = [0] * (k - 1) + [0xE] + [0] * (n - k - 1) + [0x1]
    s = Series(True, (32, 32, 32), 1, n, 1)
    s.add([[4, 4, 4]], [bonds])
    p = s.positions(0, 0)
    on_first = np.all(p == p[0], axis=1).sum()
    assert on_first == k
    q = k / n
    assert np.isclose(s.rg2(0)[0], q * (1 - q))
    st = s.statics(0)
    assert np.isclose(st["slrat"], (n - 2) / n)


def test_contact_probability_brute_force():
    rng = np.random.default_rng(3)
    n, nchain, nframe = 40, 5, 3
    s = Series(True, (16, 16, 16), nchain, n, nframe)
    counts = np.zeros(n)
    nb = {tuple(delta(b)) for b in VALID} | {(0, 0, 0)}
    for f in range(nframe):
        starts, allb = [], []
        for c in range(nchain):
            half = rng.choice(VALID, n // 2)
            bonds = np.concatenate([half, 0xF ^ half])
            rng.shuffle(bonds)
            # stored length: replace a bond and its inverse partner by two zeros
            for _ in range(3):
                i = rng.integers(n)
                if bonds[i] == 0:
                    continue
                j = np.where(bonds == (0xF ^ bonds[i]))[0]
                if len(j):
                    bonds[i] = 0
                    bonds[j[0]] = 0
            start = rng.integers(0, 16, 3)
            starts.append(start)
            allb.append(bonds)
            pos = np.concatenate([[start], start + np.cumsum([delta(b) if b else np.zeros(3, int) for b in bonds[:-1]],
                                                             axis=0)])
            for i in range(n):
                for j in range(i + 1, n):
                    if tuple(pos[j] - pos[i]) in nb:
                        counts[j - i] += 1
        s.add(starts, allb)
    pc = s.statics(0)["pc"]
    want = counts / (nchain * nframe * (n - np.arange(n)))
    assert np.allclose(pc[1:], want[1:])
    assert pc[1] == 1.0                       # consecutive monomers are always in contact


def test_msd_rigid_translation_and_periodic_unwrap():
    # two straight chains; chain 0 moves by d per frame across the periodic
    # boundary, chain 1 stays: the system drift is d/2, so each chain's
    # displacement relative to it is d/2 per frame
    L, n = 16, 8
    s = Series(False, (L, L, L), 2, n, 5)
    d = np.array([3, 1, 0])
    for f in range(5):
        st0 = (np.array([14, 2, 5]) + f * d) % L
        s.add([st0, [1, 1, 1]], [[1] * (n - 1) + [0xF], [2] * (n - 1) + [0xF]])
    g1, g3 = s.msd(0, 4)
    dd = np.sum(xyz(d / 2.0) ** 2)
    for k in range(4):
        assert np.isclose(g3[k], k * k * dd), (k, g3[k])
        assert np.isclose(g1[k], k * k * dd), (k, g1[k])
    p = s.positions(4, 0)
    assert np.array_equal(p[0], np.array([14, 2, 5]) + 4 * d)        # unwrapped across the boundary


def test_tuv_reader_matches_recovered_chains(tmp_path):
    h = elp.RingHost(N=100, L=32, density=1.2, seed=4)
    path = str(tmp_path / "t=0_dev=0.res")
    h.write_tuv(path)
    a = Series(True, h.L, h.n_pol, h.N, 1)
    a.add_tuv(path)
    b = Series(True, h.L, h.n_pol, h.N, 1)
    b.add_host(h)
    for c in range(0, h.n_pol, 37):
        assert np.array_equal(a.positions(0, c), b.positions(0, c))
    sa, sb = a.statics(0), b.statics(0)
    assert sa["rg2"] == sb["rg2"] and sa["slrat"] == sb["slrat"]
    with pytest.raises(elp.ElpError):
        Series(True, (32, 32, 32), h.n_pol + 1, h.N, 1).add_tuv(path)


def _summary_of(full):
    s = Series(full.ring, full.L, full.n_chain, full.n_mono, full.frames, summary=True)
    for f in range(full.frames):
        s.add_summary(full.summary(f))
    return s


def test_summary_series_rigid_translation():
    """Per-chain reductions unwrap across the periodic boundary exactly as the positions do."""
    L, n = 16, 8
    full = Series(False, (L, L, L), 2, n, 5)
    d = np.array([3, 1, 0])
    for f in range(5):
        st0 = (np.array([14, 2, 5]) + f * d) % L
        full.add([st0, [1, 1, 1]], [[1] * (n - 1) + [0xF], [2] * (n - 1) + [0xF]])
    s = _summary_of(full)
    a, b = full.msd(0, 4), s.msd(0, 4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert full.summary(4)["start3"][0].tolist() == ((np.array([14, 2, 5]) + 4 * d) % L).tolist()
    with pytest.raises(elp.ElpError):
        s.summary(0)                                   # positions are not kept in summary mode


def test_summary_series_equals_full_series_on_a_melt():
    """Ring melt frames from the C replay (L=32, N=100): statics and mean-square
    displacements of the summary series equal those of the full series."""
    from ringcheck import replay_ring
    h = elp.RingHost(N=100, L=32, density=1.2, seed=6)
    tile = (16, 16, 16)
    h.load_shift(tile)
    full = Series(True, h.L, h.n_pol, h.N, 6)
    full.add_host(h)
    lat, tau = h.lattice(), 0
    for f in range(1, 6):
        org = h.origins(tile, 25)
        lat = replay_ring(lat, h.L, tile, elp.SUBLAT_DENSE, 13, org, tau, h.seed)
        tau += 25
        h.set_lattice(lat)
        full.add_host(h)
    s = _summary_of(full)
    a, b = full.statics(1), s.statics(1)
    assert a["rg2"] == b["rg2"] and a["slrat"] == b["slrat"] and np.array_equal(a["pc"], b["pc"])
    assert a["pc"][1] > 0.5 and a["slrat"] > 0
    ga, gb = full.msd(0, 5), s.msd(0, 5)
    assert np.array_equal(ga[0], gb[0]) and np.array_equal(ga[1], gb[1])
    assert ga[0][4] > ga[0][1] > 0
    sm = full.summary(3)
    assert np.array_equal(s.rg2(3), full.rg2(3))
    assert sm["contacts"].sum() > 0 and sm["nsl"].sum() > 0


def test_msd_drift_convention_of_the_reference_differs_negligibly():
    """The reference analysis subtracts the drift of the system's mean in mixed
    units: avg_pos.dat holds the mean over chains of cms/sqrt(2)
    (analysis/mode_compute.c:244-251, mode_io.c:640-652), which g3 subtracts from
    a centre-of-mass displacement in bond units (:548-566) and g1 from a
    middle-monomer displacement in lattice units before dividing by 2 (:506-528).
    So g3 removes 1/sqrt(2) of the drift and g1 1/2 of it, where this library
    removes all of it.  For g3 the difference is exactly (1 - 1/sqrt(2))^2 times
    the mean-square displacement of the system mean (the chains' mean displacement
    is the system's, so the cross term vanishes); on a replay melt (L=32, N=100,
    393 rings, lags 40..160 time units from the start) g3 differs by 0.14-0.23 %
    and g1 by 0.2-1.0 %."""
    from ringcheck import replay_ring
    h = elp.RingHost(N=100, L=32, density=1.2, seed=6)
    tile = (16, 16, 16)
    h.load_shift(tile)
    nf = 6
    full = Series(True, h.L, h.n_pol, h.N, nf)
    full.add_host(h)
    lat, tau = h.lattice(), 0
    for f in range(1, nf):
        org = h.origins(tile, 40)
        lat = replay_ring(lat, h.L, tile, elp.SUBLAT_DENSE, 13, org, tau, h.seed)
        tau += 40
        h.set_lattice(lat)
        full.add_host(h)
    pos = np.array([[xyz(full.positions(f, c).astype(float)) for c in range(h.n_pol)] for f in range(nf)])
    cms = pos.mean(axis=2)                                  # (frame, chain, 3), bond units
    mid = pos[:, :, h.N // 2]
    sysm = cms.mean(axis=1)                                 # system mean per frame

    def msd(x, s):
        out = []
        for k in range(nf - 1):
            d = x[:nf - k] - x[k:] - s * (sysm[:nf - k] - sysm[k:])[:, None, :]
            out.append(np.mean(np.sum(d * d, axis=2)))
        return np.array(out)
    g1, g3 = full.msd(0, nf - 1)
    assert np.allclose(msd(mid, 1.0), g1, rtol=1e-12) and np.allclose(msd(cms, 1.0), g3, rtol=1e-12)
    g1_ref, g3_ref = msd(mid, 0.5), msd(cms, 1 / np.sqrt(2))
    assert g3[1] > 0.05                                     # the chains moved
    # the library's reference-convention entry point gives exactly these
    r1, r3 = full.msd(0, nf - 1, convention="reference")
    assert np.allclose(r1, g1_ref, rtol=1e-12) and np.allclose(r3, g3_ref, rtol=1e-12)
    # g3: the chains' mean displacement is the system's, so the cross term vanishes and the
    # reference's g3 exceeds this library's by exactly (1 - 1/sqrt(2))^2 x the system mean's MSD
    sys_msd = np.array([np.mean(np.sum((sysm[:nf - k] - sysm[k:]) ** 2, axis=1)) for k in range(nf - 1)])
    assert np.allclose(g3_ref - g3, (1 - 1 / np.sqrt(2)) ** 2 * sys_msd, rtol=1e-9, atol=1e-15)
    rel3, rel1 = g3_ref[1:] / g3[1:] - 1, g1_ref[1:] / g1[1:] - 1
    print(f"reference drift convention: g3 +{np.round(100 * rel3, 3)} %, g1 {np.round(100 * rel1, 3)} % at lags "
          f"40..160 time units")
    assert np.all(np.abs(rel3) < 2e-2) and np.all(np.abs(rel1) < 2e-2), (rel3, rel1)
