"""The learned proposal on the GPU (elp_net_* through the C ABI): MFMA forward
and backward kernels against the FP64 oracle, the site moves against the
oracle's ring moves, the feature rows against their definition, training,
and the Metropolis-Hastings chain against the exact uniform measure of a
six-monomer ring."""
import sys

import numpy as np
import pytest

import elp
from learncheck import (NPARAM, bf16_round, features_of, lib, mlp_forward, mlp_grad, patches_of, random_features,
                        random_params, ring_decode, ring_lattice, ring_observables, ring_reachable)
from ringcheck import replay_ring

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_forward_matches_fp64_oracle():
    _say("[learned] forward vs FP64")
    rng = np.random.default_rng(2)
    prm = bf16_round(random_params(rng))            # the kernels compute with bf16 weights
    x = random_features(rng, 1000)
    net = elp.Net(seed=1)
    net.set_params(prm)
    got = net.forward(x)
    want = mlp_forward(prm, x)
    err = np.abs(got - want)
    # remaining difference: bf16 rounding of the two hidden activations
    assert err.max() < 0.02 and err.mean() < 2e-3, (err.max(), err.mean())
    assert np.allclose(got.sum(axis=1), 1, atol=1e-5)
    net.close()


def test_gradient_matches_fp64_oracle():
    _say("[learned] gradient vs FP64")
    rng = np.random.default_rng(3)
    prm = bf16_round(random_params(rng))
    x = random_features(rng, 777)                    # not a multiple of the 64-row batch
    y = rng.integers(0, 11, 777)
    net = elp.Net(seed=1)
    net.set_params(prm)
    g, loss = net.gradient(x, y)
    gw, lw = mlp_grad(prm, x, y)
    assert abs(loss - lw) < 0.01 * lw
    for name, sl in (("W1", slice(0, 9216)), ("W2", slice(9216, 13312)), ("b2", slice(13312, 13376)),
                     ("W3", slice(13376, 14400)), ("b3", slice(14400, 14416))):
        a, b = g[sl].astype(np.float64), gw[sl]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        _say(f"  {name}: relative L2 error {rel:.4f}, cosine {cos:.6f}")
        assert rel < 0.03 and cos > 0.999, (name, rel, cos)
    net.close()


def _melt_patches(n, seed=4):
    h = elp.RingHost(N=100, L=32, density=1.2, seed=seed)
    lat0 = h.lattice()
    h.load_shift((16, 16, 16))
    lat = replay_ring(lat0, h.L, (16, 16, 16), elp.SUBLAT_DENSE, 13, h.origins((16, 16, 16), 20), 0, h.seed,
                      rules=elp.RULES_BALANCED)
    rng = np.random.default_rng(seed)
    return patches_of(lat, h.L, rng.choice(np.flatnonzero(lat), n, replace=False))


def _oracle_move(patch, m):
    """the plain move of the oracle that move m of the learned proposal names"""
    o = lib()
    a = patch.copy()
    s = int(patch[62])
    sl = bool(s & 0x40)
    if m < 8:
        typ, r = 0, (m & 3) | (4 if m < 4 else 0)
    elif m < 10:
        typ, r = 2, (1 if m == 8 else 0)
    else:
        typ, r = 1, 0
    o.orc_ring_move(a.ctypes.data, 5, 5, 5, typ, 2, 2, 2, r, 0)
    return a, sl


def test_site_moves_match_oracle_and_reverse():
    _say("[learned] site moves vs oracle")
    P = _melt_patches(600)
    moves = np.tile(np.arange(11), len(P))
    patches = np.repeat(P, 11, axis=0)
    out, ok, rev = elp.patch_moves(patches, moves)
    n_ok = 0
    for i in range(len(patches)):
        m = moves[i]
        want, sl = _oracle_move(patches[i], m)
        changed = not np.array_equal(want, patches[i])
        if ok[i]:
            n_ok += 1
            assert np.array_equal(out[i], want), (i, m)
            continue
        assert np.array_equal(out[i], patches[i]), (i, m)
        if not changed:
            continue
        # the plain move happened but names another learned move: m 4..7 with
        # stored length at one end only (the plain move is then forced, as
        # m - 4), a fold into P while P holds stored length, or a fold into
        # the partner while the partner holds it (the plain move folds into P)
        s = int(patches[i][62])
        b = s & 0xF
        w = (b >> 3) & 1
        nb = int(patches[i][62 + (b & 1) - w + 5 * (((b >> 1) & 1) - w) + 25 * (((b >> 2) & 1) - w)])
        if 4 <= m < 8:
            assert not (sl and nb & 0x40), (i, m)
        elif m == 8:
            assert sl, (i, m)
        else:
            assert m == 9 and not sl and (want[62] & 0x40), (i, m)
    assert n_ok > 300
    # every possible move is undone by its reverse
    sel = np.flatnonzero(ok)
    back, ok2, rev2 = elp.patch_moves(out[sel], rev[sel])
    assert np.all(ok2 == 1)
    assert np.array_equal(back, patches[sel])
    assert np.array_equal(rev2, moves[sel])


def test_tile_path_matches_patch_functions():
    """The learned engine's kernel works on the tile itself: features from an
    occupancy bitmap, moves written to the tile at once and restored on
    rejection.  For 16^3 windows of a melt, sites anywhere in the tile (faces
    included, where patch sites fall outside it) and every move, the kernel's
    functions must equal the patch functions checked above: features, move
    outcome, reverse move, written bytes, bitmap after the flips, and the
    tile and bitmap after the revert."""
    _say("[learned] tile path vs patch functions")
    h = elp.RingHost(N=100, L=32, density=1.2, seed=6)
    lat0 = h.lattice()
    h.load_shift((16, 16, 16))
    lat = replay_ring(lat0, h.L, (16, 16, 16), elp.SUBLAT_DENSE, 13, h.origins((16, 16, 16), 20), 0, h.seed,
                      rules=elp.RULES_BALANCED).reshape(32, 32, 32)     # [v][u][t]
    rng = np.random.default_rng(8)
    tiles = []
    for _ in range(48):
        o = rng.integers(0, 32, 3)
        w = np.roll(lat, (-o[2], -o[1], -o[0]), axis=(0, 1, 2))[:16, :16, :16]
        tiles.append(np.ascontiguousarray(w).reshape(-1))
    tiles = np.stack(tiles)
    cases = []
    for k in range(1400):
        ti = int(rng.integers(0, len(tiles)))
        occ = np.flatnonzero(tiles[ti])
        site = int(rng.choice(occ)) if k % 10 else int(rng.integers(0, 4096))
        t, u, v = site & 15, (site >> 4) & 15, site >> 8
        for m in range(11):
            cases.append((ti, t, u, v, m))
    cases = np.array(cases, np.int32)
    flags, ok = elp.tile_check(tiles, cases)
    bad = np.flatnonzero(flags)
    assert len(bad) == 0, (len(bad), cases[bad[:5]].tolist(), flags[bad[:5]].tolist())
    face = np.any((cases[:, 1:4] == 0) | (cases[:, 1:4] == 15), axis=1)
    _say(f"  {len(cases)} cases, {int(ok.sum())} possible moves, {int(ok[face].sum())} of them at a tile face")
    assert ok.sum() > 800 and ok[face].sum() > 150


def test_features_match_definition():
    P = _melt_patches(300, seed=5)
    P[:20, :25] = 0xFF                                  # outside-the-tile rows of a patch
    x = elp.patch_features(P)
    want = np.stack([features_of(p) for p in P])
    assert np.array_equal(x, want)


def test_training_lowers_loss_and_proposes_possible_moves():
    _say("[learned] harvest and training")
    h = elp.RingHost(N=100, L=32, density=1.2, seed=6)
    eng = elp.Engine("ring", h.L, (16, 16, 16), elp.SUBLAT_DENSE, 13)
    eng.set_rules(elp.RULES_BALANCED)
    eng.upload(h.lattice())
    h.load_shift((16, 16, 16))
    eng.run(h.origins((16, 16, 16), 200), tau0=0, seed=h.seed)
    net = elp.Net(seed=2)
    net.harvest(eng, 200000, seed=7)
    x, y = net.samples()
    assert len(y) == 200000 and np.all((y >= 0) & (y < 11))
    # a harvested label is always a possible move of its site
    p0 = net.forward(x[:2000])
    loss0 = float(-np.mean(np.log(p0[np.arange(2000), y[:2000]])))
    loss = net.train(epochs=4, batch=4096, lr=3e-3)
    p1 = net.forward(x[:2000])
    loss1 = float(-np.mean(np.log(p1[np.arange(2000), y[:2000]])))
    _say(f"  cross-entropy {loss0:.3f} -> {loss1:.3f} (epoch mean {loss:.3f})")
    assert loss1 < 0.8 * loss0
    # the trained network puts most of its mass on moves that are possible
    P = _melt_patches(400, seed=8)
    X = np.stack([features_of(p) for p in P])
    q = net.forward(X)
    possible = np.zeros((len(P), 11))
    for m in range(11):
        _, ok, _ = elp.patch_moves(P, np.full(len(P), m))
        possible[:, m] = ok
    mass = np.sum(q * possible, axis=1)
    has = possible.sum(axis=1) > 0
    _say(f"  proposal mass on possible moves: {mass[has].mean():.3f} "
         f"(uniform proposal: {np.mean(possible[has].sum(axis=1) / 11):.3f})")
    assert mass[has].mean() > 1.5 * np.mean(possible[has].sum(axis=1) / 11)
    net.close()
    eng.close()


def _chi2_uniform(samples, states):
    from scipy import stats
    nsl, rg = ring_observables(np.array(samples))
    ensl, erg = ring_observables(states)
    keys = sorted(set(zip(ensl.tolist(), erg.tolist())))
    expect = np.array([np.sum((ensl == a) & (erg == b)) for a, b in keys], float) / len(states)
    got = np.array([np.sum((nsl == a) & (rg == b)) for a, b in keys], float)
    e = expect * len(samples)
    big = e >= 10
    chi2 = float(np.sum((got[big] - e[big]) ** 2 / e[big]))
    dof = int(big.sum()) - 1
    return chi2, dof, stats.chi2.sf(chi2, dof)


def _six_ring_start():
    states = ring_reachable(6)
    start = next(tuple(int(b) for b in s) for s in states if ring_observables(s)[0][0] == 2)
    return states, start


def test_learned_chain_samples_uniform_measure_of_six_ring():
    """One 6-monomer ring in a 16^3 box; the learned proposal (a random
    network, far from uniform: Glorot weights) with the MH test; one sample
    every 4 time units.  The joint histogram of (stored-length bonds, Rg^2)
    must match the uniform measure over the reachable states (chi^2,
    p > 1e-4)."""
    _say("[learned] six-ring stationary test")
    states, start = _six_ring_start()
    lat = ring_lattice(start)
    eng = elp.Engine("ring", (16, 16, 16), (16, 16, 16), elp.SUBLAT_CUBIC, 51)
    eng.upload(lat)
    net = elp.Net(seed=3)
    rng = np.random.default_rng(9)
    net.set_params(random_params(rng, scale=1.0))
    samples = []
    prop = acc = 0
    for k in range(1200):
        org = rng.integers(0, 16, (4, 3))
        p_, a_ = net.run(eng, org, tau0=4 * k, seed=11)
        prop += p_
        acc += a_
        d = ring_decode(eng.download(), 6)
        assert d is not None
        samples.append(d)
    chi2, dof, p = _chi2_uniform(samples, states)
    _say(f"  learned: chi2 {chi2:.1f} / {dof} dof, p = {p:.3g}; {acc} moves accepted, "
         f"{len(set(samples))} distinct states")
    assert p > 1e-4
    eng.close()
    net.close()


def test_plain_engine_six_ring_measure():
    """The same test for the plain engine (64 independent boxes as one batch):
    with balanced rates both sub-lattices sample the uniform measure; with the
    reference rates the histogram departs from it (the move set is not
    reversible), which also shows the test can tell."""
    states, start = _six_ring_start()
    lat = ring_lattice(start)
    for sub, n_iter, rules in ((elp.SUBLAT_CUBIC, 51, elp.RULES_BALANCED), (elp.SUBLAT_DENSE, 13, elp.RULES_BALANCED),
                               (elp.SUBLAT_CUBIC, 51, elp.RULES_REFERENCE)):
        _say(f"[learned] six-ring stationary test, plain engine, sublattice {sub}, rules {rules}")
        bat = elp.Batch("ring", [(16, 16, 16)] * 64, (16, 16, 16), sub, n_iter)
        bat.set_rules(rules)
        for r in range(64):
            bat.upload(r, lat)
        rng = np.random.default_rng(10 + sub)
        samples = []
        for k in range(60):
            org = rng.integers(0, 16, (4, 64, 3))
            bat.run(org, tau0=4 * k, seeds=np.arange(64) + 100, units_per_graph=0)
            for r in range(64):
                samples.append(ring_decode(bat.download(r), 6))
        chi2, dof, p = _chi2_uniform(samples, states)
        _say(f"  plain: chi2 {chi2:.1f} / {dof} dof, p = {p:.3g}")
        if rules == elp.RULES_BALANCED:
            assert p > 1e-4
        else:
            assert p < 1e-6
        bat.close()


def test_learned_statics_match_plain_balanced_engine():
    """Ring melts L=32, N=100, 8 seeds: equilibrated by the plain engine
    (balanced rates, 20000 time units), then 1500 more time units of either
    the plain engine or the learned proposal (trained on moves harvested from
    the plain engine), frames every 100.  Rg^2, stored-length ratio and P(s)
    agree within 3 sigma."""
    from elp.stats import Series
    tile = (16, 16, 16)
    net = elp.Net(seed=5)
    res = {"plain": [], "learned": []}
    acc = prop = 0
    for seed in range(8):
        _say(f"[learned statics] seed {seed}")
        h = elp.RingHost(N=100, L=32, density=1.2, seed=seed)
        eng = elp.Engine("ring", h.L, tile, elp.SUBLAT_CUBIC, 51)
        eng.set_rules(elp.RULES_BALANCED)
        eng.upload(h.lattice())
        h.load_shift(tile)
        eng.run(h.origins(tile, 20000), tau0=0, seed=h.seed)
        lat_eq = eng.download()
        if seed == 0:
            net.harvest(eng, 300000, seed=17)
            net.train(epochs=4, batch=4096, lr=3e-3)
        for kind in ("plain", "learned"):
            eng.upload(lat_eq)
            h.set_lattice(lat_eq)
            ser = Series(True, h.L, h.n_pol, h.N, 16)
            tau = 20000
            for k in range(15):
                org = h.origins(tile, 100)
                if kind == "plain":
                    eng.run(org, tau0=tau, seed=h.seed)
                else:
                    p_, a_ = net.run(eng, org, tau0=tau, seed=h.seed ^ 0x5A5A)
                    prop, acc = prop + p_, acc + a_
                tau += 100
                h.set_lattice(eng.download())
                ser.add_host(h)
            st = ser.statics(0)
            res[kind].append((st["rg2"], st["slrat"], *st["pc"][[2, 5, 10, 25, 50]]))
        eng.close()
    a, b = np.array(res["plain"]), np.array(res["learned"])
    z = (a.mean(0) - b.mean(0)) / np.sqrt(a.var(0, ddof=1) / len(a) + b.var(0, ddof=1) / len(b))
    _say(f"  plain   {np.round(a.mean(0), 4)}\n  learned {np.round(b.mean(0), 4)}\n  z {np.round(z, 2)}; "
         f"learned acceptance {acc / prop:.4f}")
    assert np.all(np.abs(z) < 3), z
    net.close()


def test_trained_network_raises_accepted_moves_per_attempt():
    """What the trained network gains over the same chain without it: on an
    equilibrated melt (L=32, N=100), the fraction of proposals accepted with
    the trained network against a uniform proposal over the 11 moves (all
    weights zero)."""
    tile = (16, 16, 16)
    h = elp.RingHost(N=100, L=32, density=1.2, seed=21)
    eng = elp.Engine("ring", h.L, tile, elp.SUBLAT_CUBIC, 51)
    eng.set_rules(elp.RULES_BALANCED)
    eng.upload(h.lattice())
    h.load_shift(tile)
    eng.run(h.origins(tile, 20000), tau0=0, seed=h.seed)
    lat_eq = eng.download()
    trained = elp.Net(seed=6)
    trained.harvest(eng, 300000, seed=23)
    trained.train(epochs=4, batch=4096, lr=3e-3)
    uniform = elp.Net(seed=7)
    uniform.set_params(np.zeros(NPARAM, np.float32))
    rates = {}
    for name, net in (("uniform", uniform), ("trained", trained)):
        eng.upload(lat_eq)
        p_, a_ = net.run(eng, h.origins(tile, 200), tau0=30000, seed=99)
        rates[name] = a_ / p_
    _say(f"  accepted per proposal: uniform {rates['uniform']:.4f}, trained {rates['trained']:.4f} "
         f"({rates['trained'] / rates['uniform']:.2f}x)")
    assert rates["trained"] > 2.0 * rates["uniform"]
    trained.close()
    uniform.close()
    eng.close()


def test_learned_run_at_c5_size():
    """The learned-proposal configuration at its full size: ring melt L=128,
    N=4000, density 1.2 (629 rings) from the compact start, with a network
    trained on 10^6 moves harvested from the plain engine on the L=96, N=1000
    melt.  Over 20 learned time units: one proposal per slot and substep is
    counted, some but not all are accepted, every ring stays intact with its
    label sequence and is recovered by the host, the lattice moves, and the
    features of sampled full-size patches equal their definition."""
    from ringcheck import ring_label_sequences, site_invariants
    _say("[learned C5] harvest on L=96 N=1000")
    tile = (16, 16, 16)
    h2 = elp.RingHost(N=1000, L=96, density=1.2, seed=0)
    e2 = elp.Engine("ring", h2.L, tile, elp.SUBLAT_DENSE, 13)
    e2.set_rules(elp.RULES_BALANCED)
    e2.upload(h2.lattice())
    h2.load_shift(tile)
    e2.run(h2.origins(tile, 300), tau0=0, seed=h2.seed)
    net = elp.Net(seed=31)
    net.harvest(e2, 1_000_000, seed=37)
    loss = net.train(epochs=2, batch=4096, lr=3e-3)
    e2.close()
    _say(f"  trained: last-epoch loss {loss:.3f}")
    _say("[learned C5] L=128 N=4000, 20 learned time units")
    h = elp.RingHost(N=4000, L=128, density=1.2, seed=0)
    assert h.n_pol == 629
    lat0 = h.lattice()
    before = ring_label_sequences(lat0, h.L)
    eng = elp.Engine("ring", h.L, tile, elp.SUBLAT_CUBIC, 51)
    eng.set_rules(elp.RULES_BALANCED)
    eng.upload(lat0)
    h.load_shift(tile)
    n_time = 20
    prop, acc = net.run(eng, h.origins(tile, n_time), tau0=0, seed=h.seed)
    ntile = (128 // 16) ** 3
    assert prop == n_time * ntile * 64 * 4 * 51
    assert 0 < acc < prop
    got = eng.download()
    _say(f"  {acc} of {prop} proposals accepted; {np.count_nonzero(got != lat0)} sites changed")
    assert np.count_nonzero(got != lat0) > 10000
    assert site_invariants(got, h.L) == 629 * 4000
    assert ring_label_sequences(got, h.L) == before
    h.set_lattice(got)                                   # host recovery of every labelled ring
    rng = np.random.default_rng(12)
    P = patches_of(got, h.L, rng.choice(np.flatnonzero(got), 3000, replace=False))
    assert np.array_equal(elp.patch_features(P), np.stack([features_of(p) for p in P]))
    eng.close()
    net.close()
