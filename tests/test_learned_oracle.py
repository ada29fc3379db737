"""The FP64 network oracle against central finite differences, and the exact
state space of a 6-monomer ring that the stationary tests compare with."""
import numpy as np

from learncheck import NPARAM, mlp_forward, mlp_grad, random_features, random_params, ring_observables, ring_states


def test_fp64_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    prm = random_params(rng).astype(np.float64)
    x = random_features(rng, 24)
    y = rng.integers(0, 11, 24)
    g, loss = mlp_grad(prm, x, y, double=True)
    assert loss > 0
    idx = np.concatenate([rng.choice(9216, 40, replace=False), 9216 + rng.choice(4096, 30, replace=False),
                          13312 + rng.choice(64, 10, replace=False), 13376 + rng.choice(1024, 20, replace=False),
                          14400 + np.arange(11)])
    h = 1e-6
    for i in idx:
        p1, p2 = prm.copy(), prm.copy()
        p1[i] += h
        p2[i] -= h
        fd = (mlp_grad(p1, x, y, double=True)[1] - mlp_grad(p2, x, y, double=True)[1]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
    # logits 11..15 are not used: their parameters have zero gradient
    assert np.all(g[14400 + 11:] == 0)
    assert np.all(g[13376:14400].reshape(64, 16)[:, 11:] == 0)


def test_fp64_forward_is_a_distribution():
    rng = np.random.default_rng(1)
    p = mlp_forward(random_params(rng), random_features(rng, 50))
    assert p.shape == (50, 11) and np.all(p > 0) and np.allclose(p.sum(axis=1), 1.0)
    # a network with zero last layer proposes uniformly
    prm = random_params(rng)
    prm[13376:] = 0
    assert np.allclose(mlp_forward(prm, random_features(rng, 5)), 1 / 11)


def test_six_ring_state_space():
    s = ring_states(6)
    # without stored length: closed 6-step self-avoiding polygons through
    # monomer 0 on the FCC lattice (no site twice)
    nsl, rg = ring_observables(s)
    assert len(s) == len({tuple(r) for r in s})
    assert np.all(nsl <= 3)
    # three stored-length bonds leave an equilateral triangle (Rg^2 = 1/3 per
    # distinct site, 6 monomers on 3 sites): 8 triangle orientations x 2
    # senses per labelled placement
    tri = s[nsl == 3]
    assert np.all(rg[nsl == 3] == 12)                     # Rg^2 = 1/3 -> x36 = 12
    # every valid sequence with one stored-length bond is a closed 5-step
    # polygon with one repeated monomer: count = 6 positions x #5-polygons
    five = ring_states(5)
    n5 = int(np.sum(ring_observables(five)[0] == 0))
    assert int(np.sum(nsl == 1)) == 6 * n5
    assert len(tri) > 0


def test_six_ring_move_set_connects_the_reachable_states():
    """Breadth-first search over the moves of the oracle (every site, every
    move, every value of its random bits) from a state with a corner reaches
    exactly the states that have stored length or a corner (408 of the 24192
    six-monomer states have neither: no move applies to them and none leads
    to them).  The stationary tests compare with the uniform measure over
    the reachable class."""
    from learncheck import lib, ring_decode, ring_lattice, ring_reachable
    o = lib()
    states = {tuple(int(b) for b in s) for s in ring_reachable(6)}
    assert len(ring_states(6)) - len(states) == 408
    hexagon = None
    for s in sorted(states):
        if ring_observables(np.array(s))[0][0] == 0:
            hexagon = s
            break
    seen = {hexagon}
    frontier = [hexagon]
    while frontier:
        nxt = []
        for st in frontier:
            lat = ring_lattice(st)
            assert ring_decode(lat, 6) == st
            for x in np.flatnonzero(lat):
                t, u, v = x % 16, (x // 16) % 16, x // 256
                for typ, nr in ((0, 8), (1, 1), (2, 2)):
                    for r in range(nr):
                        a = lat.copy()
                        o.orc_ring_move(a.ctypes.data, 16, 16, 16, typ, int(t), int(u), int(v), r, 0)
                        d = ring_decode(a, 6)
                        assert d is not None and d in states, d
                        if d not in seen:
                            seen.add(d)
                            nxt.append(d)
        frontier = nxt
    assert seen == states, (len(seen), len(states))
