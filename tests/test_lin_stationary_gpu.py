"""Stationary law of the linear-chain engine on the GPU, against the exact
law of one short chain.

The linear move set (gpupol/linpol.cu:106-363) is in detailed balance with
the weight 2^(stored-length monomers) (every forward/reverse rate ratio is
2^(change in stored length), tests/test_lin_balance.py).  With one chain of
four monomers in a 16^3 box and the reference's cubic sub-lattice, no two
monomers ever sit on one coset (they are at most three bonds apart, coset
sites four), so each substep is a mixture of single-site moves and the law
2^(n_SL) over the 1812 conformations (all connected) is exactly stationary.
The dense sub-lattice and a five-monomer chain allow two monomers of the
chain to move in one substep; they are checked against the same law.

64 boxes run as one batch on the GPU (tile = box, fresh uniform tile origin
per time unit, reference rates); after 400 time units of burn-in a
conformation is read from every box every 8 time units.  The joint
histogram of (stored-length bonds, Rg^2) must match the law (chi^2,
p > 1e-4), and must not match the uniform law over the same states (p < 1e-6:
the test can tell the 2^(n_SL) weight).
"""
import sys

import numpy as np
import pytest

import elp
from lincheck import chi2_weighted, lin_decode, lin_lattice, lin_states

pytestmark = pytest.mark.gpu
NBOX, BURN, EVERY, ROUNDS = 64, 400, 8, 400


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,sub,n_iter", [(4, elp.SUBLAT_CUBIC, 64), (4, elp.SUBLAT_DENSE, 16),
                                          (5, elp.SUBLAT_CUBIC, 64)])
def test_short_chain_samples_two_to_the_stored_length(n, sub, n_iter):
    _say(f"[lin stationary] n={n} sublattice={sub}")
    states = lin_states(n)
    lat = lin_lattice(tuple(int(b) for b in states[0]))
    bat = elp.Batch("lin", [(16, 16, 16)] * NBOX, (16, 16, 16), sub, n_iter)
    for r in range(NBOX):
        bat.upload(r, lat)
    rng = np.random.default_rng(40 + 3 * n + sub)
    seeds = np.arange(NBOX) + 1000 * n + 100 * sub
    bat.run(rng.integers(0, 16, (BURN, NBOX, 3)), tau0=np.zeros(NBOX, np.int64), seeds=seeds, units_per_graph=0)
    samples = []
    for k in range(ROUNDS):
        tau = np.full(NBOX, BURN + EVERY * k, np.int64)
        bat.run(rng.integers(0, 16, (EVERY, NBOX, 3)), tau0=tau, seeds=seeds, units_per_graph=0)
        for r in range(NBOX):
            d = lin_decode(bat.download(r), n)
            assert d is not None, (k, r)
            samples.append(d[0])
    bat.close()
    chi2, dof, p = chi2_weighted(samples, states, lambda s: 2.0 ** s)
    c_u, d_u, p_u = chi2_weighted(samples, states, lambda s: np.ones_like(s))
    _say(f"  2^n_SL: chi2 {chi2:.1f} / {dof} dof, p = {p:.3g}; uniform: chi2 {c_u:.0f} / {d_u}, p = {p_u:.3g}; "
         f"{len(set(samples))} distinct of {len(states)}")
    assert p > 1e-4
    assert p_u < 1e-6
