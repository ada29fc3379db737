"""efpol's transition table (CPU, oracle).

efpol (cpupol/efpol_init.c:314-420, SetTransTable; restated byte-exact in
oracle/efpol_port.c) moves one monomer per attempt: the bond pair (a, b)
around it becomes (c, d) with probability n(a,b->c,d)/16, n = the slots of
the pair that hold (c, d).  Excluded volume is symmetric, so the table fixes
the stationary measure:
  * sl_equal: n(a,b->c,d) = n(c,d->a,b) for every transition, so the
    uniform measure over conformations is stationary.  With at most two
    monomers per site (efpol_port from ... 2) the chain samples the same
    state space as the GPU engines, uniformly: the equilibrium of the GPU
    engine with balanced rates.
  * sl_double / sl_quad: n(a,b->c,d) / n(c,d->a,b) = w^(z(c,d) - z(a,b)),
    z = stored-length bonds of the pair, w = 2 / 4: detailed balance with
    weight w per stored-length bond (the models' names).
"""
import collections
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "bin", "efpol_port_ring")


def _table(model):
    out = subprocess.run([EXE, "table", model], capture_output=True, text=True, check=True).stdout
    n = collections.Counter()
    slots = collections.Counter()
    for line in out.splitlines():
        a, b, _, c, d = (int(x, 16) if i != 2 else int(x) for i, x in enumerate(line.split()))
        slots[(a, b)] += 1
        if (c, d) != (0, 0):
            n[(a, b, c, d)] += 1
    assert set(slots.values()) == {16}
    return n


def _zeros(*bonds):
    return sum(1 for x in bonds if x == 0)


@pytest.mark.parametrize("model,w", [("sl_equal", 1), ("sl_double", 2), ("sl_quad", 4)])
def test_table_detailed_balance_weight(model, w):
    n = _table(model)
    moves = [k for k in n if k[:2] != k[2:]]
    assert len(moves) > 300
    for a, b, c, d in moves:
        back = n[(c, d, a, b)]
        assert back > 0, (model, a, b, c, d)
        assert n[(a, b, c, d)] * w ** _zeros(a, b) == back * w ** _zeros(c, d), (model, a, b, c, d)


def test_sl_equal_is_symmetric_and_the_others_are_not():
    eq, dbl = _table("sl_equal"), _table("sl_double")
    assert all(v == eq[(k[2], k[3], k[0], k[1])] for k, v in eq.items())
    assert any(v != dbl[(k[2], k[3], k[0], k[1])] for k, v in dbl.items())
