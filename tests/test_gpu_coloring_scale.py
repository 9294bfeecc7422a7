"""Coloring at scale: the GPU coloring of a large scene equals first-fit
greedy (index order for 'jacobi', priority order cut at 24 rounds for
Jones-Plassmann) over the GPU's own contact graph of the same substep.
This is size-independent (no oracle substep needed), so it runs on scenes
with tens of thousands of bodies and long dependency chains."""
import numpy as np
import pytest

from avbd_amd import scenes
from tests.parity import one_substep
from tests.test_oracle_coloring import contact_graph, greedy

pytestmark = pytest.mark.gpu


def _jp_reference(adj, prio, nb):
    order = sorted(range(nb), key=lambda i: (-float(prio[i]), i))
    ref = greedy(order, adj, nb)
    rnd = {}
    for i in order:
        r = 0
        for j in adj[i]:
            if (prio[j] > prio[i]) or (prio[j] == prio[i] and j < i):
                r = max(r, rnd[j] + (1 if j > i else 0))
        rnd[i] = r
    return [c if rnd[i] < 24 else 0 for i, c in enumerate(ref)]


@pytest.mark.parametrize("layers", [40])
def test_jacobi_large_pyramid_equals_index_greedy(layers):
    s = scenes.pyramid(layers, seed=11)
    one_substep(s)
    nb = len(s._x)
    adj = contact_graph(s)
    assert sum(len(a) for a in adj) > 4 * nb
    got = [int(c) for c in s.body_color.numpy()]
    assert got == greedy(range(nb), adj, nb)
    assert s.count_color_conflicts() == 0


def test_jacobi_dense_pile_equals_index_greedy():
    s = scenes.random_pile(3000, seed=12)
    for _ in range(320):      # 80 frames: the lower layers have landed and touch
        one_substep(s)
    nb = len(s._x)
    adj = contact_graph(s)
    assert sum(len(a) for a in adj) > nb
    got = [int(c) for c in s.body_color.numpy()]
    assert got == greedy(range(nb), adj, nb)


@pytest.mark.parametrize("layers", [30])
def test_jones_plassmann_large_pyramid(layers):
    s = scenes.pyramid(layers, seed=13, coloring_mode="jones_plassmann")
    one_substep(s)
    nb = len(s._x)
    adj = contact_graph(s)
    prio = s._flush_snapshot["priority"]
    got = [int(c) for c in s.body_color.numpy()]
    assert got == _jp_reference(adj, prio, nb)
