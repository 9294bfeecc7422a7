"""The checker's coloring against textbook greedy coloring: the reference's
serial 'jacobi' pass is first-fit greedy in body order; serial
Jones-Plassmann rounds are first-fit greedy in (priority desc, index asc)
order, cut at 24 rounds."""
import numpy as np
import pytest

from avbd_amd import scenes
from oracle.oracle import oracle_solver_class
from tests.parity import one_substep


def contact_graph(s):
    """Body adjacency from the emitted box-box rows of the last substep."""
    n = s._h.read_counters()[0]
    rows = s._h.read_rows(min(n, s._gpu_pool_n_capacity))
    nb = len(s._x)
    adj = [set() for _ in range(nb)]
    for t, a, b in zip(rows["type"], rows["body_a"], rows["body_b"]):
        if t == 2 or a < 0 or b < 0 or a == b:
            continue
        adj[a].add(int(b))
        adj[b].add(int(a))
    return adj


def greedy(order, adj, nb, blocker=None):
    color = [-1] * nb
    for i in order:
        used = {color[j] for j in adj[i] if color[j] >= 0}
        c = 0
        while c in used and c < 31:
            c += 1
        color[i] = c
    return color


@pytest.mark.parametrize("layers", [3, 5])
def test_jacobi_equals_index_greedy(layers):
    s = scenes.pyramid(layers, seed=5, solver_cls=oracle_solver_class())
    one_substep(s)
    nb = len(s._x)
    adj = contact_graph(s)
    assert sum(len(a) for a in adj) > 0
    got = list(s.body_color.numpy())
    assert got == greedy(range(nb), adj, nb)
    assert s.count_color_conflicts() == 0


@pytest.mark.parametrize("layers", [3, 5])
def test_jones_plassmann_equals_priority_greedy(layers):
    s = scenes.pyramid(layers, seed=6, solver_cls=oracle_solver_class(),
                       coloring_mode="jones_plassmann")
    one_substep(s)
    nb = len(s._x)
    adj = contact_graph(s)
    prio = s._flush_snapshot["priority"]
    order = sorted(range(nb), key=lambda i: (-float(prio[i]), i))
    ref = greedy(order, adj, nb)
    # round index of each body in the serial round loop
    rnd = {}
    for i in order:
        r = 0
        for j in adj[i]:
            if (prio[j] > prio[i]) or (prio[j] == prio[i] and j < i):
                r = max(r, rnd[j] + (1 if j > i else 0))
        rnd[i] = r
    ref = [c if rnd[i] < 24 else 0 for i, c in enumerate(ref)]
    assert list(s.body_color.numpy()) == ref


def test_priorities_are_numpy_default_rng_0xA5BD():
    """solver_6dof.py:1004-1009: default_rng(0xA5BD).uniform(0, 1, n) in f32;
    the first six are the ones the 6-box coloring scene uses."""
    S = oracle_solver_class()
    s = S()
    for i in range(6):
        s.add_box((0, i, 0), (0.5, 0.5, 0.5), mass=1.0)
    s._flush()
    p = s._flush_snapshot["priority"]
    assert np.allclose(p, [0.27599, 0.15845, 0.73974, 0.61207, 0.34449, 0.80416], atol=1e-5)
