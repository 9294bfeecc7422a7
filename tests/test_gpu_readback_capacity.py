"""The readback wire format and the fixed-capacity path on the HIP library.

* read_state_batched (solver_6dof.py:1952-2052; packing kernels_6dof.py:
  2154-2177, 2540-2555) against the checker's own readers after lockstep
  substeps: 10 floats per body [x, q, w], 3 per live row
  [lambda, active, was_static * 16 + type].
* unsafe_fixed_capacity (solver_6dof.py:1178-1230): a dense cluster whose pairs
  overflow the pool steps bit for bit with the checker, grows nothing and
  warns once.
"""
import warnings

import numpy as np
import pytest

from avbd_amd import scenes
from avbd_amd.solver_6dof import Solver6DOF
from oracle.oracle import oracle_solver_class
from tests.parity import assert_bit_exact, diff_report, one_substep, pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("include_rows", [True, False])
def test_read_state_batched_matches_checker(include_rows):
    g, o = pair(lambda **kw: scenes.pyramid(4, seed=11, **kw))
    for _ in range(12):
        one_substep(g)
        one_substep(o)
    assert_bit_exact(diff_report(g, o), "before the packed read")
    st = g.read_state_batched(include_rows=include_rows)
    nb = len(o._x)
    ob = o._h.read_bodies(nb)
    assert np.array_equal(st["positions"], ob["x"])
    assert np.array_equal(st["orientations"], ob["q"])
    assert np.array_equal(st["angular_velocities"], ob["w"])
    if not include_rows:
        assert st["n_rows"] == 0 and st["lambdas"].size == 0 and st["c_type"].size == 0
        return
    n_live = o._h.read_counters()[0]
    assert st["n_rows"] == n_live and n_live > len(o._rows)   # static and contact rows
    rows = o._h.read_rows(n_live)
    assert np.array_equal(st["lambdas"], rows["lambda"])
    assert np.array_equal(st["active"], rows["active"])
    assert np.array_equal(st["was_static"], rows["was_static"])
    assert np.array_equal(st["c_type"], rows["type"])
    assert set(np.unique(st["c_type"])) <= {0, 1, 3}
    assert st["was_static"].any()   # the resting pyramid has sticking contacts


def _dense_cluster(cls, n=60):
    s = cls(gravity=(0, -9.8, 0), iterations=4, unsafe_fixed_capacity=True)
    for i in range(n):
        s.add_box(position=(0.01 * i, 2.0, 0.0), half_extents=(0.5, 0.5, 0.5), mass=1.0,
                  friction=0.3)
    s.enable_self_collision(True)
    return s


def test_fixed_capacity_lockstep_and_single_warning():
    g, o = _dense_cluster(Solver6DOF), _dense_cluster(oracle_solver_class())
    g._flush()
    o._flush()
    cap = g._h.read_capacities()[1:]
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        for k in range(6):
            g.step()
            o.step()
            assert_bit_exact(diff_report(g, o), f"fixed-capacity frame {k}")
    msgs = [str(x.message) for x in w if issubclass(x.category, RuntimeWarning)]
    # one warning per solver (GPU and checker), never a regrow
    assert sum("fixed-capacity: broadphase pair pool overflow" in m for m in msgs) == 2
    assert not any("growing for next substep" in m for m in msgs)
    assert g._h.read_capacities()[1:] == cap and g._bp_max_pairs == 960
    st = g.read_state_batched()
    assert st["n_rows"] == g._h.read_counters()[0] and st["lambdas"].size == st["n_rows"]
