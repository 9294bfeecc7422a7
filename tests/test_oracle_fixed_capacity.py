"""The fixed-capacity path (unsafe_fixed_capacity, solver_6dof.py:1178-1230,
1277-1309) on the CPU checker: pools are never grown, contacts beyond them
are dropped, and the overflow is reported once per solver from a
once-per-frame check of the reservation counters."""
import warnings

import numpy as np
import pytest

from oracle.oracle import oracle_solver_class


def _dense_cluster(fixed, n=60):
    # the reference's dense cluster (tests/test_robustness.py:26-46)
    s = oracle_solver_class()(gravity=(0, -9.8, 0), iterations=4,
                              unsafe_fixed_capacity=fixed)
    for i in range(n):
        s.add_box(position=(0.01 * i, 2.0, 0.0), half_extents=(0.5, 0.5, 0.5), mass=1.0,
                  friction=0.3)
    s.enable_self_collision(True)
    s._flush()
    return s


def test_fixed_capacity_drops_and_warns_once():
    s = _dense_cluster(True)
    cap0 = s._h.read_capacities()
    row_cap0 = s._gpu_pool_n_capacity
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        s.step()
        n_raw, n_pairs, _ = s._h.read_counters()
        for _ in range(4):
            s.step()
    msgs = [str(x.message) for x in w if issubclass(x.category, RuntimeWarning)]
    # 60 coincident boxes: 1,770 pairs > 960 pair slots. The pairs past the
    # pool are dropped before emission, and the row pool (4 contacts x 3
    # rows per pair slot) then holds every row of the kept pairs, so only the
    # pair warning fires, once over all frames.
    assert sum("fixed-capacity: broadphase pair pool overflow" in m for m in msgs) == 1
    assert not any("row pool overflow" in m for m in msgs)
    # nothing grew
    mp, row_cap, pool_max, hash_cap = s._h.read_capacities()
    assert (row_cap, pool_max, hash_cap) == cap0[1:] and s._gpu_pool_n_capacity == row_cap0
    assert mp == 960 and s._bp_max_pairs == 960
    # after the first frame: every pair counted, the rows of the kept ones
    assert n_pairs == 1770 and n_raw == row_cap0
    assert np.isfinite(s.positions()).all()


def test_growing_path_regrows_instead():
    s = _dense_cluster(False)
    with pytest.warns(RuntimeWarning, match="row pool overflow"):
        s.step()
    for _ in range(4):
        s.step()
    assert int(s.n_active_rows.numpy()[0]) <= s._gpu_pool_n_capacity


def test_resident_device_string_implies_fixed_capacity():
    # solver_6dof.py:663-678: a GPU device string with gpu_resident and the
    # fused shuffled group primal selects the resident, fixed-capacity loop
    cls = oracle_solver_class()
    assert cls(device="cuda:0")._fixed_capacity_mode("cuda:0")
    assert not cls()._fixed_capacity_mode("cpu")
    for kw in ({"gpu_resident": False}, {"primal_group_size": 1}, {"primal_fused": False},
               {"primal_shuffle": False}):
        s = cls(device="cuda:0", **kw)
        assert not s._resident_on(s.device), kw
    s = _dense_cluster(False)
    s.device = "cuda:0"
    s.step()   # a whole frame through step_frame on the checker
    assert np.isfinite(s.positions()).all()
