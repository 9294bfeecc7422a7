"""Batched scenes on the GPU: every scene of a batch steps bit for bit as it
does alone (bodies, static-row lambdas, colors), in both coloring modes."""
import numpy as np
import pytest

from avbd_amd import scenes
from avbd_amd.batched import BatchedSolver6DOF
from avbd_amd.solver_6dof import Solver6DOF
from tests.parity import one_substep

pytestmark = pytest.mark.gpu


def _builders(mode):
    kw = dict(coloring_mode=mode, solver_cls=Solver6DOF)

    def single(**_):
        s = Solver6DOF(dt=1.0 / 60.0, iterations=10, substeps=4, friction_static_mult=1.25,
                       coloring_mode=mode)
        s.enable_self_collision(True, default_friction=0.4)
        b = s.add_box((0.0, 0.3, 0.0), (0.2, 0.3, 0.4), mass=3.0, friction=0.4)
        s.add_floor_contact_box(b, friction=0.4)
        return s

    return [lambda: scenes.random_pile(64, seed=2, lattice=(4, 4, 4), **kw),
            lambda: scenes.random_pile(27, seed=7, lattice=(3, 3, 3), **kw),
            lambda: single(),
            lambda: scenes.random_pile(120, seed=9, lattice=(5, 5, 5), **kw)]


@pytest.mark.parametrize("mode", ["jacobi", "jones_plassmann"])
def test_batch_steps_each_scene_as_alone(mode):
    builders = _builders(mode)
    alone = [b() for b in builders]
    batch = BatchedSolver6DOF([b() for b in builders])
    for s in alone:
        s._flush()
    batch._flush()
    n_rows = [len(s._rows) for s in alone]
    r0 = np.cumsum([0] + n_rows)
    for k in range(160):
        for s in alone:
            one_substep(s)
        one_substep(batch)
        bb = batch._h.read_bodies(len(batch._x))
        rows_b = batch._h.read_rows(int(r0[-1]))["lambda"]
        colors_b = batch._h.read_colors(len(batch._x))
        for j, s in enumerate(alone):
            sl = batch.scene_slice(j)
            ba = s._h.read_bodies(len(s._x))
            for key in ("x", "q", "v", "w"):
                assert np.array_equal(ba[key], bb[key][sl]), (mode, k, j, key)
            assert np.array_equal(s._h.read_rows(n_rows[j])["lambda"],
                                  rows_b[r0[j]:r0[j + 1]]), (mode, k, j)
            assert np.array_equal(s._h.read_colors(len(s._x)), colors_b[sl]), (mode, k, j)
    print(mode, "substeps", 160, "bodies", len(batch._x), flush=True)
