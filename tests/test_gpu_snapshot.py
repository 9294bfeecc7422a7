"""One substep from an identical state: the HIP solver runs a scene for a
while on its own, its full carried-over state (bodies, static rows, warm-start
cache, coloring, capacities) is loaded into the checker, both take one substep
and every observable is compared. Bar: bit-exact on coloring, pair set, SAT
axes and contact counts; x, q, lambda equal within 1e-5 relative (FP32)."""
import numpy as np
import pytest

from avbd_amd import scenes
from avbd_amd.solver_6dof import Solver6DOF
from oracle.oracle import oracle_solver_class
from tests.parity import diff_report, load_snapshot, one_substep, snapshot

pytestmark = pytest.mark.gpu

CASES = [
    ("pyramid12_f30", lambda cls: scenes.pyramid(12, seed=21, solver_cls=cls), 30),
    ("pile300_f45", lambda cls: scenes.random_pile(300, seed=22, lattice=(7, 7, 7), solver_cls=cls), 45),
    ("pyramid12_jp_f20", lambda cls: scenes.pyramid(12, seed=23, solver_cls=cls,
                                                    coloring_mode="jones_plassmann"), 20),
    # the full-size scenes of the throughput measurement: the 98,021-cube
    # pyramid (66 layers, 20 iterations) after 2 frames, and the 10,000-box
    # pile after 60 frames (its lower layers landed)
    ("pyramid66_f2", lambda cls: scenes.pyramid(66, seed=0, solver_cls=cls), 2),
    ("pile10k_f60", lambda cls: scenes.random_pile(10000, seed=0, solver_cls=cls), 60),
]


@pytest.mark.parametrize("name,build,frames", CASES, ids=[c[0] for c in CASES])
def test_one_substep_from_identical_state(name, build, frames):
    g = build(Solver6DOF)
    for _ in range(frames):
        g.step()
    o = build(oracle_solver_class())
    load_snapshot(o, snapshot(g))
    one_substep(g)
    one_substep(o)
    rep = diff_report(g, o)
    print(name, rep, flush=True)
    for key in ("pairs_equal", "sat_equal", "contact_counts_equal", "colors_equal",
                "counters_equal", "rows_active_equal", "rows_type_equal"):
        assert rep.get(key, True), (key, rep)
    nb = len(g._x)
    bg, bo = g._h.read_bodies(nb), o._h.read_bodies(nb)
    for key in ("x", "q"):
        a, b = bg[key].astype(np.float64), bo[key].astype(np.float64)
        assert np.all(np.abs(a - b) <= 1e-5 * np.maximum(np.abs(b), 1.0)), key
    assert rep["rows_lambda_maxrel"] <= 1e-5, rep["rows_lambda_maxrel"]
