"""Broad-phase traversal with 1, 8, 16 and 64 lanes per body
(AVBD_OPT_BROAD_LANES), each in lockstep with the CPU checker.

The checker's pair list is the brute-force AABB test in canonical order
(kernels_6dof.py:1946-1978 restated); every lane count must give the same
pairs, and so the same trajectory, bit for bit. The parallel rods give
bodies more partners than the count pass keeps in its 64 slots, so the write
pass walks the tree again for them."""
import numpy as np
import pytest

from avbd_amd import _native as N
from avbd_amd import scenes
from avbd_amd.solver_6dof import Solver6DOF
from oracle.oracle import oracle_solver_class
from tests.parity import assert_bit_exact, diff_report, one_substep

pytestmark = pytest.mark.gpu

LANES = (1, 8, 16, 64)


def _pile(cls, **kw):
    return scenes.random_pile(150, seed=9, lattice=(6, 6, 5), solver_cls=cls, **kw)


def _rods(cls, **kw):
    # 100 parallel rods along (1, 1, 1), 2 m long, on a 10 x 10 grid 0.1 m apart
    # across their axis: every AABB overlaps every other (99 partners, more than
    # 64), the rods themselves stay 0.06 m apart, so no contacts between them
    d = np.ones(3) / np.sqrt(3.0)
    axis = np.cross((1.0, 0.0, 0.0), d)
    ang = np.arccos(d[0])
    axis /= np.linalg.norm(axis)
    q = (*(np.sin(ang / 2) * axis), np.cos(ang / 2))
    e1 = np.cross(d, (0.0, 1.0, 0.0))
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(d, e1)
    s = cls(dt=1.0 / 60.0, iterations=4, substeps=2, **kw)
    s.enable_self_collision(True, default_friction=0.5)
    for u in range(10):
        for v in range(10):
            p = np.array((0.0, 3.0, 0.0)) + 0.1 * (u - 4.5) * e1 + 0.1 * (v - 4.5) * e2
            b = s.add_box(tuple(float(c) for c in p), (1.0, 0.02, 0.02), mass=1.0,
                          orientation=tuple(float(c) for c in q), friction=0.5)
            s.add_floor_contact_box(b, friction=0.5)
    return s


def _with_lanes(s, lanes):
    s._flush()
    s._h.set_option(N.OPT_BROAD_LANES, lanes)
    return s


@pytest.mark.parametrize("lanes", LANES)
@pytest.mark.parametrize("build,nsub", [(_pile, 16), (_rods, 6)], ids=["pile150", "rods100"])
def test_lanes_lockstep(build, nsub, lanes):
    g = _with_lanes(build(Solver6DOF), lanes)
    o = build(oracle_solver_class())
    for k in range(nsub):
        one_substep(g)
        one_substep(o)
        rep = diff_report(g, o)
        print(build.__name__, lanes, k, rep["counters"], flush=True)
        assert_bit_exact(rep, f"{build.__name__} lanes {lanes} substep {k}")
    if build is _rods:
        assert g._h.read_capacities()[0] >= 4950   # all 100 x 99 / 2 AABB pairs kept


@pytest.mark.parametrize("lanes", (1, 16))
def test_lanes_in_resident_frames(lanes):
    """The lane count is part of the captured substep: frames replayed from the
    graph match the checker stepped frame by frame."""
    g = _with_lanes(_pile(Solver6DOF, device="cuda:0"), lanes)
    o = _pile(oracle_solver_class(), device="cuda:0")
    assert g._resident_on(g.device)
    for f in range(6):
        g.step()
        o.step()
        bg, bo = g._h.read_bodies(len(g._x)), o._h.read_bodies(len(o._x))
        for key in ("x", "q", "v", "w"):
            assert np.array_equal(bg[key], bo[key]), (f, key)
    assert g._h.sync_count()[1] > 0


def test_lanes_option_is_validated():
    s = _pile(Solver6DOF)
    s._flush()
    for bad in (-1, 2, 4, 32, 128):
        with pytest.raises(N.AvbdError):
            s._h.set_option(N.OPT_BROAD_LANES, bad)
    for good in (0, 1, 8, 16, 64):
        s._h.set_option(N.OPT_BROAD_LANES, good)


def test_batched_scenes_pair_only_within_a_scene_at_every_lane_count():
    """Batched scenes share one tree over per-scene shifted boxes; each lane
    count keeps pairs inside a scene and gives the one-lane trajectory."""
    from avbd_amd.batched import BatchedSolver6DOF

    def batch():
        kw = dict(solver_cls=Solver6DOF)
        return BatchedSolver6DOF([scenes.random_pile(64, seed=2, lattice=(4, 4, 4), **kw),
                                  scenes.random_pile(27, seed=7, lattice=(3, 3, 3), **kw),
                                  scenes.random_pile(120, seed=9, lattice=(5, 5, 5), **kw)])

    runs = {lanes: _with_lanes(batch(), lanes) for lanes in LANES}
    for k in range(40):
        for s in runs.values():
            one_substep(s)
        ref = runs[1]._h.read_bodies(len(runs[1]._x))
        for lanes, s in runs.items():
            b = s._h.read_bodies(len(s._x))
            for key in ("x", "q", "v", "w"):
                assert np.array_equal(ref[key], b[key]), (k, lanes, key)
            assert s._h.read_counters() == runs[1]._h.read_counters(), (k, lanes)
