"""Particle solver on the GPU against the CPU checker, step by step: every
array the step updates (positions, velocities, previous velocities, row
lambda / penalty / active, tet lambda / penalty / active) must agree bit for
bit after every step."""
import math

import numpy as np
import pytest

from avbd_amd import particle_scenes as PS
from avbd_amd.solver import Solver
from oracle.oracle import particle_solver_class

pytestmark = pytest.mark.gpu

KEYS = ("x", "v", "prev_v", "lambda", "penalty", "active", "tet_lambda", "tet_penalty",
        "tet_active")


def _bits(a):
    """Bit patterns, every NaN mapped to one pattern (the payload of a NaN
    made by an invalid operation differs between x86 and gfx950)."""
    a = np.array(a)
    if a.dtype != np.float32:
        return a
    a[np.isnan(a)] = np.float32("nan")
    return a.view(np.int32)


def _compare(g, o, where):
    for k in KEYS:
        if not np.array_equal(_bits(g[k]), _bits(o[k])):
            bad = np.argwhere(_bits(g[k]) != _bits(o[k]))
            i = tuple(bad[0])
            raise AssertionError(f"{where}: {k} differs at {i} (gpu {g[k][i]!r}, "
                                 f"checker {o[k][i]!r}; {len(bad)} entries)")


def _lockstep(build, steps, hook=None, variant=None, **kw):
    print(f"lockstep {build.__name__} steps={steps} variant={variant} {kw}", flush=True)
    sg, so = Solver(**kw), particle_solver_class()(**kw)
    build(sg)
    build(so)
    if variant is not None:
        sg._flush()
        sg._h.set_option(1, variant)
    for f in range(steps):
        if hook is not None:
            hook(sg, f)
            hook(so, f)
        sg.step()
        so.step()
        _compare(sg.state(), so.state(), f"step {f}")
    return sg, so


def chain_fracture(s):
    PS.chain(s, 6, 0.5, fracture=30.0)


@pytest.mark.parametrize("variant", [2, 4])
def test_chain_with_fracture_lockstep(variant):
    sg, _ = _lockstep(chain_fracture, 150, variant=variant, dt=1 / 60, iterations=10)
    assert sg.active()[3:].min() == 0  # a link broke during the run


def soft_block(s):
    PS.soft_block(s, n=(3, 3, 3), size=(0.4, 0.4, 0.4), center=(0.0, 0.5, 0.0))


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5])
def test_soft_block_on_floor_lockstep(variant):
    """tets, soft distance rows, floor rows with friction rows, post-stab and
    the finalize after it (solver.py:1374-1390); every primal kernel."""
    _lockstep(soft_block, 90, variant=variant, dt=1 / 60, iterations=12)


def cloth(s):
    PS.cloth(s, 20, 16, spacing=0.05)


def test_cloth_without_post_stab_lockstep():
    _lockstep(cloth, 60, dt=1 / 60, iterations=8, post_stabilize=False, max_speed=2.0)


def plates(s):
    for k in range(8):
        b = s.add_particle((0.1 * k, 0.3, 0.05 * k), 1.0 + 0.1 * k, collide=False, friction=0.3)
        s.add_floor_contact(b, floor_y=0.0)
        s.add_plane_contact(b, (1.0, 0.2, 0.0), -0.5)
    a, b = s.add_particle((2.0, 1.0, 0.0), 1.0), s.add_particle((2.0, 0.7, 0.0), 1.0)
    s.add_pin_axes(a, (2.0, 1.0, 0.0), (False, True, True))
    s.add_distance(a, b, 0.3, stiffness=2.0e4)


def plates_hook(s, f):
    if f == 10:
        s.set_plane_contact_bulk([type("H", (), {"index": 1 + 4 * k})() for k in range(8)],
                                 offset=0.2)
    if f == 20:
        s.set_velocity(type("B", (), {"index": 3})(), (1.0, 2.0, 0.0))
        s.add_impulse(type("B", (), {"index": 8})(), (0.5, 0.0, 0.5))
    if f == 30:  # a body added mid-run: re-flush with state kept
        c = s.add_particle((3.0, 0.5, 0.0), 1.0, collide=False)
        s.add_floor_contact(c, friction=0.6)


def test_plates_perturbations_and_reflush_lockstep():
    _lockstep(plates, 60, hook=plates_hook, dt=1 / 60, iterations=10)


@pytest.mark.parametrize("variant", [2, 4])
def test_many_soft_blocks_bit_exact(variant):
    """64 soft blocks of 6^3 cells (21,952 particles, 160,512 rows, 82,944
    tets), 4 steps, per color launches and the dataflow sweep."""
    def many(s):
        PS.soft_blocks(s, 64)
    sg, _ = _lockstep(many, 4, variant=variant, dt=1 / 60, iterations=20)
    assert np.isfinite(sg.positions()).all()


def test_empty_scene_and_static_only():
    s = Solver()
    s.step()
    assert s.positions().shape == (0, 3)
    s.add_particle((1.0, 2.0, 3.0), 0.0)
    for _ in range(3):
        s.step()
    assert s.positions()[0].tolist() == [1.0, 2.0, 3.0]


def test_flush_rejects_bad_rows():
    """A row naming a missing body fails in the host's coloring, as in the
    reference (coloring.py:31-37 indexes the adjacency list); a scene handed
    to the library with out-of-range entries is refused before any launch."""
    from avbd_amd import _native as N
    s = Solver()
    a = s.add_particle((0, 0, 0), 1.0)
    s.add_particle((0, 1, 0), 1.0)
    s.add_distance(a, type("B", (), {"index": 7})(), 1.0)
    with pytest.raises(IndexError):
        s.step()
    good = Solver()
    PS.chain(good, 3, 0.5)
    arrays = good._scene_arrays(None)
    for key, val in (("con_idx", 99), ("body_b", 9), ("sibling", 5), ("color", 40)):
        bad = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in arrays.items()}
        bad[key][-1] = val
        if key == "sibling":
            bad["type"][-1] = 6  # a friction row whose normal row is not before it
        h = N.ParticleHandle(0)
        with pytest.raises(N.AvbdError):
            h.flush(bad)
        h.close()


def test_graph_replay_matches_direct_launches():
    """A step replayed from its captured graph gives the state of the same
    step launched kernel by kernel, replay after replay, with host reads and
    writes between replays and a re-capture when the parameters change."""
    print("graph vs direct: 9 soft blocks", flush=True)
    sa, sb = Solver(dt=1 / 60, iterations=12), Solver(dt=1 / 60, iterations=12)
    for s in (sa, sb):
        PS.soft_blocks(s, 9, n=(4, 4, 4))
        s._flush()
    sb._h.set_graph(False)
    for f in range(12):
        if f == 5:
            for s in (sa, sb):
                s.set_velocity(type("B", (), {"index": 3})(), (0.0, 3.0, 0.0))
        if f == 8:
            for s in (sa, sb):
                s.iterations = 7
        sa.step()
        sb.step()
        _compare(sa.state(), sb.state(), f"step {f}")


def contacts_mixed(s):
    """Spheres, cubes and a pillar falling onto each other and the floor,
    with friction: sphere, sphere-box and box-box rows and their friction
    rows, generated by the host every step."""
    from avbd_amd.scene import Shape
    r = 0.15
    bodies = []
    for k in range(3):
        bodies.append(s.add_particle((0.02 * k, r + 0.32 * k, 0.0), 2.0 - 0.5 * k,
                                     shape=Shape("cube", (r, r, r)), friction=0.6))
    for k in range(4):
        bodies.append(s.add_particle((0.5 + 0.05 * k, 0.2 + 0.3 * k, 0.03 * k), 1.0,
                                     shape=Shape("sphere", (0.1,)), friction=0.4))
    bodies.append(s.add_particle((-0.45, 0.35, 0.0), 3.0, shape=Shape("pillar", (0.1, 0.35)),
                                 velocity=(1.5, 0.0, 0.0), friction=0.5))
    bodies.append(s.add_particle((0.2, 1.5, 0.0), 1.0, shape=Shape("sphere", (0.12,)),
                                 velocity=(0.0, -2.0, 0.0)))
    for b in bodies:
        s.add_floor_contact(b, floor_y=0.1)
    s.enable_self_collision(True, default_friction=0.3)


def test_body_contacts_lockstep():
    sg, _ = _lockstep(contacts_mixed, 120, dt=1 / 60, iterations=15)
    types = set(int(t) for t in sg.c_type.numpy().tolist())
    assert {7, 8} <= types  # sphere-box and box-box rows were exercised


def broken_pair(s):
    from avbd_amd.scene import Shape
    a = s.add_particle((0.0, 2.0, 0.0), 5.0, shape=Shape("sphere", (0.1,)))
    s.add_floor_contact(a, floor_y=0.1)
    s.add_pin(a, (0.0, 2.0, 0.0))
    b = s.add_particle((0.0, 1.5, 0.0), 1.0, shape=Shape("sphere", (0.1,)))
    s.add_floor_contact(b, floor_y=0.1)
    s.add_distance(a, b, rest=0.5, fracture=5.0)
    s.enable_self_collision(True)


def broken_pair_hook(s, f):
    if f == 40:
        s.set_position(type("B", (), {"index": 1})(), (0.0, 2.3, 0.0))
        s.set_velocity(type("B", (), {"index": 1})(), (0.0, 0.0, 0.0))


def test_fracture_then_contact_lockstep():
    sg, _ = _lockstep(broken_pair, 120, hook=broken_pair_hook, dt=1 / 60, iterations=25)
    assert int((sg.c_type.numpy() == 5).sum()) >= 0


def test_bench_scene_two_steps_bit_exact():
    """The measured workload at full size: 256 soft blocks (87,808
    particles, 642,048 rows, 331,776 tets), 2 steps against the checker."""
    def bench_scene(s):
        PS.soft_blocks(s, 256)
    sg, _ = _lockstep(bench_scene, 2, dt=1 / 60, iterations=20)
    assert len(sg._bodies_x) == 87808


def test_autotune_keeps_the_trajectory():
    """Autotuning steps the scene with each variant; the result equals the
    checker stepped the same number of times."""
    sg, so = Solver(dt=1 / 60, iterations=10), particle_solver_class()(dt=1 / 60, iterations=10)
    for s in (sg, so):
        PS.soft_blocks(s, 4, n=(3, 3, 3))
    times = sg.autotune_primal(steps=2, variants=(2, 4))
    assert set(times) == {2, 4} and sg.primal_variant in (2, 4)
    for _ in range(2 * (1 + 2)):
        so.step()
    _compare(sg.state(), so.state(), "after autotune")


def test_many_sweeps_fall_back_to_per_color_launches():
    """With 4,096 or more sweeps a step the dataflow stamps would repeat
    within a step, so the step uses the per-color launches; the result
    still matches the checker."""
    _lockstep(chain_fracture, 1, variant=4, dt=1 / 60, iterations=4100)
