"""The reference's particle Solver tests (tests/test_solver.py) that need no
body-body contact generation, written against a solver class fixture ``S``:
test_oracle_particle_suite.py binds it to the CPU checker,
test_gpu_particle_suite.py to the HIP solver. Assertions and tolerances are
the reference's. The tet-volume tests follow tests/test_deformable.py:140-162
and :205-228 on a tet-lattice block in place of the downloaded bunny mesh."""
import math

import numpy as np
import pytest

from avbd_amd import particle_scenes as PS
from avbd_amd.scene import Shape


def test_free_fall_matches_bdf1(S):
    dt, g, n_steps = 1.0 / 60.0, -9.81, 60
    s = S(dt=dt, iterations=2, gravity=(0.0, g, 0.0), post_stabilize=False)
    b = s.add_particle(position=(0.0, 10.0, 0.0), mass=1.0)
    for _ in range(n_steps):
        s.step()
    expected_y = 10.0 + g * dt * dt * n_steps * (n_steps + 1) / 2.0
    pos, vel = s.positions()[b.index], s.velocities()[b.index]
    assert pos[0] == pytest.approx(0.0, abs=1e-5)
    assert pos[2] == pytest.approx(0.0, abs=1e-5)
    assert pos[1] == pytest.approx(expected_y, abs=2e-3)
    assert vel[1] == pytest.approx(g * n_steps * dt, rel=1e-3)


def test_static_body_stays_put(S):
    s = S(dt=1.0 / 60.0, iterations=2, gravity=(0.0, -9.81, 0.0))
    b = s.add_particle(position=(1.0, 2.0, 3.0), mass=0.0)
    for _ in range(30):
        s.step()
    assert s.positions()[b.index] == pytest.approx([1.0, 2.0, 3.0], abs=1e-6)


def test_pin_holds_against_gravity(S):
    s = S(dt=1.0 / 60.0, iterations=10, gravity=(0.0, -9.81, 0.0))
    b = s.add_particle(position=(0.0, 5.0, 0.0), mass=1.0)
    s.add_pin(b, world_point=(0.0, 5.0, 0.0), stiffness=math.inf)
    for _ in range(120):
        s.step()
    assert s.positions()[b.index][1] == pytest.approx(5.0, abs=5e-3)
    assert s.lambdas()[1] == pytest.approx(-9.81, abs=0.5)


def test_distance_constraint_settles(S):
    rest = 0.5
    s = S(dt=1.0 / 60.0, iterations=15, gravity=(0.0, -9.81, 0.0))
    a = s.add_particle(position=(0.0, 5.0, 0.0), mass=1.0)
    b = s.add_particle(position=(0.0, 5.0 - rest, 0.0), mass=1.0)
    s.add_pin(a, world_point=(0.0, 5.0, 0.0), stiffness=math.inf)
    s.add_distance(a, b, rest=rest, stiffness=math.inf)
    for _ in range(600):
        s.step()
    p = s.positions()
    assert float(np.linalg.norm(p[a.index] - p[b.index])) == pytest.approx(rest, abs=2e-3)


def test_chain_distance_errors_small(S):
    n, link = 5, 0.5
    s = S(dt=1.0 / 60.0, iterations=15, gravity=(0.0, -9.81, 0.0))
    PS.chain(s, n, link)
    errs = []
    for f in range(900):
        s.step()
        if f >= 450:
            pos = s.positions()
            errs.append(max(abs(np.linalg.norm(pos[i + 1] - pos[i]) - link) for i in range(n)))
    assert float(np.mean(errs)) < link * 0.05


def test_warmstart_reduces_iteration_load(S):
    s = S(dt=1.0 / 60.0, iterations=10, gravity=(0.0, -9.81, 0.0))
    b = s.add_particle((0.0, 5.0, 0.0), 1.0)
    s.add_pin(b, (0.0, 5.0, 0.0), stiffness=math.inf)
    for _ in range(200):
        s.step()
    assert float(np.linalg.norm(s.lambdas())) > 5.0


def test_fracture_breaks_top_link_first(S):
    s = S(dt=1.0 / 60.0, iterations=10, gravity=(0.0, -9.81, 0.0))
    _, handles = PS.chain(s, 5, 0.5, fracture=30.0)
    broken_first = None
    for _ in range(300):
        s.step()
        act = s.active()
        for li, h in enumerate(handles):
            if act[h.index] == 0:
                broken_first = li
                break
        if broken_first is not None:
            break
    assert broken_first == 0
    assert int(s.active()[:3].sum()) == 3


def test_no_fracture_when_threshold_is_inf(S):
    s = S(dt=1.0 / 60.0, iterations=10, gravity=(0.0, -9.81, 0.0))
    a = s.add_particle((0.0, 5.0, 0.0), 1.0)
    b = s.add_particle((0.0, 4.0, 0.0), 1.0)
    s.add_pin(a, (0.0, 5.0, 0.0), stiffness=math.inf)
    s.add_distance(a, b, rest=1.0, stiffness=math.inf)
    for _ in range(200):
        s.step()
    assert int(s.active().sum()) == 4


def test_coloring_is_valid_for_chain(S):
    n = 8
    s = S(dt=1.0 / 60.0, iterations=2, gravity=(0.0, 0.0, 0.0))
    bodies = [s.add_particle((float(i), 0.0, 0.0), 1.0) for i in range(n)]
    s.add_pin(bodies[0], (0.0, 0.0, 0.0), stiffness=math.inf)
    for i in range(n - 1):
        s.add_distance(bodies[i], bodies[i + 1], rest=1.0, stiffness=math.inf)
    s.step()
    colors = s.body_color.numpy()
    for ca, cb in zip(s.c_body_a.numpy(), s.c_body_b.numpy()):
        if cb >= 0:
            assert colors[ca] != colors[cb]
    assert s.num_colors == 2


def test_coloring_complete_graph(S):
    n = 5
    s = S(dt=1.0 / 60.0, iterations=2, gravity=(0.0, 0.0, 0.0))
    bodies = [s.add_particle((float(i), 0.0, 0.0), 1.0) for i in range(n)]
    s.add_pin(bodies[0], (0.0, 0.0, 0.0), stiffness=math.inf)
    for i in range(n):
        for j in range(i + 1, n):
            s.add_distance(bodies[i], bodies[j], rest=1.0, stiffness=math.inf)
    s.step()
    assert s.num_colors == n


def test_no_nans_in_chain_simulation(S):
    s = S(dt=1.0 / 60.0, iterations=10, gravity=(0.0, -9.81, 0.0))
    PS.chain(s, 6, 0.5)
    for _ in range(300):
        s.step()
        assert np.isfinite(s.positions()).all()
        assert np.isfinite(s.velocities()).all()
        assert np.isfinite(s.lambdas()).all()


def test_friction_sliding_ball_decelerates_to_rest(S):
    mu, v0, g = 0.5, 3.0, 9.81
    expected_stop = v0 * v0 / (2.0 * mu * g)
    s = S(dt=1.0 / 60.0, iterations=20, gravity=(0.0, -g, 0.0), post_stabilize=True)
    r = 0.2
    b = s.add_particle((0.0, r, 0.0), 1.0, velocity=(v0, 0.0, 0.0),
                       shape=Shape("sphere", (r,)), friction=mu)
    s.add_floor_contact(b, floor_y=r, friction=mu)
    s.enable_self_collision(True, default_friction=mu)
    for _ in range(600):
        s.step()
    p, v = s.positions()[b.index], s.velocities()[b.index]
    assert abs(float(v[0])) < 1e-3
    assert p[0] == pytest.approx(expected_stop, rel=0.1)


def test_friction_zero_means_frictionless(S):
    v0 = 3.0
    s = S(dt=1.0 / 60.0, iterations=20, gravity=(0.0, -9.81, 0.0), post_stabilize=True)
    r = 0.2
    b = s.add_particle((0.0, r, 0.0), 1.0, velocity=(v0, 0.0, 0.0),
                       shape=Shape("sphere", (r,)), friction=0.0)
    s.add_floor_contact(b, floor_y=r, friction=0.0)
    s.enable_self_collision(True, default_friction=0.0)
    for _ in range(120):
        s.step()
    assert abs(float(s.velocities()[b.index][0]) - v0) < 0.05


def _tet_volumes(pos, tets):
    v0, v1, v2, v3 = (pos[tets[:, k]].astype(np.float64) for k in range(4))
    return np.einsum("ij,ij->i", np.cross(v1 - v0, v2 - v0), v3 - v0) / 6.0


def test_volume_preservation_prevents_tet_inversion_on_floor(S):
    s = S(dt=1.0 / 60.0, iterations=20, gravity=(0, -9.81, 0))
    blk = PS.soft_block(s, n=(3, 3, 3), size=(0.4, 0.4, 0.4), center=(0.0, 0.7, 0.0),
                        edge_stiffness=5.0e4, volume_stiffness=1.0e4, friction=0.5,
                        floor_y=0.0)
    for _ in range(180):
        s.step()
    pos = s.positions()[blk.indices]
    assert np.all(np.isfinite(pos))
    V = _tet_volumes(pos, blk.lattice.tets)
    assert int((np.sign(V) != np.sign(blk.lattice.volumes)).sum()) == 0
    ratio = V / blk.lattice.volumes
    assert ratio.min() > 0.7 and ratio.max() < 1.3
    assert pos[:, 1].min() > -0.02  # rests on the floor


def test_tet_pool_has_one_row_per_tet(S):
    s = S(dt=1.0 / 60.0, iterations=5)
    blk = PS.soft_block(s, n=(2, 2, 2), volume_stiffness=1.0e4, floor_y=None)
    assert len(blk.volume_constraints) == len(blk.lattice.tets) == len(s._tets) == 48
    s2 = S(dt=1.0 / 60.0, iterations=5)
    blk2 = PS.soft_block(s2, n=(2, 2, 2), volume_stiffness=0.0, floor_y=None)
    assert len(blk2.volume_constraints) == 0 and len(s2._tets) == 0
