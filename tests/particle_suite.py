"""The reference's particle Solver tests (tests/test_solver.py), written
against a solver class fixture ``S``:
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
    V = _tet_volumes(pos, blk.tet.tets)
    assert int((np.sign(V) != np.sign(blk.tet.volumes)).sum()) == 0
    ratio = V / blk.tet.volumes
    assert ratio.min() > 0.7 and ratio.max() < 1.3
    assert pos[:, 1].min() > -0.02  # rests on the floor


def test_tet_pool_has_one_row_per_tet(S):
    s = S(dt=1.0 / 60.0, iterations=5)
    blk = PS.soft_block(s, n=(2, 2, 2), volume_stiffness=1.0e4, floor_y=None)
    assert len(blk.volume_constraints) == len(blk.tet.tets) == len(s._tets) == 48
    s2 = S(dt=1.0 / 60.0, iterations=5)
    blk2 = PS.soft_block(s2, n=(2, 2, 2), volume_stiffness=0.0, floor_y=None)
    assert len(blk2.volume_constraints) == 0 and len(s2._tets) == 0


# ---- body-body contact (tests/test_solver.py:264-649) ----------------------
def test_sphere_contact_prevents_interpenetration(S):
    s = S(dt=1.0 / 60.0, iterations=20, gravity=(0.0, -9.81, 0.0), post_stabilize=True)
    r = 0.2
    a = s.add_particle((-0.05, 0.5, 0.0), mass=1.0, shape=Shape("sphere", (r,)))
    b = s.add_particle((+0.05, 1.4, 0.0), mass=1.0, shape=Shape("sphere", (r,)))
    s.add_floor_contact(a, floor_y=r)
    s.add_floor_contact(b, floor_y=r)
    s.enable_self_collision(True)
    for _ in range(400):
        s.step()
    pos = s.positions()
    assert float(np.linalg.norm(pos[0] - pos[1])) >= 2 * r * 0.95


def test_sphere_contact_stack_of_three(S):
    s = S(dt=1.0 / 60.0, iterations=25, gravity=(0.0, -9.81, 0.0), post_stabilize=True)
    r = 0.2
    a = s.add_particle((0.0, r, 0.0), mass=1.0, shape=Shape("sphere", (r,)))
    b = s.add_particle((0.0, 3 * r, 0.0), mass=1.0, shape=Shape("sphere", (r,)))
    c = s.add_particle((0.0, 5 * r, 0.0), mass=1.0, shape=Shape("sphere", (r,)))
    for body in (a, b, c):
        s.add_floor_contact(body, floor_y=r)
    s.enable_self_collision(True)
    for _ in range(600):
        s.step()
    pos = s.positions()
    assert pos[a.index, 1] < pos[b.index, 1] < pos[c.index, 1]
    for p in pos:
        assert p[1] >= r - 0.02


def test_sphere_contact_pair_skipped_when_connected(S):
    s = S(dt=1.0 / 60.0, iterations=10, gravity=(0.0, -9.81, 0.0))
    r = 0.2
    a = s.add_particle((0.0, 1.0, 0.0), 1.0, shape=Shape("sphere", (r,)))
    b = s.add_particle((0.0, 0.5, 0.0), 1.0, shape=Shape("sphere", (r,)))
    s.add_pin(a, (0.0, 1.0, 0.0), stiffness=math.inf)
    s.add_distance(a, b, rest=0.5, stiffness=math.inf)
    s.enable_self_collision(True)
    s.step()
    assert int((s.c_type.numpy() == 5).sum()) == 0


def test_broken_distance_pair_collides(S):
    s = S(dt=1.0 / 60.0, iterations=25, gravity=(0.0, -9.81, 0.0), post_stabilize=True)
    r = 0.1
    a = s.add_particle((0.0, 2.0, 0.0), 5.0, shape=Shape("sphere", (r,)))
    s.add_floor_contact(a, floor_y=r)
    s.add_pin(a, (0.0, 2.0, 0.0), stiffness=math.inf)
    b = s.add_particle((0.0, 1.5, 0.0), 1.0, shape=Shape("sphere", (r,)))
    s.add_floor_contact(b, floor_y=r)
    dh = s.add_distance(a, b, rest=0.5, stiffness=math.inf, fracture=5.0)
    s.enable_self_collision(True)
    for _ in range(300):
        s.step()
        if s.active()[dh.index] == 0:
            s.set_position(b, (0.0, 3.0, 0.0))
            s.set_velocity(b, (0.0, 0.0, 0.0))
            break
    else:
        pytest.fail("distance constraint never broke")
    for _ in range(400):
        s.step()
    pa, pb = s.positions()[a.index], s.positions()[b.index]
    assert float(np.linalg.norm(pa - pb)) >= 2 * r * 0.95


def test_pillar_collision_uses_stacked_subspheres(S):
    pillar_r, pillar_h = 0.10, 0.35
    s = S(dt=1.0 / 60.0, iterations=25, gravity=(0.0, -9.81, 0.0), post_stabilize=True)
    p = s.add_particle((0.0, pillar_h, 0.0), 10.0, shape=Shape("pillar", (pillar_r, pillar_h)))
    s.add_floor_contact(p, floor_y=pillar_h)
    s.add_pin(p, (0.0, pillar_h, 0.0), stiffness=math.inf)
    assert len(s._bodies_collision_spheres[p.index]) >= 2
    sphere_r = 0.08
    b = s.add_particle((1.0, 0.10, 0.0), 1.0, shape=Shape("sphere", (sphere_r,)),
                       velocity=(-3.0, 0.0, 0.0))
    s.add_floor_contact(b, floor_y=sphere_r)
    s.enable_self_collision(True)
    for _ in range(360):
        s.step()
    pp, bp = s.positions()[p.index], s.positions()[b.index]
    radial = float(np.sqrt((pp[0] - bp[0]) ** 2 + (pp[2] - bp[2]) ** 2))
    assert radial >= (pillar_r + sphere_r) * 0.95


def test_cube_corner_contact_no_penetration(S):
    r_c, r_b = 0.18, 0.10
    s = S(dt=1.0 / 60.0, iterations=25, gravity=(0.0, -9.81, 0.0), post_stabilize=True)
    c = s.add_particle((0.0, r_c, 0.0), 100.0, shape=Shape("cube", (r_c, r_c, r_c)))
    s.add_floor_contact(c, floor_y=r_c)
    s.add_pin(c, (0.0, r_c, 0.0), stiffness=math.inf)
    b = s.add_particle((0.0, 1.0, 0.0), 1.0, shape=Shape("sphere", (r_b,)))
    s.add_floor_contact(b, floor_y=r_b)
    s.enable_self_collision(True)
    for _ in range(600):
        s.step()
    assert abs(float(s.positions()[b.index, 1]) - (2 * r_c + r_b)) < 1e-3
    assert len(s._bodies_collision_spheres[c.index]) == 9


def _corner_start(r_c, r_b):
    d = np.array([1.0, 1.0, 1.0]) / math.sqrt(3.0)
    return tuple(float(v) for v in d * (r_c + r_b) * 0.5)


def _penetration(s, b, c, r_c, r_b):
    rel = s.positions()[b.index] - s.positions()[c.index]
    return max(0.0, r_b - float(np.linalg.norm(rel - np.clip(rel, -r_c, r_c))))


def test_cube_corner_subspheres_resolve_corner_overlap(S):
    r_c, r_b = 0.18, 0.10
    s = S(dt=1.0 / 60.0, iterations=25, gravity=(0.0, 0.0, 0.0), post_stabilize=True)
    c = s.add_particle((0.0, 0.0, 0.0), 100.0, shape=Shape("cube", (r_c, r_c, r_c)))
    s.add_pin(c, (0.0, 0.0, 0.0), stiffness=math.inf)
    b = s.add_particle(_corner_start(r_c, r_b), 1.0, shape=Shape("sphere", (r_b,)))
    s.enable_self_collision(True)
    for _ in range(600):
        s.step()
    assert _penetration(s, b, c, r_c, r_b) < 0.005


def test_sphere_box_contact_analytical_face(S):
    r_c, r_b = 0.18, 0.10
    s = S(dt=1.0 / 60.0, iterations=25, gravity=(0.0, -9.81, 0.0), post_stabilize=True)
    c = s.add_particle((0.0, r_c, 0.0), 100.0, shape=Shape("cube", (r_c, r_c, r_c)))
    s.add_floor_contact(c, floor_y=r_c)
    s.add_pin(c, (0.0, r_c, 0.0), stiffness=math.inf)
    b = s.add_particle((0.0, 1.0, 0.0), 1.0, shape=Shape("sphere", (r_b,)))
    s.add_floor_contact(b, floor_y=r_b)
    s.enable_self_collision(True)
    for _ in range(800):
        s.step()
    assert abs(float(s.positions()[b.index, 1]) - (2 * r_c + r_b)) < 5e-4
    assert 7 in set(int(t) for t in s.c_type.numpy().tolist())


def test_sphere_box_no_penetration_when_pushed_into_corner(S):
    r_c, r_b = 0.18, 0.10
    s = S(dt=1.0 / 60.0, iterations=25, gravity=(0.0, 0.0, 0.0), post_stabilize=True)
    c = s.add_particle((0.0, 0.0, 0.0), 100.0, shape=Shape("cube", (r_c, r_c, r_c)))
    s.add_pin(c, (0.0, 0.0, 0.0), stiffness=math.inf)
    b = s.add_particle(_corner_start(r_c, r_b), 1.0, shape=Shape("sphere", (r_b,)))
    s.enable_self_collision(True)
    for _ in range(600):
        s.step()
    assert _penetration(s, b, c, r_c, r_b) < 1e-3


def test_box_box_two_cube_stack_settles(S):
    r = 0.18
    s = S(dt=1.0 / 60.0, iterations=25, gravity=(0.0, -9.81, 0.0), post_stabilize=True)
    bot = s.add_particle((0.0, r, 0.0), 5.0, shape=Shape("cube", (r, r, r)))
    s.add_floor_contact(bot, floor_y=r)
    top = s.add_particle((0.0, 0.7, 0.0), 1.0, shape=Shape("cube", (r, r, r)))
    s.add_floor_contact(top, floor_y=r)
    s.enable_self_collision(True)
    for _ in range(800):
        s.step()
    p = s.positions()
    assert abs(p[0, 1] - r) < 1e-3
    assert abs(p[1, 1] - 3 * r) < 1e-3
    types, active = s.c_type.numpy(), s.c_active.numpy()
    assert sum(1 for t, a in zip(types, active) if t == 8 and a == 1) == 4


def test_box_box_friction_decelerates_top_cube(S):
    r, mu, v0 = 0.18, 0.5, 2.0
    s = S(dt=1.0 / 60.0, iterations=25, gravity=(0.0, -9.81, 0.0), post_stabilize=True)
    bot = s.add_particle((0.0, r, 0.0), 100.0, shape=Shape("cube", (r, r, r)), friction=mu)
    s.add_floor_contact(bot, floor_y=r, friction=mu)
    s.add_pin(bot, (0.0, r, 0.0), stiffness=math.inf)
    top = s.add_particle((0.0, 3 * r, 0.0), 1.0, shape=Shape("cube", (r, r, r)),
                         velocity=(v0, 0.0, 0.0), friction=mu)
    s.add_floor_contact(top, floor_y=r, friction=mu)
    s.enable_self_collision(True, default_friction=mu)
    for _ in range(600):
        s.step()
    assert abs(float(s.velocities()[top.index][0])) < 1e-3
    assert float(s.positions()[top.index, 0]) == pytest.approx(v0 * v0 / (2.0 * mu * 9.81),
                                                               rel=0.15)
