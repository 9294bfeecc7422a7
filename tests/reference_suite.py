"""The reference's own Solver6DOF tests (tests/test_solver_6dof.py,
tests/test_robustness.py, tests/test_kernel_face_clip.py) written against a
solver class fixture ``S``: test_gpu_reference_suite.py binds it to the HIP
solver, test_oracle_reference_suite.py to the CPU checker. Assertions and
tolerances are the reference's."""
import math

import numpy as np
import pytest


# ---- tests/test_solver_6dof.py ------------------------------------------------
def test_free_fall_constant_acceleration(S):
    s = S(dt=1 / 60, iterations=5)
    s.add_box((0., 5., 0.), (0.5, 0.5, 0.5), mass=1.0)
    for _ in range(60):
        s.step()
    assert s.velocities()[0][1] == pytest.approx(-9.81, rel=1e-3)


def test_free_spin_preserves_angular_velocity(S):
    s = S(dt=1 / 60, iterations=5, gravity=(0., 0., 0.))
    s.add_box((0., 0., 0.), (0.5, 0.5, 0.5), mass=1.0, angular_velocity=(0., 0., 2.0))
    for _ in range(60):
        s.step()
    w = s.angular_velocities()[0]
    assert np.linalg.norm(w - np.array([0., 0., 2.0])) < 1e-3
    q = s.orientations()[0]
    expected = np.array([0., 0., math.sin(1.0), math.cos(1.0)], dtype=np.float32)
    assert min(np.linalg.norm(q - expected), np.linalg.norm(q + expected)) < 5e-3


def test_box_drops_on_floor_settles_at_half_extent(S):
    h = 0.3
    s = S(dt=1 / 60, iterations=20)
    b = s.add_box((0., 2.0, 0.), (h, h, h), mass=1.0, friction=0.5)
    s.add_floor_contact_box(b, floor_y=0.0, friction=0.5)
    for _ in range(300):
        s.step()
    assert abs(s.positions()[0][1] - h) < 0.01
    assert abs(s.velocities()[0][1]) < 0.1


def test_spinning_box_on_floor_friction_brakes_omega(S):
    s = S(dt=1 / 60, iterations=25)
    h = 0.3
    b = s.add_box((0., h + 0.15, 0.), (h, h, h), mass=1.0, friction=0.8,
                  angular_velocity=(0., 5.0, 0.))
    s.add_floor_contact_box(b, friction=0.8)
    for _ in range(300):
        s.step()
    assert np.linalg.norm(s.angular_velocities()[0]) < 0.1
    assert abs(s.positions()[0][1] - h) < 0.02


def test_tilted_box_drops_and_settles(S):
    s = S(dt=1 / 60, iterations=15, substeps=8)
    h = 0.3
    ang = math.radians(30)
    b = s.add_box((0., 1.5, 0.), (h, h, h), mass=1.0, friction=0.5,
                  orientation=(math.sin(ang / 2), 0., 0., math.cos(ang / 2)))
    s.add_floor_contact_box(b, friction=0.5)
    for _ in range(300):
        s.step()
    assert np.linalg.norm(s.angular_velocities()[0]) < 0.05
    assert np.linalg.norm(s.velocities()[0]) < 0.05
    assert h - 0.02 < s.positions()[0][1] < h + 0.02


def test_box_corner_pin_holds_against_gravity(S):
    s = S(dt=1 / 60, iterations=30)
    h = 0.3
    world_pin = (h, 2.0, h)
    b = s.add_box((0., 2.0 - h, 0.), (h, h, h), mass=1.0)
    s.add_pin_corner(b, body_local=(h, h, h), world_point=world_pin)
    for _ in range(180):
        s.step()
    pos = s.positions()[0]
    qx, qy, qz, qw = (float(v) for v in s.orientations()[0])
    R = np.array([
        [1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qw * qz), 2 * (qx * qz + qw * qy)],
        [2 * (qx * qy + qw * qz), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qw * qx)],
        [2 * (qx * qz - qw * qy), 2 * (qy * qz + qw * qx), 1 - 2 * (qx * qx + qy * qy)],
    ], dtype=np.float32)
    corner = pos + R @ np.array([h, h, h], dtype=np.float32)
    assert np.linalg.norm(corner - np.array(world_pin)) < 0.02


def test_two_cubes_stack_axis_aligned(S):
    h = 0.20
    s = S(dt=1 / 60, iterations=15, substeps=8)
    bot = s.add_box((0., 0.20, 0.), (h, h, h), mass=1.0, friction=0.5)
    top = s.add_box((0., 0.55, 0.), (h, h, h), mass=1.0, friction=0.5)
    for b in (bot, top):
        s.add_floor_contact_box(b, friction=0.5)
    s.enable_self_collision(True, default_friction=0.5)
    for _ in range(300):
        s.step()
    ys = s.positions()[:, 1]
    assert abs(ys[0] - h) < 0.01
    assert abs(ys[1] - 3 * h) < 0.02
    assert abs(ys[1] - ys[0] - 2 * h) < 0.01
    assert np.abs(s.velocities()[:, 1]).max() < 0.05


def test_three_cubes_stack(S):
    h = 0.20
    s = S(dt=1 / 60, iterations=15, substeps=8)
    bs = [s.add_box((0., y, 0.), (h, h, h), mass=1.0, friction=0.5) for y in (0.20, 0.62, 1.04)]
    for b in bs:
        s.add_floor_contact_box(b, friction=0.5)
    s.enable_self_collision(True, default_friction=0.5)
    for _ in range(360):
        s.step()
    ys = s.positions()[:, 1]
    for i, e in enumerate([h, 3 * h, 5 * h]):
        assert abs(ys[i] - e) < 0.03


def test_cube_on_floor_friction_brakes_translation(S):
    h, v0, mu = 0.20, 2.0, 0.4
    s = S(dt=1 / 60, iterations=15, substeps=8)
    b = s.add_box((0., h, 0.), (h, h, h), mass=1.0, friction=mu, velocity=(v0, 0., 0.))
    s.add_floor_contact_box(b, friction=mu)
    x0 = s.positions()[0][0]
    for _ in range(240):
        s.step()
    expected = v0 * v0 / (2 * mu * 9.81)
    assert abs((s.positions()[0][0] - x0) - expected) / expected < 0.10


def test_static_body_does_not_move(S):
    s = S(dt=1 / 60, iterations=5)
    s.add_box((0.5, 1.0, 0.5), (0.3, 0.3, 0.3), mass=0.0)
    x0, q0 = s.positions()[0].copy(), s.orientations()[0].copy()
    for _ in range(60):
        s.step()
    assert np.allclose(s.positions()[0], x0)
    assert np.allclose(s.orientations()[0], q0)


def test_pinned_box_with_spin_stays_bounded(S):
    s = S(dt=1 / 60, iterations=20, max_angular_speed=50.0)
    h = 0.2
    b = s.add_box((0., 2.0 - h, 0.), (h, h, h), mass=1.0, angular_velocity=(0., 6.0, 0.))
    s.add_pin_corner(b, body_local=(h, h, h), world_point=(h, 2.0, h))
    max_w = 0.0
    for _ in range(180):
        s.step()
        max_w = max(max_w, float(np.linalg.norm(s.angular_velocities()[0])))
    assert max_w < 12.0


def _push(S, force):
    s = S(dt=1 / 60, iterations=20, substeps=4, friction_static_mult=1.5)
    h = 0.2
    b = s.add_box((0., h, 0.), (h, h, h), mass=1.0, friction=0.4)
    s.add_floor_contact_box(b, friction=0.4)
    for _ in range(60):
        s.step()
    x0 = float(s.positions()[0][0])
    for _ in range(60):
        v = s.velocities()[0]
        s.set_velocity(b, (float(v[0] + force * s.dt), float(v[1]), float(v[2])))
        s.step()
    return float(s.positions()[0][0]) - x0


def test_static_friction_holds_under_small_push(S):
    assert abs(_push(S, 3.0)) < 0.05


def test_kinetic_friction_slips_under_large_push(S):
    assert _push(S, 10.0) > 1.0


def test_bvh_broadphase_emits_correct_pair_count(S):
    h = 0.15
    s = S(dt=1 / 60, iterations=10, substeps=4)
    s.enable_self_collision(True, default_friction=0.3)
    for k in range(3):
        b = s.add_box((0., h + 2 * h * k + 0.02 * k, 0.), (h, h, h), mass=1.0, friction=0.3)
        s.add_floor_contact_box(b, friction=0.3)
    s.step()
    assert s.broadphase_ms > 0.0
    for _ in range(120):
        s.step()
    ys = s.positions()[:, 1]
    for i in range(3):
        assert abs(ys[i] - (h + 2 * h * i)) < 0.03


def test_bvh_broadphase_scales_to_27_bodies(S):
    h, spacing = 0.12, 0.6
    s = S(dt=1 / 60, iterations=15, substeps=6)
    s.enable_self_collision(True, default_friction=0.5)
    for ix in range(3):
        for iz in range(3):
            for k in range(3):
                b = s.add_box(((ix - 1) * spacing, h + 2 * h * k, (iz - 1) * spacing), (h, h, h),
                              mass=1.0, friction=0.5)
                s.add_floor_contact_box(b, friction=0.5)
    for _ in range(120):
        s.step()
    p = s.positions()
    for tower in range(9):
        assert 4.5 * h < p[tower * 3 + 2][1] < 5.5 * h


def test_edge_edge_contact_two_rotated_boxes_separate(S):
    s = S(dt=1.0 / 240.0, iterations=20, substeps=2, gravity=(0., 0., 0.))
    s.enable_self_collision(True, default_friction=0.0)
    h = 0.25
    s.add_box((0., 0., 0.), (h, h, h), mass=1.0,
              orientation=(0., 0., math.sin(math.pi / 8), math.cos(math.pi / 8)), friction=0.0)
    s.add_box((0.0, 0.6, 0.0), (h, h, h), mass=1.0,
              orientation=(math.sin(math.pi / 8), 0., 0., math.cos(math.pi / 8)), friction=0.0,
              velocity=(0., -2.0, 0.))
    for _ in range(120):
        s.step()
    p = s.positions()
    assert abs(p[1][1] - p[0][1]) > 2 * h - 0.05
    assert np.all(np.isfinite(s.positions()))
    assert np.all(np.isfinite(s.velocities()))


def test_eq14_hessian_rescaling_settles_box_on_floor(S):
    h = 0.3
    s = S(dt=1.0 / 60.0, iterations=15, gravity=(0., -9.81, 0.))
    b = s.add_box((0., 1.5, 0.), (h, h, h), mass=1.0, friction=0.5)
    s.add_floor_contact_box(b, friction=0.5)
    for _ in range(180):
        s.step()
    y_settled = s.positions()[0][1]
    ys = []
    for _ in range(120):
        s.step()
        ys.append(s.positions()[0][1])
    assert max(ys) - min(ys) < 2.0e-3
    assert abs(y_settled - h) < 2.0e-3


def test_set_velocity_writes_buffer_in_place(S):
    """Counterpart of the reference's captured-graph buffer test
    (test_solver_6dof.py:467-527): after a mid-sim add_box + flush and a
    kick, the next step integrates the live velocity buffer."""
    h = 0.15
    s = S(dt=1 / 60, iterations=15, substeps=4, gravity=(0., -9.81, 0.),
                   post_stabilize=True)
    s.enable_self_collision(True, default_friction=0.5)
    for k in range(2):
        b = s.add_box((0., h + 2 * h * k + 0.02 * k, 0.), (h, h, h), mass=1.0, friction=0.5)
        s.add_floor_contact_box(b, friction=0.5)
    for _ in range(60):
        s.step()
    nb = s.add_box((0.3, 2.0, 0.3), (h, h, h), mass=1.0, friction=0.5)
    s.add_floor_contact_box(nb, friction=0.5)
    s._flush()
    for _ in range(60):
        s.step()
    target = (1.7, 2.3, -1.1)
    s.set_velocity(type("B", (), {"index": 0})(), target)
    s.step()
    assert not np.allclose(s.velocities()[0], target, atol=1e-4)
    assert np.all(np.isfinite(s.positions()))


# ---- tests/test_kernel_face_clip.py ----------------------------------------
def test_kick_all_after_midsim_add_box_settles_under_gravity(S):
    """tests/test_solver_6dof.py:531-580 with the reference's assertions."""
    h = 0.12
    s = S(dt=1 / 60, iterations=20, substeps=6, gravity=(0., -9.81, 0.),
                   post_stabilize=True)
    s.enable_self_collision(True, default_friction=0.5)
    for ix in range(2):
        for iz in range(2):
            for k in range(3):
                b = s.add_box((0.5 * ix, h + 2 * h * k, 0.5 * iz), (h, h, h), mass=1.0,
                              friction=0.5)
                s.add_floor_contact_box(b, friction=0.5)
    for _ in range(80):
        s.step()
    nb = s.add_box((0.25, 2.0, 0.25), (h, h, h), mass=1.0, friction=0.5)
    s.add_floor_contact_box(nb, friction=0.5)
    s._flush()
    for _ in range(80):
        s.step()
    rng = np.random.default_rng(0)
    for i in range(s.positions().shape[0]):
        body = type("B", (), {"index": i})()
        v = s.velocities()[i]
        w = s.angular_velocities()[i]
        s.set_velocity(body, tuple(float(v[j] + rng.uniform(-2.5, 2.5)) for j in range(3)))
        s.set_angular_velocity(body, tuple(float(w[j] + rng.uniform(-3., 3.)) for j in range(3)))
    for _ in range(120):
        s.step()
    pos, vel = s.positions(), s.velocities()
    assert np.all(np.isfinite(pos)) and np.all(np.isfinite(vel))
    assert float(pos[:, 1].max()) < 1.2
    assert float(np.linalg.norm(vel, axis=1).mean()) < 0.6


def test_kernel_face_clip_settles_3_cube_stack(S):
    h = 0.20
    s = S(dt=1.0 / 60.0, iterations=15, substeps=8)
    s.enable_self_collision(True, default_friction=0.5)
    bs = [s.add_box((0., y, 0.), (h, h, h), mass=1.0, friction=0.5) for y in (0.20, 0.62, 1.04)]
    for b in bs:
        s.add_floor_contact_box(b, friction=0.5)
    for _ in range(360):
        s.step()
    ys, vs = s.positions()[:, 1], s.velocities()[:, 1]
    for i, e in enumerate([0.20, 0.60, 1.00]):
        assert abs(float(ys[i]) - e) < 0.03
    assert float(np.abs(vs).max()) < 0.05


# ---- tests/test_robustness.py -----------------------------------------------
def test_dense_cluster_no_crash(S):
    s = S(gravity=(0, -9.8, 0), iterations=4)
    for i in range(60):
        s.add_box(position=(0.01 * i, 2.0, 0.0), half_extents=(0.5, 0.5, 0.5), mass=1.0,
                  friction=0.3)
    s.enable_self_collision(True)
    s._flush()
    with pytest.warns(RuntimeWarning, match="row pool overflow"):
        s.step()
    for _ in range(4):
        s.step()
    assert int(s.n_active_rows.numpy()[0]) <= s._gpu_pool_n_capacity


def test_coloring_safety_for_touching_bodies(S):
    s = S(gravity=(0, -9.8, 0), iterations=4)
    s.add_box(position=(0.0, 2.0, 0.0), half_extents=(0.5, 0.5, 0.5), mass=1.0, friction=0.3)
    s.add_box(position=(0.6, 2.0, 0.0), half_extents=(0.5, 0.5, 0.5), mass=1.0, friction=0.3)
    s.enable_self_collision(True)
    s.step()
    c = s.body_color.numpy()
    assert c[0] != c[1]


def test_coloring_safety_for_stack(S):
    s = S(gravity=(0, -9.8, 0), iterations=4)
    for i in range(5):
        s.add_box(position=(0.0, 0.5 + i * 1.0, 0.0), half_extents=(0.5, 0.5, 0.5), mass=1.0,
                  friction=0.3)
    s.enable_self_collision(True)
    s.step()
    c = s.body_color.numpy()
    for i in range(4):
        assert c[i] != c[i + 1]


def _overlap_stack(S, mode):
    s = S(gravity=(0, -9.8, 0), iterations=4, coloring_mode=mode)
    for i in range(6):
        s.add_box(position=(0.0, 0.5 + i * 0.85, 0.0), half_extents=(0.5, 0.5, 0.5), mass=1.0,
                  friction=0.3)
    s.enable_self_collision(True)
    return s


@pytest.mark.parametrize("mode", ["jones_plassmann", "jacobi"])
def test_coloring_mode_conflict_free(S, mode):
    s = _overlap_stack(S, mode)
    s.step()
    assert s.coloring_mode == mode
    assert s.count_color_conflicts() == 0
    assert 1 <= s.num_active_colors <= s._max_colors


def test_coloring_known_answer_overlapping_stack(S):
    """The docstring of test_robustness.py:106-108 gives JP = 3 colors and
    jacobi = 2 for this scene; the assignments follow from the priorities
    default_rng(0xA5BD) draws."""
    s = _overlap_stack(S, "jacobi")
    s.step()
    assert list(s.body_color.numpy()) == [0, 1, 0, 1, 0, 1]
    s = _overlap_stack(S, "jones_plassmann")
    s.step()
    assert list(s.body_color.numpy()) == [0, 1, 0, 1, 2, 0]


def test_coloring_mode_switch_takes_effect_at_runtime(S):
    s = _overlap_stack(S, "jones_plassmann")
    s.step()
    jp = s.num_active_colors
    assert s._color_dirty is False
    s.coloring_mode = "jacobi"
    assert s._color_dirty is True
    s.step()
    assert s.coloring_mode == "jacobi"
    assert s.count_color_conflicts() == 0
    assert s.num_active_colors <= jp
    s.step()
    assert s._color_dirty is False
    s.coloring_mode = "jacobi"
    assert s._color_dirty is False


def test_jacobi_coloring_packs_no_worse_than_jp(S):
    def colors(mode):
        s = _overlap_stack(S, mode)
        s.step()
        return s.num_active_colors
    assert colors("jacobi") <= colors("jones_plassmann")


def test_coloring_after_dynamic_contact(S):
    s = S(gravity=(0, -9.8, 0), iterations=4, dt=1 / 240)
    s.enable_self_collision(True)
    a = s.add_box(position=(0.0, 2.0, 0.0), half_extents=(0.2, 0.2, 0.2), mass=1.0, friction=0.3)
    b = s.add_box(position=(5.0, 2.0, 0.0), half_extents=(0.2, 0.2, 0.2), mass=1.0, friction=0.3)
    s.set_velocity(b, (-40.0, 0.0, 0.0))
    for _ in range(80):
        s.step()
        if int(s._bp_pair_count.numpy()[0]) > 0:
            break
    assert int(s._bp_pair_count.numpy()[0]) > 0
    c = s.body_color.numpy()
    assert c[a.index] != c[b.index]


def test_graph_invalidates_on_iterations(S):
    s = S(gravity=(0, -9.8, 0), iterations=4)
    s.add_box(position=(0.0, 2.0, 0.0), half_extents=(0.5, 0.5, 0.5), mass=1.0, friction=0.3)
    s.add_box(position=(0.6, 2.0, 0.0), half_extents=(0.5, 0.5, 0.5), mass=1.0, friction=0.3)
    s.enable_self_collision(True)
    s.step()
    before = s._graph_signature
    assert before is not None and before[0] == 4
    s.iterations = 9
    s.step()
    assert s._graph_signature[0] == 9 and s._graph_signature != before


def test_primal_group_size_invalidates_graph(S):
    s = S(gravity=(0, -9.8, 0), iterations=4)
    s.add_box(position=(0.0, 2.0, 0.0), half_extents=(0.5, 0.5, 0.5), mass=1.0)
    s.add_box(position=(0.6, 2.0, 0.0), half_extents=(0.5, 0.5, 0.5), mass=1.0)
    s.enable_self_collision(True)
    s.step()
    before = s._graph_signature
    s.primal_group_size = 8
    assert s.primal_group_size == 8
    s.step()
    assert s._graph_signature != before
