"""Host logic of the particle solver and the checker's row arithmetic, on the
CPU: constraint gradients against finite differences, the coloring and the
body -> row lists against direct restatements of the reference's loops,
flush-time state preservation, and the error paths."""
import math

import numpy as np
import pytest

from avbd_amd import particle_scenes as PS
from avbd_amd import solver as PSol
from avbd_amd.scene import Shape
from oracle.oracle import particle_solver_class, probe_particle_row, probe_tet

RNG = np.random.default_rng(7)


def _num_grad(f, x, h=1e-4):
    g = np.zeros(3)
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        g[k] = (f(x + e) - f(x - e)) / (2 * h)
    return g


# C(x) of each row type in float64, written from the constraint definitions
# (kernels.py:28-112): the checker's J must be its gradient in x_i.
def _C64(t, xa, xb, anchor, rest):
    if t in (0, 1, 2):
        return xa[t] - anchor[t]
    if t == PSol.DISTANCE:
        return np.linalg.norm(xa - xb) - rest
    if t == PSol.FLOOR_CONTACT:
        return xa[1] - anchor[1]
    if t == PSol.PLANE_CONTACT:
        return anchor @ xa - rest
    if t == PSol.CONTACT_TANGENT:
        return anchor @ (xa - xb) if xb is not None else anchor @ xa
    raise ValueError(t)


@pytest.mark.parametrize("t", [0, 1, 2, 3, 4, 6, 10])
@pytest.mark.parametrize("i_is_a", [True, False])
def test_row_gradient_matches_finite_difference(t, i_is_a):
    for _ in range(20):
        xi = RNG.normal(size=3)
        xo = RNG.normal(size=3)
        anchor = RNG.normal(size=3)
        if t in (PSol.PLANE_CONTACT, PSol.CONTACT_TANGENT):
            anchor /= np.linalg.norm(anchor)
        rest = float(abs(RNG.normal()))
        two = t in (PSol.DISTANCE,) or (t == PSol.CONTACT_TANGENT and RNG.random() < 0.5)
        if not two and not i_is_a:
            continue
        xi32, xo32, an32 = (v.astype(np.float32) for v in (xi, xo, anchor))
        C, J = probe_particle_row(t, i_is_a, xi32, xo32, two, an32, np.float32(rest))
        xi64, xo64, an64 = (v.astype(np.float64) for v in (xi32, xo32, an32))
        if i_is_a:
            f = lambda x: _C64(t, x, xo64 if two else None, an64, float(np.float32(rest)))
        else:
            f = lambda x: _C64(t, xo64, x, an64, float(np.float32(rest)))
        assert C == pytest.approx(f(xi64), abs=1e-5)
        np.testing.assert_allclose(J, _num_grad(f, xi64), atol=2e-4)


def _box_gap64(ps, pb, h, r):
    rel = ps - pb
    return np.linalg.norm(rel - np.clip(rel, -h, h)) - r


@pytest.mark.parametrize("t", [5, 7, 8])
@pytest.mark.parametrize("i_is_a", [True, False])
def test_contact_row_gradient_matches_finite_difference(t, i_is_a):
    """Sphere, sphere-box (sphere outside the box) and box-box rows with
    body-local offsets: C and J against float64 (kernels.py:269-317, 423-475)."""
    for _ in range(20):
        xa, xb = RNG.normal(size=3), RNG.normal(size=3)
        oa, ob = 0.2 * RNG.normal(size=3), 0.2 * RNG.normal(size=3)
        anchor = RNG.normal(size=3)
        rest = float(abs(RNG.normal()))
        if t == 7:
            ob[:] = 0.0
            anchor = np.abs(anchor) * 0.3          # half-extents
            xa = xb + np.sign(RNG.normal(size=3)) * (anchor + 0.3 + np.abs(RNG.normal(size=3)))
            xa = xa - oa
        if t == 8:
            anchor /= np.linalg.norm(anchor)
        c32 = [v.astype(np.float32) for v in (xa, xb, oa, ob, anchor)]
        xa, xb, oa, ob, anchor = (v.astype(np.float64) for v in c32)
        r64 = float(np.float32(rest))

        def C(pa, pb):
            if t == 5:
                return np.linalg.norm((pa + oa) - (pb + ob)) - r64
            if t == 7:
                return _box_gap64(pa + oa, pb, anchor, r64)
            return anchor @ ((pa + oa) - (pb + ob))
        xi, xo = (xa, xb) if i_is_a else (xb, xa)
        Cg, J = probe_particle_row(t, i_is_a, xi.astype(np.float32), xo.astype(np.float32), True,
                                   anchor.astype(np.float32), np.float32(rest),
                                   oa.astype(np.float32), ob.astype(np.float32))
        f = (lambda x: C(x, xb)) if i_is_a else (lambda x: C(xa, x))
        assert Cg == pytest.approx(C(xa, xb), abs=1e-5)
        np.testing.assert_allclose(J, _num_grad(f, xi), atol=2e-4)


def test_sphere_box_inside_takes_the_shallowest_face():
    """Centre inside the box: normal along the face of least depth, C =
    -(depth + r) (kernels.py:293-317)."""
    h = np.array([0.3, 0.2, 0.4], np.float32)
    ps = np.array([0.05, 0.15, -0.1], np.float32)       # y face is 0.05 deep
    C, J = probe_particle_row(7, True, ps, np.zeros(3, np.float32), True, h, np.float32(0.1))
    assert C == pytest.approx(-(0.05 + 0.1), abs=1e-6)
    np.testing.assert_array_equal(J, [0.0, 1.0, 0.0])
    C, J = probe_particle_row(7, False, np.zeros(3, np.float32), -ps, True, h, np.float32(0.1))
    np.testing.assert_array_equal(J, [0.0, 1.0, 0.0])  # box side: -n with n = -y


@pytest.mark.parametrize("slot", [0, 1, 2, 3])
def test_tet_gradient_and_force_match_finite_difference(slot):
    """rhs = k C J and lhs = k J J^T with C = V/V0 - 1 - ac0 and J its
    gradient in the slot's vertex (kernels.py:552-599)."""
    for _ in range(10):
        P = RNG.normal(size=(4, 3)).astype(np.float32)
        V0 = float(np.float32(0.7 * RNG.choice([-1, 1]) * (0.5 + RNG.random())))
        ac0, pen = 0.01, 37.0
        lhs, rhs = probe_tet(P, slot, V0, ac0, pen)

        def C(x):
            Q = P.astype(np.float64).copy()
            Q[slot] = x
            V = np.dot(Q[1] - Q[0], np.cross(Q[2] - Q[0], Q[3] - Q[0])) / 6.0
            return V / V0 - 1.0 - ac0
        J = _num_grad(C, P[slot].astype(np.float64))
        np.testing.assert_allclose(rhs, pen * C(P[slot].astype(np.float64)) * J, rtol=1e-3,
                                   atol=1e-3)
        np.testing.assert_allclose(lhs, pen * np.outer(J, J), rtol=1e-3, atol=1e-3)


def _loop_adjacency(rows, n_bodies):
    """solver.py:922-939 as written there: counts, prefix sum, cursor fill."""
    counts = np.zeros(n_bodies, np.int32)
    for c in rows:
        counts[c.body_a] += 1
        if c.body_b >= 0:
            counts[c.body_b] += 1
    starts = np.zeros(n_bodies + 1, np.int32)
    starts[1:] = np.cumsum(counts)
    idx = np.zeros(starts[-1], np.int32)
    cur = starts[:-1].copy()
    for ci, c in enumerate(rows):
        idx[cur[c.body_a]] = ci
        cur[c.body_a] += 1
        if c.body_b >= 0:
            idx[cur[c.body_b]] = ci
            cur[c.body_b] += 1
    return starts, idx


def test_flush_arrays_match_reference_loops():
    S = particle_solver_class()
    s = S(dt=1 / 60, iterations=4)
    blk = PS.soft_block(s, n=(2, 3, 2), floor_y=0.0)
    PS.chain(s, 4, 0.3, top=(2.0, 3.0, 0.0))
    a = s._scene_arrays(None)
    st, idx = _loop_adjacency(s._constraints, len(s._bodies_x))
    np.testing.assert_array_equal(a["con_start"], st)
    np.testing.assert_array_equal(a["con_idx"], idx)
    # tet lists: (tid << 2) | slot in tet order
    nb = len(s._bodies_x)
    lists = [[] for _ in range(nb)]
    for tid, t in enumerate(s._tets):
        for slot, b in enumerate(t[:4]):
            lists[b].append((tid << 2) | slot)
    np.testing.assert_array_equal(a["tet_idx"], np.concatenate([np.array(l, np.int32)
                                                                for l in lists]))
    # coloring: proper on the row graph and on every tet's six edges
    col = a["color"]
    for c in s._constraints:
        if c.body_b >= 0:
            assert col[c.body_a] != col[c.body_b]
    for t in blk.tet.tets:
        assert len({int(col[v]) for v in t}) == 4
    assert a["n_colors"] == col.max() + 1


def test_welsh_powell_order_and_bounds():
    # star: the hub (highest degree) is colored first and gets color 0
    adj = PSol.body_edges(6, [5, 5, 5, 5, 5], [0, 1, 2, 3, 4])
    col = PSol.welsh_powell(adj)
    assert col[5] == 0 and set(col[:5].tolist()) == {1}
    # brute force on small random graphs: proper and within max degree + 1
    for _ in range(30):
        n = 9
        ea, eb = RNG.integers(0, n, 14), RNG.integers(0, n, 14)
        adj = PSol.body_edges(n, ea.tolist(), eb.tolist())
        col = PSol.welsh_powell(adj)
        for i in range(n):
            assert all(col[i] != col[j] for j in adj[i])
        assert col.max() <= max(len(a) for a in adj)


def test_reflush_keeps_state_and_fresh_rows_start_clean():
    S = particle_solver_class()
    s = S(dt=1 / 60, iterations=8)
    bodies, links = PS.chain(s, 3, 0.5)
    for _ in range(20):
        s.step()
    before = s.state()
    extra = s.add_particle((3.0, 1.0, 0.0), 2.0)
    s.add_pin(extra, (3.0, 1.0, 0.0))
    s._flush()
    after = s.state()
    np.testing.assert_array_equal(after["x"][:4], before["x"])
    np.testing.assert_array_equal(after["v"][:4], before["v"])
    np.testing.assert_array_equal(after["lambda"][:6], before["lambda"])
    np.testing.assert_array_equal(after["penalty"][:6], before["penalty"])
    assert after["lambda"][6:].tolist() == [0.0] * 3
    assert after["penalty"][6:].tolist() == [1.0] * 3
    np.testing.assert_array_equal(after["x"][4], [3.0, 1.0, 0.0])


def test_runtime_perturbations():
    S = particle_solver_class()
    s = S(dt=1 / 60, iterations=4, gravity=(0, 0, 0))
    b = s.add_particle((0.0, 1.0, 0.0), 1.0)
    s.set_velocity(b, (1.0, 0.0, 0.0))
    s.add_impulse(b, (0.0, 2.0, 0.0))
    np.testing.assert_array_equal(s.velocities()[0], [1.0, 2.0, 0.0])
    s.set_position(b, (5.0, 5.0, 5.0))
    np.testing.assert_array_equal(s.positions()[0], [5.0, 5.0, 5.0])
    s.step()
    np.testing.assert_allclose(s.positions()[0], [5.0 + 1 / 60, 5.0 + 2 / 60, 5.0], rtol=1e-6)


def test_plane_contact_bulk_moves_plates():
    """A plate pushed along +x by set_plane_contact_bulk pushes the
    particles with it (solver.py:868-897)."""
    S = particle_solver_class()
    s = S(dt=1 / 60, iterations=10, gravity=(0, 0, 0))
    bs = [s.add_particle((0.5, float(k), 0.0), 1.0) for k in range(3)]
    hs = [s.add_plane_contact(b, (2.0, 0.0, 0.0), 0.0) for b in bs]
    s.step()
    for k in range(30):
        s.set_plane_contact_bulk(hs, offset=0.5 + 0.01 * (k + 1))
        s.step()
    x = s.positions()[:, 0]
    assert np.all(x >= 0.5 + 0.3 - 2e-3)
    assert s._c_world_anchor_host[0].tolist() == [1.0, 0.0, 0.0]


def test_contact_rows_of_each_shape_pair():
    """The broad phase's rows for sphere-sphere, sphere-box and box-box
    (solver.py:400-783): types, bodies, offsets and friction rows."""
    S = particle_solver_class()
    s = S(gravity=(0, 0, 0))
    a = s.add_particle((0.0, 0.0, 0.0), 1.0, shape=Shape("sphere", (0.2,)), friction=0.5)
    b = s.add_particle((0.35, 0.0, 0.0), 1.0, shape=Shape("sphere", (0.2,)), friction=0.5)
    s.enable_self_collision(True)
    s.step()
    types = [r.type for r in s._constraints]
    assert types == [5, 6, 6]
    r0 = s._constraints[0]
    assert (r0.body_a, r0.body_b) == (a.index, b.index) and r0.rest == pytest.approx(0.4)
    assert s._constraints[1].sibling == 0 and s._constraints[1].friction == pytest.approx(0.5)
    # tangents are unit and orthogonal to the contact normal (-x)
    t = np.array(s._constraints[1].world_anchor)
    bt = np.array(s._constraints[2].world_anchor)
    assert abs(np.linalg.norm(t) - 1) < 1e-6 and abs(t @ [1, 0, 0]) < 1e-6
    assert abs(t @ bt) < 1e-6

    s = S(gravity=(0, 0, 0))
    c = s.add_particle((0.0, 0.0, 0.0), 1.0, shape=Shape("cube", (0.2, 0.2, 0.2)))
    d = s.add_particle((0.0, 0.35, 0.0), 1.0, shape=Shape("cube", (0.2, 0.2, 0.2)))
    e = s.add_particle((0.0, -0.25, 0.0), 1.0, shape=Shape("sphere", (0.1,)))
    s.enable_self_collision(True)
    s.step()
    bb = [r for r in s._constraints if r.type == 8]
    assert len(bb) == 4 and all((r.body_a, r.body_b) == (c.index, d.index) for r in bb)
    assert all(r.world_anchor == (0.0, -1.0, 0.0) for r in bb)  # from d to c
    sb = [r for r in s._constraints if r.type == 7]
    assert len(sb) == 1 and (sb[0].body_a, sb[0].body_b) == (e.index, c.index)
    assert sb[0].world_anchor == pytest.approx((0.2, 0.2, 0.2)) and sb[0].rest == pytest.approx(0.1)


def test_contact_rows_persist_and_sphere_rows_are_rebuilt():
    """Sphere rows are rebuilt every step; box rows stay and exclude their
    pair from later broad phases (the reference's behaviour)."""
    S = particle_solver_class()
    s = S(gravity=(0, 0, 0))
    s.add_particle((0.0, 0.0, 0.0), 1.0, shape=Shape("cube", (0.2, 0.2, 0.2)))
    s.add_particle((0.0, 0.39, 0.0), 1.0, shape=Shape("cube", (0.2, 0.2, 0.2)))
    s.add_particle((2.0, 0.0, 0.0), 1.0, shape=Shape("sphere", (0.2,)))
    s.add_particle((2.0, 0.39, 0.0), 1.0, shape=Shape("sphere", (0.2,)))
    s.enable_self_collision(True)
    s.step()
    types = [r.type for r in s._constraints]
    assert types.count(8) == 4 and types.count(5) == 1
    for _ in range(5):  # box rows remain; the sphere row is rebuilt, never accumulated
        s.step()
        types = [r.type for r in s._constraints]
        assert types.count(8) == 4 and types.count(5) <= 1


def test_default_friction_backfill_and_floor_rows():
    S = particle_solver_class()
    s = S()
    b = s.add_particle((0, 1, 0), 1.0, shape=Shape("sphere", (0.1,)))
    s.enable_self_collision(True, default_friction=0.4)
    h = s.add_floor_contact(b)
    assert h.rows == 3
    rows = s._constraints
    assert [r.type for r in rows] == [PSol.FLOOR_CONTACT, PSol.CONTACT_TANGENT,
                                      PSol.CONTACT_TANGENT]
    assert rows[1].friction == pytest.approx(0.4) and rows[1].sibling == 0
    assert (rows[0].fmin, rows[0].fmax) == (-math.inf, 0.0)


def test_device_strings():
    # the reference's default "cpu" runs on HIP device 0; GPU strings name one
    from avbd_amd.solver_6dof import _parse_device
    assert PSol.Solver().device == "cpu"
    assert _parse_device("cpu") == 0 and _parse_device("cuda") == 0
    assert _parse_device("cuda:3") == 3 and _parse_device("hip:1") == 1
    with pytest.raises(ValueError):
        PSol.Solver(device="tpu")


def test_box_lattice_is_consistent():
    lat = PS.box_lattice((2, 3, 2), (0.2, 0.3, 0.2), (0, 0, 0))
    assert len(lat.tets) == 6 * 12
    assert np.all(np.abs(lat.volumes) > 0)
    assert abs(np.abs(lat.volumes).sum() - 0.2 * 0.3 * 0.2) < 1e-9
    for t in lat.tets:
        assert len(set(t.tolist())) == 4
    assert len({tuple(e) for e in lat.edges.tolist()}) == len(lat.edges)


def test_box_tet_mesh_surface_and_binding():
    """Surface triangles are the faces owned by one tet (2 per boundary cell
    face); binding gives one particle per vertex, one row per edge, floor
    rows with friction on the surface, one tet row per tet, and vertex
    masses summing to density x volume (deformable.py:552-659)."""
    from avbd_amd.deformable import bind_tet_mesh_to_solver, box_tet_mesh
    m = box_tet_mesh((2, 3, 2), (0.2, 0.3, 0.2), (0.0, 1.0, 0.0))
    assert len(m.surface_tris) == 2 * 2 * (2 * 3 + 3 * 2 + 2 * 2)
    assert set(m.surface_tris.reshape(-1).tolist()) <= set(m.surface_verts.tolist())
    S = particle_solver_class()
    s = S()
    d = bind_tet_mesh_to_solver(s, m, density=500.0, friction=0.4)
    assert len(s._bodies_x) == len(m.vertices)
    assert len(d.edge_constraints) == len(m.edges) and len(d.volume_constraints) == len(m.tets)
    assert len(d.floor_constraints) == len(m.surface_verts)
    assert sum(s._bodies_mass) == pytest.approx(500.0 * 0.2 * 0.3 * 0.2, rel=1e-5)
    assert all(h.rows == 3 for h in d.floor_constraints)


def test_reference_device_array_attributes():
    """x / v / c_lambda / c_active ... are None before the first flush and
    read the device state after it, as the reference's Warp arrays."""
    S = particle_solver_class()
    s = S()
    assert s.x is None and s.c_lambda is None
    b = s.add_particle((0.0, 1.0, 0.0), 1.0)
    s.add_pin(b, (0.0, 1.0, 0.0))
    s.step()
    np.testing.assert_array_equal(s.x.numpy(), s.positions())
    np.testing.assert_array_equal(s.c_lambda.numpy(), s.lambdas())
    assert s.c_active.numpy().tolist() == [1, 1, 1]
    assert s.v.numpy().shape == (1, 3) and s.c_penalty.numpy().shape == (3,)
