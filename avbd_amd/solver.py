"""The 3-DOF particle solver (reference ``avbd3d.Solver``,
src/avbd3d/solver.py) on the MI355X.

Same scene-building API as the reference: particles with pin (per axis),
distance, floor (with Coulomb friction rows), plane and tet-volume
constraints, fracture by |lambda|, Welsh-Powell coloring of the body graph
and the reference's step order (predict, warm start, alpha C0, iterations
of per-color primal + dual, optional post-stabilisation pass, BDF1
velocity). The host keeps the scene and builds the coloring and the
body -> row lists at flush time (solver.py:922-1122); every step runs on the
GPU (libavbd_hip.so, avbd_p_*), one primal launch per color over that
color's bodies only.

With ``enable_self_collision`` the host generates body-body contact rows
once per step before the flush, as the reference does (solver.py:603-783,
restated in particle_contacts.py), and carries each contact's lambda and
penalty to the next step by its pair key (solver.py:1160-1185, 1392-1421).
"""
from __future__ import annotations

import math
import time
import warnings
from dataclasses import dataclass

import numpy as np

from . import _native as N
from . import particle_contacts as PC
from .scene import Body, ConstraintHandle, Shape
from .solver_6dof import DeviceArray, _parse_device

# Constraint type codes (kernels.py:28-70)
PIN_X = 0
PIN_Y = 1
PIN_Z = 2
DISTANCE = 3
FLOOR_CONTACT = 4
SPHERE_CONTACT = 5
CONTACT_TANGENT = 6
SPHERE_BOX_CONTACT = 7
BOX_BOX_CONTACT = 8
TET_VOLUME = 9
PLANE_CONTACT = 10

PENALTY_INIT = 1.0       # fresh row penalty at flush (solver.py:1054)
PENALTY_INIT_TET = 10.0  # fresh tet penalty at flush (solver.py:1090)


@dataclass
class _Row:
    """One constraint row (solver.py:106-131); off_a / off_b are the
    body-local contact points of the body-body contact rows."""
    type: int
    body_a: int
    body_b: int
    world_anchor: tuple[float, float, float] = (0.0, 0.0, 0.0)
    rest: float = 0.0
    stiffness: float = math.inf
    fracture: float = math.inf
    fmin: float = -math.inf
    fmax: float = math.inf
    sibling: int = -1
    friction: float = 0.0
    off_a: tuple[float, float, float] = (0.0, 0.0, 0.0)
    off_b: tuple[float, float, float] = (0.0, 0.0, 0.0)


def collision_radius(shape: Shape | None) -> float:
    """Largest centre-to-surface reach of a body's collision spheres
    (solver.py:101-108)."""
    return PC.reach(PC.collision_spheres(shape))


def body_edges(n_bodies: int, body_a, body_b) -> list[set[int]]:
    """Bodies sharing a row (coloring.py:22-38): world-anchored rows and
    rows on one body add no edge."""
    adj: list[set[int]] = [set() for _ in range(n_bodies)]
    for a, b in zip(body_a, body_b):
        if a >= 0 and b >= 0 and a != b:
            adj[a].add(b)
            adj[b].add(a)
    return adj


def welsh_powell(adj: list[set[int]]) -> np.ndarray:
    """Greedy coloring in order of decreasing degree (ties by body index),
    each body taking the smallest color its colored neighbours do not use
    (coloring.py:41-55)."""
    n = len(adj)
    deg = np.fromiter((len(a) for a in adj), np.int64, n)
    color = np.full(n, -1, np.int32)
    for i in np.argsort(-deg, kind="stable").tolist():
        used = {int(color[j]) for j in adj[i]}
        c = 0
        while c in used:
            c += 1
        color[i] = c
    return color


def _csr(n: int, owners: np.ndarray, values: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Per-owner lists keeping the order of `values` (a stable bucket sort),
    as the cursor fill of solver.py:922-962 gives them."""
    starts = np.zeros(n + 1, np.int32)
    if owners.size:
        starts[1:] = np.cumsum(np.bincount(owners, minlength=n))
        order = np.argsort(owners, kind="stable")
        return starts, values[order].astype(np.int32)
    return starts, np.zeros(0, np.int32)


class Solver:
    """3-DOF AVBD particle solver (solver.py:133-1443) on a HIP device."""

    def __init__(self, dt: float = 1.0 / 60.0, iterations: int = 10,
                 gravity: tuple[float, float, float] = (0.0, -9.81, 0.0),
                 alpha: float = 0.99, beta: float = 1.0e5, gamma: float = 0.99,
                 post_stabilize: bool = True, device: str = "cpu",
                 max_speed: float = 30.0):
        self._device_index = _parse_device(device)
        self.device = device
        self.dt = float(dt)
        self.iterations = int(iterations)
        self.gravity = tuple(float(g) for g in gravity)
        self.alpha = float(alpha)
        self.beta = float(beta)
        self.gamma = float(gamma)
        self.post_stabilize = bool(post_stabilize)
        self.max_speed = float(max_speed)

        self._bodies_x: list[tuple[float, float, float]] = []
        self._bodies_v: list[tuple[float, float, float]] = []
        self._bodies_mass: list[float] = []
        self._bodies_collide: list[bool] = []
        self._bodies_friction: list[float] = []
        self._bodies_radius: list[float] = []
        self._bodies_kind: list[str] = []
        self._bodies_half_extents: list[tuple[float, float, float]] = []
        self._bodies_collision_spheres: list = []
        self._contact_state_cache: dict = {}
        self._pool_pair_rows: list = []
        self._constraints: list[_Row] = []
        self._tets: list[tuple[int, int, int, int, float, float, float]] = []
        self._dirty = True
        self._contact_pool_start: int | None = None
        self._enable_self_collision = False
        self._contact_margin = 0.0
        self._default_friction = 0.0
        self._c_rest_host: np.ndarray | None = None
        self._c_world_anchor_host: np.ndarray | None = None

        self._h = None
        self._flushed: dict | None = None
        self.num_colors = 0
        self.color_counts: dict[int, int] = {}
        self.body_color = self.c_type = self.c_body_a = self.c_body_b = None
        self.c_active = self.c_lambda = self.c_penalty = None
        self.primal_variant = 4  # the library's default primal kernel
        self.x = self.v = self.prev_v = self.tet_lambda = self.tet_active = None

    # ---- scene building ------------------------------------------------------
    def add_particle(self, position, mass: float, velocity=(0.0, 0.0, 0.0),
                     shape: Shape | None = None, collide: bool = True,
                     radius: float | None = None, friction: float | None = None) -> Body:
        """solver.py:228-261"""
        idx = len(self._bodies_x)
        self._bodies_x.append(tuple(float(c) for c in position))
        self._bodies_v.append(tuple(float(c) for c in velocity))
        self._bodies_mass.append(float(mass))
        sh = shape or Shape()
        self._bodies_collision_spheres.append(PC.collision_spheres(sh, radius))
        kind = sh.kind if radius is None else "sphere"
        self._bodies_kind.append(kind if kind in ("sphere", "cube", "pillar") else "")
        self._bodies_half_extents.append(
            (float(sh.size[0]), float(sh.size[1]), float(sh.size[2]))
            if sh.kind == "cube" and radius is None else (0.0, 0.0, 0.0))
        r_max = collision_radius(sh) if radius is None else float(radius)
        self._bodies_radius.append(r_max)
        self._bodies_collide.append(bool(collide) and r_max > 0.0)
        mu = float(friction) if friction is not None else self._default_friction
        self._bodies_friction.append(max(0.0, mu))
        self._dirty = True
        return Body(index=idx, shape=sh)

    def add_pin(self, body: Body, world_point, stiffness: float = math.inf,
                fracture: float = math.inf) -> ConstraintHandle:
        """All three axes (solver.py:263-274)."""
        return self.add_pin_axes(body, world_point, (True, True, True), stiffness=stiffness,
                                 fracture=fracture)

    def add_pin_axes(self, body: Body, world_point, axes, stiffness: float = math.inf,
                     fracture: float = math.inf) -> ConstraintHandle:
        """One PIN row per selected world axis (solver.py:276-311)."""
        start = len(self._constraints)
        rows = 0
        anchor = tuple(float(p) for p in world_point)
        for axis, want in zip((PIN_X, PIN_Y, PIN_Z), axes):
            if want:
                self._constraints.append(_Row(type=axis, body_a=body.index, body_b=-1,
                                              world_anchor=anchor, stiffness=stiffness,
                                              fracture=fracture))
                rows += 1
        self._dirty = True
        return ConstraintHandle(index=start, rows=rows)

    def add_distance(self, body_a: Body, body_b: Body, rest: float,
                     stiffness: float = math.inf, fracture: float = math.inf) -> ConstraintHandle:
        """solver.py:313-333"""
        idx = len(self._constraints)
        self._constraints.append(_Row(type=DISTANCE, body_a=body_a.index, body_b=body_b.index,
                                      rest=float(rest), stiffness=stiffness, fracture=fracture))
        self._dirty = True
        return ConstraintHandle(index=idx, rows=1)

    def add_tet_volume(self, v0: Body, v1: Body, v2: Body, v3: Body, stiffness: float = 1.0e8,
                       fracture: float = math.inf) -> ConstraintHandle:
        """Signed-volume constraint on four bodies; the rest volume comes from
        the bodies' scene positions in float64 (solver.py:335-365). The
        handle's index is the tet's index in the tet pool."""
        tid = len(self._tets)
        p = [np.asarray(self._bodies_x[b.index], np.float64) for b in (v0, v1, v2, v3)]
        rest_vol = float(np.dot(p[1] - p[0], np.cross(p[2] - p[0], p[3] - p[0])) / 6.0)
        self._tets.append((int(v0.index), int(v1.index), int(v2.index), int(v3.index), rest_vol,
                           float(stiffness), float(fracture)))
        self._dirty = True
        return ConstraintHandle(index=tid, rows=1)

    def enable_self_collision(self, enabled: bool = True, margin: float = 0.0,
                              default_friction: float | None = None) -> None:
        """Body-body contact generation at every step (solver.py:367-398);
        `default_friction` also back-fills bodies added with mu = 0."""
        self._enable_self_collision = bool(enabled)
        self._contact_margin = float(margin)
        if default_friction is not None:
            self._default_friction = max(0.0, float(default_friction))
            for k in range(len(self._bodies_friction)):
                if self._bodies_friction[k] == 0.0:
                    self._bodies_friction[k] = self._default_friction
        if enabled and self._contact_pool_start is None:
            self._contact_pool_start = len(self._constraints)

    def add_floor_contact(self, body: Body, floor_y: float = 0.0, stiffness: float = math.inf,
                          friction: float | None = None) -> ConstraintHandle:
        """Push-only floor row (fmin=-inf, fmax=0) and, for mu > 0, two
        friction rows along x and z whose bound is mu |lambda_n|
        (solver.py:784-832)."""
        idx = len(self._constraints)
        self._constraints.append(_Row(type=FLOOR_CONTACT, body_a=body.index, body_b=-1,
                                      world_anchor=(0.0, float(floor_y), 0.0),
                                      stiffness=stiffness, fmin=-math.inf, fmax=0.0))
        rows = 1
        mu = float(friction) if friction is not None else self._bodies_friction[body.index]
        if mu > 0.0:
            for tvec in ((1.0, 0.0, 0.0), (0.0, 0.0, 1.0)):
                self._constraints.append(_Row(type=CONTACT_TANGENT, body_a=body.index, body_b=-1,
                                              world_anchor=tvec, stiffness=stiffness,
                                              sibling=idx, friction=mu))
                rows += 1
        self._dirty = True
        return ConstraintHandle(index=idx, rows=rows)

    @staticmethod
    def _unit(normal) -> np.ndarray:
        n = np.asarray(normal, dtype=np.float32)
        return n / (np.linalg.norm(n) + 1e-20)

    def add_plane_contact(self, body: Body, normal, offset: float,
                          stiffness: float = math.inf) -> ConstraintHandle:
        """Push-only half-space n.x >= offset, n normalised in float32
        (solver.py:834-866)."""
        n = self._unit(normal)
        idx = len(self._constraints)
        self._constraints.append(_Row(type=PLANE_CONTACT, body_a=body.index, body_b=-1,
                                      world_anchor=(float(n[0]), float(n[1]), float(n[2])),
                                      rest=float(offset), stiffness=stiffness, fmin=-math.inf,
                                      fmax=0.0))
        self._dirty = True
        return ConstraintHandle(index=idx, rows=1)

    def set_plane_contact_bulk(self, handles: list[ConstraintHandle], normal=None,
                               offset: float | None = None) -> None:
        """Move plane rows on the device without a re-flush; a host mirror of
        the row anchors/offsets is kept across calls (solver.py:868-897)."""
        if not handles:
            return
        self._flush()
        f = self._flushed
        if self._c_rest_host is None:
            self._c_rest_host = f["rest"].copy()
        if self._c_world_anchor_host is None:
            self._c_world_anchor_host = f["anchor"].reshape(-1, 3).copy()
        nc = int(f["n_rows"])
        if self._c_rest_host.shape[0] != nc or self._c_world_anchor_host.shape[0] != nc:
            raise ValueError("plane mirror no longer matches the flushed row count")
        idxs = np.fromiter((h.index for h in handles), dtype=np.int32)
        if normal is not None:
            self._c_world_anchor_host[idxs] = self._unit(normal)
            self._h.write_rows(0, nc, anchor=self._c_world_anchor_host)
        if offset is not None:
            self._c_rest_host[idxs] = float(offset)
            self._h.write_rows(0, nc, rest=self._c_rest_host)

    # ---- runtime perturbations (solver.py:899-920) -----------------------------
    def _body_row(self, field: str, index: int) -> np.ndarray:
        return self._h.read((field,))[field][index]

    def set_velocity(self, body: Body, v) -> None:
        self._flush()
        self._h.write_bodies(1, body.index, np.asarray(v, np.float32).reshape(1, 3))

    def add_impulse(self, body: Body, dv) -> None:
        self._flush()
        cur = self._body_row("v", body.index)
        self._h.write_bodies(1, body.index, (cur + np.asarray(dv, np.float32)).reshape(1, 3))

    def set_position(self, body: Body, p) -> None:
        self._flush()
        self._h.write_bodies(0, body.index, np.asarray(p, np.float32).reshape(1, 3))

    # ---- flush -----------------------------------------------------------------
    def _make_handle(self):
        return N.ParticleHandle(self._device_index)

    def _params(self) -> N.PParams:
        p = N.PParams()
        p.dt = self.dt
        p.iterations = self.iterations
        for k in range(3):
            p.gravity[k] = self.gravity[k]
        p.alpha, p.beta, p.gamma = self.alpha, self.beta, self.gamma
        p.post_stabilize = 1 if self.post_stabilize else 0
        p.max_speed = self.max_speed
        return p

    def _scene_arrays(self, prev: dict | None) -> dict:
        """The flush arrays of solver.py:964-1122: bodies (simulated state
        kept for bodies that existed), coloring with the tets' six edges,
        rows with lambda/penalty/active kept below the contact pool, tets
        with their state kept, and the body -> row / body -> tet lists."""
        n_b, n_c, n_t = len(self._bodies_x), len(self._constraints), len(self._tets)
        x = np.array(self._bodies_x, np.float32).reshape(-1, 3)
        v = np.array(self._bodies_v, np.float32).reshape(-1, 3)
        prev_v = np.zeros((n_b, 3), np.float32)
        n_b_prev = 0 if prev is None else int(prev["x"].shape[0])
        if 0 < n_b_prev <= n_b:
            x[:n_b_prev] = prev["x"][:n_b_prev]
            v[:n_b_prev] = prev["v"][:n_b_prev]
            prev_v[:n_b_prev] = prev["prev_v"][:n_b_prev]

        rows = self._constraints
        ra = np.array([c.body_a for c in rows], np.int32)
        rb = np.array([c.body_b for c in rows], np.int32)
        tv = np.array([t[:4] for t in self._tets], np.int32).reshape(-1, 4)
        ea = np.concatenate([ra, tv[:, [0, 0, 0, 1, 1, 2]].reshape(-1)])
        eb = np.concatenate([rb, tv[:, [1, 2, 3, 2, 3, 3]].reshape(-1)])
        color = welsh_powell(body_edges(n_b, ea.tolist(), eb.tolist())) if n_b else \
            np.zeros(0, np.int32)
        n_colors = int(color.max() + 1) if n_b else 0

        lam = np.zeros(n_c, np.float32)
        pen = np.full(n_c, PENALTY_INIT, np.float32)
        act = np.ones(n_c, np.int32)
        upper = n_c if self._contact_pool_start is None else min(n_c, self._contact_pool_start)
        n_keep = 0 if prev is None else min(int(prev["lambda"].shape[0]), upper)
        if n_keep > 0:
            lam[:n_keep] = prev["lambda"][:n_keep]
            pen[:n_keep] = prev["penalty"][:n_keep]
            act[:n_keep] = prev["active"][:n_keep]

        tet_lam = np.zeros(n_t, np.float32)
        tet_pen = np.full(n_t, PENALTY_INIT_TET, np.float32)
        tet_act = np.ones(n_t, np.int32)
        t_keep = 0 if prev is None else min(int(prev["tet_lambda"].shape[0]), n_t)
        if t_keep > 0:
            tet_lam[:t_keep] = prev["tet_lambda"][:t_keep]
            tet_pen[:t_keep] = prev["tet_penalty"][:t_keep]
            tet_act[:t_keep] = prev["tet_active"][:t_keep]

        # body -> rows in row order, a row listed under body_a then body_b
        has_b = rb >= 0
        owners = np.stack([ra, np.where(has_b, rb, -1)], axis=1).reshape(-1)
        vals = np.repeat(np.arange(n_c, dtype=np.int64), 2)
        keep = owners >= 0
        con_start, con_idx = _csr(n_b, owners[keep], vals[keep])
        # body -> (tet << 2 | slot), tets in order, slots 0..3
        t_owners = tv.reshape(-1).astype(np.int64)
        t_vals = (np.repeat(np.arange(n_t, dtype=np.int64), 4) << 2) | np.tile(np.arange(4), n_t)
        tet_start, tet_idx = _csr(n_b, t_owners, t_vals)

        f32 = lambda seq: np.array(list(seq), np.float32)
        return {
            "n_bodies": n_b, "x": x, "v": v, "prev_v": prev_v,
            "mass": np.array(self._bodies_mass, np.float32), "color": color,
            "n_colors": n_colors,
            "n_rows": n_c, "type": np.array([c.type for c in rows], np.int32),
            "body_a": ra, "body_b": rb,
            "anchor": np.array([c.world_anchor for c in rows], np.float32).reshape(-1, 3),
            "rest": f32(c.rest for c in rows), "stiffness": f32(c.stiffness for c in rows),
            "fmin": f32(c.fmin for c in rows), "fmax": f32(c.fmax for c in rows),
            "fracture": f32(c.fracture for c in rows),
            "sibling": np.array([c.sibling for c in rows], np.int32),
            "friction": f32(c.friction for c in rows),
            "lambda": lam, "penalty": pen, "active": act,
            "con_start": con_start, "con_idx": con_idx,
            "n_tets": n_t, "tet_v": tv,
            "tet_rest_vol": f32(t[4] for t in self._tets),
            "tet_stiffness": f32(t[5] for t in self._tets),
            "tet_fracture": f32(t[6] for t in self._tets),
            "tet_lambda": tet_lam, "tet_penalty": tet_pen, "tet_active": tet_act,
            "tet_start": tet_start, "tet_idx": tet_idx,
            "off_a": np.array([c.off_a for c in rows], np.float32).reshape(-1, 3),
            "off_b": np.array([c.off_b for c in rows], np.float32).reshape(-1, 3),
        }

    def _flush(self) -> None:
        if not self._dirty:
            return
        prev = self._h.read() if self._h is not None and self._flushed is not None else None
        arrays = self._scene_arrays(prev)
        if self._enable_self_collision and self._pool_pair_rows:
            self._seed_pool(arrays)
        if self._h is None:
            self._h = self._make_handle()
        self._h.flush(arrays)
        self._flushed = arrays
        color = arrays["color"]
        self.num_colors = int(arrays["n_colors"])
        self.color_counts = {int(c): int(k) for c, k in zip(*np.unique(color, return_counts=True))}
        self.body_color = DeviceArray(lambda: color.copy())
        for name, key in (("c_type", "type"), ("c_body_a", "body_a"), ("c_body_b", "body_b")):
            setattr(self, name, DeviceArray(lambda a=arrays[key]: a.copy()))
        self.c_active = DeviceArray(lambda: self._h.read(("active",))["active"])
        # the reference's device arrays the viewer and tests read (solver.py:213-225)
        self.x = DeviceArray(lambda: self._h.read(("x",))["x"])
        self.v = DeviceArray(lambda: self._h.read(("v",))["v"])
        self.prev_v = DeviceArray(lambda: self._h.read(("prev_v",))["prev_v"])
        self.c_lambda = DeviceArray(lambda: self._h.read(("lambda",))["lambda"])
        self.c_penalty = DeviceArray(lambda: self._h.read(("penalty",))["penalty"])
        self.tet_lambda = DeviceArray(lambda: self._h.read(("tet_lambda",))["tet_lambda"])
        self.tet_active = DeviceArray(lambda: self._h.read(("tet_active",))["tet_active"])
        self._dirty = False

    def _seed_pool(self, arrays: dict) -> None:
        """Contact rows start from the state their pair key carried out of
        the last step, when it is finite (solver.py:1160-1185)."""
        lam, pen = arrays["lambda"], arrays["penalty"]
        for key, n_idx, t_idx, b_idx in self._pool_pair_rows:
            st = self._contact_state_cache.get(key)
            if st is None or not all(math.isfinite(v) for v in st):
                continue
            lam[n_idx], pen[n_idx] = st[0], st[3]
            if t_idx >= 0:
                lam[t_idx], pen[t_idx] = st[1], st[4]
            if b_idx >= 0:
                lam[b_idx], pen[b_idx] = st[2], st[5]

    def _store_pool(self) -> None:
        """Each contact whose normal row is still active keeps (lambda_n,
        lambda_t, lambda_b, k_n, k_t, k_b) under its pair key for the next
        step; missing friction rows count as (0, 1) (solver.py:1392-1421)."""
        st = self._h.read(("lambda", "penalty", "active"))
        lam, pen, act = st["lambda"], st["penalty"], st["active"]
        cache = {}
        for key, n_idx, t_idx, b_idx in self._pool_pair_rows:
            if act[n_idx] == 0:
                continue
            v = (float(lam[n_idx]), float(lam[t_idx]) if t_idx >= 0 else 0.0,
                 float(lam[b_idx]) if b_idx >= 0 else 0.0, float(pen[n_idx]),
                 float(pen[t_idx]) if t_idx >= 0 else 1.0,
                 float(pen[b_idx]) if b_idx >= 0 else 1.0)
            if all(math.isfinite(x) for x in v):
                cache[key] = v
        self._contact_state_cache = cache

    # ---- the step --------------------------------------------------------------
    def _self_collision_prelude(self) -> None:
        """solver.py:1128-1158: reset non-finite bodies to their scene
        positions (dropping their cached contacts), then rebuild the contact
        pool from the current positions; the step re-flushes the scene."""
        self._flush()
        st = self._h.read(("x", "v", "active"))
        xs = st["x"]
        bad = ~np.all(np.isfinite(xs), axis=1)
        if bad.any():
            safe = np.array(self._bodies_x, np.float32).reshape(-1, 3)
            xs[bad] = safe[bad]
            vs = st["v"]
            vs[bad] = 0.0
            self._h.write_bodies(0, 0, xs)
            self._h.write_bodies(1, 0, vs)
            gone = set(np.nonzero(bad)[0].tolist())
            self._contact_state_cache = {k: v for k, v in self._contact_state_cache.items()
                                         if k[0] not in gone and k[1] not in gone}
            warnings.warn(f"avbd3d: {int(bad.sum())} body position(s) were non-finite (solver "
                          "divergence); reset to initial scene positions and zeroed velocity.",
                          RuntimeWarning, stacklevel=3)
        self._pool_pair_rows = PC.rebuild_contact_pool(self, _Row, xs, st["active"])
        self._dirty = True

    def step(self) -> None:
        """One Solver.step (solver.py:1124-1423)."""
        if self._enable_self_collision:
            self._self_collision_prelude()
        self._flush()
        if not self._bodies_x:
            return
        self._h.step(self._params())
        if self._enable_self_collision and self._pool_pair_rows:
            self._store_pool()

    # ---- read-back (solver.py:1425-1443) ----------------------------------------
    def positions(self) -> np.ndarray:
        if self._h is None:
            return np.array(self._bodies_x, np.float32).reshape(-1, 3)
        return self._h.read(("x",))["x"]

    def velocities(self) -> np.ndarray:
        if self._h is None:
            return np.array(self._bodies_v, np.float32).reshape(-1, 3)
        return self._h.read(("v",))["v"]

    def lambdas(self) -> np.ndarray:
        if self._h is None:
            return np.zeros(len(self._constraints), np.float32)
        return self._h.read(("lambda",))["lambda"]

    def active(self) -> np.ndarray:
        if self._h is None:
            return np.ones(len(self._constraints), np.int32)
        return self._h.read(("active",))["active"]

    def autotune_primal(self, steps: int = 2, variants=(2, 4)) -> dict[int, float]:
        """Time `steps` real steps with each primal kernel variant (one
        launch per color, one dataflow launch per sweep, ...) and keep the
        fastest; every variant gives bit-identical results, so the steps
        are ordinary simulation steps. Returns ms per step by variant. Not
        in the reference."""
        self._flush()
        if self._h is None or not self._bodies_x:
            return {}
        times = {}
        for v in variants:
            self._h.set_option(N.P_OPT_PRIMAL_VARIANT, int(v))
            self.step()
            self._h.read(("x",))
            t0 = time.perf_counter()
            for _ in range(steps):
                self.step()
            self._h.read(("x",))
            times[int(v)] = 1e3 * (time.perf_counter() - t0) / steps
        best = min(times, key=times.get)
        self._h.set_option(N.P_OPT_PRIMAL_VARIANT, best)
        self.primal_variant = best
        return times

    def state(self) -> dict[str, np.ndarray]:
        """Every device array the step updates (bodies, rows, tets)."""
        self._flush()
        return self._h.read()
