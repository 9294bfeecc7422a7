"""Solver6DOF — the reference's 6-DOF AVBD rigid-body API, on MI355X.

Host-side mirror of ``avbd3d.solver_6dof`` (src/avbd3d/solver_6dof.py):
the same constructor keywords (:180-201), scene builders (:446-578),
properties (:597-661), ``step()`` (:1178-1197), readers (:1920-2052),
setters (:878-900), warnings and diagnostic attributes. Every substep runs
on the GPU through libavbd_hip.so (one ``avbd_substep`` call per
``_step_one``); this module builds the flushed scene arrays, applies the
reference's host-side policies (substep dt, overflow warnings, color-dirty
flag, graph signature) and exposes device arrays with a ``.numpy()``
method the way Warp arrays are used by the reference's callers.

Device semantics: the substep follows the reference's conflict-checked,
in-place Gauss-Seidel path (the one its CPU device runs). ``gpu_resident``,
``primal_group_size``, ``primal_shuffle``, ``primal_fused`` and
``recolor_every_substep`` are accepted and kept in the graph signature;
they do not change results.
"""
from __future__ import annotations

import math
import warnings
from dataclasses import dataclass

import numpy as np

from . import _native as N

# Constraint type codes (solver_6dof.py:31-34).
FLOOR_CONTACT_6DOF = 0
CONTACT_TANGENT_6DOF = 1
PIN_6DOF = 2
BOX_BOX_CONTACT_6DOF = 3

MAX_COLORS = 32          # kernels_6dof.py:2582
CONTACT_MARGIN = 0.005   # solver_6dof.py:1789


def _obb_sat(c_A, R_A, e_A, c_B, R_B, e_B, margin: float = 0.005):
    """15-axis separating-axis test of two oriented boxes (solver_6dof.py:37-102).

    Returns ``(axis_idx, n_hat, overlap)`` with ``n_hat`` pointing from B to A,
    or ``None`` when separated by more than ``margin`` on a face axis or at
    all on an edge-edge axis. Axis order: A's faces 0-2, B's faces 3-5, edge
    pairs 6 + 3 i + j; the first strictly smallest overlap wins.
    """
    c_A, c_B = np.asarray(c_A, np.float64), np.asarray(c_B, np.float64)
    R_A, R_B = np.asarray(R_A, np.float64), np.asarray(R_B, np.float64)
    e_A, e_B = np.asarray(e_A, np.float64), np.asarray(e_B, np.float64)
    t = c_B - c_A

    def radius(L, R, e):
        return float(sum(e[m] * abs(L @ R[:, m]) for m in range(3)))

    best = (None, math.inf, None)
    axes = [(k, R_A[:, k], e_A[k], None) for k in range(3)]
    axes += [(3 + k, R_B[:, k], None, e_B[k]) for k in range(3)]
    for k, L, ra, rb in axes:
        ra = radius(L, R_A, e_A) if ra is None else ra
        rb = radius(L, R_B, e_B) if rb is None else rb
        tl = float(t @ L)
        ov = ra + rb - abs(tl)
        if ov < -margin:
            return None
        if ov < best[1]:
            best = (k, ov, -L if tl > 0 else L)
    for i in range(3):
        for j in range(3):
            L = np.cross(R_A[:, i], R_B[:, j])
            nl = np.linalg.norm(L)
            if nl < 1e-6:
                continue
            L = L / nl
            tl = float(t @ L)
            ov = radius(L, R_A, e_A) + radius(L, R_B, e_B) - abs(tl)
            if ov < 0:
                return None
            if ov < best[1]:
                best = (6 + 3 * i + j, ov, -L if tl > 0 else L)
    return (best[0], np.asarray(best[2], np.float32), float(best[1]))


@dataclass
class RigidBody:
    """Handle returned by add_box (solver_6dof.py:105-110)."""
    index: int
    half_extents: tuple[float, float, float]


@dataclass
class _Row:
    """Static constraint row record (solver_6dof.py:113-129)."""
    type: int
    body_a: int
    body_b: int = -1                  # PIN_6DOF: axis 0/1/2
    world_anchor: tuple[float, float, float] = (0.0, 0.0, 0.0)
    off_a: tuple[float, float, float] = (0.0, 0.0, 0.0)
    off_b: tuple[float, float, float] = (0.0, 0.0, 0.0)
    rest: float = 0.0
    stiffness: float = math.inf
    fracture: float = math.inf
    fmin: float = -math.inf
    fmax: float = math.inf
    sibling: int = -1
    friction: float = 0.0
    friction_static: float = 0.0
    partner: int = -1


def box_inertia_local(mass: float, hx: float, hy: float, hz: float) -> np.ndarray:
    """Solid-box inertia m/3 (h_j^2 + h_k^2) on the diagonal; zero for m <= 0
    (solver_6dof.py:132-142)."""
    if mass <= 0.0:
        return np.zeros((3, 3), dtype=np.float32)
    d = [mass / 3.0 * (hy * hy + hz * hz), mass / 3.0 * (hx * hx + hz * hz),
         mass / 3.0 * (hx * hx + hy * hy)]
    return np.diag(d).astype(np.float32)


def box_inv_inertia_local(mass: float, hx: float, hy: float, hz: float) -> np.ndarray:
    """Inverse of box_inertia_local; zero for static bodies (solver_6dof.py:145-152)."""
    if mass <= 0.0:
        return np.zeros((3, 3), dtype=np.float32)
    return np.linalg.inv(box_inertia_local(mass, hx, hy, hz)).astype(np.float32)


def box_inertia_local_or_zero(mass: float, hx: float, hy: float, hz: float) -> np.ndarray:
    """solver_6dof.py:155-162."""
    if mass <= 0.0:
        return np.zeros((3, 3), dtype=np.float32)
    return box_inertia_local(mass, hx, hy, hz)


def _snap_group_size(g: int) -> int:
    """Largest power of two <= g within [1, 32] (solver_6dof.py:165-174)."""
    g = max(1, min(32, int(g)))
    p = 1
    while p * 2 <= g:
        p *= 2
    return p


def _parse_device(device) -> int:
    """HIP device index of a device string. The substep always runs on a HIP
    device: "cpu" (the reference's default) selects device 0 and keeps the
    reference's CPU-device behaviour (growing pools, a readback per substep);
    "cuda[:N]" / "hip[:N]" / "gpu[:N]" select device N with the CUDA-device
    behaviour (the resident, fixed-capacity path when gpu_resident is set)."""
    d = str(device).strip().lower()
    if d == "cpu" or d.startswith("cpu:"):
        return 0
    for pre in ("cuda", "hip", "gpu"):
        if d == pre:
            return 0
        if d.startswith(pre + ":"):
            return int(d.split(":", 1)[1])
    raise ValueError(f"unrecognised device {device!r}")


class DeviceArray:
    """A device buffer of the solver, read back with ``.numpy()``."""

    def __init__(self, reader, writer=None):
        self._reader = reader
        self._writer = writer

    def numpy(self) -> np.ndarray:
        return self._reader()

    def assign(self, values) -> None:
        if self._writer is None:
            raise TypeError("array is read-only")
        self._writer(np.asarray(values))

    @property
    def shape(self):
        return self._reader().shape

    def zero_(self) -> None:
        self.assign(np.zeros(self.shape, dtype=self._reader().dtype))


class Solver6DOF:
    """6-DOF AVBD rigid body solver (solver_6dof.py:177)."""

    _COLORING_MODES = ("jones_plassmann", "jacobi")

    def __init__(
        self,
        dt: float = 1.0 / 60.0,
        iterations: int = 10,
        gravity: tuple[float, float, float] = (0.0, -9.81, 0.0),
        alpha: float = 0.99,
        beta: float = 1.0e5,
        gamma: float = 0.99,
        post_stabilize: bool = True,
        device: str = "cpu",
        max_linear_speed: float = 30.0,
        max_angular_speed: float = 50.0,
        substeps: int = 1,
        friction_static_mult: float = 1.5,
        coloring_mode: str = "jacobi",
        unsafe_fixed_capacity: bool = False,
        primal_group_size: int = 16,
        primal_shuffle: bool = True,
        primal_fused: bool = True,
        gpu_resident: bool = True,
        recolor_every_substep: bool = False,
    ):
        self._device_index = _parse_device(device)
        self.device = str(device)
        N.load_library()   # fail now, loudly, if the native library is missing
        self._primal_group_size = _snap_group_size(primal_group_size)
        self._primal_shuffle = bool(primal_shuffle)
        self._primal_fused = bool(primal_fused)
        self._gpu_resident = bool(gpu_resident)
        self._recolor_every_substep = bool(recolor_every_substep)
        self._resident_recolor_rounds = 6
        self._unsafe_fixed_capacity = bool(unsafe_fixed_capacity)
        # not in the reference: fixed-capacity frames replay a captured HIP
        # graph for their substeps after the first (False: launch directly)
        self.hip_graph = True
        self._coloring_mode = None
        self.coloring_mode = coloring_mode
        self.dt = float(dt)
        self.iterations = int(iterations)
        self.gravity = tuple(float(g) for g in gravity)
        self.alpha = float(alpha)
        self.beta = float(beta)
        self.gamma = float(gamma)
        self.post_stabilize = bool(post_stabilize)
        self.max_linear_speed = float(max_linear_speed)
        self.max_angular_speed = float(max_angular_speed)
        self.substeps = max(1, int(substeps))
        self.friction_static_mult = max(1.0, float(friction_static_mult))

        self._x: list = []
        self._q: list = []
        self._v: list = []
        self._omega: list = []
        self._mass: list = []
        self._inv_I_local: list = []
        self._I_local: list = []
        self._half_extents: list = []
        self._he_dirty = True
        self._friction: list = []
        self._rows: list[_Row] = []
        self._dirty = True
        self._self_collide = False
        self._self_friction = 0.0

        self._h = None
        self._h_graph_state = None
        self.x = self.q = self.v = self.omega = None
        self.prev_v = self.prev_omega = None
        self.body_color = None
        self.c_lambda = self.c_active = None
        self.n_active_rows = None
        self.pool_count = None
        self._bp_pair_count = None
        self.num_colors = 0
        self.num_active_colors = 0
        self._n_active_colors = 0
        self._max_colors = MAX_COLORS
        self._color_dirty = True
        self._bp_max_pairs = 0
        self.broadphase_ms = 0.0
        self._gpu_pool_ready = False
        self._gpu_pool_n_static = 0
        self._gpu_pool_n_capacity = 0
        self._gpu_pool_n_dyn_capacity = 0
        self._gpu_pool_max_pairs = 0
        self._gpu_pool_hash_cap = 0
        self._gpu_pool_pool_max = 0
        self._row_overflow_warned = False
        self._fixed_cap_overflow_warned = False
        self._pair_overflow_warned = False
        self._graph = None
        self._graph_signature: tuple | None = None
        self._graph_supported: bool | None = None
        self._flush_snapshot: dict | None = None
        self._warm_start_mode = 0          # learned warm start: 0 plain, 1 log, 2 learned
        self._warm_start_params: np.ndarray | None = None
        self._warm_start_scope = 0
        self.primal_variant = 6            # the library's default primal sweep kernel
        self.dual_placement = 0            # the dual in a launch of its own (autotune_dual)

    # ---- scene building (solver_6dof.py:446-578) ----------------------------
    def add_box(self, position, half_extents, mass: float, orientation=(0.0, 0.0, 0.0, 1.0),
                velocity=(0.0, 0.0, 0.0), angular_velocity=(0.0, 0.0, 0.0),
                friction: float = 0.0) -> RigidBody:
        """Add a box; ``mass <= 0`` marks it static. Orientation is xyzw."""
        idx = len(self._x)
        self._x.append(tuple(float(v) for v in position))
        self._q.append(tuple(float(v) for v in orientation))
        self._v.append(tuple(float(v) for v in velocity))
        self._omega.append(tuple(float(v) for v in angular_velocity))
        self._mass.append(float(mass))
        hx, hy, hz = float(half_extents[0]), float(half_extents[1]), float(half_extents[2])
        self._half_extents.append((hx, hy, hz))
        self._he_dirty = True
        self._inv_I_local.append(box_inv_inertia_local(mass, hx, hy, hz))
        self._I_local.append(box_inertia_local_or_zero(mass, hx, hy, hz))
        self._friction.append(max(0.0, float(friction)))
        self._dirty = True
        return RigidBody(index=idx, half_extents=(hx, hy, hz))

    def add_floor_contact_box(self, body: RigidBody, floor_y: float = 0.0,
                              friction: float | None = None,
                              stiffness: float = math.inf) -> list[int]:
        """One FLOOR row per corner plus world-x / world-z tangent rows when
        mu > 0 (solver_6dof.py:474-541). Returns the normal-row indices."""
        mu = float(friction) if friction is not None else self._friction[body.index]
        hx, hy, hz = body.half_extents
        normals: list[int] = []
        for sx in (-1.0, 1.0):
            for sy in (-1.0, 1.0):
                for sz in (-1.0, 1.0):
                    off = (sx * hx, sy * hy, sz * hz)
                    n_idx = len(self._rows)
                    self._rows.append(_Row(type=FLOOR_CONTACT_6DOF, body_a=body.index,
                                           world_anchor=(0.0, float(floor_y), 0.0), off_a=off,
                                           stiffness=stiffness, fmin=-math.inf, fmax=0.0))
                    normals.append(n_idx)
                    if mu > 0.0:
                        mu_s = mu * self.friction_static_mult
                        tx = len(self._rows)
                        self._rows.append(_Row(type=CONTACT_TANGENT_6DOF, body_a=body.index,
                                               world_anchor=(1.0, 0.0, 0.0), off_a=off,
                                               stiffness=stiffness, sibling=n_idx, friction=mu,
                                               friction_static=mu_s))
                        tz = len(self._rows)
                        self._rows.append(_Row(type=CONTACT_TANGENT_6DOF, body_a=body.index,
                                               world_anchor=(0.0, 0.0, 1.0), off_a=off,
                                               stiffness=stiffness, sibling=n_idx, friction=mu,
                                               friction_static=mu_s))
                        self._rows[tx].partner = tz
                        self._rows[tz].partner = tx
        self._dirty = True
        return normals

    def add_pin_corner(self, body: RigidBody, body_local, world_point,
                       stiffness: float = math.inf, fracture: float = math.inf) -> list[int]:
        """Pin a body-local point to a world point, one row per axis
        (solver_6dof.py:543-568)."""
        out = []
        for axis in (0, 1, 2):
            out.append(len(self._rows))
            self._rows.append(_Row(type=PIN_6DOF, body_a=body.index, body_b=axis,
                                   world_anchor=tuple(float(p) for p in world_point),
                                   off_a=tuple(float(v) for v in body_local),
                                   stiffness=stiffness, fracture=fracture))
        self._dirty = True
        return out

    def enable_self_collision(self, enabled: bool = True, default_friction: float = 0.0) -> None:
        """Per-substep box-box contact generation (solver_6dof.py:570-578)."""
        self._self_collide = bool(enabled)
        self._self_friction = max(0.0, float(default_friction))

    # ---- properties (solver_6dof.py:597-678) ----------------------------------
    @property
    def coloring_mode(self) -> str:
        return self._coloring_mode

    @coloring_mode.setter
    def coloring_mode(self, mode: str) -> None:
        m = str(mode).strip().lower()
        if m in ("jp", "jones-plassmann", "jones plassmann"):
            m = "jones_plassmann"
        elif m in ("speculative", "speculative_greedy", "greedy"):
            m = "jacobi"
        if m not in self._COLORING_MODES:
            raise ValueError(f"coloring_mode must be one of {self._COLORING_MODES}, got {mode!r}")
        if m != self._coloring_mode:
            self._coloring_mode = m
            self._color_dirty = True

    @property
    def primal_group_size(self) -> int:
        return self._primal_group_size

    @primal_group_size.setter
    def primal_group_size(self, g: int) -> None:
        self._primal_group_size = _snap_group_size(g)

    @property
    def gpu_resident(self) -> bool:
        return self._gpu_resident

    @gpu_resident.setter
    def gpu_resident(self, on: bool) -> None:
        self._gpu_resident = bool(on)

    @property
    def recolor_every_substep(self) -> bool:
        return self._recolor_every_substep

    @recolor_every_substep.setter
    def recolor_every_substep(self, on: bool) -> None:
        self._recolor_every_substep = bool(on)

    def _resident_on(self, dev) -> bool:
        """Whether the readback-free resident loop runs (solver_6dof.py:663-672):
        a GPU device string, gpu_resident, the fused shuffled group primal.
        It implies fixed-capacity pools, and each frame runs as one call whose
        substeps after the first replay a captured HIP graph, with one counter
        readback per frame. Unlike the reference's double-buffered resident
        primal, coloring stays conflict-checked on the device, so the resident
        loop gives the same bits as the per-substep path."""
        d = str(dev).strip().lower()
        return (self._gpu_resident
                and d.startswith(("cuda", "hip", "gpu"))
                and self._primal_shuffle
                and self._primal_fused
                and self._primal_group_size > 1)

    def _fixed_capacity_mode(self, dev) -> bool:
        return self._unsafe_fixed_capacity or self._resident_on(dev)

    # ---- native plumbing ------------------------------------------------------
    def _make_handle(self):
        return N.NativeHandle(self._device_index)

    def _params(self, dt: float) -> N.Params:
        p = N.Params()
        p.dt = float(dt)
        p.iterations = int(self.iterations)
        for k in range(3):
            p.gravity[k] = float(self.gravity[k])
        p.alpha, p.beta, p.gamma = float(self.alpha), float(self.beta), float(self.gamma)
        p.post_stabilize = 1 if self.post_stabilize else 0
        if (self.max_linear_speed > 0.0 and math.isfinite(self.max_linear_speed)
                and self.max_angular_speed > 0.0 and math.isfinite(self.max_angular_speed)):
            p.max_linear_speed = float(self.max_linear_speed)
            p.max_angular_speed = float(self.max_angular_speed)
        else:
            p.max_linear_speed = math.inf
            p.max_angular_speed = math.inf
        p.friction_static_mult = float(self.friction_static_mult)
        p.coloring_mode = N.COLOR_JACOBI if self._coloring_mode == "jacobi" else N.COLOR_JONES_PLASSMANN
        p.self_collide = 1 if self._self_collide else 0
        p.self_friction = float(self._self_friction)
        p.fixed_capacity = 1 if self._fixed_capacity_mode(self.device) else 0
        p.margin = CONTACT_MARGIN
        return p

    def _scene_arrays(self) -> dict:
        """The arrays _flush uploads (solver_6dof.py:904-1129)."""
        n_b = len(self._x)
        n_static = len(self._rows)
        f32 = np.float32
        prev = None
        if self._h is not None and self.x is not None and self._gpu_pool_ready:
            prev = self._h.read_bodies(self._n_b_flushed)
        x = np.array(self._x, f32).reshape(-1, 3)
        q = np.array(self._q, f32).reshape(-1, 4)
        v = np.array(self._v, f32).reshape(-1, 3)
        w = np.array(self._omega, f32).reshape(-1, 3)
        prev_v = np.zeros((n_b, 3), f32)
        prev_w = np.zeros((n_b, 3), f32)
        if prev is not None:
            k = self._n_b_flushed
            if 0 < k <= n_b:
                x[:k], q[:k], v[:k], w[:k] = prev["x"], prev["q"], prev["v"], prev["w"]
                prev_v[:k], prev_w[:k] = prev["prev_v"], prev["prev_w"]
        if self._self_collide:
            max_pairs = max(64, 16 * max(1, n_b))
            n_dyn = max_pairs * 4 * 3
            hash_cap = max(1024, max_pairs * 8)
            pool_max = max(64, max_pairs * 4)
        else:
            max_pairs = n_dyn = hash_cap = pool_max = 0
        prio = (np.random.default_rng(seed=0xA5BD).uniform(0.0, 1.0, size=n_b).astype(f32)
                if n_b > 0 else np.zeros(0, f32))
        rows = self._rows
        if rows:
            ints = np.array([(r.type, r.body_a, r.body_b, r.sibling, r.partner) for r in rows],
                            dtype=np.int32)
            flts = np.array([(*r.world_anchor, *r.off_a, *r.off_b, r.stiffness, r.fmin, r.fmax,
                              r.fracture, r.friction, r.friction_static) for r in rows],
                            dtype=np.float64).astype(f32)
        else:
            ints = np.zeros((0, 5), np.int32)
            flts = np.zeros((0, 15), f32)
        return {
            "n_bodies": n_b,
            "x": x, "q": q, "v": v, "w": w, "prev_v": prev_v, "prev_w": prev_w,
            "mass": np.array(self._mass, f32),
            "inertia_local": (np.stack(self._I_local).astype(f32).reshape(-1, 9)
                              if n_b else np.zeros((0, 9), f32)),
            "half_extents": np.array(self._half_extents, f32).reshape(-1, 3),
            "friction": np.array(self._friction, f32),
            "priority": prio,
            "n_static_rows": n_static,
            "type": ints[:, 0].copy(),
            "body_a": ints[:, 1].copy(),
            "body_b": ints[:, 2].copy(),
            "sibling": ints[:, 3].copy(),
            "partner": ints[:, 4].copy(),
            "anchor": flts[:, 0:3].copy(),
            "off_a": flts[:, 3:6].copy(),
            "off_b": flts[:, 6:9].copy(),
            "stiffness": flts[:, 9].copy(),
            "fmin": flts[:, 10].copy(),
            "fmax": flts[:, 11].copy(),
            "fracture": flts[:, 12].copy(),
            "friction_d": flts[:, 13].copy(),
            "friction_s": flts[:, 14].copy(),
            "lambda": np.zeros(n_static, f32),
            "penalty": np.ones(n_static, f32),
            "active": np.ones(n_static, np.int32),
            "was_static": np.zeros(n_static, np.int32),
            "max_pairs": int(self._bp_max_pairs),
            "n_dyn_capacity": n_dyn,
            "hash_cap": hash_cap,
            "pool_max": pool_max,
            "flush_max_pairs": max_pairs,
        }

    @staticmethod
    def make_scene_struct(arrs: dict) -> tuple[N.Scene, list]:
        """ctypes view of a scene dict; the list keeps the arrays alive."""
        keep = []

        def f(name, shape_last=None):
            a = np.ascontiguousarray(arrs[name], dtype=np.float32)
            if a.size == 0:
                a = np.zeros(max(1, shape_last or 1), np.float32)
            keep.append(a)
            return N.fptr(a)

        def i(name):
            a = np.ascontiguousarray(arrs[name], dtype=np.int32)
            if a.size == 0:
                a = np.zeros(1, np.int32)
            keep.append(a)
            return N.iptr(a)

        sc = N.Scene()
        sc.n_bodies = int(arrs["n_bodies"])
        for name in ("x", "q", "v", "w", "prev_v", "prev_w", "mass", "inertia_local",
                     "half_extents", "friction", "priority"):
            setattr(sc, name, f(name))
        sc.n_static_rows = int(arrs["n_static_rows"])
        for name in ("type", "body_a", "body_b", "sibling", "partner"):
            setattr(sc, name, i(name))
        for name in ("anchor", "off_a", "off_b", "stiffness", "fmin", "fmax", "fracture",
                     "friction_d", "friction_s", "penalty"):
            setattr(sc, name, f(name))
        sc.lambda_ = f("lambda")
        sc.active = i("active")
        sc.was_static = i("was_static")
        sc.max_pairs = int(arrs["max_pairs"])
        sc.n_dyn_capacity = int(arrs["n_dyn_capacity"])
        sc.hash_cap = int(arrs["hash_cap"])
        sc.pool_max = int(arrs["pool_max"])
        return sc, keep

    def _flush(self) -> None:
        """Upload bodies and static rows, size every pool (solver_6dof.py:904-1174)."""
        if not self._dirty:
            return
        arrs = self._scene_arrays()
        if self._h is None:
            self._h = self._make_handle()
        sc, keep = self.make_scene_struct(arrs)
        self._h.flush(sc)
        del keep
        if self._warm_start_mode or self._warm_start_params is not None:
            self._apply_warm_start()
        self._flush_snapshot = arrs
        n_b = arrs["n_bodies"]
        n_static = arrs["n_static_rows"]
        self._n_b_flushed = n_b
        self._max_colors = MAX_COLORS
        self.num_colors = MAX_COLORS
        self.num_active_colors = MAX_COLORS
        self._n_active_colors = MAX_COLORS
        self.color_counts = {}
        self._gpu_pool_n_static = n_static
        self._gpu_pool_n_dyn_capacity = arrs["n_dyn_capacity"]
        self._gpu_pool_n_capacity = n_static + arrs["n_dyn_capacity"]
        self._gpu_pool_max_pairs = arrs["flush_max_pairs"]
        self._gpu_pool_hash_cap = arrs["hash_cap"]
        self._gpu_pool_pool_max = arrs["pool_max"]
        self._gpu_pool_ready = True
        self._dirty = False
        self._color_dirty = True
        self._graph = None
        self._bind_views()

    def _bind_views(self) -> None:
        h = self

        def body(field, key):
            return DeviceArray(lambda: h._h.read_bodies(h._n_b_flushed)[key],
                               lambda vals: h._h.write_bodies(field, 0, vals))

        self.x = body(0, "x")
        self.q = body(1, "q")
        self.v = body(2, "v")
        self.omega = body(3, "w")
        self.prev_v = body(4, "prev_v")
        self.prev_omega = body(5, "prev_w")
        self.body_color = DeviceArray(lambda: h._h.read_colors(h._n_b_flushed))
        self.c_lambda = DeviceArray(lambda: h._h.read_rows(h._gpu_pool_n_capacity)["lambda"])
        self.c_active = DeviceArray(lambda: h._h.read_rows(h._gpu_pool_n_capacity)["active"])
        self.c_was_static = DeviceArray(lambda: h._h.read_rows(h._gpu_pool_n_capacity)["was_static"])
        self.c_type = DeviceArray(lambda: h._h.read_rows(h._gpu_pool_n_capacity)["type"])
        self.c_penalty = DeviceArray(lambda: h._h.read_rows(h._gpu_pool_n_capacity)["penalty"])
        self.n_active_rows = DeviceArray(lambda: np.array([h._h.read_counters()[0]], np.int32))
        self._bp_pair_count = DeviceArray(lambda: np.array([h._h.read_counters()[1]], np.int32))
        self.pool_count = DeviceArray(lambda: np.array([h._h.read_counters()[2]], np.int32))

    # ---- setters (solver_6dof.py:878-900), in place --------------------------
    def set_position(self, body: RigidBody, p) -> None:
        self._flush()
        self._h.write_bodies(0, body.index, np.asarray(p, np.float32).reshape(1, 3))

    def set_orientation(self, body: RigidBody, q_xyzw) -> None:
        self._flush()
        self._h.write_bodies(1, body.index, np.asarray(q_xyzw, np.float32).reshape(1, 4))

    def set_velocity(self, body: RigidBody, v) -> None:
        self._flush()
        self._h.write_bodies(2, body.index, np.asarray(v, np.float32).reshape(1, 3))

    def set_angular_velocity(self, body: RigidBody, w) -> None:
        self._flush()
        self._h.write_bodies(3, body.index, np.asarray(w, np.float32).reshape(1, 3))

    # ---- step (solver_6dof.py:1178-1464) --------------------------------------
    def step(self) -> None:
        fixed_cap = self._fixed_capacity_mode(self.device)
        if fixed_cap:
            self._step_frame()
            self._check_fixed_cap_overflow()
            return
        if self.substeps <= 1:
            self._step_one()
            if fixed_cap:
                self._check_fixed_cap_overflow()
            return
        full_dt = self.dt
        self.dt = full_dt / self.substeps
        try:
            for _ in range(self.substeps):
                self._step_one()
        finally:
            self.dt = full_dt
        if fixed_cap:
            self._check_fixed_cap_overflow()

    def step_to_residual(self, tol: float = 1e-4, max_iters: int = 200,
                         on_substep=None) -> list[tuple[int, float]]:
        """One frame like step(), each substep iterated until the dual
        residual r = max over active rows of |lambda_new - lambda| / k (metres)
        is <= tol, at most `max_iters` iterations. Not in the reference: the
        "iterations to residual" measurement. Returns (iterations, r) per
        substep; a substep equals a fixed-iteration one with that count.
        `on_substep(k)`, if given, runs after substep k."""
        if self._dirty:
            self._flush()
        n_b = len(self._x)
        if n_b == 0:
            return []
        out = []
        full_dt = self.dt
        sub_dt = full_dt / self.substeps if self.substeps > 1 else full_dt
        for _ in range(max(1, self.substeps)):
            params = self._params(sub_dt)
            params.iterations = int(max_iters)
            rep, iters, r = self._h.substep_to_residual(params, self._color_dirty, tol)
            self._apply_report(rep, n_b)
            out.append((iters, r))
            if on_substep is not None:
                on_substep(len(out) - 1)
        return out

    # ---- primal kernel autotuner (not in the reference) -----------------------
    PRIMAL_VARIANTS = (6, 7, 8, 9, 10)   # the dataflow sweeps, one to eight bodies per tile

    def autotune_primal(self, candidates=None, frames: int = 1) -> dict:
        """Pick the fastest primal sweep kernel for this scene by timing real
        frames. Every variant gives bit-identical results, so the frames
        spent timing are ordinary simulation frames. Returns {variant: ms of
        primal sweeps per frame} and keeps the fastest."""
        if self._dirty:
            self._flush()
        cand = tuple(candidates) if candidates is not None else self.PRIMAL_VARIANTS
        times = {}
        for v in cand:
            self._h.set_option(N.OPT_PRIMAL_VARIANT, int(v))
            self._h.profile_enable(True)
            for _ in range(max(1, int(frames))):
                self.step()
            ms, _launches, _bytes = self._h.profile_read(1)
            self._h.profile_enable(False)
            times[int(v)] = float(ms[0]) / max(1, int(frames))
        best = min(times, key=times.get)
        self._h.set_option(N.OPT_PRIMAL_VARIANT, best)
        self.primal_variant = best
        return times

    DUAL_PLACEMENTS = (0, 2)

    def autotune_dual(self, frames: int = 3) -> dict:
        """Pick where the dual update of each iteration runs by timing real
        frames (wall clock, each frame synchronised by a read of the poses):
        in a launch of its own after the sweep (0), or in trailing workgroups
        of the sweep's launch, each group once both of its bodies have moved
        (2). Both give bit-identical results, so the frames spent timing are
        ordinary simulation frames; the trailing placement saves one launch
        per iteration, which pays on small scenes. Returns {placement: ms per
        frame} and keeps the fastest. No counterpart in the reference."""
        import time
        if self._dirty:
            self._flush()
        n = max(1, int(frames))
        times = {}
        for mode in self.DUAL_PLACEMENTS:
            self._h.set_option(N.OPT_FUSE_DUAL, int(mode))
            self.step()          # untimed: captures the graph of this placement
            self.positions()
            t0 = time.perf_counter()
            for _ in range(n):
                self.step()
            self.positions()
            times[int(mode)] = (time.perf_counter() - t0) * 1e3 / n
        best = min(times, key=times.get)
        self._h.set_option(N.OPT_FUSE_DUAL, best)
        self.dual_placement = best
        return times

    # ---- learned warm start (not in the reference) ---------------------------
    def set_learned_warm_start(self, mode: str = "plain", params=None,
                               scope: str = "all") -> None:
        """Initial dual state of box-box and floor contact groups: "plain" =
        the persistent-lambda cache (the reference's warm start), "log" =
        plain, and each substep's contact features and converged dual state
        are kept for learned_samples(), "learned" = the network with `params`
        (layout of avbd_amd.learned) corrects the plain warm start; with
        scope="empty" only groups whose cache holds no normal lambda (new
        contacts, floor corners not pressed) take the network's start."""
        from .learned import MODES
        if mode not in MODES:
            raise ValueError(f"unknown warm-start mode {mode!r}; expected one of {tuple(MODES)}")
        if scope not in ("all", "empty"):
            raise ValueError(f"unknown learned scope {scope!r}; expected 'all' or 'empty'")
        self._warm_start_scope = 1 if scope == "empty" else 0
        if params is not None:
            self._warm_start_params = np.ascontiguousarray(params, np.float32)
        if mode == "learned" and self._warm_start_params is None:
            raise ValueError("the learned warm start needs network parameters")
        self._warm_start_mode = MODES[mode]
        if self._h is not None and not self._dirty:
            self._apply_warm_start()

    def _apply_warm_start(self) -> None:
        if self._warm_start_params is not None:
            self._h.ml_set_params(self._warm_start_params)
        self._h.set_option(N.OPT_LEARNED_SCOPE, self._warm_start_scope)
        self._h.ml_set_mode(self._warm_start_mode)

    def learned_samples(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """The last substep's logged samples (mode "log"): features [n, 24],
        targets [n, 4] and output masks [n]."""
        if self._h is None:
            raise RuntimeError("no substep has run")
        return self._h.ml_read_samples()

    def _check_fixed_cap_overflow(self) -> None:
        """Once-per-frame overflow warnings of the fixed-capacity path
        (solver_6dof.py:1199-1230)."""
        if self._h is None:
            return
        n, p, _ = self._h.read_counters()
        if n > self._gpu_pool_n_capacity and not self._fixed_cap_overflow_warned:
            warnings.warn(
                f"avbd3d fixed-capacity: row pool overflow ({n} reserved > "
                f"{self._gpu_pool_n_capacity} capacity) — contacts were dropped. Increase "
                "solver._gpu_pool_n_dyn_capacity, lower --max-bodies, or disable "
                "gpu_resident/unsafe_fixed_capacity.", RuntimeWarning, stacklevel=2)
            self._fixed_cap_overflow_warned = True
        if self._bp_max_pairs > 0 and p > self._bp_max_pairs and not self._pair_overflow_warned:
            warnings.warn(
                f"avbd3d fixed-capacity: broadphase pair pool overflow ({p} pairs > "
                f"{self._bp_max_pairs} capacity) — contacts were dropped before row emission. "
                "Increase the pair budget or disable gpu_resident/unsafe_fixed_capacity.",
                RuntimeWarning, stacklevel=2)
            self._pair_overflow_warned = True

    def _step_frame(self) -> None:
        """All substeps of a frame in one library call (avbd_step_frame): in
        fixed-capacity mode no host decision depends on a count read back
        within the frame."""
        if self._dirty:
            self._flush()
        n_b = len(self._x)
        if n_b == 0:
            return
        want = (id(self._h), bool(self.hip_graph))
        if self._h_graph_state != want:
            self._h.set_option(N.OPT_GRAPH, int(want[1]))
            self._h_graph_state = want
        sub_dt = self.dt / self.substeps if self.substeps > 1 else self.dt
        rep = self._h.step_frame(self._params(sub_dt), self.substeps, self._color_dirty)
        self._apply_report(rep, n_b)

    def _step_one(self) -> None:
        if self._dirty:
            self._flush()
        n_b = len(self._x)
        if n_b == 0:
            return
        rep = self._h.substep(self._params(self.dt), self._color_dirty)
        self._apply_report(rep, n_b)

    def _apply_report(self, rep: N.Report, n_b: int) -> None:
        fixed = self._fixed_capacity_mode(self.device)
        if rep.row_overflow and not fixed and not self._row_overflow_warned:
            warnings.warn(
                f"avbd3d row pool overflow: reserved {rep.n_active_rows} rows > capacity "
                f"{rep.row_capacity}; growing for next substep. Tune "
                "solver._gpu_pool_n_dyn_capacity if frequent.", RuntimeWarning, stacklevel=3)
            self._row_overflow_warned = True
        self._gpu_pool_n_capacity = rep.new_row_capacity
        self._gpu_pool_n_dyn_capacity = rep.new_row_capacity - self._gpu_pool_n_static
        if self._self_collide and n_b >= 2:
            self._bp_max_pairs = rep.pair_capacity
            self.broadphase_ms = float(rep.broadphase_ms)
            if rep.pair_grew:
                self._graph = None
        _, _, pool_max, _ = self._h.read_capacities()
        self._gpu_pool_pool_max = pool_max
        if rep.recolored:
            self._color_dirty = False
            self._n_active_colors = int(rep.n_active_colors)
            self.num_active_colors = self._n_active_colors
        sig = self._current_graph_signature()
        if sig != self._graph_signature:
            self._graph = None
            self._graph_signature = sig

    def _current_graph_signature(self) -> tuple:
        """Scalars baked into a captured substep (solver_6dof.py:1466-1491)."""
        return (
            int(self.iterations),
            int(self.post_stabilize),
            int(self._max_colors),
            int(self._n_active_colors),
            int(self._primal_group_size),
            int(self._primal_shuffle),
            int(self._primal_fused),
            int(self._gpu_resident),
            int(self._recolor_every_substep),
            float(self.dt),
            float(self.alpha),
            float(self.beta),
            float(self.gamma),
            float(self.max_linear_speed),
            float(self.max_angular_speed),
        )

    def _graph_cuda_supported(self) -> bool:
        return False

    def count_color_conflicts(self) -> int:
        """Active body-body rows whose bodies share a color (solver_6dof.py:846-862)."""
        if self._h is None or not self._gpu_pool_ready:
            return 0
        return self._h.count_color_conflicts()

    # ---- readback (solver_6dof.py:1920-2052) ----------------------------------
    # ---- device-side boundary (not in the reference's API) -------------------
    def step_staged(self) -> None:
        """One frame like step() on the per-substep path, each substep issued
        stage by stage through the library's stage entry points in the
        reference's launch order (solver_6dof.py:1269-1464); the same bits as
        step() with the same settings."""
        if self._dirty:
            self._flush()
        n_b = len(self._x)
        if n_b == 0:
            return
        sub_dt = self.dt / self.substeps if self.substeps > 1 else self.dt
        for _ in range(max(1, self.substeps)):
            params = self._params(sub_dt)
            rep = self._h.stage_contacts(params)
            self._h.stage_coloring(params, self._color_dirty, rep)
            self._h.stage_csr()
            self._h.stage_prelude(params)
            for k in range(params.iterations + params.post_stabilize):
                self._h.stage_sweep(params, k)
            self._h.stage_end(params)
            self._apply_report(rep, n_b)
        if self._fixed_capacity_mode(self.device):
            self._check_fixed_cap_overflow()

    @property
    def stream(self) -> int:
        """The hipStream_t (as an int) the solver launches on."""
        self._flush()
        return self._h.get_stream()

    def set_stream(self, stream) -> None:
        """Launch on `stream` from now on: a torch.cuda.Stream, a raw
        hipStream_t int, or None for the solver's own stream. Work queued on
        the previous stream is waited for first."""
        self._flush()
        raw = getattr(stream, "cuda_stream", stream)
        self._h.set_stream(int(raw) if raw is not None else None)

    def torch_views(self) -> dict:
        """Zero-copy torch tensors over the live device state, valid until
        the next substep or flush: x [n, 3], q [n, 4] (xyzw), v [n, 3],
        omega [n, 3]; per constraint group (static rows' groups, then this
        substep's contacts) lam [g, 3], penalty [g, 3], state_bits [g]
        (bit r: row r active, bit 3 + r: row r sticking), header [g, 4] =
        (kind | (rows - 1) << 2, body_a, body_b, first flat row)."""
        import torch
        self._flush()
        v = self._h.device_views()
        dev = torch.device("cuda", self._device_index)

        def view(ptr, shape, strides, typestr):
            class _Mem:
                __cuda_array_interface__ = {
                    "data": (int(ptr or 0), False), "shape": tuple(shape),
                    "strides": tuple(4 * st for st in strides), "typestr": typestr,
                    "version": 2}
            if 0 in shape or not ptr:
                return torch.empty(shape, dtype=torch.float32 if typestr == "<f4"
                                   else torch.int32, device=dev)
            return torch.as_tensor(_Mem(), device=dev)

        n, g = v.n_bodies, v.n_groups
        bs, gs, hs = v.body_stride, v.group_stride, v.header_stride
        return {
            "x": view(v.x, (n, 3), (bs, 1), "<f4"),
            "q": view(v.q, (n, 4), (bs, 1), "<f4"),
            "v": view(v.v, (n, 3), (bs, 1), "<f4"),
            "omega": view(v.w, (n, 3), (bs, 1), "<f4"),
            "lam": view(v.lam, (g, 3), (gs, 1), "<f4"),
            "penalty": view(v.penalty, (g, 3), (gs, 1), "<f4"),
            "state_bits": view(v.state_bits, (g,), (gs,), "<i4"),
            "header": view(v.header, (g, 4), (hs, 1), "<i4"),
        }

    def positions(self) -> np.ndarray:
        if self.x is None:
            return np.array(self._x, dtype=np.float32).reshape(-1, 3)
        return self.x.numpy().reshape(-1, 3)

    def orientations(self) -> np.ndarray:
        if self.q is None:
            return np.array(self._q, dtype=np.float32).reshape(-1, 4)
        return self.q.numpy().reshape(-1, 4)

    def velocities(self) -> np.ndarray:
        if self.v is None:
            return np.array(self._v, dtype=np.float32).reshape(-1, 3)
        return self.v.numpy().reshape(-1, 3)

    def angular_velocities(self) -> np.ndarray:
        if self.omega is None:
            return np.array(self._omega, dtype=np.float32).reshape(-1, 3)
        return self.omega.numpy().reshape(-1, 3)

    def lambdas(self) -> np.ndarray:
        if self.c_lambda is None:
            return np.zeros(len(self._rows), dtype=np.float32)
        return self.c_lambda.numpy()

    def active(self) -> np.ndarray:
        if self.c_active is None:
            return np.ones(len(self._rows), dtype=np.int32)
        return self.c_active.numpy()

    def read_state_batched(self, include_rows: bool = True) -> dict[str, np.ndarray]:
        """Bodies packed as 10 floats and live rows as 3 floats
        (solver_6dof.py:1952-2052, kernels_6dof.py:2154-2177, 2540-2555)."""
        n_b = len(self._x)
        n_c_static = len(self._rows)
        if not include_rows:
            n_c = 0
        elif self.n_active_rows is not None:
            n_c = int(self.n_active_rows.numpy()[0])
        else:
            n_c = n_c_static
        if n_b == 0 or self.x is None:
            return {
                "positions": np.array(self._x, dtype=np.float32).reshape(-1, 3),
                "orientations": np.array(self._q, dtype=np.float32).reshape(-1, 4),
                "angular_velocities": np.array(self._omega, dtype=np.float32).reshape(-1, 3),
                "lambdas": np.zeros(n_c, dtype=np.float32),
                "active": np.ones(n_c, dtype=np.int32),
                "was_static": np.zeros(n_c, dtype=np.int32),
                "c_type": np.zeros(n_c, dtype=np.int32),
                "n_rows": n_c,
            }
        n_rows = min(n_c, self._gpu_pool_n_capacity) if n_c > 0 else 0
        bod, row = self._h.pack_state(self._n_b_flushed, n_rows)
        if n_rows > 0:
            lam = row[:, 0].astype(np.float32, copy=True)
            act = row[:, 1].astype(np.int32)
            ws_type = row[:, 2].astype(np.int32)
            was = (ws_type // 16).astype(np.int32)
            ctype = (ws_type % 16).astype(np.int32)
        else:
            lam = np.zeros(0, np.float32)
            act = was = ctype = np.zeros(0, np.int32)
        return {
            "positions": bod[:, 0:3].astype(np.float32, copy=True),
            "orientations": bod[:, 3:7].astype(np.float32, copy=True),
            "angular_velocities": bod[:, 7:10].astype(np.float32, copy=True),
            "lambdas": lam,
            "active": act,
            "was_static": was,
            "c_type": ctype,
            "n_rows": n_c,
        }
