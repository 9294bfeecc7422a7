"""ctypes binding of libavbd_hip.so (the C ABI declared in include/avbd.h).

The structs below mirror ``avbd_params``, ``avbd_scene`` and ``avbd_report``
field for field. ``NativeHandle`` wraps one solver handle; every call checks
the status code and raises ``AvbdError`` with the library's message.

Loading fails loudly: there is no CPU fallback for the solver.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("AVBD_HIP_LIB", str(_PKG / "libavbd_hip.so")))

F32P = C.POINTER(C.c_float)
I32P = C.POINTER(C.c_int)
F64P = C.POINTER(C.c_double)
I64P = C.POINTER(C.c_longlong)

COLOR_JACOBI = 0
COLOR_JONES_PLASSMANN = 1
OPT_PRIMAL_VARIANT = 0
OPT_FUSE_DUAL = 1
OPT_LEARNED_SCOPE = 2
OPT_FLOW_GIVE_UP_TEST = 3
OPT_GRAPH = 4
OPT_BROAD_LANES = 5


class AvbdError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [
        ("dt", C.c_float),
        ("iterations", C.c_int),
        ("gravity", C.c_float * 3),
        ("alpha", C.c_float),
        ("beta", C.c_float),
        ("gamma", C.c_float),
        ("post_stabilize", C.c_int),
        ("max_linear_speed", C.c_float),
        ("max_angular_speed", C.c_float),
        ("friction_static_mult", C.c_float),
        ("coloring_mode", C.c_int),
        ("self_collide", C.c_int),
        ("self_friction", C.c_float),
        ("fixed_capacity", C.c_int),
        ("margin", C.c_float),
    ]


class Scene(C.Structure):
    _fields_ = [
        ("n_bodies", C.c_int),
        ("x", F32P), ("q", F32P), ("v", F32P), ("w", F32P), ("prev_v", F32P), ("prev_w", F32P),
        ("mass", F32P), ("inertia_local", F32P), ("half_extents", F32P), ("friction", F32P),
        ("priority", F32P),
        ("n_static_rows", C.c_int),
        ("type", I32P), ("body_a", I32P), ("body_b", I32P), ("sibling", I32P), ("partner", I32P),
        ("anchor", F32P), ("off_a", F32P), ("off_b", F32P), ("stiffness", F32P), ("fmin", F32P),
        ("fmax", F32P), ("fracture", F32P), ("friction_d", F32P), ("friction_s", F32P),
        ("lambda_", F32P), ("penalty", F32P),
        ("active", I32P), ("was_static", I32P),
        ("max_pairs", C.c_int),
        ("n_dyn_capacity", C.c_int),
        ("hash_cap", C.c_int),
        ("pool_max", C.c_int),
    ]


class Report(C.Structure):
    _fields_ = [
        ("recolored", C.c_int),
        ("n_active_colors", C.c_int),
        ("n_active_rows", C.c_int),
        ("row_capacity", C.c_int),
        ("row_overflow", C.c_int),
        ("new_row_capacity", C.c_int),
        ("pair_count", C.c_int),
        ("pair_capacity", C.c_int),
        ("pair_grew", C.c_int),
        ("pool_count", C.c_int),
        ("n_contacts", C.c_int),
        ("color_conflicts", C.c_int),
        ("broadphase_ms", C.c_float),
    ]


class Views(C.Structure):
    """avbd_views: device addresses of the live state."""
    _fields_ = [
        ("n_bodies", C.c_int),
        ("x", C.c_void_p), ("q", C.c_void_p), ("v", C.c_void_p), ("w", C.c_void_p),
        ("body_stride", C.c_int),
        ("n_groups", C.c_int),
        ("lam", C.c_void_p), ("penalty", C.c_void_p), ("state_bits", C.c_void_p),
        ("group_stride", C.c_int),
        ("header", C.c_void_p),
        ("header_stride", C.c_int),
    ]


class PParams(C.Structure):
    """avbd_p_params: one particle Solver.step's settings."""
    _fields_ = [
        ("dt", C.c_float),
        ("iterations", C.c_int),
        ("gravity", C.c_float * 3),
        ("alpha", C.c_float),
        ("beta", C.c_float),
        ("gamma", C.c_float),
        ("post_stabilize", C.c_int),
        ("max_speed", C.c_float),
    ]


class PScene(C.Structure):
    """avbd_p_scene: the particle solver's flushed scene (pointers stay owned
    by the caller for the duration of avbd_p_flush)."""
    _fields_ = [
        ("n_bodies", C.c_int),
        ("x", F32P), ("v", F32P), ("prev_v", F32P), ("mass", F32P),
        ("color", I32P), ("n_colors", C.c_int),
        ("n_rows", C.c_int),
        ("type", I32P), ("body_a", I32P), ("body_b", I32P), ("anchor", F32P),
        ("rest", F32P), ("stiffness", F32P), ("fmin", F32P), ("fmax", F32P),
        ("fracture", F32P), ("sibling", I32P), ("friction", F32P),
        ("lambda_", F32P), ("penalty", F32P), ("active", I32P),
        ("con_start", I32P), ("con_idx", I32P),
        ("n_tets", C.c_int),
        ("tet_v", I32P), ("tet_rest_vol", F32P), ("tet_stiffness", F32P),
        ("tet_fracture", F32P), ("tet_lambda", F32P), ("tet_penalty", F32P),
        ("tet_active", I32P), ("tet_start", I32P), ("tet_idx", I32P),
        ("off_a", F32P), ("off_b", F32P),
    ]


def fptr(a: np.ndarray):
    return a.ctypes.data_as(F32P)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(I32P)


# Every exported symbol and its ctypes signature (argtypes, restype).
SIGNATURES = {
    "last_error": ([], C.c_char_p),
    "device_count": ([I32P], C.c_int),
    "create": ([C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "destroy": ([C.c_void_p], C.c_int),
    "flush": ([C.c_void_p, C.POINTER(Scene)], C.c_int),
    "substep": ([C.c_void_p, C.POINTER(Params), C.c_int, C.POINTER(Report)], C.c_int),
    "substep_to_residual": ([C.c_void_p, C.POINTER(Params), C.c_int, C.c_float,
                             C.POINTER(C.c_int), C.POINTER(C.c_float), C.POINTER(Report)], C.c_int),
    "step_frame": ([C.c_void_p, C.POINTER(Params), C.c_int, C.c_int, C.POINTER(Report)], C.c_int),
    "sync_count": ([C.c_void_p, I64P, I64P], C.c_int),
    "stage_contacts": ([C.c_void_p, C.POINTER(Params), C.POINTER(Report)], C.c_int),
    "stage_coloring": ([C.c_void_p, C.POINTER(Params), C.c_int, C.POINTER(Report)], C.c_int),
    "stage_csr": ([C.c_void_p], C.c_int),
    "stage_prelude": ([C.c_void_p, C.POINTER(Params)], C.c_int),
    "stage_sweep": ([C.c_void_p, C.POINTER(Params), C.c_int], C.c_int),
    "stage_end": ([C.c_void_p, C.POINTER(Params)], C.c_int),
    "get_stream": ([C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    "set_stream": ([C.c_void_p, C.c_void_p], C.c_int),
    "device_views": ([C.c_void_p, C.POINTER(Views)], C.c_int),
    "read_bodies": ([C.c_void_p, F32P, F32P, F32P, F32P, F32P, F32P], C.c_int),
    "write_bodies": ([C.c_void_p, C.c_int, C.c_int, C.c_int, F32P], C.c_int),
    "read_rows": ([C.c_void_p, C.c_int, F32P, F32P, I32P, I32P, I32P, I32P, I32P, F32P], C.c_int),
    "write_static_rows": ([C.c_void_p, C.c_int, F32P, F32P, I32P, I32P], C.c_int),
    "read_colors": ([C.c_void_p, I32P], C.c_int),
    "write_colors": ([C.c_void_p, I32P, C.c_int], C.c_int),
    "read_counters": ([C.c_void_p, I32P, I32P, I32P], C.c_int),
    "count_color_conflicts": ([C.c_void_p, I32P], C.c_int),
    "pack_state": ([C.c_void_p, F32P, C.c_int, F32P], C.c_int),
    "export_hash": ([C.c_void_p, C.c_int, I32P, F32P, I32P], C.c_int),
    "import_hash": ([C.c_void_p, C.c_int, I32P, F32P], C.c_int),
    "read_capacities": ([C.c_void_p, I32P, I32P, I32P, I32P], C.c_int),
    "read_pairs": ([C.c_void_p, C.c_int, I32P, I32P, I32P, I32P, F32P, F32P, I32P], C.c_int),
    "read_manifold": ([C.c_void_p, C.c_int, I32P, I32P, F32P, F32P, F32P, F32P, F32P], C.c_int),
    "set_option": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "profile_enable": ([C.c_void_p, C.c_int], C.c_int),
    "profile_read": ([C.c_void_p, C.c_int, F64P, I64P, F64P], C.c_int),
    # batched scenes (no counterpart in the reference)
    "set_scenes": ([C.c_void_p, C.c_int, I32P, C.c_float], C.c_int),
    # learned warm start (no counterpart in the reference)
    "mlp_num_params": ([], C.c_int),
    "ml_set_mode": ([C.c_void_p, C.c_int], C.c_int),
    "ml_set_params": ([C.c_void_p, F32P, C.c_int], C.c_int),
    "ml_read_samples": ([C.c_void_p, C.c_int, F32P, F32P, I32P, I32P], C.c_int),
    "mlp_forward": ([C.c_int, C.c_int, F32P, F32P, F32P], C.c_int),
    "mlp_grad": ([C.c_int, C.c_int, F32P, F32P, I32P, F32P, C.c_float, F32P, F32P], C.c_int),
    "mlp_grad_noalloc": ([C.c_int, C.c_int, F32P, F32P, I32P, F32P, C.c_float, F32P, F32P],
                         C.c_int),
    "mlp_train": ([C.c_int, C.c_int, F32P, F32P, I32P, F32P, C.c_int, C.c_int, C.c_float,
                   C.c_float, C.c_uint, F32P], C.c_int),
    # 3-DOF particle solver (Solver, src/avbd3d/solver.py)
    "p_last_error": ([], C.c_char_p),
    "p_create": ([C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "p_destroy": ([C.c_void_p], C.c_int),
    "p_flush": ([C.c_void_p, C.POINTER(PScene)], C.c_int),
    "p_step": ([C.c_void_p, C.POINTER(PParams)], C.c_int),
    "p_read": ([C.c_void_p, F32P, F32P, F32P, F32P, F32P, I32P, F32P, F32P, I32P], C.c_int),
    "p_write_bodies": ([C.c_void_p, C.c_int, C.c_int, C.c_int, F32P], C.c_int),
    "p_write_rows": ([C.c_void_p, C.c_int, C.c_int, F32P, F32P], C.c_int),
    "p_set_option": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "p_profile_enable": ([C.c_void_p, C.c_int], C.c_int),
    "p_profile_read": ([C.c_void_p, F64P, I64P, F64P], C.c_int),
}

# entry points of batched scenes and the learned warm start (the checker has none)
ML_NAMES = ("set_scenes", "mlp_num_params", "ml_set_mode", "ml_set_params", "ml_read_samples", "mlp_forward",
            "mlp_grad", "mlp_grad_noalloc", "mlp_train")

# entry points of the particle solver; the checker has all but the last
# three (graph replay and launch timing are properties of the GPU path)
P_NAMES = tuple(k for k in SIGNATURES if k.startswith("p_"))
P_GPU_ONLY = ("p_set_option", "p_profile_enable", "p_profile_read")
P_OPT_GRAPH = 0
P_OPT_PRIMAL_VARIANT = 1

_LIB = None


def bind(lib: C.CDLL, prefix: str, names=None) -> C.CDLL:
    for name, (argt, rest) in SIGNATURES.items():
        if names is not None and name not in names:
            continue
        fn = getattr(lib, prefix + name)
        fn.argtypes = argt
        fn.restype = rest
    return lib


def load_library() -> C.CDLL:
    """Load libavbd_hip.so once; raise with build instructions if it is absent."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"avbd_amd: native library {LIB_PATH} is missing. Build it with "
                "`python avbd_amd/build.py` (or __graft_entry__.build()). There is no "
                "CPU fallback for the solver.")
        # torch first: the library then binds to the HIP runtime torch
        # loaded, so device addresses and streams are shared with torch
        # (with the library loaded first, torch would load a second runtime
        # that cannot open the GPU beside it)
        import torch  # noqa: F401
        _LIB = bind(C.CDLL(str(LIB_PATH)), "avbd_")
    return _LIB


def device_count() -> int:
    lib = load_library()
    n = C.c_int(0)
    lib.avbd_device_count(C.byref(n))
    return int(n.value)


class NativeHandle:
    """One solver handle of libavbd_hip.so; all state lives on the GPU."""

    prefix = "avbd_"

    def __init__(self, device: int = 0, lib: C.CDLL | None = None):
        self.lib = lib if lib is not None else load_library()
        self.h = C.c_void_p()
        self._create(device)

    def _create(self, device: int) -> None:
        self._check(self._fn("create")(int(device), C.byref(self.h)))

    def _fn(self, name: str):
        return getattr(self.lib, self.prefix + name)

    def _check(self, rc: int) -> None:
        if rc != 0:
            msg = self._fn("last_error")()
            raise AvbdError(f"{self.prefix}: {msg.decode() if msg else 'error'} (code {rc})")

    def close(self) -> None:
        if self.h:
            self._fn("destroy")(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- calls ----
    def flush(self, scene: Scene) -> None:
        self._check(self._fn("flush")(self.h, C.byref(scene)))

    def substep(self, params: Params, color_dirty: bool) -> Report:
        rep = Report()
        self._check(self._fn("substep")(self.h, C.byref(params), int(bool(color_dirty)), C.byref(rep)))
        return rep

    def substep_to_residual(self, params: Params, color_dirty: bool, tol: float):
        """One substep iterated until the dual residual is <= tol (params.iterations
        is the cap). Returns (report, iterations used, final residual)."""
        rep = Report()
        it = C.c_int(0)
        r = C.c_float(0.0)
        self._check(self._fn("substep_to_residual")(self.h, C.byref(params), int(bool(color_dirty)),
                                                   float(tol), C.byref(it), C.byref(r),
                                                   C.byref(rep)))
        return rep, int(it.value), float(r.value)

    def step_frame(self, params: Params, n_substeps: int, color_dirty: bool) -> Report:
        """One frame of n_substeps substeps (avbd_step_frame)."""
        rep = Report()
        self._check(self._fn("step_frame")(self.h, C.byref(params), int(n_substeps),
                                           int(bool(color_dirty)), C.byref(rep)))
        return rep

    def sync_count(self) -> tuple[int, int]:
        """(counter synchronisations, substep graph launches) since creation."""
        a, b = C.c_longlong(0), C.c_longlong(0)
        self._check(self._fn("sync_count")(self.h, C.byref(a), C.byref(b)))
        return int(a.value), int(b.value)

    # ---- one substep stage by stage (avbd_stage_*) ----
    def stage_contacts(self, params: Params) -> Report:
        rep = Report()
        self._check(self._fn("stage_contacts")(self.h, C.byref(params), C.byref(rep)))
        return rep

    def stage_coloring(self, params: Params, color_dirty: bool, rep: Report) -> Report:
        self._check(self._fn("stage_coloring")(self.h, C.byref(params), int(bool(color_dirty)),
                                               C.byref(rep)))
        return rep

    def stage_csr(self) -> None:
        self._check(self._fn("stage_csr")(self.h))

    def stage_prelude(self, params: Params) -> None:
        self._check(self._fn("stage_prelude")(self.h, C.byref(params)))

    def stage_sweep(self, params: Params, sweep: int) -> None:
        self._check(self._fn("stage_sweep")(self.h, C.byref(params), int(sweep)))

    def stage_end(self, params: Params) -> None:
        self._check(self._fn("stage_end")(self.h, C.byref(params)))

    # ---- streams and device views ----
    def get_stream(self) -> int:
        st = C.c_void_p()
        self._check(self._fn("get_stream")(self.h, C.byref(st)))
        return int(st.value or 0)

    def set_stream(self, stream: int | None) -> None:
        self._check(self._fn("set_stream")(self.h, C.c_void_p(stream or None)))

    def device_views(self) -> Views:
        v = Views()
        self._check(self._fn("device_views")(self.h, C.byref(v)))
        return v

    def read_bodies(self, n: int) -> dict[str, np.ndarray]:
        out = {k: np.zeros((n, 3), np.float32) for k in ("x", "v", "w", "prev_v", "prev_w")}
        out["q"] = np.zeros((n, 4), np.float32)
        self._check(self._fn("read_bodies")(self.h, fptr(out["x"]), fptr(out["q"]), fptr(out["v"]),
                                            fptr(out["w"]), fptr(out["prev_v"]), fptr(out["prev_w"])))
        return out

    def write_bodies(self, field: int, first: int, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, dtype=np.float32)
        n = v.shape[0] if v.ndim > 1 else 1
        self._check(self._fn("write_bodies")(self.h, int(field), int(first), int(n), fptr(v)))

    def read_rows(self, n: int) -> dict[str, np.ndarray]:
        out = {
            "lambda": np.zeros(n, np.float32), "penalty": np.zeros(n, np.float32),
            "active": np.zeros(n, np.int32), "was_static": np.zeros(n, np.int32),
            "type": np.zeros(n, np.int32), "body_a": np.zeros(n, np.int32),
            "body_b": np.zeros(n, np.int32), "alpha_c0": np.zeros(n, np.float32),
        }
        self._check(self._fn("read_rows")(
            self.h, int(n), fptr(out["lambda"]), fptr(out["penalty"]), iptr(out["active"]),
            iptr(out["was_static"]), iptr(out["type"]), iptr(out["body_a"]), iptr(out["body_b"]),
            fptr(out["alpha_c0"])))
        return out

    def write_static_rows(self, lam, pen, act, was) -> None:
        lam = np.ascontiguousarray(lam, np.float32)
        pen = np.ascontiguousarray(pen, np.float32)
        act = np.ascontiguousarray(act, np.int32)
        was = np.ascontiguousarray(was, np.int32)
        self._check(self._fn("write_static_rows")(self.h, int(lam.shape[0]), fptr(lam), fptr(pen),
                                                  iptr(act), iptr(was)))

    def read_colors(self, n: int) -> np.ndarray:
        c = np.zeros(max(n, 1), np.int32)
        self._check(self._fn("read_colors")(self.h, iptr(c)))
        return c[:n]

    def write_colors(self, colors: np.ndarray, n_active_colors: int) -> None:
        c = np.ascontiguousarray(colors, np.int32)
        self._check(self._fn("write_colors")(self.h, iptr(c), int(n_active_colors)))

    def read_counters(self) -> tuple[int, int, int]:
        a, b, c = C.c_int(0), C.c_int(0), C.c_int(0)
        self._check(self._fn("read_counters")(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return int(a.value), int(b.value), int(c.value)

    def count_color_conflicts(self) -> int:
        n = C.c_int(0)
        self._check(self._fn("count_color_conflicts")(self.h, C.byref(n)))
        return int(n.value)

    def pack_state(self, nb: int, n_rows: int) -> tuple[np.ndarray, np.ndarray]:
        bod = np.zeros(max(nb, 1) * 10, np.float32)
        rows = np.zeros(max(n_rows, 1) * 3, np.float32)
        self._check(self._fn("pack_state")(self.h, fptr(bod), int(n_rows), fptr(rows)))
        return bod[: nb * 10].reshape(nb, 10), rows[: n_rows * 3].reshape(n_rows, 3)

    def export_hash(self) -> tuple[np.ndarray, np.ndarray]:
        n = C.c_int(0)
        _, _, _, hc = self.read_capacities()
        keys = np.zeros(max(hc, 1), np.int32)
        st = np.zeros((max(hc, 1), 8), np.float32)
        self._check(self._fn("export_hash")(self.h, int(hc), iptr(keys), fptr(st), C.byref(n)))
        m = int(n.value)
        return keys[:m].copy(), st[:m].copy()

    def import_hash(self, keys: np.ndarray, state: np.ndarray) -> None:
        k = np.ascontiguousarray(keys, np.int32)
        s = np.ascontiguousarray(state, np.float32).reshape(-1, 8)
        self._check(self._fn("import_hash")(self.h, int(k.shape[0]), iptr(k), fptr(s)))

    def read_capacities(self) -> tuple[int, int, int, int]:
        v = [C.c_int(0) for _ in range(4)]
        self._check(self._fn("read_capacities")(self.h, *[C.byref(x) for x in v]))
        return tuple(int(x.value) for x in v)

    def read_pairs(self) -> dict[str, np.ndarray]:
        mp, _, _, _ = self.read_capacities()
        cap = max(mp, 1)
        out = {"a": np.zeros(cap, np.int32), "b": np.zeros(cap, np.int32),
               "overlap": np.zeros(cap, np.int32), "sat_idx": np.zeros(cap, np.int32),
               "n_hat": np.zeros((cap, 3), np.float32), "depth": np.zeros(cap, np.float32)}
        cnt = C.c_int(0)
        self._check(self._fn("read_pairs")(self.h, cap, iptr(out["a"]), iptr(out["b"]),
                                           iptr(out["overlap"]), iptr(out["sat_idx"]),
                                           fptr(out["n_hat"]), fptr(out["depth"]), C.byref(cnt)))
        k = min(int(cnt.value), mp)
        res = {key: val[:k] for key, val in out.items()}
        res["count"] = int(cnt.value)
        return res

    def read_manifold(self) -> dict[str, np.ndarray]:
        mp, _, _, _ = self.read_capacities()
        cap = max(mp, 1)
        out = {"count": np.zeros(cap, np.int32), "ref_is_a": np.zeros(cap, np.int32),
               "n_hat": np.zeros((cap, 3), np.float32), "t_hat": np.zeros((cap, 3), np.float32),
               "b_hat": np.zeros((cap, 3), np.float32),
               "off_ref": np.zeros((cap, 4, 3), np.float32),
               "off_inc": np.zeros((cap, 4, 3), np.float32)}
        self._check(self._fn("read_manifold")(self.h, cap, iptr(out["count"]), iptr(out["ref_is_a"]),
                                              fptr(out["n_hat"]), fptr(out["t_hat"]),
                                              fptr(out["b_hat"]), fptr(out["off_ref"]),
                                              fptr(out["off_inc"])))
        _, pc, _ = self.read_counters()
        k = min(pc, mp)
        return {key: val[:k] for key, val in out.items()}

    def set_option(self, option: int, value: int) -> None:
        self._check(self._fn("set_option")(self.h, int(option), int(value)))

    def profile_enable(self, on: bool) -> None:
        self._check(self._fn("profile_enable")(self.h, int(bool(on))))

    def set_scenes(self, first_body: np.ndarray, pitch: float) -> None:
        f = np.ascontiguousarray(first_body, np.int32)
        self._check(self._fn("set_scenes")(self.h, int(f.size - 1), iptr(f), float(pitch)))

    # ---- learned warm start ----
    def ml_set_mode(self, mode: int) -> None:
        self._check(self._fn("ml_set_mode")(self.h, int(mode)))

    def ml_set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, np.float32)
        self._check(self._fn("ml_set_params")(self.h, fptr(p), int(p.size)))

    def ml_read_samples(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        n = C.c_int(0)
        self._check(self._fn("ml_read_samples")(self.h, 0, None, None, None, C.byref(n)))
        m = int(n.value)
        feat = np.zeros((m, 24), np.float32)
        targ = np.zeros((m, 4), np.float32)
        mask = np.zeros(m, np.int32)
        if m:
            self._check(self._fn("ml_read_samples")(self.h, m, fptr(feat), fptr(targ), iptr(mask),
                                                    C.byref(n)))
        return feat, targ, mask

    def profile_read(self, n: int = 7):
        ms = np.zeros(n, np.float64)
        launches = np.zeros(n, np.int64)
        nbytes = np.zeros(n, np.float64)
        self._check(self._fn("profile_read")(self.h, n, ms.ctypes.data_as(F64P),
                                             launches.ctypes.data_as(I64P),
                                             nbytes.ctypes.data_as(F64P)))
        return ms, launches, nbytes


class ParticleHandle:
    """One particle-solver handle (avbd_p_*); all state lives on the GPU."""

    prefix = "avbd_"

    def __init__(self, device: int = 0, lib: C.CDLL | None = None):
        self.lib = lib if lib is not None else load_library()
        self.h = C.c_void_p()
        self._check(self._fn("p_create")(int(device), C.byref(self.h)))
        self.nb = self.nc = self.nt = 0

    def _fn(self, name: str):
        return getattr(self.lib, self.prefix + name)

    def _check(self, rc: int) -> None:
        if rc != 0:
            msg = self._fn("p_last_error")()
            raise AvbdError(f"{self.prefix}p: {msg.decode() if msg else 'error'} (code {rc})")

    def close(self) -> None:
        if self.h:
            self._fn("p_destroy")(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def flush(self, arrays: dict) -> None:
        """Upload a scene given as a dict of numpy arrays named like the
        avbd_p_scene fields ("lambda" for lambda_)."""
        keep = {}
        sc = PScene()
        for name, ctype in PScene._fields_:
            key = "lambda" if name == "lambda_" else name
            val = arrays[key]
            if ctype is C.c_int:
                setattr(sc, name, int(val))
            elif ctype is F32P:
                a = np.ascontiguousarray(val, np.float32)
                keep[name] = a
                setattr(sc, name, fptr(a))
            else:
                a = np.ascontiguousarray(val, np.int32)
                keep[name] = a
                setattr(sc, name, iptr(a))
        self._check(self._fn("p_flush")(self.h, C.byref(sc)))
        self.nb, self.nc, self.nt = sc.n_bodies, sc.n_rows, sc.n_tets

    def step(self, params: PParams) -> None:
        self._check(self._fn("p_step")(self.h, C.byref(params)))

    _READ = (("x", 3, np.float32), ("v", 3, np.float32), ("prev_v", 3, np.float32),
             ("lambda", 0, np.float32), ("penalty", 0, np.float32), ("active", 0, np.int32),
             ("tet_lambda", 0, np.float32), ("tet_penalty", 0, np.float32),
             ("tet_active", 0, np.int32))

    def read(self, keys=None) -> dict[str, np.ndarray]:
        """Device state; `keys` limits it to some of x, v, prev_v, lambda,
        penalty, active, tet_lambda, tet_penalty, tet_active."""
        out, args = {}, []
        for key, width, dt in self._READ:
            n = self.nb if width else (self.nt if key.startswith("tet") else self.nc)
            if keys is not None and key not in keys:
                args.append(None)
                continue
            a = np.zeros((n, 3) if width else n, dt)
            out[key] = a
            args.append(fptr(a) if dt == np.float32 else iptr(a))
        self._check(self._fn("p_read")(self.h, *args))
        return out

    def write_bodies(self, field: int, first: int, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, np.float32).reshape(-1, 3)
        self._check(self._fn("p_write_bodies")(self.h, int(field), int(first), int(v.shape[0]),
                                               fptr(v)))

    def write_rows(self, first: int, n: int, anchor=None, rest=None) -> None:
        a = None if anchor is None else np.ascontiguousarray(anchor, np.float32).reshape(-1, 3)
        r = None if rest is None else np.ascontiguousarray(rest, np.float32).reshape(-1)
        self._check(self._fn("p_write_rows")(self.h, int(first), int(n),
                                             None if a is None else fptr(a),
                                             None if r is None else fptr(r)))

    def set_graph(self, on: bool) -> None:
        self.set_option(P_OPT_GRAPH, int(bool(on)))

    def set_option(self, option: int, value: int) -> None:
        self._check(self._fn("p_set_option")(self.h, int(option), int(value)))

    def profile_enable(self, on: bool) -> None:
        self._check(self._fn("p_profile_enable")(self.h, int(bool(on))))

    def profile_read(self):
        ms = np.zeros(3, np.float64)
        n = np.zeros(3, np.int64)
        b = np.zeros(3, np.float64)
        self._check(self._fn("p_profile_read")(self.h, ms.ctypes.data_as(F64P),
                                               n.ctypes.data_as(I64P), b.ctypes.data_as(F64P)))
        return ms, n, b
