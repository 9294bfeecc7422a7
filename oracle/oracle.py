"""Python driver of the CPU checker (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the avbd_amd package.

``OracleHandle`` exposes the same methods as ``avbd_amd._native.NativeHandle``
so a scene can be stepped through both side by side.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from avbd_amd import _native as N

HERE = Path(__file__).resolve().parent
LIB = HERE / "build" / "liboracle.so"
_L = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def load() -> C.CDLL:
    global _L
    if _L is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        N.bind(lib, "orc_", names=[k for k in N.SIGNATURES if k not in ("create", "device_count",
                                                                        "pack_state",
                                                                        "set_option",
                                                                        "profile_enable",
                                                                        "profile_read",
                                                                        "step_frame",
                                                                        "sync_count",
                                                                        "get_stream",
                                                                        "set_stream",
                                                                        "device_views")
                                   and not k.startswith("stage_")
                                   and k not in N.ML_NAMES and k not in N.P_GPU_ONLY])
        lib.orc_create.argtypes = [C.POINTER(C.c_void_p)]
        lib.orc_create.restype = C.c_int
        lib.orc_set_capacities.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        lib.orc_set_capacities.restype = C.c_int
        F = np.ctypeslib.ndpointer(dtype=np.float32, flags="C")
        lib.orc_fp_sincos.argtypes = [C.c_int, F, F, F]
        lib.orc_fp_atan2.argtypes = [C.c_int, F, F, F]
        lib.orc_fp_geom_stiffness.argtypes = [F, F, F]
        lib.orc_fp_quat.argtypes = [C.c_int, F, F, F, F]
        lib.orc_fp_rotvec.argtypes = [C.c_int, F, F, F]
        lib.orc_fp_inverse.argtypes = [C.c_int, F, F]
        lib.orc_fp_hash.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]
        lib.orc_fp_hash.restype = C.c_int
        lib.orc_probe_pair.argtypes = [F, F, F, F, F, F, C.c_float, C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), F, F, C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), F, F]
        lib.orc_probe_pair.restype = C.c_int
        lib.orc_p_probe_row.argtypes = [C.c_int, C.c_int, F, F, C.c_int, F, C.c_float, F, F,
                                        C.POINTER(C.c_float), F]
        lib.orc_p_probe_row.restype = None
        lib.orc_p_probe_tet.argtypes = [F, C.c_int, C.c_float, C.c_float, C.c_float, F, F]
        lib.orc_p_probe_tet.restype = None
        _L = lib
    return _L


class OracleHandle(N.NativeHandle):
    """The CPU restatement behind the NativeHandle method set."""

    prefix = "orc_"

    def __init__(self, device: int = 0):
        super().__init__(device, lib=load())

    def _create(self, device: int) -> None:
        self._check(self.lib.orc_create(C.byref(self.h)))

    def set_capacities(self, max_pairs: int, row_cap: int, pool_max: int) -> None:
        self._check(self.lib.orc_set_capacities(self.h, int(max_pairs), int(row_cap),
                                                int(pool_max)))

    def pack_state(self, nb, n_rows):
        b = self.read_bodies(nb)
        bod = np.concatenate([b["x"], b["q"], b["w"]], axis=1)
        r = self.read_rows(n_rows)
        rows = np.stack([r["lambda"], r["active"].astype(np.float32),
                         (r["was_static"] * 16 + r["type"]).astype(np.float32)], axis=1)
        return bod, rows

    def profile_enable(self, on):
        pass

    def set_option(self, option, value):
        pass

    def step_frame(self, params, n_substeps, color_dirty):
        """A frame as consecutive substeps, the first one taking the dirty
        coloring flag (the library's avbd_step_frame contract)."""
        rep, recolored = None, 0
        for k in range(max(1, int(n_substeps))):
            rep = self.substep(params, bool(color_dirty) and k == 0)
            recolored |= rep.recolored
        rep.recolored = recolored
        return rep

    def sync_count(self):
        return 0, 0


def oracle_solver_class():
    """Solver6DOF whose handle is the CPU restatement (for side-by-side runs)."""
    from avbd_amd.solver_6dof import Solver6DOF

    class OracleSolver6DOF(Solver6DOF):
        def _make_handle(self):
            return OracleHandle()

    return OracleSolver6DOF


class ParticleOracleHandle(N.ParticleHandle):
    """The CPU restatement of the particle step behind the ParticleHandle
    method set (orc_p_*)."""

    prefix = "orc_"

    def __init__(self, device: int = 0):
        super().__init__(device, lib=load())


def particle_solver_class():
    """avbd_amd.solver.Solver whose handle is the CPU restatement."""
    from avbd_amd.solver import Solver

    class OracleSolver(Solver):
        def _make_handle(self):
            return ParticleOracleHandle()

    return OracleSolver


def probe_pair(xa, qa, ea, xb, qb, eb, margin=0.005):
    """SAT + manifold of one explicit box pair through the restatement."""
    lib = load()
    f = lambda v: np.ascontiguousarray(v, np.float32)
    ov, sat, cnt, refa = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
    n = np.zeros(3, np.float32)
    depth = np.zeros(1, np.float32)
    offr = np.zeros((4, 3), np.float32)
    offi = np.zeros((4, 3), np.float32)
    lib.orc_probe_pair(f(xa), f(qa), f(ea), f(xb), f(qb), f(eb), float(margin), C.byref(ov),
                       C.byref(sat), n, depth, C.byref(cnt), C.byref(refa), offr, offi)
    return {"overlap": ov.value, "sat_idx": sat.value, "n_hat": n, "depth": float(depth[0]),
            "count": cnt.value, "ref_is_a": refa.value, "off_ref": offr, "off_inc": offi}


def probe_particle_row(t, i_is_a, xi, xo, has_b, anchor, rest, oa=(0, 0, 0), ob=(0, 0, 0)):
    """(C, dC/dx_i) of one particle-solver row through the restatement."""
    lib = load()
    f = lambda v: np.ascontiguousarray(v, np.float32)
    c = C.c_float(0.0)
    J = np.zeros(3, np.float32)
    lib.orc_p_probe_row(int(t), int(bool(i_is_a)), f(xi), f(xo), int(bool(has_b)), f(anchor),
                        float(rest), f(oa), f(ob), C.byref(c), J)
    return float(c.value), J


def probe_tet(p, slot, V0, ac0, pen):
    """One tet's (lhs, rhs) contribution to vertex `slot` through the restatement."""
    lib = load()
    lhs = np.zeros(9, np.float32)
    rhs = np.zeros(3, np.float32)
    lib.orc_p_probe_tet(np.ascontiguousarray(p, np.float32).reshape(-1), int(slot), float(V0),
                        float(ac0), float(pen), lhs, rhs)
    return lhs.reshape(3, 3), rhs
