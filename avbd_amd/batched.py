"""Many independent scenes stepped as one: the batched configuration.

`BatchedSolver6DOF(scenes)` takes Solver6DOF scenes built with the usual
API (same solver settings, not yet stepped) and runs them as one solver on
one GPU. The bodies of scene k are bodies [first[k], first[k+1]) of the
batch, and each scene's static rows follow its bodies. One launch per
kernel covers every scene: the primal dataflow sweep is one launch per
sweep whatever the number of scenes.

Three things keep the scenes apart, and with them each scene steps exactly
as it would alone, bit for bit (avbd_set_scenes):
  * the broad phase pairs bodies of one scene only; its tree is built over
    each scene's boxes moved to its own grid cell;
  * the recolor gate recolors only the scenes with a color conflict;
  * per-scene priorities and capacities are the ones the scene would use
    alone (solver_6dof.py:937-943, 1153-1163).
Pair order, coloring order and per-body accumulation order are relative
orders inside a scene; none of them sees another scene.

Not in the reference, which steps one scene per Solver6DOF.
"""
from __future__ import annotations

import dataclasses
from collections.abc import Sequence

import numpy as np

from .solver_6dof import Solver6DOF

# settings that must agree between the scenes of a batch
_SHARED = ("dt", "iterations", "gravity", "alpha", "beta", "gamma", "post_stabilize",
           "max_linear_speed", "max_angular_speed", "substeps", "friction_static_mult",
           "coloring_mode", "_self_collide", "_self_friction", "_unsafe_fixed_capacity",
           "_device_index")


class BatchedSolver6DOF(Solver6DOF):
    """Independent scenes as one solver. Readers return the whole batch;
    `scene_slice(k)` gives scene k's bodies, `positions(k)` etc. its part."""

    def __init__(self, scenes: list[Solver6DOF], pitch: float | None = None):
        if not scenes:
            raise ValueError("a batch needs at least one scene")
        s0 = scenes[0]
        for k, s in enumerate(scenes):
            for a in _SHARED:
                if getattr(s, a) != getattr(s0, a):
                    raise ValueError(f"scene {k}: {a} = {getattr(s, a)!r} differs from scene 0 "
                                     f"({getattr(s0, a)!r})")
            if s._h is not None:
                raise ValueError(f"scene {k} has already been stepped; batch fresh scenes")
        super().__init__(dt=s0.dt, iterations=s0.iterations, gravity=s0.gravity,
                         alpha=s0.alpha, beta=s0.beta, gamma=s0.gamma,
                         post_stabilize=s0.post_stabilize, device=s0.device,
                         max_linear_speed=s0.max_linear_speed,
                         max_angular_speed=s0.max_angular_speed, substeps=s0.substeps,
                         friction_static_mult=s0.friction_static_mult,
                         coloring_mode=s0.coloring_mode,
                         unsafe_fixed_capacity=s0._unsafe_fixed_capacity)
        self._scenes = list(scenes)
        self.enable_self_collision(s0._self_collide, s0._self_friction)
        first, rfirst = [0], [0]
        for s in scenes:
            for lst in ("_x", "_q", "_v", "_omega", "_mass", "_inv_I_local", "_I_local",
                        "_half_extents", "_friction"):
                getattr(self, lst).extend(getattr(s, lst))
            first.append(len(self._x))
            rfirst.append(rfirst[-1] + len(s._rows))
        self._first = np.asarray(first, np.int32)
        self._rows = _RowsView(self._scenes, self._first, np.asarray(rfirst, np.int64))
        self._pitch = float(pitch) if pitch is not None else self._auto_pitch()

    @property
    def n_scenes(self) -> int:
        return len(self._scenes)

    def scene_slice(self, k: int) -> slice:
        return slice(int(self._first[k]), int(self._first[k + 1]))

    def _auto_pitch(self) -> float:
        """Grid pitch: twice the widest scene's initial extent (its bodies'
        centres +- their half-diagonals) plus 20 m, so scenes stay in their
        own cells while they spread. Only the broad phase's speed depends
        on it."""
        x = np.asarray(self._x, np.float64).reshape(-1, 3)
        he = np.linalg.norm(np.asarray(self._half_extents, np.float64).reshape(-1, 3), axis=1)
        ext = 0.0
        for k in range(self.n_scenes):
            sl = self.scene_slice(k)
            if sl.stop > sl.start:
                lo = (x[sl] - he[sl, None]).min(axis=0)
                hi = (x[sl] + he[sl, None]).max(axis=0)
                ext = max(ext, float((hi - lo).max()))
        return 2.0 * ext + 20.0

    def add_box(self, *a, **kw):
        if getattr(self, "_first", None) is not None:
            raise NotImplementedError("add bodies to a scene before batching it")
        return super().add_box(*a, **kw)

    def _scene_arrays(self) -> dict:
        """The scenes' own flush arrays end to end: body indices of rows
        shifted by the scene's first body, sibling/partner row indices by
        its first row; priorities and capacities as each scene has them
        alone (capacities summed)."""
        parts = [s._scene_arrays() for s in self._scenes]
        nb = np.array([p["n_bodies"] for p in parts], np.int64)
        nr = np.array([p["n_static_rows"] for p in parts], np.int64)
        b0 = np.repeat(np.concatenate([[0], np.cumsum(nb)[:-1]]), nr).astype(np.int32)
        r0 = np.repeat(np.concatenate([[0], np.cumsum(nr)[:-1]]), nr).astype(np.int32)
        out = {"n_bodies": int(nb.sum()), "n_static_rows": int(nr.sum())}
        for key, val in parts[0].items():
            if isinstance(val, np.ndarray):
                out[key] = np.concatenate([p[key] for p in parts])
        out["body_a"] = out["body_a"] + b0
        for key in ("sibling", "partner"):
            out[key] = np.where(out[key] >= 0, out[key] + r0, out[key]).astype(np.int32)
        for key in ("n_dyn_capacity", "hash_cap", "pool_max", "flush_max_pairs"):
            out[key] = int(sum(p[key] for p in parts))
        out["max_pairs"] = int(self._bp_max_pairs)
        if self._h is not None and self.x is not None and self._gpu_pool_ready:
            prev = self._h.read_bodies(self._n_b_flushed)   # state of a re-flush
            for key in ("x", "q", "v", "w", "prev_v", "prev_w"):
                out[key] = prev[key]
        return out

    def _flush(self) -> None:
        was = self._dirty
        super()._flush()
        if was:
            self._h.set_scenes(self._first, self._pitch)

    # per-scene readers
    def positions(self, scene: int | None = None) -> np.ndarray:
        x = super().positions()
        return x if scene is None else x[self.scene_slice(scene)]

    def orientations(self, scene: int | None = None) -> np.ndarray:
        q = super().orientations()
        return q if scene is None else q[self.scene_slice(scene)]

    def velocities(self, scene: int | None = None) -> np.ndarray:
        v = super().velocities()
        return v if scene is None else v[self.scene_slice(scene)]

    def angular_velocities(self, scene: int | None = None) -> np.ndarray:
        w = super().angular_velocities()
        return w if scene is None else w[self.scene_slice(scene)]


class _RowsView(Sequence):
    """The batch's static rows, re-indexed on access (no copy of the
    scenes' row lists)."""

    def __init__(self, scenes, first_body, first_row):
        self._scenes, self._fb, self._fr = scenes, first_body, first_row

    def __len__(self) -> int:
        return int(self._fr[-1])

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[j] for j in range(*i.indices(len(self)))]
        if i < 0:
            i += len(self)
        if not 0 <= i < len(self):
            raise IndexError(i)
        k = int(np.searchsorted(self._fr, i, side="right")) - 1
        r = self._scenes[k]._rows[i - int(self._fr[k])]
        r0 = int(self._fr[k])
        return dataclasses.replace(
            r, body_a=r.body_a + int(self._fb[k]),
            sibling=r.sibling + r0 if r.sibling >= 0 else r.sibling,
            partner=r.partner + r0 if r.partner >= 0 else r.partner)
