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
        first = [0]
        for s in scenes:
            b0, r0 = len(self._x), len(self._rows)
            for lst in ("_x", "_q", "_v", "_omega", "_mass", "_inv_I_local", "_I_local",
                        "_half_extents", "_friction"):
                getattr(self, lst).extend(getattr(s, lst))
            for r in s._rows:
                self._rows.append(dataclasses.replace(
                    r, body_a=r.body_a + b0,
                    sibling=r.sibling + r0 if r.sibling >= 0 else r.sibling,
                    partner=r.partner + r0 if r.partner >= 0 else r.partner))
            first.append(len(self._x))
        self._first = np.asarray(first, np.int32)
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
        arrs = super()._scene_arrays()
        parts = [s._scene_arrays() for s in self._scenes]
        arrs["priority"] = np.concatenate([p["priority"] for p in parts]).astype(np.float32)
        for key in ("n_dyn_capacity", "hash_cap", "pool_max", "flush_max_pairs"):
            arrs[key] = int(sum(p[key] for p in parts))
        return arrs

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
