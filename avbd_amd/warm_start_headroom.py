"""Headroom of the warm start: how many iterations to residual a perfect
starting dual state would save.

    python -m avbd_amd.warm_start_headroom pyramid 8 [--frames 60 --every 5]

Every `every`-th substep (after two frames) is run with the plain warm start
and read back; then the same substep is re-run from the same state
(bodies, static rows, contact cache, coloring) with the converged state
written in as the start, divided by the prelude's decay (lambda / (alpha
gamma), k / gamma):
  box       converged lambda of every box-box contact
  box+floor the same plus the floor rows' converged lambda
  new       converged lambda of the contacts that were not cached
  k         converged penalties, lambda plain
  k_n=1e8   every cached contact's normal penalty at 1e8, lambda plain
The plain trajectory then continues from its own result. Not in the
reference (it has no residual measurement).
"""
from __future__ import annotations

import argparse
import time

import numpy as np

from . import learned as L
from . import scenes

VARIANTS = ("plain", "box", "box+floor", "new", "k", "k_n=1e8")


def _substep(s, tol: float, cap: int) -> int:
    p = s._params(s.dt / s.substeps if s.substeps > 1 else s.dt)
    p.iterations = int(cap)
    rep, iters, _r = s._h.substep_to_residual(p, s._color_dirty, tol)
    s._apply_report(rep, len(s._x))
    return int(iters)


def _snapshot(s) -> dict:
    nb = len(s._x)
    keys, st8 = s._h.export_hash()
    return {"bodies": s._h.read_bodies(nb), "rows": s._h.read_rows(s._gpu_pool_n_static),
            "keys": keys, "st8": st8, "colors": s._h.read_colors(nb),
            "n_colors": s._n_active_colors, "color_dirty": s._color_dirty}


def _restore(s, snap: dict, keys=None, st8=None, floor_lambda=None) -> None:
    for field, k in enumerate(("x", "q", "v", "w", "prev_v", "prev_w")):
        s._h.write_bodies(field, 0, snap["bodies"][k])
    r = snap["rows"]
    s._h.write_static_rows(r["lambda"] if floor_lambda is None else floor_lambda, r["penalty"],
                           r["active"], r["was_static"])
    s._h.import_hash(snap["keys"] if keys is None else keys,
                     snap["st8"] if st8 is None else st8)
    s._h.write_colors(snap["colors"], snap["n_colors"])
    s._n_active_colors = snap["n_colors"]
    s._color_dirty = snap["color_dirty"]


def headroom(s, frames: int, every: int, tol: float = 1e-4, cap: int = 200,
             alpha: float = 0.99, gamma: float = 0.99, out=print) -> dict:
    ag = alpha * gamma
    s._flush()
    res = {v: [] for v in VARIANTS}
    ns = max(1, s.substeps)
    t0 = time.perf_counter()
    for f in range(frames):
        for j in range(ns):
            if (f * ns + j) % every or f < 2:
                _substep(s, tol, cap)
                continue
            before = _snapshot(s)
            its = [_substep(s, tol, cap)]
            k2, h2 = s._h.export_hash()
            rows2 = s._h.read_rows(s._gpu_pool_n_static)
            after = _snapshot(s)
            old = {k: i for i, k in enumerate(before["keys"].tolist())}
            # the converged key set with the cached (plain) start of each key
            start = np.zeros((len(k2), 8), np.float32)
            start[:, 3:6] = (1.0e6, 1.0, 1.0)
            cached = np.zeros(len(k2), bool)
            for i, k in enumerate(k2.tolist()):
                if k in old:
                    start[i] = before["st8"][old[k]]
                    cached[i] = True
            perfect = start.copy()
            perfect[:, 0:3] = h2[:, 0:3] / ag
            new_only = start.copy()
            new_only[~cached, 0:3] = h2[~cached, 0:3] / ag
            perfect_k = start.copy()
            perfect_k[:, 3:6] = h2[:, 3:6] / gamma
            high_k = before["st8"].copy()
            high_k[:, 3] = 1.0e8 / gamma
            for keys, st8, floor in ((k2, perfect, None), (k2, perfect, rows2["lambda"] / ag),
                                     (k2, new_only, None), (k2, perfect_k, None),
                                     (before["keys"], high_k, None)):
                _restore(s, before, keys, st8, floor)
                its.append(_substep(s, tol, cap))
            for v, it in zip(VARIANTS, its):
                res[v].append(it)
            _restore(s, after)
        if f % 20 == 0:
            out(f"frame {f}: " + ", ".join(f"{v} {np.mean(res[v]):.1f}" for v in VARIANTS
                                            if res[v]) + f" ({time.perf_counter() - t0:.1f} s)")
    return {v: float(np.mean(res[v])) if res[v] else float("nan") for v in VARIANTS} | {
        "samples": len(res["plain"])}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("scene", choices=("pyramid", "pile", "stack"))
    ap.add_argument("size", type=int)
    ap.add_argument("--seed", type=int, default=9500)
    ap.add_argument("--frames", type=int, default=60)
    ap.add_argument("--every", type=int, default=5)
    ap.add_argument("--settle", type=int, default=0, help="frames run first, plain")
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--cap", type=int, default=200)
    a = ap.parse_args(argv)
    build = {"pyramid": lambda: scenes.pyramid(a.size, seed=a.seed),
             "pile": lambda: L.pile(a.size, seed=a.seed),
             "stack": lambda: scenes.cube_stack(a.size, seed=a.seed)}[a.scene]
    s = build()
    for _ in range(a.settle):
        s.step_to_residual(a.tol, a.cap)
    r = headroom(s, a.frames, a.every, a.tol, a.cap, s.alpha, s.gamma,
                 out=lambda m: print(m, flush=True))
    print(f"RESULT {a.scene} {a.size}: " + ", ".join(
        f"{v} {r[v]:.2f}" for v in VARIANTS) + f" ({r['samples']} substeps)", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
