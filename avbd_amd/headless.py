"""The reference's benchmark scene and headless driver on the MI355X solver.

`build_scene` builds the same rigid scene as the reference's interactive
viewer (examples/viewer.py:91-194): an N x N grid of cube towers (half-extent
0.12 m, 1 kg, <= 1 mm seeded x/z jitter) plus a row of standing domino slabs,
all with floor rows and self-collision. It draws from the seeded generator in
the same order (one color draw per tower, two jitter draws per cube), so the
bodies are identical for the same arguments. `run_headless` reproduces the
printout of the reference's headless mode (examples/viewer.py:1691-1737).

    python -m avbd_amd.headless headless [--stress] [--iterations 4] ...

Rendering, drag gizmos and the viser server are not part of this module.
"""
from __future__ import annotations

import argparse
import time
from dataclasses import dataclass

import numpy as np

from .solver_6dof import RigidBody, Solver6DOF


@dataclass
class SceneBox:
    """One body of the scene with its display color (the viewer's
    ViewerBox without the rendering handles)."""
    body: RigidBody
    color: tuple[float, float, float]
    is_static: bool = False


def build_scene(args, solver_cls=Solver6DOF) -> tuple[Solver6DOF, list[SceneBox], list[int]]:
    """The viewer's rigid scene (examples/viewer.py:91-194). Returns the
    solver, the scene boxes in creation order and the (empty) pin-row list."""
    s = solver_cls(
        dt=1.0 / 60.0,
        iterations=int(args.iterations),
        gravity=(0.0, -9.81, 0.0),
        post_stabilize=True,
        device=args.device,
        substeps=int(args.substeps),
        friction_static_mult=float(args.static_mult),
        coloring_mode=str(getattr(args, "coloring", "jacobi")),
        unsafe_fixed_capacity=bool(getattr(args, "fixed_capacity", False)),
        primal_group_size=int(getattr(args, "primal_group", 16)),
        primal_shuffle=bool(getattr(args, "primal_shuffle", True)),
        primal_fused=bool(getattr(args, "primal_fused", True)),
        gpu_resident=bool(getattr(args, "gpu_resident", True)),
        recolor_every_substep=bool(getattr(args, "recolor_every_substep", False)),
    )
    s.enable_self_collision(True, default_friction=args.friction)
    boxes: list[SceneBox] = []
    rng = np.random.default_rng(seed=args.seed)

    def tint():
        # one draw per tower keeps the generator's draw order of the reference
        # scene (viewer.py:125-126); the value only tints the tower
        t = float(rng.random())
        return (0.35 + 0.5 * t, 0.55, 0.85 - 0.5 * t)

    # grid of cube towers
    stress = bool(getattr(args, "stress", False))
    h_cube = 0.12
    tower_spacing = 0.55
    tower_height = int(getattr(args, "stress_height", 6)) if stress else 3
    grid_n = int(getattr(args, "stress_grid", 8)) if stress else 3
    grid_origin = -(grid_n - 1) * tower_spacing * 0.5
    for ix in range(grid_n):
        for iz in range(grid_n):
            cx = grid_origin + ix * tower_spacing
            cz = grid_origin + iz * tower_spacing
            tower_color = tint()
            for k in range(tower_height):
                jitter = rng.uniform(-1e-3, 1e-3, size=2)
                cy = h_cube + 2.0 * h_cube * k
                b = s.add_box((cx + float(jitter[0]), cy, cz + float(jitter[1])),
                              (h_cube, h_cube, h_cube), mass=1.0, friction=args.friction)
                s.add_floor_contact_box(b, friction=args.friction)
                shade = 1.0 - 0.06 * k
                boxes.append(SceneBox(body=b, color=tuple(min(1.0, c * shade) for c in tower_color)))

    # row of standing domino slabs
    domino_he = (0.025, 0.12, 0.06)
    domino_spacing = 0.10
    domino_count = int(getattr(args, "stress_dominoes", 30)) if stress else 6
    domino_z = (-grid_origin + tower_spacing) if stress else 1.8
    domino_x0 = -(domino_count - 1) * domino_spacing * 0.5
    for i in range(domino_count):
        cx = domino_x0 + i * domino_spacing
        b = s.add_box((cx, domino_he[1], domino_z), domino_he, mass=0.5, friction=args.friction)
        s.add_floor_contact_box(b, friction=args.friction)
        c = (0.95, 0.55, 0.25) if i % 2 == 0 else (0.25, 0.55, 0.95)
        boxes.append(SceneBox(body=b, color=c))
    return s, boxes, []


def run_headless(args, out=print, solver_cls=Solver6DOF) -> dict:
    """Time `headless_frames` steps after `headless_warmup` warm-up steps and
    print what the reference's headless mode prints; returns the figures."""
    solver, boxes, _ = build_scene(args, solver_cls=solver_cls)
    n_b = len(boxes)
    out(f"bodies:           {n_b}")
    out(f"static rows:      {len(solver._rows)}")
    out(f"device:           {solver.device}")
    out(f"substeps:         {solver.substeps}")
    out(f"iterations:       {solver.iterations}")
    out(f"post_stabilize:   {solver.post_stabilize}")
    out(f"coloring mode:    {solver.coloring_mode}")
    for _ in range(args.headless_warmup):
        solver.step()
    if getattr(args, "autotune", False) and hasattr(solver, "autotune_dual"):
        tuned = solver.autotune_primal(frames=1)
        out(f"primal autotune:  {tuned} -> {solver.primal_variant}")
        tuned = solver.autotune_dual(frames=3)
        out(f"dual autotune:    {tuned} -> {solver.dual_placement}")
    out(f"colors (cap):     {solver.num_colors}")
    out(f"colors (active):  {solver.num_active_colors}")
    out(f"total rows (cap): {solver._gpu_pool_n_capacity}")
    n = max(1, int(args.headless_frames))
    step_times, bp_times = [], []
    t0_all = time.perf_counter()
    for _ in range(n):
        t0 = time.perf_counter()
        solver.step()
        solver.positions()  # the step has finished on the device
        step_times.append((time.perf_counter() - t0) * 1000.0)
        bp_times.append(solver.broadphase_ms)
    total_ms = (time.perf_counter() - t0_all) * 1000.0
    step_times.sort()
    avg = sum(step_times) / len(step_times)
    p50 = step_times[len(step_times) // 2]
    p95 = step_times[min(len(step_times) - 1, int(0.95 * len(step_times)))]
    bp_avg = sum(bp_times) / len(bp_times)
    n_active = int(solver.n_active_rows.numpy()[0])
    out(f"active rows:      {n_active}")
    out(f"avg step time:    {avg:.2f} ms")
    out(f"p50 step time:    {p50:.2f} ms")
    out(f"p95 step time:    {p95:.2f} ms")
    out(f"avg broadphase:   {bp_avg:.2f} ms")
    out(f"total {n} frames: {total_ms:.1f} ms")
    return {"bodies": n_b, "avg_ms": avg, "p50_ms": p50, "p95_ms": p95,
            "broadphase_ms": bp_avg, "active_rows": n_active, "total_ms": total_ms}


def make_parser() -> argparse.ArgumentParser:
    """The viewer's command line (examples/viewer.py:1523-1681), minus the
    rendering-only options, with the same names and defaults except the
    device, which defaults to the GPU."""
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("mode", nargs="?", default="headless", choices=("headless",))
    p.add_argument("--friction", type=float, default=0.5)
    p.add_argument("--static-mult", type=float, default=1.5)
    p.add_argument("--iterations", type=int, default=4)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--device", type=str, default="cuda:0")
    p.add_argument("--substeps", type=int, default=8)
    p.add_argument("--coloring", type=str, default="jacobi",
                   choices=("jacobi", "jones_plassmann"))
    p.add_argument("--stress", action="store_true")
    p.add_argument("--stress-grid", type=int, default=8)
    p.add_argument("--stress-height", type=int, default=6)
    p.add_argument("--stress-dominoes", type=int, default=30)
    p.add_argument("--primal-group", type=int, default=16)
    p.add_argument("--primal-shuffle", action=argparse.BooleanOptionalAction, default=True)
    p.add_argument("--primal-fused", action=argparse.BooleanOptionalAction, default=True)
    p.add_argument("--gpu-resident", action=argparse.BooleanOptionalAction, default=True)
    p.add_argument("--recolor-every-substep", action="store_true")
    p.add_argument("--fixed-capacity", action="store_true")
    p.add_argument("--headless-warmup", type=int, default=5)
    p.add_argument("--headless-frames", type=int, default=30)
    p.add_argument("--autotune", action="store_true",
                   help="after the warm-up, pick the primal sweep kernel and the dual's "
                        "placement by timing frames (identical results; not in the reference)")
    return p


def main(argv=None) -> int:
    run_headless(make_parser().parse_args(argv))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
