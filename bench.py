"""Throughput of the Solver6DOF substep on one MI355X.

Workload (default): the square pyramid of 98,021 cubes (66 layers,
half-extent 0.25 m, 125 kg, mu_d 0.5 / mu_s 0.6, seeded 1 mm jitter),
dt = 1/60, 4 substeps, 20 iterations, FP32. One step = one frame
(4 substeps: contacts, coloring, CSR, prelude, 21 primal sweeps, 20 dual
sweeps, finalize, cache refresh). Metric: body-substeps/s = bodies x
substeps x frames / wall time, with the scene resident in HBM before the
timed region starts. The line also carries iterations to the dual residual
1e-4 (cap 200), plain vs learned warm start, on the landed 10,000-box pile,
the 98,021-cube pyramid and a 285-cube pyramid.

--workload pile: the random pile of 10,000 boxes (half-extents U[0.1, 0.4] m,
random orientations, 22 x 22 x 21 lattice, mu_d 0.4 / mu_s 0.5), dt 1/60,
4 substeps, 10 iterations; the timed frames follow --warmup frames of the
fall (use --warmup 600 for the landed pile of the configuration).

--workload batched: 2048 independent piles of n = round(exp(U(ln 8,
ln 1024))) boxes each (the 10k-box pile's generator per scene), 10
iterations, stepped as one batch; same metric over all scenes.

--workload train: Adam steps of the warm-start network's MFMA training
kernels on 2,097,152 synthetic samples, batch 16,384. Metric: samples/s.

--workload particles: the 3-DOF particle Solver on 256 soft blocks (tet
lattices of 6^3 cells: 87,808 particles, 642,048 rows, 331,776 tets) dropped
on a floor, dt 1/60, 20 iterations + post-stabilisation. Metric:
particle-steps/s.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the primal sweep kernel measured with rocprofv3
# FETCH_SIZE / WRITE_SIZE (scaled by a copy-kernel calibration) on this
# workload; the summary file records the commands and the raw counters.
PRIMAL_KERNELS = ["k_primal_color (one launch per color)", "k_primal_tiled<1> per color",
                  "k_primal_tiled<2> per color", "k_primal_tiled<4> per color",
                  "k_primal_tiled<8> per color", "k_primal_tiled<1|2> per color",
                  "k_primal_tiled<2, flow> one launch", "k_primal_tiled<1, flow> one launch",
                  "k_primal_tiled<4, flow> one launch", "k_primal_tiled<8, flow> one launch",
                  "k_primal_tiled<3, flow> one launch"]
TRAFFIC_FILE = ROOT / "profiles" / "k_primal_sweep_pyramid66_pmc.md"


# The sources the primal sweep kernel is built from: a traffic figure
# measured on other sources does not describe this build.
KERNEL_SOURCES = ["avbd_amd/csrc/kernels_solve.hip", "avbd_amd/csrc/avbd_internal.h",
                  "avbd_amd/csrc/avbd_fp.h"]


def kernel_source_hash() -> str:
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        h.update((ROOT / f).read_bytes())
    return h.hexdigest()[:16]


def measured_traffic(layers: int, variant: int) -> float | None:
    """Bytes per primal launch from the committed PMC summary, when it was
    taken on this workload, kernel variant and kernel sources; None
    otherwise."""
    if layers != 66 or not TRAFFIC_FILE.exists():
        return None
    keys = ("traffic_bytes_per_launch:", "primal_variant:", "kernel_source_sha256:")
    fields = {k.strip(): v.strip() for k, v in
              (line.split(":", 1) for line in TRAFFIC_FILE.read_text().splitlines()
               if line.startswith(keys))}
    if int(fields.get("primal_variant", "-1")) != variant:
        return None
    if fields.get("kernel_source_sha256") != kernel_source_hash():
        return None
    return float(fields["traffic_bytes_per_launch"])


SQ_FILE = ROOT / "profiles" / "k_primal_sweep_pyramid66_sq.md"
SIMDS, SCLK_GHZ = 1024, 2.4   # MI355X: 256 CUs x 4 SIMDs, peak engine clock


def valu_issue(layers: int, variant: int, avg_launch_us: float | None) -> dict | None:
    """VALU issue of the primal sweep: its VALU instructions per launch, from
    the committed SQ counter summary when it was taken on this workload,
    kernel variant and kernel sources, against this run's launch time. A
    wave64 FP32 instruction without packed math holds its SIMD for 4 cycles,
    so the issue limit is SIMDS x clock / 4 wavefront instructions per
    second. None when the summary does not match."""
    if layers != 66 or not avg_launch_us or not SQ_FILE.exists():
        return None
    keys = ("valu_instructions_per_launch:", "primal_variant:", "kernel_source_sha256:")
    fields = {k.strip(): v.strip() for k, v in
              (line.split(":", 1) for line in SQ_FILE.read_text().splitlines()
               if line.startswith(keys))}
    if int(fields.get("primal_variant", "-1")) != variant:
        return None
    if fields.get("kernel_source_sha256") != kernel_source_hash():
        return None
    instr = float(fields["valu_instructions_per_launch"])
    limit_us = instr * 4.0 / (SIMDS * SCLK_GHZ * 1e3)
    return {"instructions_per_launch": instr, "issue_limit_us": limit_us,
            "frac": limit_us / avg_launch_us, "source": SQ_FILE.name}


def cpu_baseline(layers: int, seed: int, substeps_sample: int) -> dict:
    """The C++ restatement of the reference substep (one host thread) on the
    same scene: one untimed substep from the initial state, then
    `substeps_sample` timed substeps."""
    from avbd_amd import scenes
    from oracle.oracle import oracle_solver_class
    from tests.parity import one_substep

    s = scenes.pyramid(layers, seed=seed, solver_cls=oracle_solver_class())
    s._flush()
    n_b = len(s._x)
    one_substep(s)   # warm-up: first recolor, first touch of every array
    t0 = time.perf_counter()
    for _ in range(substeps_sample):
        one_substep(s)
    dt = time.perf_counter() - t0
    return {
        "value": n_b * substeps_sample / dt,
        "unit": "body-substeps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"substeps 2-{substeps_sample + 1} of the {n_b}-cube pyramid from its "
                   f"initial state ({dt:.1f} s; substep 1 untimed), C++ restatement of the "
                   f"reference substep on 1 thread of a {os.cpu_count()}-thread host"),
    }


def batched_scene(n_scenes: int, seed: int):
    """2048 piles of log-uniform size in [8, 1024] boxes as one batch."""
    import numpy as np
    from avbd_amd import learned
    from avbd_amd.batched import BatchedSolver6DOF

    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(8), np.log(1024), n_scenes))).astype(int)
    parts = []
    t0 = time.perf_counter()
    for k, n in enumerate(sizes):
        parts.append(learned.pile(int(n), seed=seed * 100003 + k))
        if (k + 1) % 256 == 0:
            print(f"# built {k + 1}/{n_scenes} scenes ({time.perf_counter() - t0:.0f} s)",
                  file=sys.stderr, flush=True)
    return BatchedSolver6DOF(parts), sizes


def iters_to_residual() -> dict:
    """Iterations per substep to the dual residual 1e-4 (cap 200), plain vs
    the shipped learned warm start, from identical states: the landed
    10,000-box pile (600 ordinary frames first, then 10 measured frames),
    the 98,021-cube pyramid (3 frames) and a 285-cube pyramid (30 frames).
    Where every substep stops at the cap, the median residual it stopped at
    says which start got further."""
    from avbd_amd import learned, scenes

    p = learned.load_params()
    scope = learned.load_meta().get("scope", "all")
    cases = [("pile of 10,000 boxes, landed (600 frames), 10 frames x 4 substeps",
              lambda: scenes.random_pile(10_000, seed=0), 600, 10),
             ("pyramid of 98,021 cubes (66 layers), 3 frames x 4 substeps",
              lambda: scenes.pyramid(66, seed=0), 0, 3),
             ("pyramid of 285 cubes (9 layers), 30 frames x 4 substeps",
              lambda: scenes.pyramid(9, seed=9500), 0, 30)]
    out = {"tol": 1e-4, "cap": 200, "scope": scope, "scenes": []}
    for name, build, settle, frames in cases:
        t0 = time.perf_counter()
        r = learned.compare(build, frames, p, 1e-4, 200, scope, settle)
        out["scenes"].append({
            "scene": name, "plain": r["plain_mean"], "learned": r["learned_mean"],
            "plain_capped": r["plain_capped"], "learned_capped": r["learned_capped"],
            "substeps": r["substeps"], "plain_resid_median": r["plain_resid_median"],
            "learned_resid_median": r["learned_resid_median"]})
        print(f"# iterations to residual, {name}: plain {r['plain_mean']:.2f} learned "
              f"{r['learned_mean']:.2f} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr,
              flush=True)
    return out


PARTICLE_TRAFFIC_FILE = ROOT / "profiles" / "k_p_primal_grp_softblocks256_pmc.md"
PARTICLE_KERNELS = ["k_p_primal (lane per body) per color", "k_p_primal_grp<16> per color",
                    "k_p_primal_grp<32> per color", "k_p_primal_grp<64> per color",
                    "k_p_primal_grp<32, flow> one launch per sweep",
                    "k_p_primal_grp<16, flow> one launch per sweep"]


def particle_traffic(blocks: int, variant: int) -> float | None:
    if blocks != 256 or not PARTICLE_TRAFFIC_FILE.exists():
        return None
    fields = dict(line.split(":", 1) for line in PARTICLE_TRAFFIC_FILE.read_text().splitlines()
                  if line.startswith(("traffic_bytes_per_launch:", "primal_variant:")))
    if int(fields.get("primal_variant", "-1")) != variant:
        return None
    return float(fields["traffic_bytes_per_launch"])


def particle_bench(args) -> dict:
    """The particle Solver's step on soft blocks; timed steps with the step
    graph replayed, then a few steps with per-launch event timing for the
    primal kernel's roofline line."""
    import torch

    from avbd_amd import particle_scenes as PS
    from avbd_amd.solver import Solver
    t0 = time.perf_counter()
    s = Solver(dt=1 / 60, iterations=20)
    PS.soft_blocks(s, args.blocks)
    s._flush()
    nb = len(s._bodies_x)
    print(f"# scene: {nb} particles, {len(s._constraints)} rows, {len(s._tets)} tets, "
          f"{s.num_colors} colors, build+flush {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        s.step()
    tuned = None
    if args.primal_variant is not None:
        variant = args.primal_variant
        s._h.set_option(1, variant)
    else:
        # per-color launches or the dataflow sweep, by timing real steps
        # (bit-identical results, untimed)
        tuned = s.autotune_primal(steps=3)
        variant = s.primal_variant
        print(f"# primal autotune (ms per step): {tuned} -> variant {variant}", file=sys.stderr,
              flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    s._h.profile_enable(True)
    for _ in range(3):
        s.step()
    ms, launches, nbytes = s._h.profile_read()
    s._h.profile_enable(False)
    assert bool(__import__("numpy").isfinite(s.positions()).all()), "particle state diverged"
    prim_n = int(launches[0])
    avg_s = float(ms[0]) / prim_n * 1e-3
    bytes_per = float(nbytes[0]) / prim_n
    achieved = bytes_per / avg_s / 1e9
    out = {
        "metric": f"particle-steps/s, {args.blocks} soft blocks",
        "value": nb * args.steps / elapsed,
        "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.blocks} soft blocks (6^3-cell tet lattices, edge rows "
                               f"5e4, tet rows 1e4, floor + friction rows) falling on a floor, "
                               f"dt 1/60, 20 iterations + post-stabilisation",
                   "particles": nb, "rows": len(s._constraints), "tets": len(s._tets),
                   "colors": s.num_colors, "parallelism": "single GPU",
                   "primal_variant": variant,
                   "primal_autotune_ms": ({str(k): round(v, 3) for k, v in tuned.items()}
                                          if tuned else None)},
        "roofline": {"bound": "hbm", "kernel": PARTICLE_KERNELS[variant],
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": particle_traffic(args.blocks, variant),
                     "avg_launch_us": avg_s * 1e6, "algorithmic_bytes_per_launch": bytes_per},
        "kernel_ms_3_steps": {"primal": round(float(ms[0]), 3), "dual": round(float(ms[1]), 3)},
    }
    if args.cpu_baseline:
        from oracle.oracle import particle_solver_class
        so = particle_solver_class()(dt=1 / 60, iterations=20)
        PS.soft_blocks(so, args.blocks)
        so._flush()
        t0 = time.perf_counter()
        for _ in range(args.cpu_substeps):
            so.step()
        so.positions()
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": nb * args.cpu_substeps / el, "unit": "particle-steps/s",
                               "cores": 1, "kind": "port",
                               "sample": f"the same scene, {args.cpu_substeps} steps from the "
                                         f"initial state, C++ restatement on one host thread "
                                         f"({el:.1f} s)"}
    return out


def substep_roofline(n_b: int, n_c: int, n_f: int, iters: int, ms_per_substep: float) -> dict:
    """The whole substep against HBM peak: the minimal-record byte count of
    one substep (prelude, per-contact prelude, iters + 1 primal sweeps, iters
    duals, post-stabilisation pass) for n_b bodies, n_c box contacts and n_f
    floor corners, over the measured time per substep."""
    b = (n_b * (188 + 132) + n_c * 124 + n_f * 88
         + (iters + 1) * (n_b * 120 + 2 * n_c * 112 + n_f * 76)
         + iters * (n_c * 136 + n_f * 100)
         + n_c * 80 + n_f * 44)
    gbs = b / (ms_per_substep * 1e-3) / 1e9
    return {"algorithmic_bytes_per_substep": b, "ms_per_substep": ms_per_substep,
            "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
            "body_substeps_per_s_at_peak": n_b / (b / (HBM_PEAK_GBS * 1e9))}


TRAIN_FLOP_PER_SAMPLE = (2 * (24 * 64 + 64 * 64 + 64 * 4)                        # forward
                         + 2 * (64 * 4 + 64 * 4 + 64 * 64 + 64 * 64 + 24 * 64))  # backward


def train_bench(args) -> dict:
    """The warm-start network's training kernels (avbd_mlp_train: one MFMA
    forward+backward launch and one Adam launch per step) on synthetic
    samples. The timed call includes copying the sample set to the device
    and drawing the minibatch order, so `value` is a lower bound on the
    kernels' own rate; rocprof gives the kernel durations."""
    import numpy as np
    from avbd_amd import learned as L

    rng = np.random.default_rng(args.seed)
    n = args.train_samples
    X = rng.normal(0.0, 1.0, (n, L.N_IN)).astype(np.float32)
    Y = np.stack([np.tanh(X[:, 0] + 0.5 * X[:, 1]), np.sin(X[:, 2]) * X[:, 3],
                  0.3 * X[:, 4] ** 2, X[:, 5] - X[:, 6]], axis=1).astype(np.float32)
    M = np.full(n, 15, np.int32)
    p = L.init_params(*L.normalisation(X[:65536], Y[:65536], M[:65536]), seed=1)
    L.train_gpu(p, X[:65536], Y[:65536], M[:65536], steps=max(1, args.warmup), batch=args.batch)
    t0 = time.perf_counter()
    p1, hist = L.train_gpu(p, X, Y, M, steps=args.steps, batch=args.batch, lr=1e-3, seed=2)
    el = time.perf_counter() - t0
    assert np.all(np.isfinite(hist)), "training diverged"
    samples = args.steps * min(args.batch, n)
    return {
        "metric": "warm-start network training samples/s (MFMA forward+backward+Adam)",
        "value": samples / el, "unit": "samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (fp32 accumulation)",
        "data": "synthetic",
        "config": {"workload": f"24-64-64-4 SiLU MLP, {args.steps} Adam steps of {args.batch} "
                               f"samples from {n} synthetic samples (host arrays copied in)",
                   "flop_per_sample": TRAIN_FLOP_PER_SAMPLE,
                   "parallelism": "single GPU"},
        "achieved_tflops_wall": samples * TRAIN_FLOP_PER_SAMPLE / el / 1e12,
        "loss_first_last": [float(hist[0]), float(hist[-1])],
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("pyramid", "pile", "batched", "particles", "train"),
                    default="pyramid")
    ap.add_argument("--boxes", type=int, default=10_000, help="pile: number of boxes")
    ap.add_argument("--train-samples", type=int, default=2_097_152,
                    help="train: synthetic samples")
    ap.add_argument("--batch", type=int, default=16_384, help="train: minibatch")
    ap.add_argument("--blocks", type=int, default=256, help="particles: number of soft blocks")
    ap.add_argument("--scenes", type=int, default=2048, help="batched: number of scenes")
    ap.add_argument("--residual", type=int, default=1,
                    help="1: add iterations to residual, plain vs learned warm start")
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--layers", type=int, default=66)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--cpu-substeps", type=int, default=2,
                    help="timed substeps of the CPU baseline (after one untimed one)")
    ap.add_argument("--direct", type=int, default=0,
                    help="1: direct launches per substep (the CPU-device semantics) instead "
                         "of resident frames replayed as a HIP graph")
    ap.add_argument("--profile-frames", type=int, default=2,
                    help="frames after the timed region with per-launch event timing")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="only run 5 device-to-device copies of MB MiB (the reference point "
                         "for FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--fuse-dual", type=int, default=0,
                    help="dual update with the dataflow sweep: 0 separate launch (default), "
                         "1 inside the sweep's tiles, 2 in trailing workgroups of its launch, "
                         "-1 the faster of 0 and 2 by timing frames (Solver6DOF.autotune_dual)")
    ap.add_argument("--broad-lanes", type=int, default=0,
                    help="broad-phase traversal lanes per body: 0 by body count (default), "
                         "1, 8, 16 or 64")
    ap.add_argument("--primal-variant", type=int, default=None,
                    help="0 lane per body; 1-4 one wavefront per 1, 2, 4, 8 bodies; "
                         "5 one or two bodies per wavefront by color size; "
                         "6-10 one dataflow launch per sweep with 2, 1, 4, 8, 3 bodies per tile; "
                         "default: picked by the autotuner among 6-10 after the warm-up")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the substep is not sharded across GPUs.", file=sys.stderr)
        return 2

    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"# copied {n} bytes 5 times", file=sys.stderr, flush=True)
        return 0

    if args.workload == "particles":
        print(json.dumps(particle_bench(args)), flush=True)
        return 0
    if args.workload == "train":
        print(json.dumps(train_bench(args)), flush=True)
        return 0

    from avbd_amd import scenes

    t_build = time.perf_counter()
    if args.workload == "batched":
        s, sizes = batched_scene(args.scenes, args.seed)
    elif args.workload == "pile":
        s = scenes.random_pile(args.boxes, seed=args.seed)
    else:
        s = scenes.pyramid(args.layers, seed=args.seed)
    # "cuda:0": the resident frame (fixed-capacity pools, substeps replayed
    # as a captured HIP graph, one counter readback per frame); "cpu": the
    # reference's CPU-device semantics, a direct launch sequence per substep
    s.device = "cpu" if args.direct else "cuda:0"
    s._flush()
    variant = args.primal_variant
    if variant is not None:
        s._h.set_option(0, variant)
    s._h.set_option(1, max(args.fuse_dual, 0))
    s._h.set_option(5, args.broad_lanes)
    n_b = len(s._x)
    print(f"# scene: {n_b} bodies, {len(s._rows)} static rows, build+flush "
          f"{time.perf_counter() - t_build:.1f} s", file=sys.stderr, flush=True)
    for w in range(args.warmup):
        t0 = time.perf_counter()
        s.step()
        print(f"# warmup frame {w}: {1e3 * (time.perf_counter() - t0):.1f} ms", file=sys.stderr,
              flush=True)
    tuned = None
    frames_before = args.warmup  # frames the scene has advanced before the timed ones
    if variant is None:
        # the primal sweep kernel for this scene, by timing one frame with
        # each (identical results, so these are ordinary untimed frames)
        tuned = s.autotune_primal(frames=1)
        variant = s.primal_variant
        print(f"# primal autotune (ms of primal sweeps per frame): {tuned} -> variant {variant}",
              file=sys.stderr, flush=True)
        s.step()   # untimed: captures the graph of the chosen kernel
        frames_before += len(tuned) + 1
    dual_tuned = None
    if args.fuse_dual < 0:
        dual_tuned = s.autotune_dual(frames=3)
        print(f"# dual placement autotune (ms per frame): {dual_tuned} -> {s.dual_placement}",
              file=sys.stderr, flush=True)
        s.step()   # untimed: captures the graph of the chosen placement
        frames_before += len(dual_tuned) * 4 + 1
    import torch
    syncs0, graphs0 = s._h.sync_count()
    torch.cuda.synchronize()
    frame_ms = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        tf = time.perf_counter()
        s.step()   # a resident frame ends on its counter readback
        frame_ms.append(1e3 * (time.perf_counter() - tf))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    syncs1, graphs1 = s._h.sync_count()
    assert bool(__import__("numpy").isfinite(s.positions()).all()), "state diverged"
    n_rows, n_pairs, n_contacts = s._h.read_counters()
    dropped = n_pairs > s._bp_max_pairs or n_rows > s._gpu_pool_n_capacity
    # kernel times: a few more frames with per-launch events (direct
    # launches, since event timing is not captured into the graph)
    s._h.profile_enable(True)
    for _ in range(args.profile_frames):
        s.step()
    ms, launches, nbytes = s._h.profile_read(7)
    s._h.profile_enable(False)
    substeps = s.substeps * args.steps
    value = n_b * substeps / elapsed
    prim_ms, prim_n, prim_b = float(ms[0]), int(launches[0]), float(nbytes[0])
    achieved = (prim_b / prim_n) / (prim_ms / prim_n * 1e-3) / 1e9 if prim_n else 0.0
    classes = ["primal", "dual", "prelude", "contacts", "coloring", "csr", "other"]
    breakdown = {c: {"ms": round(float(ms[i]), 3), "launches": int(launches[i])}
                 for i, c in enumerate(classes)}
    print(f"# kernel time by class over {args.profile_frames} profiled frames: {breakdown}",
          file=sys.stderr, flush=True)
    fs = sorted(frame_ms)
    pct = lambda q: fs[min(len(fs) - 1, int(round(q * (len(fs) - 1))))]
    timing = {"timed_frames": args.steps,
              # 1-based frames of the scene's life: warm-up, autotune and
              # graph-capture frames come first
              "timed_frame_range": [frames_before + 1, frames_before + args.steps],
              "launch": ("direct launches per substep" if args.direct else
                         "resident frames: substeps replayed from a captured HIP graph"),
              "graph_substeps": graphs1 - graphs0, "host_syncs": syncs1 - syncs0,
              "frame_ms_p50": round(pct(0.5), 3), "frame_ms_p99": round(pct(0.99), 3),
              "frame_ms_max": round(fs[-1], 3),
              "contacts_dropped_by_fixed_pools": bool(dropped),
              "profiled_frames": args.profile_frames}
    if args.workload == "batched":
        workload = (f"{args.scenes} independent piles of 8-1024 boxes (log-uniform, "
                    f"{n_b} boxes in all) stepped as one batch, dt 1/60, {s.substeps} substeps, "
                    f"{s.iterations} iterations")
        metric = f"body-substeps/s, {args.scenes} batched scenes"
    elif args.workload == "pile":
        workload = (f"random pile of {n_b} boxes, dt 1/60, {s.substeps} substeps, "
                    f"{s.iterations} iterations, frames {frames_before + 1}-"
                    f"{frames_before + args.steps} after the drop")
        metric = "body-substeps/s, 10k-box pile"
    else:
        workload = (f"pyramid of {n_b} cubes ({args.layers} layers), dt 1/60, "
                    f"{s.substeps} substeps, {s.iterations} iterations")
        metric = "body-substeps/s @100k boxes"
    out = {
        "metric": metric,
        "value": value,
        "unit": "body-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload,
                   "bodies": n_b, "substeps_per_step": s.substeps, "iterations": s.iterations,
                   "contacts": n_contacts, "pairs": n_pairs, "rows": n_rows,
                   "colors": s.num_active_colors, "parallelism": "single GPU",
                   "primal_variant": variant, "broad_lanes": args.broad_lanes,
                   "dual_placement": (s.dual_placement if dual_tuned else args.fuse_dual),
                   "dual_autotune_ms": ({str(k): round(v, 3) for k, v in dual_tuned.items()}
                                        if dual_tuned else None),
                   "primal_autotune_ms": ({str(k): round(v, 3) for k, v in tuned.items()}
                                          if tuned else None)},
        "roofline": {
            "bound": "hbm",
            "kernel": PRIMAL_KERNELS[variant] + " (primal sweep)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": (measured_traffic(args.layers, variant) if args.workload == "pyramid"
                    else None),
            "avg_launch_us": 1e3 * prim_ms / prim_n if prim_n else None,
            "algorithmic_bytes_per_launch": prim_b / prim_n if prim_n else None,
            # the bound the sweep actually meets: VALU issue
            "valu": (valu_issue(args.layers, variant, 1e3 * prim_ms / prim_n if prim_n else None)
                     if args.workload == "pyramid" else None),
        },
        "kernel_ms": breakdown,
        "substep_roofline": substep_roofline(n_b, n_contacts, len(s._rows) // 3, s.iterations,
                                             1e3 * elapsed / substeps),
        "timing": timing,
    }
    if args.workload == "batched":
        out["config"]["scenes"] = args.scenes
        out["config"]["scene_bodies_min_max"] = [int(sizes.min()), int(sizes.max())]
        # every scene of the batch advances in the same launches, so a
        # scene's frame latency is the batch's frame time
        out["per_scene_latency_ms"] = {"p50": timing["frame_ms_p50"],
                                       "p99": timing["frame_ms_p99"]}
    if args.residual and args.workload == "pyramid":
        out["iters_to_residual"] = iters_to_residual()
    if args.cpu_baseline and args.workload == "pyramid":
        out["cpu_baseline"] = cpu_baseline(args.layers, args.seed, args.cpu_substeps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
