"""Train the learned warm start on traces of the plain solver and compare
iterations to residual, plain vs learned, on held-out scenes.

    python -m avbd_amd.train_warm_start [--train-piles 40] [--out PATH] ...

1. Traces: piles of n boxes (n log-uniform in [min, max]) dropped from a
   lattice, every substep iterated to the dual residual tolerance with the
   plain warm start; each substep's contact features and converged dual
   state are logged (a random subset of at most --keep-cap per substep).
2. Training: Adam on the GPU (avbd_mlp_train), minibatches from seeded
   shuffles, the learning rate stepped down over --phases phases; 5 % of the
   samples held out.
3. Evaluation: the same held-out scenes run twice from identical states,
   plain and learned warm start, every substep iterated to the tolerance:
   mean iterations per substep, substeps that hit the cap.
The parameters and the figures go to --out (npz) and --report (JSON).
"""
from __future__ import annotations

import argparse
import json
import time
from pathlib import Path

import numpy as np

from . import learned as L
from . import scenes


def _log(msg: str) -> None:
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


def gather(args, rng) -> L.Trace:
    """Traces of piles (n log-uniform in [min, max]), resting pyramids and
    cube stacks."""
    xs, ys, ms = [], [], []
    sizes = np.exp(rng.uniform(np.log(args.min_boxes), np.log(args.max_boxes), args.train_piles))
    jobs = [(f"pile {n}", lambda n=int(n), k=k: _settled(L.pile(n, seed=1000 + k),
                                                          args.settle_frames), args.frames)
            for k, n in enumerate(np.round(sizes).astype(int))]
    jobs += [(f"pyramid {lay}", lambda lay=lay, k=k: scenes.pyramid(lay, seed=2000 + k),
              args.pyramid_frames) for k, lay in enumerate(args.train_pyramids)]
    jobs += [(f"stack {h}", lambda h=h, k=k: scenes.cube_stack(h, seed=3000 + k),
              args.pyramid_frames) for k, h in enumerate(args.train_stacks)]
    for k, (name, build, frames) in enumerate(jobs):
        t0 = time.perf_counter()
        tr = _log_capped(build(), args, frames, seed=k)
        xs.append(tr.X)
        ys.append(tr.Y)
        ms.append(tr.mask)
        _log(f"trace {k + 1}/{len(jobs)}: {name}, {len(tr.X)} samples, "
             f"{time.perf_counter() - t0:.1f} s")
    return L.Trace(np.concatenate(xs), np.concatenate(ys), np.concatenate(ms))


def _settled(s, frames: int):
    """The scene after `frames` ordinary frames (fixed iterations, plain
    warm start): a dropped pile has landed before its samples are logged."""
    for _ in range(frames):
        s.step()
    return s


def _log_capped(s, args, frames, seed) -> L.Trace:
    """Every substep's samples from `frames` frames of the plain solver, at
    most args.keep_cap per substep (a seeded random subset); groups with
    nothing to predict (mask 0: pins) are dropped."""
    rng = np.random.default_rng(seed)
    s.set_learned_warm_start("log")
    xs, ys, ms = [], [], []

    def grab(_k):
        X, Y, M = s.learned_samples()
        use = M != 0
        X, Y, M = X[use], Y[use], M[use]
        if len(X) > args.keep_cap:
            sel = rng.choice(len(X), args.keep_cap, replace=False)
            X, Y, M = X[sel], Y[sel], M[sel]
        if len(X):
            xs.append(X)
            ys.append(Y)
            ms.append(M)

    for _ in range(frames):
        s.step_to_residual(args.tol, args.max_iters, on_substep=grab)
    s.set_learned_warm_start("plain")
    if not xs:
        return L.Trace(np.zeros((0, L.N_IN), np.float32), np.zeros((0, L.N_OUT), np.float32),
                       np.zeros(0, np.int32))
    return L.Trace(np.concatenate(xs), np.concatenate(ys), np.concatenate(ms))


def train(tr: L.Trace, args, rng):
    n = len(tr.X)
    perm = rng.permutation(n)
    n_val = max(1, n // 20)
    val, fit = perm[:n_val], perm[n_val:]
    if args.scope == "empty":   # train on the groups the network will start
        fit = fit[tr.X[fit, 15] == 0.0]
        val = val[tr.X[val, 15] == 0.0]
    X, Y, M = tr.X[fit], tr.Y[fit], tr.mask[fit]
    if args.clip > 0:   # bound the corrections the network is asked to make
        Y = np.clip(Y, -args.clip, args.clip)
    p = L.init_params(*L.normalisation(X, Y, M), seed=args.seed)
    lr = args.lr
    hist = []
    for ph in range(args.phases):
        t0 = time.perf_counter()
        p, h = L.train_gpu(p, X, Y, M, steps=args.steps // args.phases, batch=args.batch, lr=lr,
                           seed=args.seed + ph, huber=args.huber)
        vl = L.loss_grad_np(p, tr.X[val], tr.Y[val], tr.mask[val], huber=args.huber)[0]
        hist.append({"phase": ph, "lr": lr, "last_minibatch_loss": float(np.mean(h[-50:])),
                     "val_loss": vl})
        _log(f"phase {ph}: lr {lr:g}, minibatch loss {np.mean(h[-50:]):.4f}, "
             f"held-out loss {vl:.4f} ({time.perf_counter() - t0:.1f} s)")
        lr *= 0.3
    # the held-out loss of the plain warm start (zero correction)
    p0 = p.copy()
    p0[:L.N_WEIGHTS] = 0.0
    base = L.loss_grad_np(p0, tr.X[val], tr.Y[val], tr.mask[val], huber=args.huber)[0]
    # median absolute error of the normal-row correction per kind (box-box, floor)
    yv, _ = L.forward_np(p, tr.X[val])
    for kind, name in ((0, "box-box"), (1, "floor")):
        k = tr.X[val][:, 23] == kind
        if k.any():
            e = np.abs(yv[k, 0] - tr.Y[val][k, 0])
            e0 = np.abs(tr.Y[val][k, 0])
            _log(f"held-out {name}: median |error| of the normal correction {np.median(e):.4f} "
                 f"(plain start {np.median(e0):.4f})")
    return p, hist, base


def evaluate(p, args) -> list[dict]:
    out = []
    # (name, builder, settle frames, measured frames)
    cases = [(f"pile{n}", (lambda n=n, sd=sd: L.pile(n, seed=sd)), args.settle_frames,
              args.eval_frames)
             for n, sd in zip(args.eval_sizes, range(9000, 9000 + len(args.eval_sizes)))]
    if args.eval_pyramid > 0:
        cases.append((f"pyramid{args.eval_pyramid}",
                      lambda: scenes.pyramid(args.eval_pyramid, seed=9500), 0, args.eval_frames))
    if args.eval_configs:
        # the 10,000-box pile (landed) and the 98,021-cube pyramid
        cases.append(("pile10000", lambda: scenes.random_pile(10_000, seed=0),
                      args.eval_settle_frames, args.eval_frames))
        cases.append(("pyramid66", lambda: scenes.pyramid(66, seed=0), 0,
                      args.eval_frames_large))
    for name, builder, settle, frames in cases:
        t0 = time.perf_counter()
        r = L.compare(builder, frames, p, args.tol, args.max_iters, args.scope, settle)
        r = {k: v for k, v in r.items() if not isinstance(v, np.ndarray)}
        r["scene"] = name
        r["settle_frames"] = settle
        r["frames"] = frames
        out.append(r)
        _log(f"eval {name}: iterations/substep plain {r['plain_mean']:.2f} learned "
             f"{r['learned_mean']:.2f} (capped {r['plain_capped']} / {r['learned_capped']} of "
             f"{r['substeps']}; median final residual {r['plain_resid_median']:.3g} / "
             f"{r['learned_resid_median']:.3g}), {time.perf_counter() - t0:.1f} s")
        if args.headroom_frames > 0:
            t0 = time.perf_counter()
            sv = builder()
            for _ in range(settle):
                sv.step()
            b = L.perfect_start_bound(sv, args.headroom_frames, args.tol, args.max_iters)
            r["perfect_start"] = {k: v for k, v in b.items() if not isinstance(v, np.ndarray)}
            _log(f"bound {name}: iterations/substep plain {b['plain_mean']:.2f}, from the "
                 f"state the plain run ends at {b['perfect_mean']:.2f} (capped "
                 f"{b['plain_capped']} / {b['perfect_capped']} of {b['substeps']}; median final "
                 f"residual {b['plain_resid_median']:.3g} / {b['perfect_resid_median']:.3g}), "
                 f"{time.perf_counter() - t0:.1f} s")
    return out


def make_parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--train-piles", type=int, default=40)
    p.add_argument("--min-boxes", type=int, default=1000)
    p.add_argument("--max-boxes", type=int, default=10000)
    p.add_argument("--frames", type=int, default=120, help="logged frames per pile")
    p.add_argument("--settle-frames", type=int, default=0,
                   help="ordinary frames a pile runs before it is logged or measured")
    p.add_argument("--keep-cap", type=int, default=500)
    p.add_argument("--train-pyramids", type=int, nargs="*", default=[6, 8, 10, 12, 14])
    p.add_argument("--train-stacks", type=int, nargs="*", default=[3, 5, 8])
    p.add_argument("--pyramid-frames", type=int, default=60)
    p.add_argument("--scope", choices=("all", "empty"), default="all",
                   help="groups the network starts: all, or those with an empty cache")
    p.add_argument("--huber", type=float, default=0.0,
                   help="Huber threshold on normalised errors (0 = squared error)")
    p.add_argument("--clip", type=float, default=0.0,
                   help="clip training targets to +-clip (0 = off)")
    p.add_argument("--tol", type=float, default=1e-4)
    p.add_argument("--max-iters", type=int, default=200)
    p.add_argument("--steps", type=int, default=6000)
    p.add_argument("--phases", type=int, default=3)
    p.add_argument("--batch", type=int, default=4096)
    p.add_argument("--lr", type=float, default=2e-3)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--eval-sizes", type=int, nargs="*", default=[1500, 4000])
    p.add_argument("--eval-pyramid", type=int, default=9)
    p.add_argument("--eval-frames", type=int, default=120)
    p.add_argument("--eval-configs", action="store_true",
                   help="also measure the 10,000-box pile and the 98,021-cube pyramid")
    p.add_argument("--headroom-frames", type=int, default=0,
                   help="frames of the perfect-start bound on each evaluation scene (0 = off)")
    p.add_argument("--eval-settle-frames", type=int, default=600,
                   help="ordinary frames the 10,000-box pile runs before it is measured")
    p.add_argument("--eval-frames-large", type=int, default=5,
                   help="measured frames of the 98,021-cube pyramid")
    p.add_argument("--out", type=Path, default=L.WEIGHTS_PATH)
    p.add_argument("--report", type=Path, default=None)
    p.add_argument("--eval-only", type=Path, default=None,
                   help="evaluate the network in this npz (no traces, no training)")
    p.add_argument("--trace-cache", type=Path, default=None,
                   help="npz of logged samples: loaded if present, else written")
    return p


def main(argv=None) -> int:
    args = make_parser().parse_args(argv)
    rng = np.random.default_rng(args.seed)
    t0 = time.perf_counter()
    if args.eval_only is not None:
        # no traces, no training: evaluate the network in the given file
        p = L.load_params(args.eval_only)
        ev = evaluate(p, args)
        if args.report is not None:
            args.report.parent.mkdir(parents=True, exist_ok=True)
            args.report.write_text(json.dumps({"params": str(args.eval_only), "eval": ev,
                                               "tol": args.tol, "max_iters": args.max_iters,
                                               "scope": args.scope}, indent=1))
        return 0
    if args.trace_cache is not None and args.trace_cache.exists():
        with np.load(args.trace_cache, allow_pickle=False) as z:
            tr = L.Trace(z["X"], z["Y"], z["mask"])
        _log(f"loaded {len(tr.X)} samples from {args.trace_cache}")
    else:
        tr = gather(args, rng)
        if args.trace_cache is not None:
            args.trace_cache.parent.mkdir(parents=True, exist_ok=True)
            np.savez(args.trace_cache, X=tr.X, Y=tr.Y, mask=tr.mask)
    _log(f"{len(tr.X)} samples; target RMS "
         f"{np.sqrt(np.mean(tr.Y ** 2, axis=0)).round(4).tolist()}")
    p, hist, base = train(tr, args, rng)
    _log(f"held-out loss of the plain warm start (zero correction): {base:.4f}")
    ev = evaluate(p, args)
    meta = {"train_piles": args.train_piles, "boxes": [args.min_boxes, args.max_boxes],
            "frames": args.frames, "settle_frames": args.settle_frames,
            "keep_cap": args.keep_cap, "train_pyramids": args.train_pyramids,
            "train_stacks": args.train_stacks, "samples": int(len(tr.X)), "tol": args.tol,
            "max_iters": args.max_iters, "steps": args.steps, "batch": args.batch,
            "lr": args.lr, "scope": args.scope, "huber": args.huber, "clip": args.clip, "history": hist, "plain_val_loss": base, "eval": ev,
            "seconds": time.perf_counter() - t0}
    L.save_params(p, args.out, meta=json.dumps(meta))
    if args.report is not None:
        args.report.parent.mkdir(parents=True, exist_ok=True)
        args.report.write_text(json.dumps(meta, indent=1))
    _log(f"saved {args.out}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
