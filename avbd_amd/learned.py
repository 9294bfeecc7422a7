"""Learned warm start of contact lambdas on the MI355X solver.

A 24 -> 64 -> 64 -> 4 MLP (SiLU; bf16 weights and layer inputs on MFMA,
fp32 accumulation) predicts, for every box-box contact group at emission and
every floor contact group, corrections to the plain warm start (the cached
or persistent state decayed by alpha * gamma, kernels_6dof.py:2695-2912 and
the substep prelude, kernels_6dof.py:3140-3260):

    y_r = slog(lambda_r converged) - slog(alpha gamma lambda_r cached), r = n, t, b
    y_3 = log10(k_n converged)     - log10(clamp(gamma k_n cached))

with slog(x) = sign(x) log(1 + |x|). A zero output leaves the plain warm
start untouched bit for bit, so the plain warm start is also the network's
fallback; a non-finite output falls back too. Only the initial dual state of
a substep changes: the iterations that follow are the plain solver's and are
judged by its own residual r = max over active rows of |dlambda| / k.

This module holds
  * the parameter layout and a numpy reference of the network (forward and
    backward with the same bf16 roundings the HIP kernels make), used to
    check the kernels;
  * thin wrappers of the library's standalone network entry points
    (avbd_mlp_forward / avbd_mlp_grad / avbd_mlp_train);
  * trace logging from the plain solver, training and the
    iterations-to-residual comparison of plain and learned warm starts.

No counterpart in the reference, which has no learned component.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _native as N

N_IN, N_HID, N_OUT = 24, 64, 4
# parameter layout (fp32, in order)
_SHAPES = [("W1", (N_IN, N_HID)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
           ("W3", (N_HID, N_OUT)), ("b3", (N_OUT,)), ("x_mean", (N_IN,)), ("x_std", (N_IN,)),
           ("y_mean", (N_OUT,)), ("y_std", (N_OUT,))]
_OFF = {}
_o = 0
for _name, _shape in _SHAPES:
    _OFF[_name] = (_o, _shape)
    _o += int(np.prod(_shape))
N_PARAMS = _o
N_WEIGHTS = _OFF["x_mean"][0]      # the trained part; the normalisation is fixed
del _o, _name, _shape

FEATURE_NAMES = (
    "n_x", "n_y", "n_z", "offa_x", "offa_y", "offa_z", "offb_x", "offb_y", "offb_z", "C0",
    "vrel_n", "vrel_t", "vrel_b", "inv_mass_a", "inv_mass_b", "slog_lam_n", "slog_lam_t",
    "slog_lam_b", "age", "log10_k_n", "log10_k_t", "log10_k_b", "mu", "kind")
MODES = {"plain": 0, "log": 1, "learned": 2}
WEIGHTS_PATH = Path(__file__).resolve().parent / "data" / "warm_start_mlp.npz"


# ---------------------------------------------------------------------------
# parameters
# ---------------------------------------------------------------------------
def unpack(params: np.ndarray) -> dict[str, np.ndarray]:
    """Views of the named parts; float64 input stays float64 (for finite
    differences), anything else becomes float32."""
    p = np.asarray(params)
    if p.dtype != np.float64:
        p = np.asarray(p, np.float32)
    if p.shape != (N_PARAMS,):
        raise ValueError(f"expected {N_PARAMS} parameters, got shape {p.shape}")
    return {k: p[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in _OFF.items()}


def pack(parts: dict[str, np.ndarray]) -> np.ndarray:
    out = np.zeros(N_PARAMS, np.float32)
    for k, (o, s) in _OFF.items():
        out[o:o + int(np.prod(s))] = np.asarray(parts[k], np.float32).reshape(-1)
    return out


def normalisation(X: np.ndarray, Y: np.ndarray, mask: np.ndarray):
    """Feature mean/std (std floored; constant features get std 1) and output
    scales: y_mean = 0 keeps "zero output = plain warm start", y_std = the
    RMS of the used targets."""
    X = np.asarray(X, np.float64)
    xm = X.mean(axis=0)
    xs = X.std(axis=0)
    xs = np.where(xs < 1e-6, 1.0, xs)
    ym = np.zeros(N_OUT)
    ys = np.ones(N_OUT)
    for o in range(N_OUT):
        used = (np.asarray(mask) >> o & 1).astype(bool)
        if used.any():
            ys[o] = max(float(np.sqrt(np.mean(np.asarray(Y, np.float64)[used, o] ** 2))), 1e-3)
    return xm, xs, ym, ys


def init_params(xm, xs, ym, ys, seed: int = 0) -> np.ndarray:
    """He-normal hidden layers, a small output layer (the network starts
    close to the plain warm start)."""
    rng = np.random.default_rng(seed)
    return pack({
        "W1": rng.normal(0.0, np.sqrt(2.0 / N_IN), (N_IN, N_HID)),
        "b1": np.zeros(N_HID),
        "W2": rng.normal(0.0, np.sqrt(2.0 / N_HID), (N_HID, N_HID)),
        "b2": np.zeros(N_HID),
        "W3": rng.normal(0.0, 0.01, (N_HID, N_OUT)),
        "b3": np.zeros(N_OUT),
        "x_mean": xm, "x_std": xs, "y_mean": ym, "y_std": ys})


def save_params(params: np.ndarray, path: Path = WEIGHTS_PATH, **meta) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez(path, params=np.asarray(params, np.float32),
             **{k: np.asarray(v) for k, v in meta.items()})


def load_meta(path: Path = WEIGHTS_PATH) -> dict:
    """The training record stored beside the parameters (JSON), or {}."""
    import json
    with np.load(path, allow_pickle=False) as z:
        return json.loads(str(z["meta"])) if "meta" in z.files else {}


def load_params(path: Path = WEIGHTS_PATH) -> np.ndarray:
    with np.load(path, allow_pickle=False) as z:
        p = np.asarray(z["params"], np.float32)
    unpack(p)  # shape check
    return p


# ---------------------------------------------------------------------------
# numpy reference of the network
# ---------------------------------------------------------------------------
def bf16_round(a) -> np.ndarray:
    """Round fp32 values to bfloat16 (nearest, ties to even), kept in fp32."""
    a = np.ascontiguousarray(a, np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).reshape(a.shape)


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward_np(params: np.ndarray, X: np.ndarray, bf16: bool = True):
    """De-normalised outputs [n, 4] and the intermediates. With bf16 the
    weights, the normalised input and both hidden activations are rounded to
    bfloat16 as the HIP kernels round them; sums in float64."""
    P = unpack(params)
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    if bf16:
        xn = (np.asarray(X, np.float32) - P["x_mean"].astype(np.float32)) / P["x_std"].astype(np.float32)
    else:
        xn = (np.asarray(X, np.float64) - P["x_mean"]) / P["x_std"]
    x0 = rnd(xn).astype(np.float64)
    W1, W2, W3 = (rnd(P[k]).astype(np.float64) for k in ("W1", "W2", "W3"))
    z1 = x0 @ W1 + P["b1"]
    a1 = rnd(_silu(z1).astype(np.float32)).astype(np.float64) if bf16 else _silu(z1)
    z2 = a1 @ W2 + P["b2"]
    a2 = rnd(_silu(z2).astype(np.float32)).astype(np.float64) if bf16 else _silu(z2)
    yn = a2 @ W3 + P["b3"]
    y = yn * P["y_std"] + P["y_mean"]
    return y, dict(x0=x0, z1=z1, a1=a1, z2=z2, a2=a2, yn=yn, W1=W1, W2=W2, W3=W3)


def loss_grad_np(params: np.ndarray, X: np.ndarray, Y: np.ndarray, mask: np.ndarray,
                 bf16: bool = True, huber: float = 0.0):
    """Masked loss on normalised targets (summed over the used outputs,
    averaged over samples): squared error d^2, or with huber > 0 the Huber
    loss (d^2 for |d| <= huber, huber (2|d| - huber) beyond), and its
    gradient w.r.t. the weights (N_WEIGHTS values, parameter order). With
    bf16 the backward rounds what the training kernel feeds its MFMA
    products to bfloat16 (the output error dy and the pre-activation
    gradients dz2, dz1, besides the forward's weights and activations), and
    the roundings are treated as identities for differentiation; bias
    gradients sum the unrounded values, as the kernel does."""
    P = unpack(params)
    n = len(X)
    _, c = forward_np(params, X, bf16)
    if bf16:
        yt = (np.asarray(Y, np.float32) - P["y_mean"].astype(np.float32)) / P["y_std"].astype(np.float32)
    else:
        yt = (np.asarray(Y, np.float64) - P["y_mean"]) / P["y_std"]
    m = np.stack([(np.asarray(mask) >> o & 1) for o in range(N_OUT)], axis=1).astype(np.float64)
    d = (c["yn"] - yt) * m
    if huber > 0:
        big = np.abs(d) > huber
        loss = float((np.sum(np.where(big, huber * (2 * np.abs(d) - huber), d * d))) / n)
        d = np.clip(d, -huber, huber)
    else:
        loss = float(np.sum(d * d) / n)
    dy = 2.0 * d / n
    if bf16:
        def rb(a):
            return bf16_round(np.asarray(a, np.float32)).astype(np.float64)
    else:
        def rb(a):
            return a
    g = {}
    g["W3"] = c["a2"].T @ rb(dy)
    g["b3"] = dy.sum(axis=0)
    dz2 = (dy @ c["W3"].T) * _silu_grad(c["z2"])
    g["W2"] = c["a1"].T @ rb(dz2)
    g["b2"] = dz2.sum(axis=0)
    dz1 = (rb(dz2) @ c["W2"].T) * _silu_grad(c["z1"])
    g["W1"] = c["x0"].T @ rb(dz1)
    g["b1"] = dz1.sum(axis=0)
    out = np.zeros(N_WEIGHTS)
    for k in ("W1", "b1", "W2", "b2", "W3", "b3"):
        o, s = _OFF[k]
        out[o:o + int(np.prod(s))] = g[k].reshape(-1)
    return loss, out


# ---------------------------------------------------------------------------
# the library's network entry points
# ---------------------------------------------------------------------------
def _lib():
    return N.load_library()


def _check(rc: int) -> None:
    if rc != 0:
        msg = _lib().avbd_last_error()
        raise N.AvbdError(f"avbd_mlp: {msg.decode() if msg else 'error'} (code {rc})")


def _f32(a, shape_last=None):
    a = np.ascontiguousarray(a, np.float32)
    if shape_last is not None and (a.ndim != 2 or a.shape[1] != shape_last):
        raise ValueError(f"expected an [n, {shape_last}] array, got {a.shape}")
    return a


def _mask(mask, n):
    m = np.ascontiguousarray(mask if mask is not None else np.full(n, 15), np.int32)
    if m.shape != (n,):
        raise ValueError("mask must hold one int per sample")
    return m


def forward_gpu(params: np.ndarray, X: np.ndarray, device: int = 0) -> np.ndarray:
    """De-normalised outputs of the MFMA inference kernel."""
    X = _f32(X, N_IN)
    p = _f32(params)
    if p.size != N_PARAMS:
        raise ValueError(f"expected {N_PARAMS} parameters")
    y = np.zeros((len(X), N_OUT), np.float32)
    _check(_lib().avbd_mlp_forward(int(device), len(X), N.fptr(X), N.fptr(p), N.fptr(y)))
    return y


def grad_gpu(params, X, Y, mask=None, huber: float = 0.0, device: int = 0):
    X, Y, p = _f32(X, N_IN), _f32(Y, N_OUT), _f32(params)
    m = _mask(mask, len(X))
    g = np.zeros(N_WEIGHTS, np.float32)
    loss = np.zeros(1, np.float32)
    _check(_lib().avbd_mlp_grad(int(device), len(X), N.fptr(X), N.fptr(Y), N.iptr(m), N.fptr(p),
                                C.c_float(huber), N.fptr(g), N.fptr(loss)))
    return float(loss[0]), g


def train_gpu(params, X, Y, mask=None, steps: int = 1000, batch: int = 4096, lr: float = 1e-3,
              seed: int = 0, huber: float = 0.0, device: int = 0):
    """`steps` Adam steps on the GPU; returns (new params, per-step loss)."""
    X, Y = _f32(X, N_IN), _f32(Y, N_OUT)
    p = np.array(params, np.float32, copy=True)
    m = _mask(mask, len(X))
    hist = np.zeros(max(int(steps), 1), np.float32)
    _check(_lib().avbd_mlp_train(int(device), len(X), N.fptr(X), N.fptr(Y), N.iptr(m), N.fptr(p),
                                 int(steps), int(batch), C.c_float(lr), C.c_float(huber),
                                 C.c_uint(seed & 0xFFFFFFFF), N.fptr(hist)))
    return p, hist[:int(steps)]


# ---------------------------------------------------------------------------
# traces, training and evaluation
# ---------------------------------------------------------------------------
def pile(n_boxes: int, seed: int, **kw):
    """A pile of `n_boxes` boxes dropped from a lattice (the 10k-box pile's
    generator at another size): lattice side ceil(cbrt(n)) in x and z."""
    from .scenes import random_pile
    side = int(np.ceil(n_boxes ** (1.0 / 3.0)))
    ny = int(np.ceil(n_boxes / (side * side)))
    return random_pile(n_boxes, seed=seed, lattice=(side, side, ny), **kw)


@dataclass
class Trace:
    X: np.ndarray
    Y: np.ndarray
    mask: np.ndarray


def log_trace(solver, frames: int, tol: float = 1e-4, max_iters: int = 200,
              keep: float = 1.0, seed: int = 0, skip_frames: int = 0) -> Trace:
    """Run `frames` frames with every substep iterated to the residual
    tolerance, the plain warm start, and samples logged; keep a random
    fraction `keep` of each substep's samples."""
    rng = np.random.default_rng(seed)
    solver.set_learned_warm_start("log")
    xs, ys, ms = [], [], []

    def grab(_k):
        X, Y, M = solver.learned_samples()
        if len(X) == 0:
            return
        if keep < 1.0:
            sel = rng.random(len(X)) < keep
            X, Y, M = X[sel], Y[sel], M[sel]
        xs.append(X)
        ys.append(Y)
        ms.append(M)

    for f in range(frames):
        solver.step_to_residual(tol, max_iters, on_substep=grab if f >= skip_frames else None)
    solver.set_learned_warm_start("plain")
    if not xs:
        return Trace(np.zeros((0, N_IN), np.float32), np.zeros((0, N_OUT), np.float32),
                     np.zeros(0, np.int32))
    return Trace(np.concatenate(xs), np.concatenate(ys), np.concatenate(ms))


def iterations_to_residual(solver, frames: int, mode: str = "plain", params=None,
                           tol: float = 1e-4, max_iters: int = 200,
                           scope: str = "all", residuals: list | None = None) -> np.ndarray:
    """Iterations each substep needed to reach the residual tolerance
    (capped at max_iters), under the plain or the learned warm start. The
    residual each substep ended at is appended to `residuals` if given."""
    solver.set_learned_warm_start(mode, params, scope)
    its = []
    for _ in range(frames):
        for i, r in solver.step_to_residual(tol, max_iters):
            its.append(i)
            if residuals is not None:
                residuals.append(r)
    return np.asarray(its, np.int64)


def _state(solver) -> dict:
    """Everything of a solver that carries from one substep to the next."""
    h, nb = solver._h, len(solver._x)
    keys, st = h.export_hash()
    return {"bodies": h.read_bodies(nb), "rows": h.read_rows(solver._gpu_pool_n_static),
            "keys": keys, "hash": st, "colors": h.read_colors(nb),
            "n_colors": solver._n_active_colors, "dirty": solver._color_dirty}


def _load(solver, bodies, rows, keys, hash_state, colors, n_colors, dirty) -> None:
    h = solver._h
    for field, key in enumerate(("x", "q", "v", "w", "prev_v", "prev_w")):
        h.write_bodies(field, 0, bodies[key])
    h.write_static_rows(rows["lambda"], rows["penalty"], rows["active"], rows["was_static"])
    h.import_hash(keys, hash_state)
    if not dirty:   # a dirty coloring is recomputed by the next substep
        h.write_colors(colors, n_colors)
    solver._color_dirty = dirty
    solver._n_active_colors = n_colors


def perfect_start_bound(solver, frames: int, tol: float = 1e-4, max_iters: int = 200) -> dict:
    """What any warm start could gain, measured: every substep of `frames`
    frames runs plain to the residual tolerance (cap max_iters), then again
    from the same starting state with the state that run ended at written in
    as its warm start (static rows' and cached contacts' lambda, penalty and
    stick state; penalties divided by the prelude's gamma, lambdas by its
    alpha * gamma when it decays them). The trajectory continues from the
    plain run. Returns the iterations of both runs per substep."""
    if solver._dirty:
        solver._flush()
    n_b = len(solver._x)
    sub_dt = solver.dt / solver.substeps if solver.substeps > 1 else solver.dt
    decay_lam = 1.0 if solver.post_stabilize else solver.alpha * solver.gamma
    plain, perfect, rp, rq = [], [], [], []
    for _ in range(frames):
        for _k in range(max(1, solver.substeps)):
            s0 = _state(solver)
            params = solver._params(sub_dt)
            params.iterations = int(max_iters)
            rep, it0, r0 = solver._h.substep_to_residual(params, solver._color_dirty, tol)
            solver._apply_report(rep, n_b)
            s1 = _state(solver)
            rows = dict(s1["rows"])
            rows["lambda"] = (rows["lambda"] / np.float32(decay_lam)).astype(np.float32)
            rows["penalty"] = (rows["penalty"] / np.float32(solver.gamma)).astype(np.float32)
            st = s1["hash"].copy()
            st[:, 0:3] /= np.float32(decay_lam)
            st[:, 3:6] /= np.float32(solver.gamma)
            _load(solver, s0["bodies"], rows, s1["keys"], st, s0["colors"], s0["n_colors"],
                  s0["dirty"])
            rep2, it1, r1 = solver._h.substep_to_residual(params, solver._color_dirty, tol)
            plain.append(it0)
            perfect.append(it1)
            rp.append(r0)
            rq.append(r1)
            # continue from the plain run's end state
            _load(solver, s1["bodies"], s1["rows"], s1["keys"], s1["hash"], s1["colors"],
                  s1["n_colors"], s1["dirty"])
    plain, perfect = np.asarray(plain), np.asarray(perfect)
    return {"plain_mean": float(plain.mean()), "perfect_mean": float(perfect.mean()),
            "plain_capped": int((plain >= max_iters).sum()),
            "perfect_capped": int((perfect >= max_iters).sum()), "substeps": int(len(plain)),
            "plain_resid_median": float(np.median(rp)),
            "perfect_resid_median": float(np.median(rq)), "plain": plain, "perfect": perfect}


def undo_decay_params(like: np.ndarray, alpha: float = 0.99, gamma: float = 0.99) -> np.ndarray:
    """A constant network: every normal-row correction is log(alpha gamma),
    i.e. the start is the cached lambda without the prelude's decay (to
    first order in 1/|lambda|). The trivial rule a learned start has to
    beat."""
    p = np.array(like, np.float32, copy=True)
    p[:N_WEIGHTS] = 0.0
    P = unpack(p)
    P["b3"][0] = np.log(alpha * gamma) / P["y_std"][0] - P["y_mean"][0] / P["y_std"][0]
    return p


def compare(builder, frames: int, params, tol: float = 1e-4, max_iters: int = 200,
            scope: str = "all", settle_frames: int = 0) -> dict:
    """Iterations to residual from the same scene, plain vs learned warm
    start. With settle_frames, both copies first run that many ordinary
    frames (fixed iterations, plain warm start), so the measurement starts
    from the same settled state."""
    def settled():
        s = builder()
        for _ in range(settle_frames):
            s.step()
        return s
    rp, rl = [], []
    plain = iterations_to_residual(settled(), frames, "plain", None, tol, max_iters, "all", rp)
    learned = iterations_to_residual(settled(), frames, "learned", params, tol, max_iters, scope,
                                     rl)
    return {"plain_mean": float(plain.mean()), "learned_mean": float(learned.mean()),
            # where the substeps stopped: below tol, or the residual at the cap
            "plain_resid_median": float(np.median(rp)), "learned_resid_median": float(np.median(rl)),
            "plain_total": int(plain.sum()), "learned_total": int(learned.sum()),
            "plain_capped": int((plain >= max_iters).sum()),
            "learned_capped": int((learned >= max_iters).sum()),
            "substeps": int(len(plain)), "plain": plain, "learned": learned}
