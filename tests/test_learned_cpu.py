"""The numpy reference of the warm-start network (avbd_amd.learned), which
the HIP kernels are checked against: its forward against torch's own layers,
its backward against central finite differences, its bf16 rounding against
torch's bfloat16 conversion."""
import numpy as np
import pytest
import torch

from avbd_amd import learned as L


def _data(n=96, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(0.0, 1.0, (n, L.N_IN)).astype(np.float32) * 3.0 + 1.0
    Y = rng.normal(0.0, 0.5, (n, L.N_OUT)).astype(np.float32)
    M = rng.choice([9, 15], size=n).astype(np.int32)
    return X, Y, M


def _params(X, Y, M, seed=1, w3=0.3):
    p = L.init_params(*L.normalisation(X, Y, M), seed=seed)
    P = L.unpack(p)
    rng = np.random.default_rng(seed + 7)
    P["W3"][:] = rng.normal(0.0, w3, P["W3"].shape)
    for k in ("b1", "b2", "b3"):
        P[k][:] = rng.normal(0.0, 0.1, P[k].shape)
    return p


def test_layout_sizes():
    assert L.N_PARAMS == 24 * 64 + 64 + 64 * 64 + 64 + 64 * 4 + 4 + 24 + 24 + 4 + 4 == 6076
    assert L.N_WEIGHTS == 6020
    p = np.arange(L.N_PARAMS, dtype=np.float32)
    assert np.array_equal(L.pack(L.unpack(p)), p)
    assert L.unpack(p)["W2"][1, 0] == 24 * 64 + 64 + 64   # row-major [in][out]


def test_bf16_round_matches_torch_and_ties_to_even():
    rng = np.random.default_rng(3)
    a = np.concatenate([rng.normal(0, 1, 4000), rng.normal(0, 1e4, 1000), rng.normal(0, 1e-4, 1000),
                        [0.0, -0.0, 1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(1.0 + 2.0 ** -8)]])
    a = a.astype(np.float32)
    ours = L.bf16_round(a)
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))
    # halfway cases go to the even mantissa
    assert L.bf16_round(np.float32(1.0 + 2.0 ** -8)) == np.float32(1.0)
    assert L.bf16_round(np.float32(1.0 + 3 * 2.0 ** -8)) == np.float32(1.0 + 2.0 ** -6)


def test_forward_matches_torch_layers():
    X, Y, M = _data()
    p = _params(X, Y, M)
    P = L.unpack(p)
    y, _ = L.forward_np(p, X, bf16=False)
    t = lambda a: torch.from_numpy(np.asarray(a, np.float64))
    xn = (t(X) - t(P["x_mean"])) / t(P["x_std"])
    h = torch.nn.functional.silu(xn @ t(P["W1"]) + t(P["b1"]))
    h = torch.nn.functional.silu(h @ t(P["W2"]) + t(P["b2"]))
    ref = (h @ t(P["W3"]) + t(P["b3"])) * t(P["y_std"]) + t(P["y_mean"])
    assert np.allclose(y, ref.numpy(), rtol=1e-12, atol=1e-12)


def test_zero_network_outputs_zero():
    X, Y, M = _data()
    p = L.init_params(*L.normalisation(X, Y, M))
    P = L.unpack(p)
    for k in ("W1", "b1", "W2", "b2", "W3", "b3"):
        P[k][:] = 0.0
    y, _ = L.forward_np(p, X)
    assert np.array_equal(y, np.zeros_like(y))


def test_backward_matches_finite_differences():
    X, Y, M = _data(48)
    p = _params(X, Y, M).astype(np.float64)
    loss, g = L.loss_grad_np(p, X, Y, M, bf16=False)
    assert loss > 0
    rng = np.random.default_rng(5)
    # a sample of coordinates from every weight block, plus one random direction
    idx = []
    for k in ("W1", "b1", "W2", "b2", "W3", "b3"):
        o, s = L._OFF[k]
        idx += list(o + rng.choice(int(np.prod(s)), size=min(6, int(np.prod(s))), replace=False))
    eps = 1e-5
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (L.loss_grad_np(pp, X, Y, M, bf16=False)[0] - L.loss_grad_np(pm, X, Y, M, bf16=False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    d = np.zeros_like(p)
    d[:L.N_WEIGHTS] = rng.normal(0, 1, L.N_WEIGHTS)
    fd = (L.loss_grad_np(p + eps * d, X, Y, M, bf16=False)[0]
          - L.loss_grad_np(p - eps * d, X, Y, M, bf16=False)[0]) / (2 * eps)
    assert abs(fd - g @ d[:L.N_WEIGHTS]) <= 1e-6 * max(1.0, abs(fd))


def test_mask_drops_unused_outputs():
    """A dropped output contributes neither loss nor gradient: perturbing its
    target changes nothing."""
    X, Y, M = _data(32)
    M[:] = 9
    p = _params(X, Y, M)
    l0, g0 = L.loss_grad_np(p, X, Y, M)
    Y2 = Y.copy()
    Y2[:, 1:3] += 5.0
    l1, g1 = L.loss_grad_np(p, X, Y2, M)
    assert l0 == l1 and np.array_equal(g0, g1)


def test_params_round_trip(tmp_path):
    X, Y, M = _data()
    p = _params(X, Y, M)
    f = tmp_path / "w.npz"
    L.save_params(p, f, note="x")
    assert np.array_equal(L.load_params(f), p)
    with pytest.raises(ValueError):
        L.unpack(p[:-1])


def test_huber_backward_matches_finite_differences():
    """Huber loss (threshold 0.3 on normalised errors): gradient against
    central differences, and the loss equals the squared error where every
    error is under the threshold."""
    X, Y, M = _data(48, seed=4)
    p = _params(X, Y, M, seed=2).astype(np.float64)
    loss, g = L.loss_grad_np(p, X, Y, M, bf16=False, huber=0.3)
    rng = np.random.default_rng(6)
    d = np.zeros_like(p)
    d[:L.N_WEIGHTS] = rng.normal(0, 1, L.N_WEIGHTS)
    eps = 1e-6
    fd = (L.loss_grad_np(p + eps * d, X, Y, M, bf16=False, huber=0.3)[0]
          - L.loss_grad_np(p - eps * d, X, Y, M, bf16=False, huber=0.3)[0]) / (2 * eps)
    assert abs(fd - g @ d[:L.N_WEIGHTS]) <= 1e-5 * max(1.0, abs(fd))
    big = L.loss_grad_np(p, X, Y, M, bf16=False, huber=1e6)
    mse = L.loss_grad_np(p, X, Y, M, bf16=False)
    assert big[0] == pytest.approx(mse[0], rel=1e-12) and np.allclose(big[1], mse[1], rtol=1e-12)
    assert loss < mse[0]


def test_undo_decay_network_outputs_log_alpha_gamma():
    X, Y, M = _data()
    p = L.undo_decay_params(_params(X, Y, M))
    y, _ = L.forward_np(p, X, bf16=False)
    assert np.allclose(y[:, 0], np.log(0.99 * 0.99), rtol=1e-6) and np.all(y[:, 1:] == 0)


def test_perfect_start_bound_keeps_the_plain_trajectory():
    """The warm-start bound on the checker (pyramid of 10 cubes from rest, 2
    frames, first substep with a dirty coloring): its plain iteration counts
    and the state carried to the next substep (bodies, static rows, contact
    cache) are those of an ordinary run to the residual, bit for bit, and so
    are the frames after it; the converged start needs fewer iterations in
    total. The contact rows of the pool are rebuilt from the cache by every
    substep and are not compared."""
    from avbd_amd import scenes
    from oracle.oracle import oracle_solver_class
    O = oracle_solver_class()
    a = scenes.pyramid(4, seed=3, solver_cls=O)
    b = scenes.pyramid(4, seed=3, solver_cls=O)
    r = L.perfect_start_bound(a, 2, 1e-4, 200)
    its = L.iterations_to_residual(b, 2)
    assert np.array_equal(r["plain"], its)
    assert np.array_equal(a.positions(), b.positions())
    n_static = a._gpu_pool_n_static
    assert np.array_equal(a.lambdas()[:n_static], b.lambdas()[:n_static])
    for x, y in zip(a._h.export_hash(), b._h.export_hash()):
        assert np.array_equal(x, y)
    for _ in range(2):
        a.step()
        b.step()
    assert np.array_equal(a.positions(), b.positions())
    assert r["substeps"] == 8 and r["perfect"].sum() < r["plain"].sum()
