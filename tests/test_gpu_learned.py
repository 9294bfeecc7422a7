"""Learned warm start on the GPU: the MFMA inference and the training
kernel against the numpy reference of the network (avbd_amd.learned), and
the solver-side guarantees (a zero network is the plain warm start bit for
bit; logging does not perturb the trajectory)."""
import numpy as np
import pytest

from avbd_amd import learned as L
from avbd_amd import scenes
from avbd_amd.solver_6dof import Solver6DOF
from tests.parity import assert_bit_exact, diff_report, one_substep

pytestmark = pytest.mark.gpu


def _data(n, seed=0):
    rng = np.random.default_rng(seed)
    X = (rng.normal(0.0, 1.0, (n, L.N_IN)) * rng.uniform(0.1, 5.0, L.N_IN)
         + rng.normal(0.0, 2.0, L.N_IN)).astype(np.float32)
    Y = rng.normal(0.0, 0.5, (n, L.N_OUT)).astype(np.float32)
    M = rng.choice([9, 15], size=n).astype(np.int32)
    return X, Y, M


def _params(X, Y, M, seed=1):
    p = L.init_params(*L.normalisation(X, Y, M), seed=seed)
    P = L.unpack(p)
    rng = np.random.default_rng(seed + 7)
    P["W3"][:] = rng.normal(0.0, 0.3, P["W3"].shape)
    for k in ("b1", "b2", "b3"):
        P[k][:] = rng.normal(0.0, 0.1, P[k].shape)
    return p


@pytest.mark.parametrize("n", [1, 63, 1037, 200_003])
def test_mfma_forward_matches_numpy(n):
    """bf16 operands, fp32 accumulation: the kernel and the reference round
    the same values to bf16; they may differ where an fp32 sum lands next to
    a bf16 rounding boundary. Tolerance (outputs in units of y_std): 99.9 %
    within 1e-4, all within 2e-2. n = 200,003 walks every workgroup over
    several tiles and ends on a ragged one."""
    X, Y, M = _data(n, seed=n)
    p = _params(X, Y, M)
    ys = L.unpack(p)["y_std"]
    got = L.forward_gpu(p, X)
    ref, _ = L.forward_np(p, X)
    err = np.abs(got - ref) / ys
    print(n, "max", err.max(), "q999", np.quantile(err, 0.999), flush=True)
    assert np.quantile(err, 0.999) <= 1e-4
    assert err.max() <= 2e-2


@pytest.mark.parametrize("n", [64, 777])
def test_training_gradient_matches_numpy(n):
    """Loss and weight gradient of the training kernel against the numpy
    backward on the same bf16-rounded weights and layer inputs (squared
    error, and the Huber loss on the ragged case): loss within
    1e-4 relative, gradient within 1e-3 relative L2 and 1e-2 per block."""
    X, Y, M = _data(n, seed=n + 1)
    p = _params(X, Y, M, seed=3)
    huber = 0.0 if n == 64 else 0.5   # the ragged case with the Huber loss
    loss, g = L.grad_gpu(p, X, Y, M, huber=huber)
    # the workgroups' rows are summed in a fixed order: a repeat call gives
    # the same bits
    loss2, g2 = L.grad_gpu(p, X, Y, M, huber=huber)
    assert loss2 == loss and np.array_equal(g2, g)
    loss_ref, g_ref = L.loss_grad_np(p, X, Y, M, huber=huber)
    print(n, loss, loss_ref, np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref), flush=True)
    assert abs(loss - loss_ref) <= 1e-4 * loss_ref
    assert np.linalg.norm(g - g_ref) <= 1e-3 * np.linalg.norm(g_ref)
    for k in ("W1", "b1", "W2", "b2", "W3", "b3"):
        o, s = L._OFF[k]
        a, b = g[o:o + int(np.prod(s))], g_ref[o:o + int(np.prod(s))]
        assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b) + 1e-9, k


def test_training_fits_a_smooth_target():
    """Adam on the GPU fits a smooth function of the features: the held-out
    loss falls below a tenth of the starting one."""
    rng = np.random.default_rng(11)
    X = rng.normal(0.0, 1.0, (20_000, L.N_IN)).astype(np.float32)
    Y = np.stack([np.tanh(X[:, 0] + 0.5 * X[:, 1]), np.sin(X[:, 2]) * X[:, 3],
                  0.3 * X[:, 4] ** 2, X[:, 5] - X[:, 6]], axis=1).astype(np.float32)
    M = np.full(len(X), 15, np.int32)
    tr, te = slice(0, 16_000), slice(16_000, None)
    p0 = L.init_params(*L.normalisation(X[tr], Y[tr], M[tr]), seed=2)
    l0 = L.loss_grad_np(p0, X[te], Y[te], M[te])[0]
    p1, hist = L.train_gpu(p0, X[tr], Y[tr], M[tr], steps=1500, batch=512, lr=2e-3, seed=4)
    l1 = L.loss_grad_np(p1, X[te], Y[te], M[te])[0]
    print("held-out loss", l0, "->", l1, "minibatch", hist[:3], hist[-3:], flush=True)
    assert np.all(np.isfinite(hist)) and l1 < 0.1 * l0
    # the normalisation is not trained
    assert np.array_equal(p1[L.N_WEIGHTS:], p0[L.N_WEIGHTS:])


def _zero_params():
    X, Y, M = _data(256)
    p = L.init_params(*L.normalisation(X, Y, M))
    p[:L.N_WEIGHTS] = 0.0
    return p


def _out_of_range_params():
    """Every output far beyond the trusted correction range."""
    p = _zero_params()
    L.unpack(p)["b3"][:] = 1e4
    return p


@pytest.mark.parametrize("mode", ["learned", "log", "out-of-range"])
def test_zero_network_and_logging_keep_the_plain_trajectory(mode):
    """A network that outputs zero leaves the plain warm start untouched,
    one whose outputs are all out of range falls back to it, and logging only
    reads: each trajectory equals the plain one bit for bit (pyramid of 55
    cubes, 24 substeps)."""
    a = scenes.pyramid(5, seed=13, solver_cls=Solver6DOF)
    b = scenes.pyramid(5, seed=13, solver_cls=Solver6DOF)
    a._flush()
    if mode == "out-of-range":
        b.set_learned_warm_start("learned", _out_of_range_params())
    else:
        b.set_learned_warm_start(mode, _zero_params())
    b._flush()
    for k in range(24):
        one_substep(a)
        one_substep(b)
        assert_bit_exact(diff_report(a, b), f"{mode} substep {k}")


def test_logged_samples_of_a_resting_stack():
    """Four 1 m cubes at rest: the 32 floor corner groups (normal +y, offsets
    at the corners, the four bottom ones on the floor), then one sample per
    box-box contact group (3 rows each with friction, all with mask 15);
    normals +-y, body offsets on the
    cubes' horizontal faces, inverse masses 1/1000, friction coefficient 0.5;
    the age counts the substeps a contact has persisted."""
    s = scenes.cube_stack(4, seed=0, solver_cls=Solver6DOF)
    s.set_learned_warm_start("log")
    ages, counts = [], []
    for k in range(12):
        one_substep(s)
        X, Y, M = s.learned_samples()
        # 8 floor corner groups per cube first, then the box-box groups
        assert len(X) > 32 and np.all(X[:32, 23] == 1.0) and np.all(X[32:, 23] == 0.0)
        floor = X[:32]
        assert np.array_equal(floor[:, :3], np.tile([0.0, 1.0, 0.0], (32, 1)).astype(np.float32))
        assert np.allclose(np.abs(floor[:, 3:6]), 0.5) and np.all(floor[:, 6:9] == 0)
        assert np.allclose(np.sort(floor[:, 9])[:4], 0.0, atol=2e-3)  # bottom corners on the floor
        X, Y, M = X[32:], Y[32:], M[32:]
        live = min(s._h.read_counters()[0], s._gpu_pool_n_capacity)
        dyn_rows = live - s._gpu_pool_n_static
        counts.append(len(X))
        print(k, len(X), dyn_rows, X[:, 18].tolist(), flush=True)
        assert dyn_rows == 3 * len(X) and len(X) > 0 and np.all(M == 15)
        assert np.allclose(np.abs(X[:, 1]), 1.0, atol=1e-2)
        assert np.allclose(np.abs(X[:, [4, 7]]), 0.5, atol=2e-2)
        assert np.all(np.abs(X[:, 3:9]) <= 0.52)
        assert np.allclose(X[:, 13:15], 1e-3) and np.allclose(X[:, 22], 0.5)
        assert np.all(np.isfinite(Y))
        ages.append(X[:, 18].max())
    assert ages == sorted(ages) and ages[-1] >= 8, (ages, counts)


def test_shipped_network_keeps_the_converged_answer():
    """The committed network on a resting pyramid of 91 cubes, every substep
    iterated to the residual 1e-4 (cap 200): states stay finite, every
    substep that stops early meets the tolerance, and after 30 frames the
    cubes sit within 5 mm of where the plain warm start puts them."""
    p = L.load_params()
    a = scenes.pyramid(6, seed=21, solver_cls=Solver6DOF)
    b = scenes.pyramid(6, seed=21, solver_cls=Solver6DOF)
    b.set_learned_warm_start("learned", p, scope=L.load_meta().get("scope", "all"))
    ia, ib = [], []
    for _ in range(30):
        ia += a.step_to_residual(1e-4, 200)
        rb = b.step_to_residual(1e-4, 200)
        ib += rb
        assert all(np.isfinite(r) and (r <= 1e-4 or it == 200) for it, r in rb)
    xa, xb = a.positions(), b.positions()
    assert np.all(np.isfinite(xb))
    print("iterations plain", np.mean([i for i, _ in ia]), "learned", np.mean([i for i, _ in ib]),
          "max position difference", np.abs(xa - xb).max(), flush=True)
    assert np.abs(xa - xb).max() < 5e-3
