"""avbd_mlp_grad_noalloc (no counterpart in the reference) through the C ABI:
the same gradient and loss as avbd_mlp_grad, bit for bit (both run the same
kernels, which sum the workgroups' gradient rows in a fixed order), on two
calls in a row, the second smaller, so that nothing of the first call's
staging leaks into the second; sizes past the static limit are refused."""
import ctypes as C

import numpy as np
import pytest

from avbd_amd import _native as N
from avbd_amd import learned as L

pytestmark = pytest.mark.gpu


def _case(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0.0, 1.0, (n, L.N_IN)).astype(np.float32)
    Y = rng.normal(0.0, 1.0, (n, L.N_OUT)).astype(np.float32)
    mask = rng.integers(1, 16, n).astype(np.int32)
    p = L.init_params(np.zeros(L.N_IN), np.ones(L.N_IN), np.zeros(L.N_OUT), np.ones(L.N_OUT),
                      seed=seed)
    return X, Y, mask, np.ascontiguousarray(p, np.float32)


def _grad(fn, X, Y, mask, p, huber):
    lib = N.load_library()
    g = np.zeros(L.N_WEIGHTS, np.float32)
    loss = np.zeros(1, np.float32)
    rc = getattr(lib, fn)(0, len(X), N.fptr(X), N.fptr(Y), N.iptr(mask), N.fptr(p),
                          C.c_float(huber), N.fptr(g), N.fptr(loss))
    if rc != 0:
        raise N.AvbdError(lib.avbd_last_error().decode())
    return g, float(loss[0])


@pytest.mark.parametrize("calls", [((3000, 0.0), (700, 0.5))])
def test_noalloc_grad_matches_allocating_grad(calls):
    for k, (n, huber) in enumerate(calls):
        X, Y, mask, p = _case(n, seed=17 + k)
        print(f"call {k}: n={n} huber={huber}", flush=True)
        g_ref, l_ref = _grad("avbd_mlp_grad", X, Y, mask, p, huber)
        g, l = _grad("avbd_mlp_grad_noalloc", X, Y, mask, p, huber)
        assert np.isfinite(g).all() and np.abs(g_ref).max() > 0
        np.testing.assert_array_equal(g, g_ref)
        assert l == l_ref


def test_noalloc_grad_refuses_sizes_past_its_limit():
    X, Y, mask, p = _case(4097, seed=3)
    with pytest.raises(N.AvbdError, match="static scratch limit"):
        _grad("avbd_mlp_grad_noalloc", X, Y, mask, p, 0.0)
    # a call within the limit works after the refused one
    g, _ = _grad("avbd_mlp_grad_noalloc", X[:64], Y[:64], mask[:64], p, 0.0)
    assert np.isfinite(g).all()
