"""Residual-driven iteration on the checker: a substep iterated until the
dual residual max |lambda_new - lambda| / k <= tol equals a fixed-iteration
substep with the same count, and the count responds to the tolerance."""
import numpy as np

from avbd_amd import scenes
from oracle.oracle import oracle_solver_class
from tests.parity import diff_report, assert_bit_exact


def _resid_substep(s, tol, cap):
    sub = s.dt / s.substeps
    p = s._params(sub)
    p.iterations = cap
    rep, iters, r = s._h.substep_to_residual(p, s._color_dirty, tol)
    s._apply_report(rep, len(s._x))
    return iters, r


def _fixed_substep(s, iters):
    sub = s.dt / s.substeps
    p = s._params(sub)
    p.iterations = iters
    rep = s._h.substep(p, s._color_dirty)
    s._apply_report(rep, len(s._x))


def test_residual_iterations_equal_fixed_iterations():
    S = oracle_solver_class()
    a = scenes.pyramid(4, seed=21, solver_cls=S)
    b = scenes.pyramid(4, seed=21, solver_cls=S)
    a._flush()
    b._flush()
    counts = []
    for k in range(12):
        iters, r = _resid_substep(a, 1e-4, 60)
        _fixed_substep(b, iters)
        counts.append(iters)
        assert 1 <= iters <= 60
        assert (r <= 1e-4) or iters == 60
        assert_bit_exact(diff_report(a, b), f"substep {k}")
    assert max(counts) > 1


def test_tolerance_controls_iterations():
    S = oracle_solver_class()
    loose = scenes.pyramid(3, seed=22, solver_cls=S)
    tight = scenes.pyramid(3, seed=22, solver_cls=S)
    loose._flush()
    tight._flush()
    it_loose, _ = _resid_substep(loose, 1e9, 30)
    it_tight, r = _resid_substep(tight, 0.0, 30)
    assert it_loose == 1
    assert it_tight == 30 and r > 0.0
