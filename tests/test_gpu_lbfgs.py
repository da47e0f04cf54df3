"""The device-resident L-BFGS memory (pfast_lbfgs_*: FP64 on the GPU, compensated dot
products) against the ring restatement of tests/lbfgs_ref.py, which is itself pinned to
the explicit inverse-BFGS matrix and the secant equation (tests/test_lbfgs_ring.py):
LM = 5 over iterations 2..11, through the wrap of the ring, on a small quadratic
(every direction also checked against the explicit matrix) and on N = 300000 nodes
(many partial sums per dot product).  Bound: 1e-12 relative."""
import ctypes as C

import numpy as np
import pytest

import lbfgs_ref as R

pytestmark = pytest.mark.gpu

LM = 5


def _gpu(n):
    import pfast_amd
    from pfast_amd import _lib
    ctx = pfast_amd.Context(0)
    h = C.c_void_p()
    assert _lib.lib().pfast_lbfgs_create(ctx.h, n, LM, C.byref(h)) == 0
    return ctx, h


def _direction(h, k, g1, g0, dk, alpha):
    from pfast_amd import _lib
    d = np.ascontiguousarray(dk, np.float64).copy()
    gam = C.c_double(0)
    rc = _lib.lib().pfast_lbfgs_direction(h, k, np.ascontiguousarray(g1), np.ascontiguousarray(g0),
                                          d, alpha, C.byref(gam))
    assert rc == 0, _lib.lib().pfast_last_error()
    return d, gam.value


def test_lbfgs_direction_small_quadratic_through_the_wrap():
    from pfast_amd import _lib
    n = 12
    ctx, h = _gpu(n)
    ring = R.Ring(n, LM)
    for k, g1, g0, dk, alpha in R.quadratic_run(n, LM, 11, 5):
        d_ref, gam_ref = ring.direction(k, g1, g0, dk, alpha)
        d, gam = _direction(h, k, g1, g0, dk, alpha)
        H = R.inverse_bfgs(ring.pairs(k), gam_ref)
        err = np.linalg.norm(d - d_ref) / np.linalg.norm(d_ref)
        print(f"iteration {k}: |d - d_ref| / |d_ref| = {err:.2e}", flush=True)
        assert err <= 1e-12 and abs(gam - gam_ref) <= 1e-12 * abs(gam_ref)
        assert np.linalg.norm(d - H @ g1) <= 1e-10 * np.linalg.norm(d)
    _lib.lib().pfast_lbfgs_destroy(h)
    ctx.close()


def test_lbfgs_direction_and_dot_large():
    from pfast_amd import _lib
    n = 300000
    rng = np.random.default_rng(9)
    ctx, h = _gpu(n)
    ring = R.Ring(n, LM)
    g0 = rng.standard_normal(n)
    dk = g0.copy()
    for k in range(2, 12):
        alpha = rng.uniform(0.1, 1.0)
        s = alpha * dk
        g1 = g0 - (0.8 * s + 0.1 * np.linalg.norm(s) / np.sqrt(n) * rng.standard_normal(n))
        d_ref, _ = ring.direction(k, g1, g0, dk, alpha)
        d, _ = _direction(h, k, g1, g0, dk, alpha)
        err = np.linalg.norm(d - d_ref) / np.linalg.norm(d_ref)
        print(f"iteration {k}: {err:.2e}", flush=True)
        assert err <= 1e-12
        dk, g0 = d_ref, g1
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    out = C.c_double(0)
    assert _lib.lib().pfast_lbfgs_dot(h, a, b, C.byref(out)) == 0
    exact = float(np.dot(a.astype(np.longdouble), b.astype(np.longdouble)))
    assert abs(out.value - exact) <= 1e-13 * np.abs(a * b).sum()
    _lib.lib().pfast_lbfgs_destroy(h)
    ctx.close()
