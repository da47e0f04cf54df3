"""Optimiser support of the drivers (host functions of libpfast: no GPU needed)
and the FFT-regularisation restatement used as the GPU regularisation's checker.

Pinned against the reference's own routines compiled from its sources
(oracle/_ref; tests marked `ref`, skipped where absent) bit for bit, and against
closed forms that do not need the reference.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O


def lib():
    from pfast_amd._lib import lib as L
    return L()


def vecs(seed, n):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(n)


# ---------------------------------------------------------------- closed forms
def test_beta_hs_closed_form():
    g1, g0, d = vecs(1, 37)
    y = g1 - g0
    assert lib().pfast_beta_hs(g1, g0, d, 37) == pytest.approx(g1 @ y / (-d @ y), rel=1e-13)


def test_beta_cgdescent_closed_form():
    g1, g0, d = vecs(2, 29)
    y = -g1 + g0
    dty, ny = d @ y, y @ y
    want = ((y - 2 * d * ny / dty) * (-g1)).sum() / dty
    assert lib().pfast_beta_cgdescent(g1, g0, d, 29) == pytest.approx(want, rel=1e-12)


def test_gd2sk_stores_pair_in_slot():
    n, lm, it, a1 = 23, 4, 7, 0.37
    g1, g0, d = vecs(3, n)
    rho, sk, yk = np.zeros(lm), np.zeros(lm * n), np.zeros(lm * n)
    gamma = lib().pfast_gd2sk(g1, g0, d, a1, rho, sk, yk, it, lm, n)
    slot = (it - 2) % lm
    s, y = a1 * d, -g1 + g0
    assert np.array_equal(sk[slot * n:(slot + 1) * n], s)
    assert np.array_equal(yk[slot * n:(slot + 1) * n], y)
    assert rho[slot] == pytest.approx(1 / (s @ y), rel=1e-13)
    assert gamma == pytest.approx((s @ y) / (y @ y), rel=1e-13)


def test_cubic_interp_recovers_minimum_of_a_cubic():
    # phi(a) = (a - 2)^3 / 3 - 4 (a - 2): local minimum at a = 4
    phi = lambda a: (a - 2) ** 3 / 3 - 4 * (a - 2)  # noqa: E731
    dphi = lambda a: (a - 2) ** 2 - 4  # noqa: E731
    a0, a1 = 3.0, 5.5
    got = lib().pfast_cubic_interp(a0, a1, phi(a0), phi(a1), dphi(a0), dphi(a1))
    assert got == pytest.approx(4.0, rel=1e-12)


def test_magnitude_and_mean():
    g = np.array([0.5, -4.0, 2.0, 2.5])
    assert lib().pfast_magnitude_of(g, 4) == 0.25
    assert lib().pfast_mean_of(g, 4) == 1.0 / abs(g.mean())


def test_wavenumbers_even_is_fft_order_odd_is_shifted():
    k = O.ref_wavenumbers(8, 2.0)
    assert np.allclose(k, 2 * np.pi * np.fft.fftfreq(8, 2.0))
    k5 = O.ref_wavenumbers(5, 1.0)  # -pi + 2 pi j / 5, halves of 2 swapped
    base = -np.pi + 2 * np.pi * np.arange(5) / 5
    assert np.array_equal(k5, base[[2, 3, 0, 1, 4]])
    assert not np.any(k5 == 0.0)  # no DC bin for odd n


def test_regu_np_constant_field_is_scaled_dc():
    # a constant field has all its energy in bin 0, whose wavenumber for even n is 0
    Dj = np.full((4, 6, 8), 3.0)
    out = O.regu_np(Dj, np.full(Dj.shape, 2000.0), (5.0, 7.0, 9.0), (10.0, 10.0, 10.0))
    assert np.allclose(out, 3.0, rtol=1e-13)


# ---------------------------------------------------------------- reference
ref = pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built")


@ref
@pytest.mark.ref
def test_helpers_bitexact_vs_reference():
    nr, nc = 7, 9
    n = nr * nc
    g1, g0, d = vecs(5, n)
    R2 = O.ref2d()
    assert lib().pfast_beta_hs(g1, g0, d, n) == R2.ref_beta_HS2d(g1, g0, d, nr, nc)
    assert lib().pfast_beta_cgdescent(g1, g0, d, n) == R2.ref_beta_CGDESCENT2d(g1, g0, d, nr, nc)
    assert lib().pfast_magnitude_of(g1, n) == R2.ref_magnitudeOf2d(g1, nr, nc)
    # cases on the reference's defined path (its a == 0 branch can never hold, so it
    # returns an undefined value there; this library returns the secant step)
    for args in [(0.0, 1.0, 2.0, 1.5, -3.0, 0.5), (1.0, 3.0, 5.0, 4.5, -2.0, 1.0),
                 (0.0, 2.0, 1.0, 0.8, -1.0, -0.2)]:
        assert lib().pfast_cubic_interp(*args) == R2.ref_cubicInterp(*args)
    lm = 3
    for it in (2, 3, 6):
        r1, s1, y1 = np.zeros(lm), np.zeros(lm * n), np.zeros(lm * n)
        r2, s2, y2 = np.zeros(lm), np.zeros(lm * n), np.zeros(lm * n)
        a = lib().pfast_gd2sk(g1, g0, d, 0.7, r1, s1, y1, it, lm, n)
        b = R2.ref_gd2sk2d(g1, g0, d, 0.7, r2, s2, y2, it, lm, nr, nc)
        assert a == b and np.array_equal(r1, r2) and np.array_equal(s1, s2) and np.array_equal(y1, y2)


@ref
@pytest.mark.ref
@pytest.mark.parametrize("shape", [(6, 7, 9), (5, 8, 6)])
def test_regu_np_vs_reference_3d(shape):
    rng = np.random.default_rng(11)
    Dj = rng.standard_normal(shape)
    vp = 2000 + 100 * rng.random(shape)
    vp[0, :2, :3] = 330.0
    nz, nr, nc = shape
    got = O.regu_np(Dj, vp, (25.0, 50.0, 40.0), (12.5, 20.0, 25.0))
    ref = Dj.ravel().copy()
    O.ref3d().ref_regu3d(ref, vp.ravel().copy(), 40.0, 50.0, 25.0, 25.0, 20.0, 12.5, nr, nc, nz)
    assert np.abs(got.ravel() - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.all(ref.reshape(shape)[vp == 330] == 0)


@ref
@pytest.mark.ref
@pytest.mark.parametrize("shape", [(9, 14), (12, 11)])
def test_regu_np_vs_reference_2d(shape):
    rng = np.random.default_rng(12)
    Dj = rng.standard_normal(shape)
    vp = 1800 + 100 * rng.random(shape)
    nr, nc = shape
    got = O.regu_np(Dj, vp, (20.0, 60.0), (8.0, 12.5))
    ref = Dj.ravel().copy()
    O.ref2d().ref_regu2d(ref, vp.ravel().copy(), 60.0, 20.0, 12.5, 8.0, nr, nc)
    assert np.abs(got.ravel() - ref).max() <= 1e-12 * np.abs(ref).max()
