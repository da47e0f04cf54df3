"""GPU FFT regularisation (pfast_regu3d / pfast_regu2d, hipFFT) against the numpy
restatement of the reference's regu3d / regu2d (tests/oracle_lib.regu_np, pinned to
the reference in tests/test_optim.py).  FP64 throughout; different FFT algorithms
round differently, so the bound is 1e-12 of the field's max-norm."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(6, 7, 9), (5, 8, 6), (11, 13, 16)])
def test_regu3d_gpu(shape):
    from pfast_amd._lib import check, lib
    rng = np.random.default_rng(21)
    Dj = rng.standard_normal(shape)
    vp = 2000 + 100 * rng.random(shape)
    vp[1, :3, :2] = 330.0
    nz, nr, nc = shape
    want = O.regu_np(Dj, vp, (25.0, 50.0, 40.0), (12.5, 20.0, 25.0))
    got = np.ascontiguousarray(Dj.ravel())
    check(lib().pfast_regu3d(got, np.ascontiguousarray(vp.ravel()), 40.0, 50.0, 25.0, 25.0, 20.0,
                             12.5, nr, nc, nz))
    assert np.abs(got - want.ravel()).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("shape", [(9, 14), (12, 11), (160, 600)])
def test_regu2d_gpu(shape):
    from pfast_amd._lib import check, lib
    rng = np.random.default_rng(22)
    Dj = rng.standard_normal(shape)
    vp = 1800 + 100 * rng.random(shape)
    vp[:2, :] = 330.0
    nr, nc = shape
    want = O.regu_np(Dj, vp, (20.0, 60.0), (8.0, 12.5))
    got = np.ascontiguousarray(Dj.ravel())
    check(lib().pfast_regu2d(got, np.ascontiguousarray(vp.ravel()), 60.0, 20.0, 12.5, 8.0, nr, nc))
    assert np.abs(got - want.ravel()).max() <= 1e-12 * np.abs(want).max()
