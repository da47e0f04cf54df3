"""The L-BFGS ring (LM pairs, slot i % LM) past the point where the reference's own
indexing leaves its array (Iterk >= LM + 2, src/T_PFAST3D.cpp:467, src/RT_PFAST2D.cpp:529),
checked on the CPU against two facts the mathematics fixes: the two-loop recursion equals
the explicit inverse-BFGS matrix built from the same pairs times the gradient, and that
matrix maps the newest y onto the newest s (the secant equation).  A planted slot error
must fail both."""
import numpy as np
import pytest

import lbfgs_ref as R

LM = 5


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_two_loop_equals_explicit_inverse_bfgs_past_the_wrap(seed):
    n = 12
    ring = R.Ring(n, LM)
    steps = R.quadratic_run(n, LM, 11, seed)
    assert max(k for k, *_ in steps) >= LM + 4
    for k, g1, g0, dk, alpha in steps:
        d, gamma = ring.direction(k, g1, g0, dk, alpha)
        H = R.inverse_bfgs(ring.pairs(k), gamma)
        assert np.linalg.norm(d - H @ g1) <= 1e-10 * np.linalg.norm(d), k
        s, y = ring.pairs(k)[-1]
        assert np.linalg.norm(H @ y - s) <= 1e-10 * np.linalg.norm(s), k
        assert len(ring.pairs(k)) == min(LM, k - 1)


def test_planted_slot_error_is_caught():
    n = 12
    ring = R.Ring(n, LM)
    caught = False
    for k, g1, g0, dk, alpha in R.quadratic_run(n, LM, 11, 4):
        good = R.Ring(n, LM)
        good.s, good.y, good.rho = ring.s.copy(), ring.y.copy(), ring.rho.copy()
        d_bad, gamma = ring.direction(k, g1, g0, dk, alpha, slot_shift=1)
        H = R.inverse_bfgs(ring.pairs(k), gamma)
        if k > 2 and np.linalg.norm(d_bad - H @ g1) > 1e-6 * np.linalg.norm(H @ g1):
            caught = True
    assert caught
