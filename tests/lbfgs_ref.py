"""The L-BFGS two-loop recursion with the drivers' ring of LM pairs (test
infrastructure: the checker of pfast_lbfgs_direction).

src/T_PFAST3D.cpp:450-484 and src/RT_PFAST2D.cpp:499-552: gd2sk stores the pair of
iteration k (s = alpha1 dk, y = Dj_{k-1} - Dj_k, rho = 1 / s.y) in slot (k - 2) % LM,
then the recursion runs over pairs i = Iterk-2 down to max(Iterk-LM-1, 0) (slot i % LM;
the reference indexes its ai array with i, which leaves the array once Iterk >= LM + 2)
and back, from H0 = gamma I with gamma = s.y / y.y of the newest pair.
"""
import numpy as np


class Ring:
    def __init__(self, n, lm):
        self.lm = lm
        self.s = np.zeros((lm, n))
        self.y = np.zeros((lm, n))
        self.rho = np.zeros(lm)

    def direction(self, it, g1, g0, dk, alpha1, slot_shift=0):
        """New direction at iteration `it` >= 2 (g1 = Dj_k, g0 = Dj_{k-1}, dk the
        previous direction).  slot_shift plants a wrong slot in the second loop."""
        lm = self.lm
        slot = (it - 2) % lm
        self.s[slot] = alpha1 * dk
        self.y[slot] = -g1 + g0
        sy = self.s[slot] @ self.y[slot]
        self.rho[slot] = 1.0 / sy
        gamma = sy / (self.y[slot] @ self.y[slot])
        lo = max(it - lm - 1, 0)
        q = g1.copy()
        a = np.zeros(lm)
        for i in range(it - 2, lo - 1, -1):
            m = i % lm
            a[m] = self.rho[m] * (self.s[m] @ q)
            q -= a[m] * self.y[m]
        d = gamma * q
        for i in range(lo, it - 1):
            m = (i + slot_shift) % lm
            b = self.rho[m] * (self.y[m] @ d)
            d += self.s[m] * (a[m] - b)
        return d, gamma

    def pairs(self, it):
        """The stored pairs the recursion at iteration `it` uses, oldest first."""
        lo = max(it - self.lm - 1, 0)
        return [(self.s[i % self.lm], self.y[i % self.lm]) for i in range(lo, it - 1)]


def inverse_bfgs(pairs, gamma):
    """H from H0 = gamma I by the BFGS inverse update, oldest pair first:
    H <- (I - rho s y^T) H (I - rho y s^T) + rho s s^T."""
    n = pairs[0][0].size
    H = gamma * np.eye(n)
    for s, y in pairs:
        rho = 1.0 / (s @ y)
        V = np.eye(n) - rho * np.outer(y, s)
        H = V.T @ H @ V + rho * np.outer(s, s)
    return H


def quadratic_run(n, lm, iters, seed):
    """Minimise 1/2 x^T A x - b^T x along L-BFGS directions (Dj = -gradient, as the
    drivers' descent directions) with inexact steps, 0.5 to 1.5 times the exact line
    minimum (exact steps on a quadratic make every pair but the newest drop out of the
    recursion, which would hide the ring's slots): returns, per iteration k >= 2,
    (k, Dj_k, Dj_{k-1}, previous direction, previous step)."""
    rng = np.random.default_rng(seed)
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    A = Q @ np.diag(np.geomspace(1.0, 50.0, n)) @ Q.T
    b = rng.standard_normal(n)
    x = rng.standard_normal(n)
    ring = Ring(n, lm)
    g1 = -(A @ x - b)
    dk = g1.copy()
    out = []
    for k in range(1, iters + 1):
        if k >= 2:
            out.append((k, g1.copy(), g0.copy(), dk.copy(), alpha))
            dk, _ = ring.direction(k, g1, g0, dk, alpha)
        alpha = rng.uniform(0.5, 1.5) * (g1 @ dk) / (dk @ A @ dk)
        x = x + alpha * dk
        g0 = g1
        g1 = -(A @ x - b)
    return out
