"""Pins the FP64 CPU oracle (oracle/pfast_oracle.c) before it is trusted as the checker.

1. Golden fixtures written by the compiled reference (tests/golden/gen_golden.py):
   bit-exact.
2. Reference-held data: the shipped 2D example's forward picks
   (models/RT_SynMod2D_Inv.rec cols 4 and 6, produced by the reference's own FP64
   forward run, mfiles/RT_SynMod2D.m:98-99,126-131): bit-exact; source indices as
   printed in par/RT_SynMod2D_FW.out:47-49.
3. Mathematics: local-solve closed forms, axis rays in a homogeneous model,
   monotonicity / fixed point, the analytic v(z) traveltime, adjoint zero-residual
   and linearity, and a directional finite-difference check of the gradient.
Each check is chosen so that a plausible slip (dropped term, swapped spacing,
wrong sweep direction, wrong neighbour) fails it.
"""
import os

import numpy as np
import pytest

import cases
import oracle_lib as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bits(a):
    return cases.bits(a)


# ----------------------------------------------------------------- golden vectors
def test_golden_3d():
    g = np.load(os.path.join(GOLD, "golden_3d.npz"))
    for tag in "ab":
        vp = g[f"{tag}_vp"]
        dx, dy, dz = g[f"{tag}_spacing"]
        for s in range(3):
            T, _ = O.fsm3d(vp, g[f"{tag}_src"], s, dx, dy, dz)
            assert np.array_equal(bits(T), bits(g[f"{tag}_T"][s]))
            lam, EE, _ = O.adjoint_fsm3d(T, g[f"{tag}_T_obs"], g[f"{tag}_rec"], g[f"{tag}_nxyz"],
                                         g[f"{tag}_src"], g[f"{tag}_srnum"], s, dx, dy, dz)
            assert np.array_equal(bits(lam), bits(g[f"{tag}_lam"][s]))
            assert EE == g[f"{tag}_EE"][s]


def test_golden_2d():
    g = np.load(os.path.join(GOLD, "golden_2d.npz"))
    dx, dz = g["spacing"]
    WR = float(g["WR"])
    R = g["Z"].shape[0]
    for s in range(2):
        r = O.rt_fsm2d(g["vp"], g["Z"], g["src"], s, R, dx, dz)
        assert r["status"] == 0
        assert np.array_equal(bits(r["T"]), bits(g["T"][s]))
        assert np.array_equal(bits(r["TTd"]), bits(g["TTd"][s]))
        assert np.array_equal(bits(r["TTu"]), bits(g["TTu"][s]))
        for Cr, lk, ek in ((1.0, "lamT", "EET"), (1.0 - WR, "lamT_wr", "EET_wr")):
            lam, EE, _ = O.tadjoint_fsm2d(r["T"], g["T_obs"], Cr, g["rec"], g["nxy"], g["src"],
                                          g["srnum"], s, dx, dz)
            assert np.array_equal(bits(lam), bits(g[lk][s])) and EE == g[ek][s]
        for Cr, lk, ek in ((1.0, "lamR", "EER"), (WR, "lamR_wr", "EER_wr")):
            lam, EE, _ = O.radjoint_fsm2d(r["TTu"], g["TR_obs"], Cr, g["rec"], g["nxy"],
                                          g["srnum"], s, R, dx, dz)
            assert np.array_equal(bits(lam), bits(g[lk][s])) and EE == g[ek][s]


def test_example2d_reference_picks():
    """Reference-held vectors: the shipped example's forward picks, bit for bit."""
    e = np.load(os.path.join(GOLD, "example2d.npz"))
    dx, dz = e["spacing"]
    src = np.stack([O.map_index(e["src"], 0, dz), O.map_index(e["src"], 1, dx)], 1)
    # par/RT_SynMod2D_FW.out:47-49 prints sources 0-2 at (row, col) = (41,30),(42,40),(42,50), 1-based
    assert (src[:3] + 1).tolist() == [[41, 30], [42, 40], [42, 50]]
    Z = e["ref"].reshape(2, 600, 3)
    rows = e["rec_rows"]
    ri = np.stack([O.map_index(rows, 0, dz), O.map_index(rows, 1, dx)], 1)
    for n, s in enumerate(e["src_ids"]):
        r = O.rt_fsm2d(e["vp"], Z, src, int(s), 2, dx, dz)
        sl = slice(600 * n, 600 * n + 600)
        i, j = ri[sl, 0], ri[sl, 1]
        assert np.array_equal(bits(r["T"][i, j]), bits(rows[sl, 4]))
        assert np.array_equal(bits(r["TTu"][1][i, j]), bits(rows[sl, 6]))


# ----------------------------------------------------------------- geometry
def test_map_index_ceil_minus_one():
    # node i sits at coordinate (i+1)*h; src/T_PFAST3D.cpp:250-252
    t = np.array([[8.0], [8.000001], [7.999999], [16.0], [0.5]])
    assert O.map_index(t, 0, 8.0).tolist() == [0, 1, 0, 1, 0]


# ----------------------------------------------------------------- closed forms
def test_homogeneous_axis_rays_3d():
    """In a constant model the traveltime along a grid axis through the source is
    exactly |n|*h/v (one-neighbour updates only)."""
    nz, nr, nc, h, v = 7, 9, 11, 10.0, 2000.0
    vp = np.full((nz, nr, nc), v)
    src = np.array([[4, 5, 3]], np.int32)
    T, _ = O.fsm3d(vp, src, 0, h, h, h)
    s = h / v
    for j in range(nc):
        assert T[3, 4, j] == pytest.approx(abs(j - 5) * s, rel=1e-14)
    for i in range(nr):
        assert T[3, i, 5] == pytest.approx(abs(i - 4) * s, rel=1e-14)
    for k in range(nz):
        assert T[k, 4, 5] == pytest.approx(abs(k - 3) * s, rel=1e-14)


def test_two_neighbour_closed_form_2d():
    """Diagonal neighbour of the source in a constant 2D model with dx=dz:
    q = h*s + h*s/sqrt(2)... solved as b + ... : (q-b)^2+(q-a)^2 = (h s)^2 with
    a=b=h*s gives q = h*s*(1 + 1/sqrt(2))."""
    h, v = 10.0, 2500.0
    vp = np.full((9, 9), v)
    src = np.array([[4, 4]], np.int32)
    r = O.rt_fsm2d(vp, None, src, 0, 0, h, h)
    hs = h / v
    assert r["T"][5, 5] == pytest.approx(hs * (1 + 1 / np.sqrt(2)), rel=1e-14)


def test_three_neighbour_closed_form_3d():
    """Cube-diagonal neighbour: a=b=c=h*s*(1+1/sqrt2) (edge-diagonal nodes), then
    3(q-a)^2 = (h s)^2 gives q = a + h*s/sqrt(3)."""
    h, v = 5.0, 1800.0
    vp = np.full((7, 7, 7), v)
    src = np.array([[3, 3, 3]], np.int32)
    T, _ = O.fsm3d(vp, src, 0, h, h, h)
    hs = h / v
    a = hs * (1 + 1 / np.sqrt(2))
    assert T[3, 4, 4] == pytest.approx(a, rel=1e-14)
    assert T[4, 4, 4] == pytest.approx(a + hs / np.sqrt(3), rel=1e-14)


def test_anisotropic_spacing_is_not_symmetric():
    """Swapping dx and dy must transpose the field (catches a swapped spacing)."""
    c = cases.case3d(5, shape=(6, 8, 8), spacing=(10.0, 20.0, 15.0), nsrc=1)
    vp = np.full(c["vp"].shape, 2000.0)
    src = np.array([[3, 4, 2]], np.int32)
    T1, _ = O.fsm3d(vp, src, 0, 10.0, 20.0, 15.0)
    T2, _ = O.fsm3d(vp, np.array([[4, 3, 2]], np.int32), 0, 20.0, 10.0, 15.0)
    np.testing.assert_allclose(T1, np.transpose(T2, (0, 2, 1)), rtol=1e-13)


def test_fixed_point_and_monotone():
    c = cases.case3d(9, shape=(8, 9, 10))
    T, cyc = O.fsm3d(c["vp"], c["src"], 0, c["dx"], c["dy"], c["dz"])
    T1, _ = O.fsm3d(c["vp"], c["src"], 0, c["dx"], c["dy"], c["dz"], MAXSwp=1)
    assert (T <= T1).all()  # more cycles never increase a traveltime
    T2, _ = O.fsm3d(c["vp"], c["src"], 0, c["dx"], c["dy"], c["dz"], MAXSwp=cyc + 3)
    assert np.array_equal(bits(T), bits(T2))  # converged field is a fixed point


def test_analytic_linear_gradient_3d():
    """v(z) = v0 + g z: T = acosh(1 + g^2 r^2 / (2 v_s v_r)) / g.  First-order FSM
    error is a few percent at these distances; a dropped term is off by far more."""
    nz, nr, nc, h = 31, 41, 41, 25.0
    v0, gr = 2000.0, 0.6
    z = (np.arange(nz) * h)[:, None, None]
    vp = np.broadcast_to(v0 + gr * z, (nz, nr, nc)).copy()
    src = np.array([[20, 20, 0]], np.int32)
    T, _ = O.fsm3d(vp, src, 0, h, h, h)
    k, i, j = np.meshgrid(np.arange(nz), np.arange(nr), np.arange(nc), indexing="ij")
    r2 = ((i - 20) * h) ** 2 + ((j - 20) * h) ** 2 + (k * h) ** 2
    ta = np.arccosh(1 + gr ** 2 * r2 / (2 * v0 * vp)) / gr
    far = r2 > (10 * h) ** 2
    rel = (T - ta)[far] / ta[far]
    # first-order upwind FSM with a point source overestimates, most along the
    # grid diagonals (about 10% at 10-20 cells); a dropped term is off by far more
    assert rel.min() > -0.01 and rel.max() < 0.15 and rel.mean() < 0.07
    # straight down the axis the scheme is a rectangle-rule integral of 1/v(z)
    ax = (T[1:, 20, 20] - ta[1:, 20, 20]) / ta[1:, 20, 20]
    assert np.abs(ax).max() < 0.01


def test_adjoint_zero_residual_and_linearity():
    c = cases.case3d(13, shape=(8, 9, 10), nsrc=2)
    T, _ = O.fsm3d(c["vp"], c["src"], 0, c["dx"], c["dy"], c["dz"])
    n = c["srnum"][0]
    rec = c["rec"][:n]
    Tr = T[rec[:, 2], rec[:, 0], rec[:, 1]]
    lam0, EE0, _ = O.adjoint_fsm3d(T, Tr, rec, c["nxyz"][:n], c["src"][:1], c["srnum"][:1], 0,
                                   c["dx"], c["dy"], c["dz"])
    assert EE0 == 0.0 and not lam0.any()
    r = 0.01 * np.arange(1, n + 1)
    lam1, EE1, _ = O.adjoint_fsm3d(T, Tr + r, rec, c["nxyz"][:n], c["src"][:1], c["srnum"][:1], 0,
                                   c["dx"], c["dy"], c["dz"])
    lam2, EE2, _ = O.adjoint_fsm3d(T, Tr + 2 * r, rec, c["nxyz"][:n], c["src"][:1],
                                   c["srnum"][:1], 0, c["dx"], c["dy"], c["dz"])
    np.testing.assert_allclose(lam2, 2 * lam1, rtol=1e-9, atol=1e-12)
    assert EE2 == pytest.approx(4 * EE1, rel=1e-9)
    assert np.abs(lam1).max() > 0


def test_gradient_directional_fd():
    """Directional derivatives of the misfit by central differences vs the
    adjoint-state gradient along three localized velocity perturbations.  The
    reference's gradient -sum(lambda)/v^3 is dE/dv only up to a grid-dependent
    positive scale (its boundary condition is a point flux), and the adjoint of the
    continuous eikonal is only approximately the derivative of the discrete misfit,
    so the check is that the three ratios FD/adjoint agree in sign and in size."""
    nz, nr, nc, h = 12, 16, 16, 20.0
    z = (np.arange(nz) * h)[:, None, None]
    vp = np.broadcast_to(1800 + 1.0 * z, (nz, nr, nc)).copy()
    src = np.array([[8, 3, 0], [4, 12, 0]], np.int32)
    jj, ii = np.meshgrid(np.arange(nc), np.arange(nr))
    rec1 = np.stack([ii.ravel(), jj.ravel(), np.zeros(nr * nc, int)], 1).astype(np.int32)
    rec = np.concatenate([rec1, rec1])
    srnum = np.array([nr * nc, nr * nc], np.int32)
    nxyz = np.tile([0.0, 0.0, -1.0], (rec.shape[0], 1))
    k, i, j = np.meshgrid(np.arange(nz), np.arange(nr), np.arange(nc), indexing="ij")

    def bump(ci, cj, ck, amp):
        return amp * np.exp(-(((i - ci) ** 2 + (j - cj) ** 2) / 12.0 + (k - ck) ** 2 / 6.0))

    truth = vp + bump(8, 8, 6, 40.0)
    T_obs = []
    for s in range(2):
        T, _ = O.fsm3d(truth, src, s, h, h, h)
        T_obs.append(T[rec1[:, 2], rec1[:, 0], rec1[:, 1]])
    T_obs = np.concatenate(T_obs)

    def misfit_grad(v):
        g = np.zeros(v.size)
        E = 0.0
        for s in range(2):
            T, _ = O.fsm3d(v, src, s, h, h, h)
            lam, EE, _ = O.adjoint_fsm3d(T, T_obs, rec, nxyz, src, srnum, s, h, h, h)
            O.oracle().orc_grad_accum(g, np.ascontiguousarray(lam.ravel()),
                                      np.ascontiguousarray(v.ravel()), v.size)
            E += EE
        return E, g.reshape(v.shape)

    _, g0 = misfit_grad(vp)
    ratios = []
    for dv in (bump(8, 8, 6, 40.0), bump(5, 10, 4, 30.0), bump(11, 6, 7, 30.0)):
        eps = 0.05
        Ep, _ = misfit_grad(vp + eps * dv)
        Em, _ = misfit_grad(vp - eps * dv)
        fd = (Ep - Em) / (2 * eps)
        ad = float((g0 * dv).sum())
        assert ad != 0
        ratios.append(fd / ad)
    ratios = np.array(ratios)
    assert (ratios > 0).all(), ratios
    assert ratios.max() / ratios.min() < 3.0, ratios
