"""GPU parity of the batched 2D shot sets (FP32 sweeps, FP64 reductions) against
the FP64 oracle and the reference's own picks of the shipped 2D example.

Tolerances (the path's precision contract):
  traveltimes (T, TTd, TTu): max|T32 - T64| / max|T64| <= 1e-5 over reached nodes
  gradients (gT, gR, gRT):    ||g32 - g64||_2 / ||g64||_2 <= 1e-3
  misfits (Et, Er):           relative <= 1e-4
  reflection rows k2vec:      identical (integer)
"""
import os

import numpy as np
import pytest

import cases
import oracle_lib as O

pytestmark = pytest.mark.gpu

T_TOL, G_TOL, E_TOL = 1e-5, 1e-3, 1e-4
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ctx():
    import pfast_amd
    c = pfast_amd.Context(0)
    yield c
    c.close()


def terr(a, b):
    reach = b < 1e6
    assert np.array_equal(reach, a < 1e6)
    return float(np.abs(a - b)[reach].max() / b[reach].max()) if reach.any() else 0.0


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


@pytest.mark.parametrize("seed,NumofR,shape,nsrc", [
    (61, 0, (22, 31), 4),
    (62, 1, (26, 37), 5),
    (63, 2, (30, 41), 6),
    (64, 0, (3, 40), 4),  # one interior row: the adjoint takes the plane loop
])
def test_shots2d_vs_oracle(ctx, seed, NumofR, shape, nsrc):
    import pfast_amd
    c = cases.case2d(seed, shape=shape, NumofR=max(NumofR, 1), nsrc=nsrc, nrec_per=9)
    Z = c["Z"][:NumofR] if NumofR else None
    TR = c["TR_obs"][:, :max(NumofR, 1)]
    m = pfast_amd.Model2D(ctx, c["vp"], c["dx"], c["dz"], Z)
    sh = pfast_amd.Shots2D(m, c["src"], c["srnum"], c["rec"], c["nxy"], c["T_obs"], TR)
    sh.run(adjoint=True, MAXSwp=500)
    WR = np.random.default_rng(seed).random(c["vp"].shape)
    gT, gR, gRT, (Et, Er, E) = sh.gradient(WR)
    cyc, k2 = sh.info()
    N = c["vp"].size
    v3 = c["vp"] ** 3
    oT, oR = np.zeros(c["vp"].shape), np.zeros(c["vp"].shape)
    oEt = oEr = 0.0
    worst = 0.0
    for s in range(nsrc):
        o = O.rt_fsm2d(c["vp"], Z if NumofR else np.zeros((1, shape[1], 3)), c["src"], s, NumofR,
                       c["dx"], c["dz"])
        assert o["status"] == 0
        worst = max(worst, terr(sh.field(s, "T"), o["T"]))
        for ir in range(NumofR):
            worst = max(worst, terr(sh.field(s, "TTd", ir), o["TTd"][ir]))
            worst = max(worst, terr(sh.field(s, "TTu", ir), o["TTu"][ir]))
        if NumofR:
            assert np.array_equal(k2[s], o["k2vec"]), f"k2vec of source {s}"
        lt, et, _ = O.tadjoint_fsm2d(o["T"], c["T_obs"], 1.0, c["rec"], c["nxy"], c["src"],
                                     c["srnum"], s, c["dx"], c["dz"])
        oT -= lt / v3
        oEt += et
        if NumofR:
            lr, er, _ = O.radjoint_fsm2d(o["TTu"], TR, 1.0, c["rec"], c["nxy"], c["srnum"], s,
                                         NumofR, c["dx"], c["dz"])
            oR -= lr / v3
            oEr += er
    eT, eR = rel_l2(gT, oT), rel_l2(gR, oR)
    eRT = rel_l2(gRT, oT * (1 - WR) + oR * WR)
    print(f"seed {seed} R {NumofR}: T err {worst:.2e} gT {eT:.2e} gR {eR:.2e} gRT {eRT:.2e} "
          f"Et {Et:.6g}/{oEt:.6g} Er {Er:.6g}/{oEr:.6g} cycles {cyc[:, 0].tolist()}", flush=True)
    assert worst <= T_TOL
    assert eT <= G_TOL and eR <= G_TOL and eRT <= G_TOL
    assert abs(Et - oEt) <= E_TOL * abs(oEt)
    assert abs(Er - oEr) <= E_TOL * max(abs(oEr), 1e-300)
    assert E == Et + Er
    assert not NumofR or abs(gR).max() > 0
    sh.close()
    m.close()


def test_shots2d_reference_example_picks(ctx):
    """The shipped 2D example (models/RT_SynMod2D_FW.vp, .src, .ref, 160 x 600, two
    reflectors): batched FP32 picks against the reference's own FP64 picks held in
    models/RT_SynMod2D_Inv.rec (transmission column 4, second reflector column 6)."""
    import pfast_amd
    e = np.load(os.path.join(GOLD, "example2d.npz"))
    dx, dz = e["spacing"]
    src = np.stack([O.map_index(e["src"], 0, dz), O.map_index(e["src"], 1, dx)], 1)
    Z = e["ref"].reshape(2, 600, 3)
    rows = e["rec_rows"]
    ri = np.stack([O.map_index(rows, 0, dz), O.map_index(rows, 1, dx)], 1)
    ids = [int(s) for s in e["src_ids"]]
    m = pfast_amd.Model2D(ctx, e["vp"], dx, dz, Z)
    sh = pfast_amd.Shots2D(m, src[ids], np.zeros(len(ids), np.int32))
    sh.run(adjoint=False, MAXSwp=500)
    for n in range(len(ids)):
        sl = slice(600 * n, 600 * n + 600)
        i, j = ri[sl, 0], ri[sl, 1]
        T = sh.field(n, "T")[i, j]
        Tu = sh.field(n, "TTu", 1)[i, j]
        et = np.abs(T - rows[sl, 4]).max() / rows[sl, 4].max()
        eu = np.abs(Tu - rows[sl, 6]).max() / rows[sl, 6].max()
        print(f"source {ids[n]}: transmission pick err {et:.2e}, reflector-2 pick err {eu:.2e}",
              flush=True)
        assert et <= T_TOL and eu <= T_TOL
    sh.close()
    m.close()


def test_shots2d_fatal_exit_is_reported(ctx):
    """A reflector above every grid row has no k2 in its columns: the reference exits
    (src/RT_pfsm2d.h:232); the batched call returns PFAST_ERR_FATAL."""
    import pfast_amd
    c = cases.case2d(71, shape=(20, 24), NumofR=1, nsrc=2)
    Z = c["Z"].copy()
    Z[0, :, 0] = 0.5 * c["dz"]
    m = pfast_amd.Model2D(ctx, c["vp"], c["dx"], c["dz"], Z)
    sh = pfast_amd.Shots2D(m, c["src"], c["srnum"])
    with pytest.raises(pfast_amd.FatalError):
        sh.run(adjoint=False)
    sh.close()
    m.close()


def _all_results(sh, nsrc, R):
    out = {"info": np.asarray(sh.info()[0]), "k2": np.asarray(sh.info()[1])}
    gT, gR, gRT, E = sh.gradient(0.5)
    out.update(gT=gT, gR=gR, gRT=gRT, E=np.array(E))
    for s in range(nsrc):
        out[f"T{s}"] = sh.field(s, "T")
        out[f"lT{s}"] = sh.field(s, "lambdaT")
        for r in range(R):
            out[f"Td{s}_{r}"] = sh.field(s, "TTd", r)
            out[f"Tu{s}_{r}"] = sh.field(s, "TTu", r)
        if R:
            out[f"lR{s}"] = sh.field(s, "lambdaR")
    return out


@pytest.mark.parametrize("seed,NumofR,shape,nsrc", [
    (71, 0, (23, 35), 5),
    (72, 1, (29, 44), 6),
    (73, 2, (31, 47), 7),
    (74, 0, (3, 36), 4),
    (75, 0, (4, 30), 3),
])
def test_row_window_matches_plane_loop_bitwise(ctx, seed, NumofR, shape, nsrc):
    """The row-window sweeps (a work-group walking the rows, independent solves of a
    source side by side) and the plane loop (one work-group per source stepping the
    anti-diagonals) update every node with the same operands in the same order: every
    field, reflection row, cycle count, gradient and misfit agrees bit for bit."""
    import pfast_amd
    c = cases.case2d(seed, shape=shape, NumofR=max(NumofR, 1), nsrc=nsrc, nrec_per=9)
    Z = c["Z"][:NumofR] if NumofR else None
    TR = c["TR_obs"][:, :max(NumofR, 1)]
    res = []
    for form in (0, 1):
        ctx.set_sweep2d(form)
        m = pfast_amd.Model2D(ctx, c["vp"], c["dx"], c["dz"], Z)
        sh = pfast_amd.Shots2D(m, c["src"], c["srnum"], c["rec"], c["nxy"], c["T_obs"], TR)
        sh.run(adjoint=True, MAXSwp=500)
        res.append(_all_results(sh, nsrc, NumofR))
        sh.close()
        m.close()
    ctx.set_sweep2d(0)
    a, b = res
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(np.asarray(a[k]).view(np.uint8), np.asarray(b[k]).view(np.uint8)), k


def test_model2d_size_bound(ctx):
    """A 2D grid above 2^28 nodes is refused when the model is created, before the
    velocities are read: the row-window sweeps address a source's fields and the FP64
    velocities with 32-bit byte offsets."""
    import ctypes as C
    from pfast_amd import _lib
    L = _lib.lib()
    vp = np.full((4, 4), 2000.0)
    out = C.c_void_p()
    rc = L.pfast_model2d_create(ctx.h, vp, 16385, 16384, 10.0, 10.0, None, 0, C.byref(out))
    assert rc != 0 and b"too large" in L.pfast_last_error()
