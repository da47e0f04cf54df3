"""The oracle against the compiled reference (oracle/_ref, built from the reference's
own headers by `make -C oracle ref`) on seeded random cases, bit for bit.  Skipped
where the reference is not available (it never travels with the repository)."""
import numpy as np
import pytest

import cases
import oracle_lib as O

pytestmark = [pytest.mark.ref,
              pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built")]


@pytest.mark.parametrize("seed", range(40, 52))
def test_fsm3d_adjoint3d_vs_reference(seed):
    c = cases.case3d(seed)
    nz, nr, nc = c["vp"].shape
    N = c["vp"].size
    for s in range(c["src"].shape[0]):
        T, _ = O.fsm3d(c["vp"], c["src"], s, c["dx"], c["dy"], c["dz"])
        Tr = np.empty(N)
        O.ref3d().ref_fsm3d(c["vp"].ravel().copy(), Tr, c["src"].ravel().copy(), s, 100,
                            c["dx"], c["dy"], c["dz"], nr, nc, nz)
        assert np.array_equal(cases.bits(T), cases.bits(Tr))
        lam, EE, _ = O.adjoint_fsm3d(T, c["T_obs"], c["rec"], c["nxyz"], c["src"], c["srnum"], s,
                                     c["dx"], c["dy"], c["dz"])
        lr = np.empty(N)
        EEr = O.ref3d().ref_adjoint_fsm3d(lr, Tr, c["T_obs"].copy(), c["rec"].ravel().copy(),
                                          c["nxyz"].ravel().copy(), c["src"].ravel().copy(),
                                          c["srnum"].copy(), s, 100, c["dx"], c["dy"], c["dz"],
                                          nr, nc, nz)
        assert np.array_equal(cases.bits(lam), cases.bits(lr)) and EE == EEr


@pytest.mark.parametrize("seed", range(60, 70))
@pytest.mark.parametrize("maxswp", [500, 2])
def test_rt2d_adjoint2d_vs_reference(seed, maxswp, tmp_path, monkeypatch):
    # the reference writes a debug dump into the working directory when MAXSwp is
    # exhausted (src/RT_pfsm2d.h:188); keep it out of the repository
    monkeypatch.chdir(tmp_path)
    c = cases.case2d(seed)
    nr, nc = c["vp"].shape
    N = c["vp"].size
    R = c["NumofR"]
    for s in range(c["src"].shape[0]):
        r = O.rt_fsm2d(c["vp"], c["Z"], c["src"], s, R, c["dx"], c["dz"], MAXSwp=maxswp)
        if r["status"]:
            continue  # the reference would exit(1) here
        U, TTd, TTu = np.empty(N), np.empty(R * N), np.empty(R * N)
        O.ref2d().ref_rt_fsm2d(c["vp"].ravel().copy(), U, TTd, TTu, c["Z"].ravel().copy(),
                               c["src"].ravel().copy(), c["rec"].ravel().copy(), c["srnum"].copy(),
                               s, R, maxswp, c["dx"], c["dz"], nr, nc)
        assert np.array_equal(cases.bits(r["T"]), cases.bits(U))
        assert np.array_equal(cases.bits(r["TTd"]), cases.bits(TTd))
        assert np.array_equal(cases.bits(r["TTu"]), cases.bits(TTu))
        args = (c["rec"].ravel().copy(), c["nxy"].ravel().copy(), c["src"].ravel().copy(),
                c["srnum"].copy(), s, maxswp)
        lt, Et, _ = O.tadjoint_fsm2d(r["T"], c["T_obs"], 1.0, c["rec"], c["nxy"], c["src"],
                                     c["srnum"], s, c["dx"], c["dz"], MAXSwp=maxswp)
        lr = np.empty(N)
        Er = O.ref2d().ref_tadjoint_fsm2d(lr, U.copy(), c["T_obs"].copy(), *args, c["dx"], c["dz"], nr, nc)
        assert np.array_equal(cases.bits(lt), cases.bits(lr)) and Et == Er
        lR, ER, _ = O.radjoint_fsm2d(r["TTu"], c["TR_obs"], 1.0, c["rec"], c["nxy"], c["srnum"], s,
                                     R, c["dx"], c["dz"], MAXSwp=maxswp)
        lRr = np.empty(N)
        ERr = O.ref2d().ref_radjoint_fsm2d(lRr, TTd.copy(), TTu.copy(), c["TR_obs"].ravel().copy(),
                                           *args[:5], R, maxswp, c["dx"], c["dz"], nr, nc)
        assert np.array_equal(cases.bits(lR), cases.bits(lRr)) and ER == ERr
