"""GPU parity of the single-source FP64 forms: bit-identical to the FP64 oracle
(which is itself bit-identical to the compiled reference, see test_oracle_ref.py)."""
import numpy as np
import pytest

import cases
import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pf():
    import pfast_amd
    assert pfast_amd.device_count() > 0
    return pfast_amd


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_fsm3d_bitexact(pf, seed):
    c = cases.case3d(seed)
    for s in range(c["src"].shape[0]):
        Ug, cg = pf.fsm3d(c["vp"], c["src"], s, 100, c["dx"], c["dy"], c["dz"])
        Uo, co = O.fsm3d(c["vp"], c["src"], s, c["dx"], c["dy"], c["dz"])
        print(f"seed {seed} src {s} shape {c['vp'].shape} cycles gpu {cg} oracle {co}", flush=True)
        assert cg == co
        assert np.array_equal(cases.bits(Ug), cases.bits(Uo))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_adjoint3d_bitexact(pf, seed):
    c = cases.case3d(seed)
    for s in range(c["src"].shape[0]):
        T, _ = O.fsm3d(c["vp"], c["src"], s, c["dx"], c["dy"], c["dz"])
        Lg, Eg, cg = pf.adjoint_fsm3d(T, c["T_obs"], c["rec"], c["nxyz"], c["src"], c["srnum"], s,
                                      100, c["dx"], c["dy"], c["dz"])
        Lo, Eo, co = O.adjoint_fsm3d(T, c["T_obs"], c["rec"], c["nxyz"], c["src"], c["srnum"], s,
                                     c["dx"], c["dy"], c["dz"])
        print(f"seed {seed} src {s} adj cycles gpu {cg} oracle {co} EE {Eg}", flush=True)
        assert cg == co
        assert Eg == Eo
        assert np.array_equal(cases.bits(Lg), cases.bits(Lo))


def test_fsm3d_maxswp_exhausted(pf):
    """MAXSwp smaller than the cycles needed: both stop at the same partial field."""
    c = cases.case3d(7, shape=(9, 10, 11))
    Ug, cg = pf.fsm3d(c["vp"], c["src"], 0, 1, c["dx"], c["dy"], c["dz"])
    Uo, co = O.fsm3d(c["vp"], c["src"], 0, c["dx"], c["dy"], c["dz"], MAXSwp=1)
    assert cg == co == 1
    assert np.array_equal(cases.bits(Ug), cases.bits(Uo))


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_rt_fsm2d_bitexact(pf, seed):
    c = cases.case2d(seed)
    for s in range(c["src"].shape[0]):
        g = pf.RT_fsm2d(c["vp"], c["Z"], c["src"], s, c["NumofR"], 500, c["dx"], c["dz"])
        o = O.rt_fsm2d(c["vp"], c["Z"], c["src"], s, c["NumofR"], c["dx"], c["dz"])
        assert o["status"] == 0
        print(f"seed {seed} src {s} cycles gpu {list(g['cycles'])} oracle {list(o['cycles'])}", flush=True)
        assert list(g["cycles"]) == list(o["cycles"])
        assert np.array_equal(g["k2vec"], o["k2vec"])
        assert np.array_equal(cases.bits(g["T"]), cases.bits(o["T"]))
        assert np.array_equal(cases.bits(g["TTd"]), cases.bits(o["TTd"]))
        assert np.array_equal(cases.bits(g["TTu"]), cases.bits(o["TTu"]))


@pytest.mark.parametrize("seed", [11, 12])
@pytest.mark.parametrize("WR", [None, 0.3])
def test_adjoint2d_bitexact(pf, seed, WR):
    c = cases.case2d(seed)
    for s in range(c["src"].shape[0]):
        o = O.rt_fsm2d(c["vp"], c["Z"], c["src"], s, c["NumofR"], c["dx"], c["dz"])
        Cr_t = 1.0 if WR is None else 1.0 - WR
        Cr_r = 1.0 if WR is None else WR
        Lg, Eg, cg = pf.Tadjoint_fsm2d(o["T"], c["T_obs"], c["rec"], c["nxy"], c["src"], c["srnum"],
                                       s, 500, c["dx"], c["dz"], WR=WR)
        Lo, Eo, co = O.tadjoint_fsm2d(o["T"], c["T_obs"], Cr_t, c["rec"], c["nxy"], c["src"],
                                      c["srnum"], s, c["dx"], c["dz"])
        assert cg == co and Eg == Eo
        assert np.array_equal(cases.bits(Lg), cases.bits(Lo))
        Rg, ERg, crg = pf.Radjoint_fsm2d(o["TTu"], c["TR_obs"], c["rec"], c["nxy"], c["src"],
                                         c["srnum"], s, c["NumofR"], 500, c["dx"], c["dz"], WR=WR)
        Ro, ERo, cro = O.radjoint_fsm2d(o["TTu"], c["TR_obs"], Cr_r, c["rec"], c["nxy"], c["srnum"],
                                        s, c["NumofR"], c["dx"], c["dz"])
        print(f"seed {seed} src {s} WR {WR} T-cycles {cg}/{co} R-cycles {list(crg)}/{list(cro)}", flush=True)
        assert list(crg) == list(cro) and ERg == ERo
        assert np.array_equal(cases.bits(Rg), cases.bits(Ro))
