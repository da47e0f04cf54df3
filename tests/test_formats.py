"""File formats and the serving workload generator (CPU).

pfast_amd.formats follows the reference drivers' parameter-file reading
(src/RT_PFAST2D.cpp:140-177, src/T_PFAST3D.cpp:105-136), headerless FP64 tables
(src/T_pfsm3d.h:39-57) and the node mapping ceil(c/h)-1 (src/T_PFAST3D.cpp:250-252).
"""
import os

import numpy as np
import pytest

from pfast_amd import formats as F
from pfast_amd import loadgen as LG

REF_INP = "/root/reference/par/RT_SynMod2D_Inv.inp"


def test_map_index_known_answers():
    # the shipped 2D example's first source, (z, x) = (325.97..., 375) with dz = 8, dx = 12.5,
    # is printed 1-based as row 41, column 30 by the reference (par/RT_SynMod2D_FW.out)
    assert F.map_index(325.97357475, 8.0) + 1 == 41
    assert F.map_index(375.0, 12.5) + 1 == 30
    c = np.array([0.5, 1.0, 12.5, 12.500001, 25.0, 37.4])
    assert np.array_equal(F.map_index(c, 12.5), np.ceil(c / 12.5).astype(int) - 1)


def test_bin_round_trip(tmp_path):
    a = np.random.default_rng(1).standard_normal((3, 4, 5))
    p = str(tmp_path / "a.vp")
    F.write_bin(p, a)
    assert os.path.getsize(p) == a.size * 8
    assert np.array_equal(F.read_bin(p, a.shape), a)
    with pytest.raises(ValueError):
        F.read_bin(p, (4, 4, 5))


def test_inp2d_round_trip(tmp_path):
    p = str(tmp_path / "m.inp")
    F.write_inp2d(p, "m.vp", "m.src", (56, 3), "m.rec", (33600, 7), "m.ref", (1200, 3), 160, 600,
                  12.5, 8.0, vpfile="out.vp", nux=60, nuz=20, twfile="m.wr", schemetag=1, lsrc=3)
    q = F.parse_inp2d(p)
    assert (q["vp0file"], q["srcfile"], q["recfile"], q["refgeom"], q["vpfile"], q["twfile"]) == \
        ("m.vp", "m.src", "m.rec", "m.ref", "out.vp", "m.wr")
    assert q["src_shape"] == (56, 3) and q["rec_shape"] == (33600, 7) and q["ref_shape"] == (1200, 3)
    assert (q["nrows"], q["ncols"], q["dx"], q["dz"]) == (160, 600, 12.5, 8.0)
    assert (q["nux"], q["nuz"], q["schemetag"], q["lsrc"]) == (60.0, 20.0, 1, 3)
    assert q["NumofR"] == 2  # ceil(1200 / 600), src/RT_PFAST2D.cpp:230


def test_inp3d_round_trip(tmp_path):
    p = str(tmp_path / "m.inp")
    F.write_inp3d(p, "v.vp2", "s.src", (400, 4), "r.rec", (800000, 7), 201, 201, 101, 25, 25, 25,
                  vpfile="f.vp", nux=50, nuy=50, nuz=25, schemetag=3, lsrc=2)
    q = F.parse_inp3d(p)
    assert (q["vp0file"], q["srcfile"], q["recfile"], q["vpfile"]) == ("v.vp2", "s.src", "r.rec", "f.vp")
    assert q["src_shape"] == (400, 4) and q["rec_shape"] == (800000, 7)
    assert (q["nrows"], q["ncols"], q["nz"]) == (201, 201, 101)
    assert (q["dx"], q["dy"], q["dz"], q["nux"], q["nuy"], q["nuz"]) == (25, 25, 25, 50, 50, 25)


@pytest.mark.skipif(not os.path.exists(REF_INP), reason="reference tree not present")
def test_parse_reference_parameter_file():
    """The reference's own 2D inversion parameter file, with its comment lines."""
    q = F.parse_inp2d(REF_INP)
    assert q["vp0file"].endswith("RT_SynMod2D_Inv.vp") and q["src_shape"] == (56, 3)
    assert q["rec_shape"] == (33600, 7) and q["ref_shape"] == (1200, 3)
    assert (q["nrows"], q["ncols"], q["dx"], q["dz"]) == (160, 600, 12.5, 8.0)
    assert (q["nux"], q["nuz"], q["schemetag"], q["lsrc"]) == (60.0, 20.0, 3, 2)
    assert q["NumofR"] == 2


def test_load_survey2d(tmp_path):
    nr, nc, dx, dz = 20, 30, 12.5, 8.0
    vp = np.full((nr, nc), 2000.0)
    srcA = np.array([[dz, 5 * dx, 3.0], [dz, 20 * dx, 2.0]])
    recA = np.array([[dz, (j + 1) * dx, -1.0, 0.0, 0.1 * j, 0.0, 0.2 * j] for j in range(5)])
    Z = np.stack([np.full(2 * nc, 100.0), -np.ones(2 * nc), np.zeros(2 * nc)], 1)
    for name, a in [("m.vp", vp), ("m.src", srcA), ("m.rec", recA), ("m.ref", Z)]:
        F.write_bin(str(tmp_path / name), a)
    F.write_inp2d(str(tmp_path / "m.inp"), "m.vp", "m.src", srcA.shape, "m.rec", recA.shape,
                  "m.ref", Z.shape, nr, nc, dx, dz)
    S = F.load_survey2d(str(tmp_path / "m.inp"))
    assert np.array_equal(S["src"], [[0, 4], [0, 19]]) and np.array_equal(S["srnum"], [3, 2])
    assert np.array_equal(S["rec"][:, 1], np.arange(5)) and np.array_equal(S["rec"][:, 0], np.zeros(5))
    assert S["TR_obs"].shape == (5, 2) and np.array_equal(S["TT_obs"], recA[:, 4])
    assert S["Z"].shape == (2, nc, 3) and S["params"]["NumofR"] == 2


def test_serving_workload():
    """The five serving grids and the seeded request tables of the Poisson workload."""
    assert [g[1] for g in LG.SERVING_GRIDS] == [(51, 101), (101, 401), (65, 65, 65), (65, 129, 129),
                                                 (161, 321, 321)]
    v = LG.velocity_model("3d", (12, 13, 14), 25.0, 3)
    assert v.shape == (12, 13, 14) and 0.85 * 2000 < v.min() and v.max() < 1.15 * (2000 + 0.6 * 11 * 25)
    rng = np.random.default_rng(0)
    r = LG.make_request("3d", (12, 13, 14), 25.0, "gradient", 5, rng)
    assert r["src"].shape == (5, 3) and np.all(r["src"][:, 2] == 0) and r["srnum"].sum() == 160
    assert r["rec"].shape == (160, 3) and r["nxyz"].shape == (160, 3) and np.all(r["T_obs"] > -0.02)
    assert np.all((r["rec"][:, 0] < 13) & (r["rec"][:, 1] < 14))
    t = LG.make_request("2d", (20, 30), 25.0, "table", 7, rng)
    assert t["src"].shape == (7, 2) and "rec" not in t and np.all(t["srnum"] == 0)
    recs = [dict(grid=0, kind="table", nsrc=2, latency=0.002, queue=0.0, t_submit=0.0, t_done=0.002),
            dict(grid=2, kind="gradient", nsrc=8, latency=0.050, queue=0.001, t_submit=0.5, t_done=0.55)]
    s = LG.summarize(recs)
    assert s["requests"] == 2 and s["p50_ms"] == 26.0 and set(s["per_grid"]) == {"51x101", "65x65x65"}
