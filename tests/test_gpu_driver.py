"""The tomography drivers (pfast_amd/bin/pfast_rt2d, pfast_t3d: batched FP32 GPU
gradients) against the reference's own drivers (RT_PFAST2D, T_PFAST3D, FP64 CPU)
on the same small surveys (tests/golden/driver2d.npz, driver3d.npz, written by
tests/golden/gen_driver_golden.py).

How closely can another arithmetic follow the reference's trajectory?  The
fixtures also hold the reference's own runs from starting models perturbed by
1e-5 relative (E_ens, steps_ens).  Its first trial step is 500 / max|gradient|,
which puts the first secant trials far into the nonlinear range of the misfit,
and the accepted steps are unevaluated secant extrapolations: a 1e-5 change of the
input moves its own step 1 by up to 80% in 3D and its final misfit by up to 20% in
2D.  So the bounds are:
  * the first line-search trial, which is well conditioned: misfit E0 within 3e-5,
    the trial step (500 / max|Dj|, a max-norm of the FP32 gradient) within 1e-3,
    phi0' within 1e-3, the trial misfit within 1e-3 and its directional derivative
    within 1e-2 (relative);
  * per iteration, misfits and accepted steps within max(1%, 3x the reference's
    own spread under the perturbation), while every perturbed run is still going;
  * the inversion ends no worse than the reference's runs typically end: its final
    misfit at most 1% above the median of the reference's final misfit and the last
    misfits of its perturbed runs (in 2D these end 5% to 18% above the reference's,
    in 3D two end within 3% of it and two 9x and 29x above it);
  * the models after iteration 1 and at the end differ from the reference's by at
    most 3x the reference's own spread (its perturbed runs against its run);
  * conditional parity along our own path: the FP64 oracle's misfit at every model
    the driver writes equals the misfit the driver logged for it (1e-4; in 2D the
    transmission and reflection parts each), and in 3D, at the driver's own
    first-iteration trial models (its direction, its steps), the oracle's misfit
    and directional derivative equal the driver's (1e-4, 1e-2).
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden")
BIN = os.path.join(ROOT, "pfast_amd", "bin")
E_TOL = 1e-2


def _run(exe, d, args):
    r = subprocess.run([os.path.join(BIN, exe), "m.inp", *args], cwd=d, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    return r.stdout


def _trials(log):
    """Line-search trials (alpha0, alpha1, phi0, phi1, dphi0, dphi1) of a driver log,
    grouped by iteration."""
    import re
    groups = []
    for chunk in re.split(r"Stepsize=[-0-9.eE+]+", log)[:-1]:
        rows = re.findall(r"\(alpha0,alpha1,phi0,phi1,dphi0,dphi1\)=\(([^\n]*)", chunk)
        groups.append(np.array([[float(v) for v in r.split()[:6]] for r in rows]).reshape(-1, 6))
    return groups


def _rel(a, b):
    return abs(a - b) / abs(b)


def _compare(E, steps, log, g):
    """Bounds of the module docstring against fixture g (reference run + ensemble)."""
    Eref, sref, tref = g["E"], g["steps"], g["trials"]
    Eens, sens = g["E_ens"], g["steps_ens"]
    print("E ours", E.tolist(), "\nE ref ", Eref.tolist(), "\nE ens ", Eens.tolist(),
          "\nsteps ours", steps.tolist(), "\nsteps ref ", sref.tolist(), "\nsteps ens ",
          sens.tolist(), flush=True)
    t = _trials(log)[0][0]
    r = tref[0]
    print("first trial ours", t.tolist(), "\nfirst trial ref ", r.tolist(), flush=True)
    assert _rel(E[0], Eref[0]) <= 3e-5
    assert _rel(t[1], r[1]) <= 1e-3
    assert _rel(t[4], r[4]) <= 1e-3
    assert _rel(t[3], r[3]) <= 1e-3
    assert _rel(t[5], r[5]) <= 1e-2
    alive = ~np.isnan(Eens).any(axis=0)
    n = min(len(E), len(Eref), int(np.argmin(alive)) if not alive.all() else Eens.shape[1])
    for k in range(n):
        spread = np.max(np.abs(Eens[:, k] - Eref[k])) / Eref[k]
        tol = max(E_TOL, 3 * spread)
        print(f"iteration {k}: E {E[k]:.6g} ref {Eref[k]:.6g} (tol {tol:.3g})", flush=True)
        assert _rel(E[k], Eref[k]) <= tol
        if k + 1 < n:
            sspread = np.max(np.abs(sens[:, k] - sref[k])) / abs(sref[k])
            stol = max(E_TOL, 3 * sspread)
            print(f"  step {steps[k]:.6g} ref {sref[k]:.6g} (tol {stol:.3g})", flush=True)
            assert _rel(steps[k], sref[k]) <= stol
    last = [row[~np.isnan(row)][-1] for row in Eens]
    print(f"final: E {E[-1]:.6g} ref {Eref[-1]:.6g} perturbed runs {last}", flush=True)
    assert E[-1] <= (1 + E_TOL) * np.median([Eref[-1], *last])


def _model_error(label, ours, ref, v0, ens):
    """The driver's model against the reference's FP64 one (relative L2 of the
    difference over the reference model, and over the reference's own update
    ref - v0), bounded by the reference's own spread: 3x the largest difference
    between its runs from the perturbed starting models (ens) and ref."""
    d = np.linalg.norm(ours - ref)
    r_model, r_update = d / np.linalg.norm(ref), d / np.linalg.norm(ref - v0)
    spread = max(np.linalg.norm(e - ref) for e in ens) / np.linalg.norm(ref)
    print(f"{label}: |ours - ref| / |ref| = {r_model:.3e} (/ |ref - v0| = {r_update:.3e}); "
          f"reference's own spread {spread:.3e}", flush=True)
    assert r_model <= max(3 * spread, 1e-6)


def _steps(log):
    import re
    return np.array([float(v) for v in re.findall(r"Stepsize=([-0-9.eE+]+)", log)])


def test_rt2d_driver_matches_reference_trajectory(tmp_path):
    from pfast_amd import formats as F
    g = np.load(os.path.join(GOLD, "driver2d.npz"))
    nr, nc = (int(v) for v in g["grid"])
    dx, dz = (float(v) for v in g["spacing"])
    d = str(tmp_path)
    for name, key in [("m.vp", "v0"), ("m.src", "srcA"), ("m.rec", "recA"), ("m.ref", "Z"),
                      ("m.wr", "WR")]:
        F.write_bin(os.path.join(d, name), g[key])
    F.write_inp2d(os.path.join(d, "m.inp"), "m.vp", "m.src", g["srcA"].shape, "m.rec",
                  g["recA"].shape, "m.ref", g["Z"].shape, nr, nc, dx, dz, vpfile="final.vp",
                  nux=60, nuz=20, twfile="m.wr", schemetag=3, lsrc=2)
    log = _run("pfast_rt2d", d, ["50", "3"])
    E = np.loadtxt(os.path.join(d, "m_GlobalErr_TR_T_R.txt"), ndmin=2)
    assert _rel(E[0, 1], g["Et"][0]) <= 1e-4 and _rel(E[0, 2], g["Er"][0]) <= 1e-4
    _compare(E[:, 0], _steps(log), log, g)
    _conditional_parity2d(d, E)
    _model_error("model after iteration 1", F.read_bin(os.path.join(d, "m_1"), (nr, nc)), g["vp1"],
                 g["v0"], g["vp1_ens"])
    _model_error("final model", F.read_bin(os.path.join(d, "final.vp"), (nr, nc)), g["vpF"],
                 g["v0"], g["vpF_ens"])


def _conditional_parity2d(d, E):
    """FP64 oracle misfits (joint, transmission, reflection) at the driver's models."""
    import oracle_lib as O
    from pfast_amd import formats as F
    S = F.load_survey2d(os.path.join(d, "m.inp"))
    p = S["params"]
    R = p["NumofR"]
    for k in range(1, len(E)):
        vp = F.read_bin(os.path.join(d, f"m_{k}"), (p["nrows"], p["ncols"]))
        Et = Er = 0.0
        for s in range(len(S["src"])):
            o = O.rt_fsm2d(vp, S["Z"], S["src"], s, R, p["dx"], p["dz"])
            assert o["status"] == 0
            Et += O.tadjoint_fsm2d(o["T"], S["TT_obs"], 1.0, S["rec"], S["nxy"], S["src"],
                                   S["srnum"], s, p["dx"], p["dz"])[1]
            Er += O.radjoint_fsm2d(o["TTu"], S["TR_obs"], 1.0, S["rec"], S["nxy"], S["srnum"], s,
                                   R, p["dx"], p["dz"])[1]
        print(f"model after iteration {k}: oracle E {Et + Er:.8g} ({Et:.6g} + {Er:.6g}), driver "
              f"{E[k].tolist()}", flush=True)
        assert _rel(E[k, 0], Et + Er) <= 1e-4
        assert _rel(E[k, 1], Et) <= 1e-4 and _rel(E[k, 2], Er) <= 1e-4


def test_t3d_driver_matches_reference_trajectory(tmp_path):
    from pfast_amd import formats as F
    g = np.load(os.path.join(GOLD, "driver3d.npz"))
    nr, nc, nz = (int(v) for v in g["grid"])
    h = float(g["spacing"][0])
    d = str(tmp_path)
    for name, key in [("m.vp", "v0"), ("m.src", "srcA"), ("m.rec", "recA")]:
        F.write_bin(os.path.join(d, name), g[key])
    F.write_inp3d(os.path.join(d, "m.inp"), "m.vp", "m.src", g["srcA"].shape, "m.rec",
                  g["recA"].shape, nr, nc, nz, h, h, h, vpfile="final.vp", nux=50, nuy=50,
                  nuz=25, schemetag=3, lsrc=2)
    log = _run("pfast_t3d", d, ["50"])
    E = np.loadtxt(os.path.join(d, "m_GlobalErr.txt"), ndmin=1)
    _compare(E, _steps(log), log, g)
    _conditional_parity3d(d, g, E, log)
    shape = (nz, nr, nc)
    _model_error("model after iteration 1", F.read_bin(os.path.join(d, "m.vp1"), shape), g["vp1"],
                 g["v0"], g["vp1_ens"])
    _model_error("final model", F.read_bin(os.path.join(d, "final.vp"), shape), g["vpF"],
                 g["v0"], g["vpF_ens"])


def _conditional_parity3d(d, g, E, log):
    """FP64 oracle misfits and directional derivatives at the driver's own models."""
    import oracle_lib as O
    from pfast_amd import formats as F
    nr, nc, nz = (int(v) for v in g["grid"])
    h = float(g["spacing"][0])
    shape = (nz, nr, nc)
    srcA, recA = g["srcA"], g["recA"]
    src = np.stack([F.map_index(srcA[:, a], h) for a in range(3)], 1).astype(np.int32)
    rec = np.stack([F.map_index(recA[:, a], h) for a in range(3)], 1).astype(np.int32)
    nxyz, tobs = np.ascontiguousarray(recA[:, 3:6]), np.ascontiguousarray(recA[:, 6])
    srnum = srcA[:, 3].astype(np.int32)

    def oracle_eval(vp):
        G, EE = np.zeros(shape), 0.0
        for s in range(len(src)):
            T, _ = O.fsm3d(vp, src, s, h, h, h)
            lam, e, _ = O.adjoint_fsm3d(T, tobs, rec, nxyz, src, srnum, s, h, h, h)
            G -= lam / vp ** 3
            EE += e
        return G, EE

    vp0 = g["v0"]
    for k in range(1, len(E)):
        _, Ek = oracle_eval(F.read_bin(os.path.join(d, f"m.vp{k}"), shape))
        print(f"model after iteration {k}: oracle E {Ek:.8g}, driver {E[k]:.8g}", flush=True)
        assert _rel(E[k], Ek) <= 1e-4
    dk = F.read_bin(os.path.join(d, "m_af_regu_1.vp"), shape)
    for t in _trials(log)[0]:
        G, phi = oracle_eval(vp0 + t[1] * dk)
        dphi = -(dk * O.regu_np(G, vp0, (25.0, 50.0, 50.0), (h, h, h))).sum()
        print(f"trial alpha {t[1]:.6g}: oracle phi {phi:.6g} dphi {dphi:.6g}; driver {t[3]:.6g} "
              f"{t[5]:.6g}", flush=True)
        assert _rel(t[3], phi) <= 1e-4
        assert _rel(t[5], dphi) <= 1e-2


def test_rt2d_forward_mode_writes_reference_picks(tmp_path):
    """Forward modelling (the reference's -DNoInverse build) on the shipped 2D example
    geometry: _Tr2D.bin holds T and the reflector picks of every receiver row, within
    the FP32 traveltime bound of the reference's own FP64 picks."""
    from pfast_amd import formats as F
    e = np.load(os.path.join(GOLD, "example2d.npz"))
    dx, dz = (float(v) for v in e["spacing"])
    ids = [int(s) for s in e["src_ids"]]
    srcA = e["src"][ids].copy()
    srcA[:, 2] = 600
    recA = e["rec_rows"][:, :4]
    d = str(tmp_path)
    F.write_bin(os.path.join(d, "m.vp"), e["vp"])
    F.write_bin(os.path.join(d, "m.src"), srcA)
    F.write_bin(os.path.join(d, "m.rec"), recA)
    F.write_bin(os.path.join(d, "m.ref"), e["ref"])
    F.write_inp2d(os.path.join(d, "m.inp"), "m.vp", "m.src", srcA.shape, "m.rec", recA.shape,
                  "m.ref", e["ref"].shape, 160, 600, dx, dz)
    _run("pfast_rt2d", d, ["--forward", "--no-shot-gather"])
    Tr = F.read_bin(os.path.join(d, "m_Tr2D.bin"), (len(ids) * 600, 3))
    rows = e["rec_rows"]
    et = np.abs(Tr[:, 0] - rows[:, 4]).max() / rows[:, 4].max()
    eu = np.abs(Tr[:, 2] - rows[:, 6]).max() / rows[:, 6].max()
    print(f"forward picks: transmission err {et:.2e}, reflector-2 err {eu:.2e}", flush=True)
    assert et <= 1e-5 and eu <= 1e-5


def test_t3d_driver_resume_matches_reference(tmp_path):
    """The reference's resume convention (src/T_PFAST3D.cpp:141-147,345-384): vp0file
    m.vp2 ends in the number of finished iterations; the run reloads the gradients
    m_bf_regu_2.vp / m_bf_regu_3.vp, global_E.tmp and alpha1.tmp and continues at
    iteration 3.  Its first trial comes from those files and one evaluation, so it
    must match the reference's resumed run (fixture res_*) closely."""
    from pfast_amd import formats as F
    g = np.load(os.path.join(GOLD, "driver3d.npz"))
    nr, nc, nz = (int(v) for v in g["grid"])
    h = float(g["spacing"][0])
    shape = (nz, nr, nc)
    d = str(tmp_path)
    F.write_bin(os.path.join(d, "m.vp2"), g["res_vp2"].reshape(shape))
    F.write_bin(os.path.join(d, "m_bf_regu_2.vp"), g["res_bf2"].reshape(shape))
    F.write_bin(os.path.join(d, "m_bf_regu_3.vp"), g["res_bf3"].reshape(shape))
    F.write_bin(os.path.join(d, "m.src"), g["srcA"])
    F.write_bin(os.path.join(d, "m.rec"), g["recA"])
    with open(os.path.join(d, "global_E.tmp"), "w") as fh:
        fh.write("".join(f"{e:.17g}\n" for e in g["res_gE"]))
    with open(os.path.join(d, "alpha1.tmp"), "w") as fh:
        fh.write(f"{float(g['res_a1'][0]):.17g}\n")
    F.write_inp3d(os.path.join(d, "m.inp"), "m.vp2", "m.src", g["srcA"].shape, "m.rec",
                  g["recA"].shape, nr, nc, nz, h, h, h, vpfile="final.vp", nux=50, nuy=50,
                  nuz=25, schemetag=3, lsrc=2)
    log = _run("pfast_t3d", d, ["50"])
    assert "Iterating No. 2 resumed" in log
    E = np.loadtxt(os.path.join(d, "m_GlobalErr.txt"), ndmin=1)
    ref = g["res_E"]
    t = _trials(log)[0][0]
    r = g["res_trials"][0]
    print("E ours", E.tolist(), "\nE ref ", ref.tolist(), "\nfirst trial ours", t.tolist(),
          "\nfirst trial ref ", r.tolist(), flush=True)
    assert len(E) == len(ref)
    assert np.all(np.abs(E[:3] - ref[:3]) <= 1e-5 * ref[:3])      # carried over from the files
    assert np.all(np.abs(E[3:] - ref[3:]) <= 1e-4 * ref[3:])      # one GPU evaluation each
    assert t[1] == r[1] and _rel(t[2], r[2]) <= 1e-5               # trial step, phi0 from files
    assert _rel(t[4], r[4]) <= 1e-4                                # phi0' from the file gradients
    assert _rel(t[3], r[3]) <= 1e-4                                # misfit at the trial model
    for k in range(3, len(E)):
        assert os.path.exists(os.path.join(d, f"m.vp{k}"))


def _residuals(log):
    import re
    return np.array([[float(x) for x in m] for m in re.findall(
        r"Average Time Residual \(TR\): ([-0-9.eE+]+) ms\s+\(T\): ([-0-9.eE+]+) ms\s+"
        r"\(R\): ([-0-9.eE+]+) ms", log)])


@pytest.mark.parametrize("tag", [3, 1, 2])
def test_rt2d_example_inversion_matches_reference_logs(tmp_path, tag):
    """The reference's shipped 2D inversion example (tests/golden/example2d_inv.npz,
    written by gen_example2d_inv.py from models/RT_SynMod2D_* and the logs
    par/RT_SynMod2D_Inv_{RT,Trans,Ref}.out): 160 x 600, 56 sources x 600 receivers, two
    reflectors, LM 50, RTtag 3 / 1 / 2.  Anchors at the logs' 6 digits: the residuals of
    iterations 1 and 2 (1e-4, 1e-3), and iteration 1's line search: the trial step
    500 / max|Dj| (1e-3), phi0 and phi0' (1e-4, 1e-3), the misfit and derivative at
    both trials and the second trial step (1e-2), the accepted step (1e-2)."""
    from pfast_amd import formats as F
    e = np.load(os.path.join(GOLD, "example2d_inv.npz"))
    nr, nc = (int(v) for v in e["grid"])
    dx, dz = (float(v) for v in e["spacing"])
    d = str(tmp_path)
    n = len(e["tt"])
    rec = np.concatenate([np.tile(e["geom"], (n // len(e["geom"]), 1)),
                          e["tt"].astype(np.float64)[:, None], np.zeros((n, 1)),
                          e["tr2"].astype(np.float64)[:, None]], 1)
    F.write_bin(os.path.join(d, "m.vp"), e["vp"].astype(np.float64))
    F.write_bin(os.path.join(d, "m.src"), e["src"])
    F.write_bin(os.path.join(d, "m.rec"), rec)
    F.write_bin(os.path.join(d, "m.ref"), e["ref"])
    F.write_bin(os.path.join(d, "m.wr"), np.full((nr, nc), float(e["wr"][0])))
    F.write_inp2d(os.path.join(d, "m.inp"), "m.vp", "m.src", e["src"].shape, "m.rec", rec.shape,
                  "m.ref", e["ref"].shape, nr, nc, dx, dz, vpfile="final.vp",
                  nux=float(e["nu"][0]), nuz=float(e["nu"][1]), twfile="m.wr", schemetag=3, lsrc=2)
    log = _run("pfast_rt2d", d, ["50", str(tag), "--max-iter", "2", "--no-shot-gather"])
    res = _residuals(log)
    t = _trials(log)[0]
    steps = _steps(log)
    rres, rt, rstep = e[f"res{tag}"], e[f"trials{tag}"], e[f"step{tag}"]
    print(f"RTtag {tag}\nresiduals ours {res[:2].tolist()}\nresiduals ref  {rres.tolist()}\n"
          f"trials ours {t.tolist()}\ntrials ref  {rt.tolist()}\nstep ours {steps[:1]} ref {rstep}",
          flush=True)
    assert np.all(np.abs(res[0] - rres[0]) <= 1e-4 * rres[0])
    assert len(t) >= 2
    assert _rel(t[0, 1], rt[0, 1]) <= 1e-3
    assert _rel(t[0, 2], rt[0, 2]) <= 1e-4 and _rel(t[0, 4], rt[0, 4]) <= 1e-3
    for k in range(2):
        assert _rel(t[k, 3], rt[k, 3]) <= 1e-2 and _rel(t[k, 5], rt[k, 5]) <= 1e-2
    assert _rel(t[1, 1], rt[1, 1]) <= 1e-2
    assert _rel(steps[0], rstep[0]) <= 1e-2
    assert np.all(np.abs(res[1] - rres[1]) <= 1e-3 * rres[1])


def test_rt2d_config2_tomography_lm5_through_the_ring_wrap(tmp_path):
    """The 2D first-arrival tomography workload (101 x 401, h = 10 m, 60 surface sources x
    200 receivers, picks = the true model's FP64 traveltimes + N(0, 2 ms), starting model
    the 1D gradient; pfast_amd.synth.config2d(2)) through the drop-in driver with L-BFGS
    memory 5 for 10 iterations (RTtag 1, transmission only): iterations 7..10 use the ring
    past the point where the reference's own ai[i] leaves its array.  Conditional parity
    at every model the driver writes: the FP64 oracle's misfit equals the driver's logged
    misfit within 1e-4.  The final model must be closer to the true model than the start."""
    from concurrent.futures import ThreadPoolExecutor

    import oracle_lib as O
    import pfast_amd
    from pfast_amd import formats as F
    from pfast_amd import synth
    cfg = synth.config2d(2)
    h, nr, nc = cfg["h"], cfg["nrows"], cfg["ncols"]
    nsrc, nrec = cfg["src"].shape[0], int(cfg["srnum"][0])
    rng = np.random.default_rng(20261204)
    tt = np.empty(nsrc * nrec)
    for s in range(nsrc):
        T = pfast_amd.RT_fsm2d(cfg["vp"], None, cfg["src"], s, 0, 500, h, h)["T"]
        rows = slice(s * nrec, (s + 1) * nrec)
        tt[rows] = T[cfg["rec"][rows, 0], cfg["rec"][rows, 1]]
    tt += rng.normal(0.0, 0.002, tt.size)
    d = str(tmp_path)
    rec = np.concatenate([cfg["rec_xz"], cfg["nxy"], tt[:, None]], 1)
    F.write_bin(os.path.join(d, "m.vp"), cfg["vp0"])
    F.write_bin(os.path.join(d, "m.src"), cfg["src_tab"])
    F.write_bin(os.path.join(d, "m.rec"), rec)
    F.write_bin(os.path.join(d, "m.ref"), np.zeros((0, 3)))
    F.write_bin(os.path.join(d, "m.wr"), np.zeros((nr, nc)))
    F.write_inp2d(os.path.join(d, "m.inp"), "m.vp", "m.src", cfg["src_tab"].shape, "m.rec",
                  rec.shape, "m.ref", (0, 3), nr, nc, h, h, vpfile="final.vp", nux=60, nuz=20,
                  twfile="m.wr", schemetag=3, lsrc=2)
    # the model starts at 1500 m/s: the reference's 1600 m/s clamp would lift its top rows
    # at the first update (src/RT_padjoint2d.h:17), so the clamp is lowered to 1400
    log = _run("pfast_rt2d", d, ["5", "1", "--max-iter", "10", "--no-shot-gather", "--vmin",
                                 "1400"])
    E = np.loadtxt(os.path.join(d, "m_GlobalErr_TR_T_R.txt"), ndmin=2)
    print("misfits", E[:, 0].tolist(), flush=True)
    print("\n".join(ln for ln in log.split("\n") if "Stepsize" in ln), flush=True)
    nmod = len(E) - 1
    assert nmod >= 7, "the run stopped before the ring wrapped"
    S = F.load_survey2d(os.path.join(d, "m.inp"))

    def oracle_misfit(vp):
        def one(s):
            o = O.rt_fsm2d(vp, None, S["src"], s, 0, h, h)
            return O.tadjoint_fsm2d(o["T"], S["TT_obs"], 1.0, S["rec"], S["nxy"], S["src"],
                                    S["srnum"], s, h, h)[1]
        with ThreadPoolExecutor(8) as ex:
            return sum(ex.map(one, range(nsrc)))

    for k in range(1, nmod + 1):
        vp = F.read_bin(os.path.join(d, f"m_{k}"), (nr, nc))
        Eo = oracle_misfit(vp)
        print(f"model after iteration {k}: oracle E {Eo:.8g}, driver {E[k, 0]:.8g}", flush=True)
        assert _rel(E[k, 0], Eo) <= 1e-4
    final = F.read_bin(os.path.join(d, f"m_{nmod}"), (nr, nc))
    e0 = np.linalg.norm(cfg["vp0"] - cfg["vp"]) / np.linalg.norm(cfg["vp"])
    e1 = np.linalg.norm(final - cfg["vp"]) / np.linalg.norm(cfg["vp"])
    print(f"model error vs the true model: start {e0:.4e}, after {nmod} iterations {e1:.4e}; "
          f"misfit {E[0, 0]:.6g} -> {E[-1, 0]:.6g}", flush=True)
    assert e1 < e0 and E[-1, 0] < E[0, 0]
