"""Writes the golden fixtures under tests/golden/ from the reference itself.

Run in the build container only (it needs /root/reference and oracle/_ref built by
`make -C oracle ref`):

    python tests/golden/gen_golden.py

Fixtures (numpy .npz, no pickles):
  golden_3d.npz      -- fsm3d / adjoint_fsm3d outputs of the compiled reference
                        (src/T_pfsm3d.h, src/T_padjoint3d.h) on two seeded cases.
  golden_2d.npz      -- RT_fsm2d / Tadjoint_fsm2d / Radjoint_fsm2d (both overloads)
                        of the compiled reference (src/RT_pfsm2d.h, src/RT_padjoint2d.h)
                        on a seeded case with two non-flat reflectors.
  example2d.npz      -- the reference's shipped 2D example: models/RT_SynMod2D_FW.vp,
                        models/RT_SynMod2D.src, models/RT_SynMod2D.ref and, for three
                        sources, the receiver rows of models/RT_SynMod2D_Inv.rec whose
                        columns 4 and 6 are the reference's own FP64 forward picks
                        (mfiles/RT_SynMod2D.m:98-99,126-131).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import cases  # noqa: E402
import oracle_lib as O  # noqa: E402

REF_MODELS = "/root/reference/models"


def gen3d():
    out = {}
    specs = [(101, (7, 9, 11), (10.0, 10.0, 10.0)), (102, (8, 7, 10), (12.5, 10.0, 8.0))]
    for tag, (seed, shape, spacing) in zip("ab", specs):
        c = cases.case3d(seed, shape=shape, spacing=spacing, nsrc=3, nrec_per=7)
        nz, nr, nc = shape
        N = nz * nr * nc
        Ts, Ls, EEs = [], [], []
        for s in range(3):
            U = np.empty(N)
            O.ref3d().ref_fsm3d(c["vp"].ravel().copy(), U, c["src"].ravel().copy(), s, 100,
                                c["dx"], c["dy"], c["dz"], nr, nc, nz)
            lam = np.empty(N)
            EE = O.ref3d().ref_adjoint_fsm3d(lam, U.copy(), c["T_obs"].copy(), c["rec"].ravel().copy(),
                                             c["nxyz"].ravel().copy(), c["src"].ravel().copy(),
                                             c["srnum"].copy(), s, 100, c["dx"], c["dy"], c["dz"],
                                             nr, nc, nz)
            Ts.append(U.reshape(shape))
            Ls.append(lam.reshape(shape))
            EEs.append(EE)
        for k in ("vp", "src", "srnum", "rec", "nxyz", "T_obs"):
            out[f"{tag}_{k}"] = c[k]
        out[f"{tag}_spacing"] = np.array(spacing)
        out[f"{tag}_T"] = np.array(Ts)
        out[f"{tag}_lam"] = np.array(Ls)
        out[f"{tag}_EE"] = np.array(EEs)
    np.savez_compressed(os.path.join(HERE, "golden_3d.npz"), **out)


def gen2d():
    c = cases.case2d(201, shape=(24, 31), NumofR=2, nsrc=2, nrec_per=7)
    nr, nc = c["vp"].shape
    N = nr * nc
    R = c["NumofR"]
    out = {k: c[k] for k in ("vp", "Z", "src", "srnum", "rec", "nxy", "T_obs", "TR_obs")}
    out["spacing"] = np.array([c["dx"], c["dz"]])
    outs = {k: [] for k in ("T", "TTd", "TTu", "lamT", "EET", "lamT_wr", "EET_wr", "lamR", "EER",
                            "lamR_wr", "EER_wr")}
    WR = 0.3
    for s in range(2):
        U = np.empty(N)
        TTd = np.empty(R * N)
        TTu = np.empty(R * N)
        O.ref2d().ref_rt_fsm2d(c["vp"].ravel().copy(), U, TTd, TTu, c["Z"].ravel().copy(),
                               c["src"].ravel().copy(), c["rec"].ravel().copy(), c["srnum"].copy(),
                               s, R, 500, c["dx"], c["dz"], nr, nc)
        args = (c["rec"].ravel().copy(), c["nxy"].ravel().copy(), c["src"].ravel().copy(),
                c["srnum"].copy(), s)
        lt = np.empty(N)
        et = O.ref2d().ref_tadjoint_fsm2d(lt, U.copy(), c["T_obs"].copy(), *args, 500, c["dx"],
                                          c["dz"], nr, nc)
        ltw = np.empty(N)
        etw = O.ref2d().ref_tadjoint_fsm2d_wr(ltw, U.copy(), c["T_obs"].copy(), WR, *args, 500,
                                              c["dx"], c["dz"], nr, nc)
        lr = np.empty(N)
        er = O.ref2d().ref_radjoint_fsm2d(lr, TTd.copy(), TTu.copy(), c["TR_obs"].ravel().copy(),
                                          *args, R, 500, c["dx"], c["dz"], nr, nc)
        lrw = np.empty(N)
        erw = O.ref2d().ref_radjoint_fsm2d_wr(lrw, TTd.copy(), TTu.copy(), c["TR_obs"].ravel().copy(),
                                              WR, *args, R, 500, c["dx"], c["dz"], nr, nc)
        for k, v in (("T", U), ("TTd", TTd), ("TTu", TTu), ("lamT", lt), ("EET", et),
                     ("lamT_wr", ltw), ("EET_wr", etw), ("lamR", lr), ("EER", er),
                     ("lamR_wr", lrw), ("EER_wr", erw)):
            outs[k].append(v)
    for k, v in outs.items():
        out[k] = np.array(v)
    out["WR"] = np.array(WR)
    np.savez_compressed(os.path.join(HERE, "golden_2d.npz"), **out)


def gen_example2d():
    vp = np.fromfile(os.path.join(REF_MODELS, "RT_SynMod2D_FW.vp"), dtype="<f8").reshape(160, 600)
    src = np.fromfile(os.path.join(REF_MODELS, "RT_SynMod2D.src"), dtype="<f8").reshape(56, 3)
    ref = np.fromfile(os.path.join(REF_MODELS, "RT_SynMod2D.ref"), dtype="<f8").reshape(1200, 3)
    rec = np.fromfile(os.path.join(REF_MODELS, "RT_SynMod2D_Inv.rec"), dtype="<f8").reshape(33600, 7)
    ids = np.array([0, 17, 55])
    rows = np.concatenate([rec[600 * s:600 * s + 600] for s in ids])
    np.savez_compressed(os.path.join(HERE, "example2d.npz"), vp=vp, src=src, ref=ref,
                        src_ids=ids, rec_rows=rows, grid=np.array([160, 600]),
                        spacing=np.array([12.5, 8.0]))


if __name__ == "__main__":
    if not O.have_ref():
        sys.exit("oracle/_ref is not built: make -C oracle ref")
    gen3d()
    gen2d()
    gen_example2d()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
