"""The reference's shipped 2D inversion example as a fixture (example2d_inv.npz).

Needs the reference tree under /root/reference (read as data only):

    python tests/golden/gen_example2d_inv.py

Inputs: models/RT_SynMod2D_Inv.vp (starting model, 160 x 600), models/RT_SynMod2D.src
(56 sources), models/RT_SynMod2D_Inv_rand2p.rec (56 x 600 rows: the same 600 receivers
for every source, picks TT and TR2; TR1 is zero), models/RT_SynMod2D.ref (2 reflectors),
models/RT_SynMod2D_Inv.wr (0.5 everywhere), par/RT_SynMod2D_Inv.inp (grid, nux, nuz).
The model and the picks are kept as float32 (velocity rounding <= 3e-8 relative, picks
<= 1.2e-7 s), the geometry once.

Anchors: the reference's own logs of the three runs (par/RT_SynMod2D_Inv_RT.out,
_Trans.out, _Ref.out: RTtag 3, 1, 2), the residuals of iterations 1 and 2 and the two
line-search trials of iteration 1 with the accepted step (6 significant digits).
"""
import os
import re

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def rd(name, shape):
    return np.fromfile(os.path.join(REF, "models", name), dtype="<f8").reshape(shape)


def anchors(log):
    txt = open(os.path.join(REF, "par", log)).read()
    res = [[float(x) for x in m] for m in
           re.findall(r"Average Time Residual \(TR\): ([-0-9.eE+]+) ms\s+\(T\): ([-0-9.eE+]+) ms\s+"
                      r"\(R\): ([-0-9.eE+]+) ms", txt)]
    trials = [[float(v) for v in m.split()[:6]] for m in
              re.findall(r"\(alpha0,alpha1,phi0,phi1,dphi0,dphi1\)=\(([^\n]*)", txt)]
    steps = [float(v) for v in re.findall(r"Stepsize=([-0-9.eE+]+)", txt)]
    return np.array(res[:2]), np.array(trials[:2]), np.array(steps[:1])


def main():
    rec = rd("RT_SynMod2D_Inv_rand2p.rec", (33600, 7))
    blocks = rec.reshape(56, 600, 7)
    geom = blocks[0, :, :4].copy()
    assert all(np.array_equal(geom, blocks[k, :, :4]) for k in range(56))
    assert np.all(rec[:, 5] == 0)
    wr = rd("RT_SynMod2D_Inv.wr", (160, 600))
    assert np.all(wr == 0.5)
    out = dict(vp=rd("RT_SynMod2D_Inv.vp", (160, 600)).astype("<f4"), src=rd("RT_SynMod2D.src", (56, 3)),
               geom=geom, tt=rec[:, 4].astype("<f4"), tr2=rec[:, 6].astype("<f4"),
               ref=rd("RT_SynMod2D.ref", (1200, 3)), wr=np.array([0.5]), grid=np.array([160, 600]),
               spacing=np.array([12.5, 8.0]), nu=np.array([60.0, 20.0]))
    for tag, log in [(3, "RT_SynMod2D_Inv_RT.out"), (1, "RT_SynMod2D_Inv_Trans.out"),
                     (2, "RT_SynMod2D_Inv_Ref.out")]:
        res, trials, steps = anchors(log)
        out[f"res{tag}"], out[f"trials{tag}"], out[f"step{tag}"] = res, trials, steps
        print(tag, res.tolist(), trials.tolist(), steps.tolist())
    np.savez_compressed(os.path.join(HERE, "example2d_inv.npz"), **out)


if __name__ == "__main__":
    main()
