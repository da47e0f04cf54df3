"""The 2D first-arrival forward workload (101 x 401, h = 10 m, Vp = 1500 + 0.8 z with three
seeded Gaussian anomalies, 20 evenly spaced surface sources, 200 receivers;
pfast_amd.synth.config2d(1)), the reference's forward mode (src/RT_PFAST2D.cpp with
-DNoInverse: RT_fsm2d per source, src/RT_pfsm2d.h:116-380) at its full size:

- the single-source FP64 form is bit-identical to the FP64 oracle (fields and cycle
  counts) on two sources, one at each end of the line;
- the batched FP32 shot set is within 1e-5 (max-norm, relative, over reached nodes) of the
  FP64 form for every one of the 20 sources, and so are its receiver picks.
"""
import numpy as np
import pytest

import oracle_lib as O
from pfast_amd import synth

pytestmark = pytest.mark.gpu

T_TOL = 1e-5


@pytest.fixture(scope="module")
def cfg1():
    return synth.config2d(1)


def test_config1_plain_fp64_bitexact_vs_oracle(cfg1):
    import pfast_amd
    h = cfg1["h"]
    for s in (0, 19):
        r = pfast_amd.RT_fsm2d(cfg1["vp"], None, cfg1["src"], s, 0, 500, h, h)
        o = O.rt_fsm2d(cfg1["vp"], None, cfg1["src"], s, 0, h, h)
        print(f"source {s}: cycles GPU {r['cycles'].tolist()} oracle {o['cycles'][:1].tolist()}",
              flush=True)
        assert o["status"] == 0
        assert np.array_equal(r["T"].view(np.int64), o["T"].view(np.int64)), s
        assert r["cycles"][0] == o["cycles"][0]


def test_config1_batched_fp32_all_sources_vs_fp64(cfg1):
    import pfast_amd
    h = cfg1["h"]
    nsrc = cfg1["src"].shape[0]
    nrec = int(cfg1["srnum"][0])
    rec = cfg1["rec"]
    ctx = pfast_amd.Context(0)
    m = pfast_amd.Model2D(ctx, cfg1["vp"], h, h, cfg1["Z"])
    sh = pfast_amd.Shots2D(m, cfg1["src"], cfg1["srnum"], rec, cfg1["nxy"])
    sh.run(adjoint=False, MAXSwp=500)
    worst, worst_pick = 0.0, 0.0
    for s in range(nsrc):
        T64 = pfast_amd.RT_fsm2d(cfg1["vp"], None, cfg1["src"], s, 0, 500, h, h)["T"]
        T32 = sh.field(s, "T")
        reach = T64 < 1e6
        assert np.array_equal(reach, T32 < 1e6)
        err = float(np.abs(T32 - T64)[reach].max() / T64[reach].max())
        rows = slice(s * nrec, (s + 1) * nrec)
        p64 = T64[rec[rows, 0], rec[rows, 1]]
        p32 = T32[rec[rows, 0], rec[rows, 1]]
        perr = float(np.abs(p32 - p64).max() / p64.max())
        worst, worst_pick = max(worst, err), max(worst_pick, perr)
    print(f"20 sources: max T error {worst:.2e}, max pick error {worst_pick:.2e}", flush=True)
    sh.close()
    m.close()
    ctx.close()
    assert worst <= T_TOL and worst_pick <= T_TOL
