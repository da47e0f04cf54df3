"""GPU parity of the persistent dataflow schedule of the batched 3D sweeps.

One dataflow launch per sweep cycle (pfast_amd/csrc/batch3d.hip fwd3d_flow /
adj3d_flow, schedule rules in pfast_amd/csrc/flow.h) performs the same node updates
in the same order as one launch per tile anti-diagonal and ordering, which the other
GPU tests hold to the FP64 oracle (src/T_pfsm3d.h:131-185, src/T_padjoint3d.h:73-172):
traveltimes, adjoint fields, cycle counts, gradient and misfit must agree bit for bit,
for every tile shape, with exact tile skipping on and off, with several chunks of
sources and grids whose edge tiles are partial.
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _solve(c, nsrc, dataflow, variant=0, skip=True, maxswp=100):
    import pfast_amd
    ctx = pfast_amd.Context(0)
    ctx.set_sweep(variant)
    ctx.set_tile_skip(skip)
    ctx.set_dataflow(dataflow)
    m = pfast_amd.Model3D(ctx, c["vp"], c["dx"], c["dy"], c["dz"])
    sh = pfast_amd.Shots3D(m, c["src"], c["srnum"], c["rec"], c["nxyz"], c["T_obs"])
    sh.run(adjoint=True, MAXSwp=maxswp)
    g, E = sh.gradient()
    out = (np.stack([sh.field(s) for s in range(nsrc)]),
           np.stack([sh.field(s, adjoint=True) for s in range(nsrc)]), sh.cycles(), g, E)
    sh.close()
    m.close()
    ctx.close()
    return out


def _same(a, b):
    assert all(np.array_equal(cases.bits(x), cases.bits(y)) for x, y in zip(a[:4], b[:4]))
    assert a[4] == b[4]


@pytest.mark.parametrize("seed,shape,spacing,nsrc", [
    (51, (20, 27, 35), (25.0, 25.0, 25.0), 7),
    (52, (13, 30, 17), (12.5, 10.0, 8.0), 19),
    (53, (7, 9, 41), (10.0, 10.0, 10.0), 9),
    (54, (24, 8, 8), (25.0, 25.0, 25.0), 3),
    (55, (16, 3, 21), (25.0, 25.0, 25.0), 5),
])
@pytest.mark.parametrize("skip", [True, False])
def test_dataflow_matches_diagonal_launches_bitwise(seed, shape, spacing, nsrc, skip):
    c = cases.case3d(seed, shape=shape, spacing=spacing, nsrc=nsrc, nrec_per=10)
    print(f"case seed {seed} shape {shape} nsrc {nsrc} skip {skip}", flush=True)
    ref = _solve(c, nsrc, False, skip=skip)
    out = _solve(c, nsrc, True, skip=skip)
    print(f"cycles {ref[2].tolist()}", flush=True)
    _same(out, ref)


def test_dataflow_every_variant_bitwise():
    """Each tile shape and batch depth of the dataflow kernels against the default
    per-diagonal schedule."""
    import pfast_amd
    c = cases.case3d(56, shape=(18, 29, 33), spacing=(25.0, 25.0, 25.0), nsrc=11, nrec_per=9)
    ref = _solve(c, 11, False)
    for v, name in enumerate(pfast_amd.sweep_variants()):
        print(f"variant {v} ({name})", flush=True)
        _same(_solve(c, 11, True, variant=v), ref)


def test_dataflow_exhausted_maxswp_bitwise():
    """A phase cut short by MAXSwp stops after the same cycle in both schedules."""
    c = cases.case3d(57, shape=(15, 22, 26), spacing=(20.0, 20.0, 20.0), nsrc=6, nrec_per=8)
    _same(_solve(c, 6, True, maxswp=2), _solve(c, 6, False, maxswp=2))
