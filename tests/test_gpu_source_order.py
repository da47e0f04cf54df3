"""GPU check of the batched 3D shot set's internal source order.

A shot set holds its sources in an internal order (Morton order of the source nodes, so
that a chunk of eight sources holds neighbours) and maps every per-source result back to
the caller's order.  Each source's sweep is independent of its chunk mates, so creating
the same survey with its sources listed in another order must give, source by source,
the same traveltime and adjoint fields and cycle counts bit for bit, the picks in the
caller's receiver-row order, and the same misfit and gradient up to the order of the FP64
sums over sources.
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _run(c, order):
    import pfast_amd
    off = np.concatenate([[0], np.cumsum(c["srnum"])])
    rows = np.concatenate([np.arange(off[s], off[s + 1]) for s in order])
    ctx = pfast_amd.Context(0)
    m = pfast_amd.Model3D(ctx, c["vp"], c["dx"], c["dy"], c["dz"])
    sh = pfast_amd.Shots3D(m, c["src"][order], c["srnum"][order], c["rec"][rows],
                           c["nxyz"][rows], c["T_obs"][rows])
    sh.run(adjoint=True, MAXSwp=100)
    n = len(order)
    out = dict(T=np.stack([sh.field(s) for s in range(n)]),
               L=np.stack([sh.field(s, adjoint=True) for s in range(n)]),
               cyc=sh.cycles(), picks=sh.picks(), rows=rows)
    out["g"], out["E"] = sh.gradient()
    sh.close()
    m.close()
    ctx.close()
    return out


@pytest.mark.parametrize("seed,nsrc", [(81, 21), (82, 9)])
def test_results_follow_the_callers_source_order(seed, nsrc):
    c = cases.case3d(seed, shape=(14, 27, 31), spacing=(25.0, 25.0, 25.0), nsrc=nsrc,
                     nrec_per=7)
    ident = np.arange(nsrc)
    perm = np.random.default_rng(seed).permutation(nsrc)
    a, b = _run(c, ident), _run(c, perm)
    for k, s in enumerate(perm):  # b's source k is a's source s
        assert np.array_equal(cases.bits(b["T"][k]), cases.bits(a["T"][s]))
        assert np.array_equal(cases.bits(b["L"][k]), cases.bits(a["L"][s]))
        assert np.array_equal(b["cyc"][k], a["cyc"][s])
    assert np.array_equal(b["picks"], a["picks"][b["rows"]])
    assert abs(b["E"] - a["E"]) <= 1e-14 * abs(a["E"])
    assert np.linalg.norm(b["g"] - a["g"]) <= 1e-13 * np.linalg.norm(a["g"])
