"""GPU parity of the convergence-check interval and of hipGraph replay of the batched 3D
sweep cycles.

The reference tests convergence after every sweep cycle (src/T_pfsm3d.h:181,
src/T_padjoint3d.h:165).  The batched solver keeps that test on the device
(pfast_amd/csrc/batch3d.hip cycle_end: a source's cycle count is the first cycle in which
it changed nothing, a chunk none of whose sources changed retires), so the host may check
every n cycles, and the cycles between two checks may be replayed from a captured
hipGraph: fields, adjoint fields, cycle counts, gradient and misfit must not change by a
bit, under every schedule, with several streams, with MAXSwp cutting a batch short, and
when a captured graph is replayed for a second run (after a model update).
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _solve(c, interval=1, graphs=False, dataflow="auto", streams=1, maxswp=100, runs=1,
           vp2=None):
    import pfast_amd
    ctx = pfast_amd.Context(0)
    ctx.set_check_interval(interval)
    ctx.set_graphs(graphs)
    ctx.set_dataflow(dataflow)
    ctx.set_streams(streams)
    m = pfast_amd.Model3D(ctx, c["vp"], c["dx"], c["dy"], c["dz"])
    sh = pfast_amd.Shots3D(m, c["src"], c["srnum"], c["rec"], c["nxyz"], c["T_obs"])
    outs = []
    for r in range(runs):
        if r and vp2 is not None:
            m.set_vp(vp2)
        sh.run(adjoint=True, MAXSwp=maxswp)
        g, E = sh.gradient()
        n = len(c["src"])
        outs.append((np.stack([sh.field(s) for s in range(n)]),
                     np.stack([sh.field(s, adjoint=True) for s in range(n)]), sh.cycles(), g, E))
    sh.close()
    m.close()
    ctx.close()
    return outs


def _same(a, b):
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(cases.bits(x), cases.bits(y))
    assert a[4] == b[4]


CASES = [
    (71, (20, 27, 35), (25.0, 25.0, 25.0), 21),
    (72, (13, 30, 17), (12.5, 10.0, 8.0), 9),
]


@pytest.mark.parametrize("seed,shape,spacing,nsrc", CASES)
@pytest.mark.parametrize("dataflow", [False, True, "auto"])
def test_check_interval_and_graphs_bitwise(seed, shape, spacing, nsrc, dataflow):
    c = cases.case3d(seed, shape=shape, spacing=spacing, nsrc=nsrc, nrec_per=9)
    ref = _solve(c, 1, False, dataflow)[0]
    print(f"case {seed} dataflow {dataflow}: cycles {ref[2].tolist()}", flush=True)
    for interval, graphs in ((3, False), (1, True), (4, True), (64, True)):
        print(f"interval {interval} graphs {graphs}", flush=True)
        _same(_solve(c, interval, graphs, dataflow)[0], ref)


def test_graphs_several_streams_bitwise():
    c = cases.case3d(73, shape=(18, 29, 33), spacing=(25.0, 25.0, 25.0), nsrc=35, nrec_per=8)
    ref = _solve(c, 1, False, False, streams=1)[0]
    for streams, interval in ((3, 1), (3, 5), (2, 2)):
        print(f"streams {streams} interval {interval}", flush=True)
        _same(_solve(c, interval, True, False, streams=streams)[0], ref)


def test_graphs_maxswp_cut_bitwise():
    """A batch that MAXSwp cuts short stops after the same cycle as checking every cycle."""
    c = cases.case3d(74, shape=(15, 22, 26), spacing=(20.0, 20.0, 20.0), nsrc=13, nrec_per=8)
    ref = _solve(c, 1, False, maxswp=5)[0]
    for interval in (2, 3, 8):
        _same(_solve(c, interval, True, maxswp=5)[0], ref)
        _same(_solve(c, interval, False, maxswp=5)[0], ref)


def test_graph_replay_after_model_update_bitwise():
    """The second run replays the graphs captured by the first on the updated model."""
    c = cases.case3d(75, shape=(16, 25, 21), spacing=(25.0, 25.0, 25.0), nsrc=17, nrec_per=8)
    vp2 = c["vp"] * (1.0 + 0.03 * np.cos(np.arange(c["vp"].size)).reshape(c["vp"].shape))
    ref = _solve(c, 1, False, runs=2, vp2=vp2)
    out = _solve(c, 2, True, runs=2, vp2=vp2)
    for a, b in zip(out, ref):
        _same(a, b)
    assert not np.array_equal(ref[0][0], ref[1][0])  # the update did change the fields


def test_check_interval_out_of_range_is_refused():
    import pfast_amd
    ctx = pfast_amd.Context(0)
    for bad in (0, 65):
        with pytest.raises(pfast_amd.PfastError):
            ctx.set_check_interval(bad)
    ctx.close()


def test_autotune_schedule_results_unchanged():
    """The (schedule, check interval) tuner sets one of its candidates on the context and
    the solve after it is bit-identical to the default's."""
    import pfast_amd
    c = cases.case3d(76, shape=(17, 26, 23), spacing=(25.0, 25.0, 25.0), nsrc=19, nrec_per=8)
    ref = _solve(c)[0]
    ctx = pfast_amd.Context(0)
    m = pfast_amd.Model3D(ctx, c["vp"], c["dx"], c["dy"], c["dz"])
    sh = pfast_amd.Shots3D(m, c["src"], c["srnum"], c["rec"], c["nxyz"], c["T_obs"])
    chosen, ms = sh.autotune_schedule()
    print(f"chosen {chosen}, times {ms}", flush=True)
    assert chosen in pfast_amd.Shots3D.SCHEDULE_CANDIDATES and len(ms) == 5
    assert all(t > 0 for t in ms.values())
    sh.run(adjoint=True, MAXSwp=100)
    g, E = sh.gradient()
    out = (np.stack([sh.field(s) for s in range(len(c["src"]))]),
           np.stack([sh.field(s, adjoint=True) for s in range(len(c["src"]))]), sh.cycles(), g, E)
    sh.close()
    m.close()
    ctx.close()
    _same(out, ref)
