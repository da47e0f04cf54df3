"""GPU parity of the batched FP32 3D path against the FP64 oracle.

Tolerances (stated by the path's precision contract):
  traveltime: max|T32 - T64| / max|T64| <= 1e-5 over reached nodes
  gradient  : ||g32 - g64||_2 / ||g64||_2 <= 1e-3
  misfit    : |E32 - E64| / E64 <= 1e-4
"""
import numpy as np
import pytest

import cases
import oracle_lib as O

pytestmark = pytest.mark.gpu

T_TOL = 1e-5
G_TOL = 1e-3
E_TOL = 1e-4


@pytest.fixture(scope="module")
def ctx():
    import pfast_amd
    c = pfast_amd.Context(0)
    yield c
    c.close()


def oracle_fwdadj(c, nsrc):
    N = c["vp"].size
    g = np.zeros(N)
    E = 0.0
    Ts = []
    for s in range(nsrc):
        T, _ = O.fsm3d(c["vp"], c["src"], s, c["dx"], c["dy"], c["dz"])
        lam, EE, _ = O.adjoint_fsm3d(T, c["T_obs"], c["rec"], c["nxyz"], c["src"], c["srnum"], s,
                                     c["dx"], c["dy"], c["dz"])
        O.oracle().orc_grad_accum(g, np.ascontiguousarray(lam.ravel()),
                                  np.ascontiguousarray(c["vp"].ravel()), N)
        E += EE
        Ts.append(T)
    return Ts, g.reshape(c["vp"].shape), E


@pytest.mark.parametrize("seed,shape,spacing,nsrc", [
    (21, (9, 12, 10), (10.0, 10.0, 10.0), 5),
    (22, (11, 9, 13), (12.5, 10.0, 8.0), 6),
    (23, (17, 19, 23), (25.0, 25.0, 25.0), 9),
    (24, (6, 33, 20), (10.0, 10.0, 10.0), 4),
    (25, (20, 27, 35), (25.0, 25.0, 25.0), 7),    # several tiles, ragged edges
    (26, (13, 30, 17), (12.5, 10.0, 8.0), 5),     # anisotropic spacing, ragged edges
])
def test_batch3d_vs_oracle(ctx, seed, shape, spacing, nsrc):
    import pfast_amd
    c = cases.case3d(seed, shape=shape, spacing=spacing, nsrc=nsrc, nrec_per=8)
    m = pfast_amd.Model3D(ctx, c["vp"], c["dx"], c["dy"], c["dz"])
    sh = pfast_amd.Shots3D(m, c["src"], c["srnum"], c["rec"], c["nxyz"], c["T_obs"])
    sh.run(adjoint=True, MAXSwp=100)
    g32, E32 = sh.gradient()
    Ts, g64, E64 = oracle_fwdadj(c, nsrc)
    worst = 0.0
    for s in range(nsrc):
        T32 = sh.field(s)
        T64 = Ts[s]
        reach = T64 < 1e6
        assert np.array_equal(reach, T32 < 1e6)
        err = np.abs(T32 - T64)[reach].max() / T64[reach].max()
        worst = max(worst, err)
    gerr = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    eerr = abs(E32 - E64) / E64
    print(f"seed {seed} shape {shape} T err {worst:.3e} grad err {gerr:.3e} misfit err {eerr:.3e}"
          f" cycles {sh.cycles().tolist()}", flush=True)
    assert worst <= T_TOL
    assert gerr <= G_TOL
    assert eerr <= E_TOL
    picks = sh.picks()
    ref = np.concatenate([Ts[s].ravel()[(c["rec"][c["srnum"][:s].sum():c["srnum"][:s + 1].sum(), 2]
                                          * shape[1] + c["rec"][c["srnum"][:s].sum():c["srnum"][:s + 1].sum(), 0])
                                         * shape[2] + c["rec"][c["srnum"][:s].sum():c["srnum"][:s + 1].sum(), 1]]
                          for s in range(nsrc)])
    assert np.abs(picks - ref).max() <= T_TOL * max(T.max() for T in Ts)
    sh.close()
    m.close()


def test_batch3d_plain_source_matches_oracle(ctx):
    """The resident single-source FP64 form equals the oracle bit for bit."""
    import pfast_amd
    c = cases.case3d(31, shape=(8, 9, 10), spacing=(10.0, 10.0, 10.0), nsrc=3)
    m = pfast_amd.Model3D(ctx, c["vp"], c["dx"], c["dy"], c["dz"])
    sh = pfast_amd.Shots3D(m, c["src"], c["srnum"], c["rec"], c["nxyz"], c["T_obs"])
    for s in range(3):
        EE, cyc = sh.plain_source(s)
        T, cf = O.fsm3d(c["vp"], c["src"], s, c["dx"], c["dy"], c["dz"])
        _, Eo, ca = O.adjoint_fsm3d(T, c["T_obs"], c["rec"], c["nxyz"], c["src"], c["srnum"], s,
                                    c["dx"], c["dy"], c["dz"])
        assert EE == Eo and cyc[0] == cf and cyc[1] == ca


@pytest.mark.parametrize("seed,shape,spacing", [
    (41, (20, 27, 35), (25.0, 25.0, 25.0)),
    (42, (13, 30, 17), (12.5, 10.0, 8.0)),
    (43, (7, 9, 41), (10.0, 10.0, 10.0)),
])
def test_sweep_variants_bitwise_identical(seed, shape, spacing):
    """Every sweep kernel variant (schedule, tile shape, batch depth) performs the
    same per-node arithmetic in the same order, so traveltimes, adjoint fields,
    cycle counts, gradient and misfit agree bit for bit."""
    import pfast_amd
    c = cases.case3d(seed, shape=shape, spacing=spacing, nsrc=7, nrec_per=10)
    names = pfast_amd.sweep_variants()
    assert len(names) >= 2
    ref = None
    for v, name in enumerate(names):
        ctx = pfast_amd.Context(0)
        ctx.set_sweep(v)
        m = pfast_amd.Model3D(ctx, c["vp"], c["dx"], c["dy"], c["dz"])
        sh = pfast_amd.Shots3D(m, c["src"], c["srnum"], c["rec"], c["nxyz"], c["T_obs"])
        sh.run(adjoint=True, MAXSwp=100)
        g, E = sh.gradient()
        out = (np.stack([sh.field(s) for s in range(7)]),
               np.stack([sh.field(s, adjoint=True) for s in range(7)]), sh.cycles(), g, E)
        sh.close()
        m.close()
        ctx.close()
        print(f"variant {v} ({name}): cycles {out[2].tolist()}", flush=True)
        if ref is None:
            ref = out
            continue
        assert np.array_equal(out[0], ref[0]), name
        assert np.array_equal(out[1], ref[1]), name
        assert np.array_equal(out[2], ref[2]), name
        assert np.array_equal(out[3], ref[3]) and out[4] == ref[4], name


def test_stream_split_bitwise_identical():
    """Splitting a cycle's chunks over 1-4 streams changes only which launch carries
    a chunk, so every result is bit-identical; 0 and 5 streams are refused."""
    import pfast_amd
    c = cases.case3d(44, shape=(14, 23, 19), spacing=(25.0, 25.0, 25.0), nsrc=23, nrec_per=8)
    ref = None
    for k in (1, 2, 3, 4):
        ctx = pfast_amd.Context(0)
        ctx.set_streams(k)
        m = pfast_amd.Model3D(ctx, c["vp"], c["dx"], c["dy"], c["dz"])
        sh = pfast_amd.Shots3D(m, c["src"], c["srnum"], c["rec"], c["nxyz"], c["T_obs"])
        sh.run(adjoint=True, MAXSwp=100)
        g, E = sh.gradient()
        out = (np.stack([sh.field(s) for s in range(23)]),
               np.stack([sh.field(s, adjoint=True) for s in range(23)]), sh.cycles(), g, E)
        sh.close()
        m.close()
        if k == 4:
            for bad in (0, 5):
                with pytest.raises(pfast_amd.PfastError):
                    ctx.set_streams(bad)
        ctx.close()
        if ref is None:
            ref = out
            continue
        assert all(np.array_equal(cases.bits(a), cases.bits(b)) for a, b in zip(out[:4], ref[:4])), k
        assert out[4] == ref[4], k


@pytest.mark.parametrize("seed,shape,spacing", [
    (45, (21, 37, 29), (25.0, 25.0, 25.0)),
    (46, (16, 25, 41), (12.5, 10.0, 8.0)),
])
def test_tile_skip_bitwise_identical(seed, shape, spacing):
    """Exact tile skipping only leaves out sweeps, and depth ranges only the levels of a
    sweep, that would change nothing: fields, cycle counts, gradient and misfit equal
    those with every tile swept, in the per-diagonal launches (where depth ranges
    apply) and in the dataflow launches."""
    import pfast_amd
    c = cases.case3d(seed, shape=shape, spacing=spacing, nsrc=9, nrec_per=12)
    outs = {}
    for flow, mode in ((0, 0), (0, 1), (0, 2), (1, 2)):
        ctx = pfast_amd.Context(0)
        ctx.set_dataflow(flow)
        ctx.set_tile_skip(mode)
        m = pfast_amd.Model3D(ctx, c["vp"], c["dx"], c["dy"], c["dz"])
        sh = pfast_amd.Shots3D(m, c["src"], c["srnum"], c["rec"], c["nxyz"], c["T_obs"])
        sh.run(adjoint=True, MAXSwp=100)
        g, E = sh.gradient()
        outs[flow, mode] = (np.stack([sh.field(s) for s in range(9)]),
                            np.stack([sh.field(s, adjoint=True) for s in range(9)]), sh.cycles(),
                            g, E)
        sh.close()
        m.close()
        ctx.close()
    ref = outs[0, 0]
    print(f"cycles {ref[2].tolist()}", flush=True)
    for key, out in outs.items():
        assert all(np.array_equal(cases.bits(x), cases.bits(y))
                   for x, y in zip(out[:4], ref[:4])), key
        assert out[4] == ref[4], key


def test_start_levels_across_schedule_switch():
    """A phase that starts in per-diagonal launches and switches to dataflow launches as
    chunks converge (auto schedule: 12 chunks x 41 x 41 tiles of 4 x 4 columns are above
    the switch, 7 below it), and back in the next phase: the depth ranges kept by
    the per-diagonal launches are reset after each dataflow cycle, so every mode gives
    the same bits."""
    import pfast_amd
    names = pfast_amd.sweep_variants()
    v44 = [i for i, n in enumerate(names) if " 4x4 " in n][0]
    c = cases.case3d(61, shape=(12, 161, 161), spacing=(25.0, 25.0, 25.0), nsrc=96,
                     nrec_per=10)
    outs = {}
    for mode in (0, 1, 2):
        ctx = pfast_amd.Context(0)
        ctx.set_sweep(v44)
        ctx.set_tile_skip(mode)
        m = pfast_amd.Model3D(ctx, c["vp"], c["dx"], c["dy"], c["dz"])
        sh = pfast_amd.Shots3D(m, c["src"], c["srnum"], c["rec"], c["nxyz"], c["T_obs"])
        sh.run(adjoint=True, MAXSwp=100)
        g, E = sh.gradient()
        outs[mode] = (np.stack([sh.field(s) for s in range(0, 96, 7)]),
                      np.stack([sh.field(s, adjoint=True) for s in range(0, 96, 7)]),
                      sh.cycles(), g, E)
        sh.close()
        m.close()
        ctx.close()
    print(f"cycles {np.unique(outs[0][2], axis=0).tolist()}", flush=True)
    for mode in (1, 2):
        assert all(np.array_equal(cases.bits(x), cases.bits(y))
                   for x, y in zip(outs[mode][:4], outs[0][:4])), mode
        assert outs[mode][4] == outs[0][4], mode


@pytest.mark.parametrize("variant", [0, 5, 6, 8])
def test_paired_launches_bitwise_identical(variant):
    """Per-diagonal launches that pair orderings 2p and 2p+1 (the second trailing by
    ntj + 1 tile diagonals) sweep every tile with the values the one-ordering launches
    give it: fields, cycle counts, gradient and misfit are the same bits, without tile
    skipping, with whole tiles skipped and with depth ranges, for tile shapes with
    different ntj."""
    import pfast_amd
    c = cases.case3d(47 + variant, shape=(19, 35, 41), spacing=(25.0, 25.0, 25.0), nsrc=11,
                     nrec_per=12)
    outs = {}
    for skip in (0, 1, 2):
        for paired in (False, True):
            ctx = pfast_amd.Context(0)
            ctx.set_sweep(variant)
            ctx.set_dataflow(0)
            ctx.set_tile_skip(skip)
            ctx.set_paired(paired)
            m = pfast_amd.Model3D(ctx, c["vp"], c["dx"], c["dy"], c["dz"])
            sh = pfast_amd.Shots3D(m, c["src"], c["srnum"], c["rec"], c["nxyz"], c["T_obs"])
            sh.run(adjoint=True, MAXSwp=100)
            g, E = sh.gradient()
            outs[skip, paired] = (np.stack([sh.field(s) for s in range(11)]),
                                  np.stack([sh.field(s, adjoint=True) for s in range(11)]),
                                  sh.cycles(), g, E)
            sh.close()
            m.close()
            ctx.close()
    ref = outs[False, False]
    print(f"variant {variant} cycles {ref[2].tolist()}", flush=True)
    for key, out in outs.items():
        assert all(np.array_equal(cases.bits(x), cases.bits(y))
                   for x, y in zip(out[:4], ref[:4])), key
        assert out[4] == ref[4], key


def test_sweep_variant_rejects_unknown():
    import pfast_amd
    ctx = pfast_amd.Context(0)
    with pytest.raises(pfast_amd.PfastError):
        ctx.set_sweep(len(pfast_amd.sweep_variants()))
    ctx.close()


def test_parent_close_closes_children():
    """A context or model is never freed under a live child: the C ABI refuses,
    and the Python wrappers close children first (no dangling stream pointer when
    a test fails before its own cleanup)."""
    import pfast_amd
    from pfast_amd._lib import PFAST_ERR_ARG, lib
    c = cases.case3d(51, shape=(6, 7, 8), spacing=(10.0, 10.0, 10.0), nsrc=2)
    ctx = pfast_amd.Context(0)
    m = pfast_amd.Model3D(ctx, c["vp"], c["dx"], c["dy"], c["dz"])
    sh = pfast_amd.Shots3D(m, c["src"], c["srnum"], c["rec"], c["nxyz"], c["T_obs"])
    assert lib().pfast_ctx_destroy(ctx.h) == PFAST_ERR_ARG
    assert lib().pfast_model3d_destroy(m.h) == PFAST_ERR_ARG
    sh.run(adjoint=False)
    ctx.close()
    assert sh.h is None and m.h is None and ctx.h is None
    sh.close()


def test_chunk_field_bound():
    """The sweep kernels address a chunk's field with 32-bit byte offsets, levels up to
    64 past either end of a column included: a grid whose chunk field passes that
    bound (3400 x 200 x 200 nodes of 8 FP32 sources: 4.43 GB with the margin) is
    refused when the shot set is created, before any field is allocated."""
    import pfast_amd
    nz, nr, nc = 3400, 200, 200
    ctx = pfast_amd.Context(0)
    m = pfast_amd.Model3D(ctx, np.full((nz, nr, nc), 2000.0), 25.0, 25.0, 25.0)
    with pytest.raises(pfast_amd.PfastError, match="too large"):
        pfast_amd.Shots3D(m, np.array([[nr // 2, nc // 2, 0]]), np.array([0]))
    m.close()
    # the adjoint field holds (lambda, T) pairs, 8 bytes per node and source: 1700 x 200 x
    # 200 nodes pass for traveltimes alone (2.26 GB with the margin) and are refused for a
    # shot set with picks (4.52 GB)
    nz = 1700
    m = pfast_amd.Model3D(ctx, np.full((nz, nr, nc), 2000.0), 25.0, 25.0, 25.0)
    src = np.array([[nr // 2, nc // 2, 0]])
    pfast_amd.Shots3D(m, src, np.array([0])).close()
    with pytest.raises(pfast_amd.PfastError, match="too large for the batched adjoint"):
        pfast_amd.Shots3D(m, src, np.array([1]), np.array([[nr // 2, nc // 2 + 4, 0]]),
                          np.array([[0.0, 0.0, 1.0]]), np.array([0.1]))
    m.close()
    ctx.close()


def test_autotune_picks_fastest_candidate_results_unchanged():
    """pfast_shots3d_autotune times the listed variants, leaves the fastest set on the
    context (the variant in use, 0, unless another is at least 2% faster), and the shot
    set run afterwards gives the same bits as under the default."""
    import pfast_amd
    c = cases.case3d(47, shape=(17, 29, 33), spacing=(25.0, 25.0, 25.0), nsrc=11, nrec_per=9)
    ctx = pfast_amd.Context(0)
    m = pfast_amd.Model3D(ctx, c["vp"], c["dx"], c["dy"], c["dz"])
    sh = pfast_amd.Shots3D(m, c["src"], c["srnum"], c["rec"], c["nxyz"], c["T_obs"])
    sh.run(adjoint=True, MAXSwp=100)
    ref = (np.stack([sh.field(s) for s in range(11)]),
           np.stack([sh.field(s, adjoint=True) for s in range(11)])) + sh.gradient()
    chosen, ms = sh.autotune([0, 1, 4])
    print(f"autotune: {ms} -> {chosen}", flush=True)
    assert sorted(ms) == [0, 1, 4] and all(t > 0 for t in ms.values())
    fastest = min(ms, key=ms.get)
    assert chosen == (0 if ms[0] <= 1.02 * ms[fastest] else fastest)
    sh.run(adjoint=True, MAXSwp=100)
    out = (np.stack([sh.field(s) for s in range(11)]),
           np.stack([sh.field(s, adjoint=True) for s in range(11)])) + sh.gradient()
    assert all(np.array_equal(a, b) for a, b in zip(out[:3], ref[:3])) and out[3] == ref[3]
    with pytest.raises(pfast_amd.PfastError):
        sh.autotune([len(pfast_amd.sweep_variants())])
    sh.close()
    m.close()
    ctx.close()
