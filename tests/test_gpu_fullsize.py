"""Parity at the full workload size (3D 201 x 201 x 101, h = 25 m, the bench's model and
survey generator, 16 sources x 2000 receivers) through properties that hold at any size:

- the batched FP32 fields against the single-source FP64 form on the same resident inputs
  (that form is bit-identical to the oracle, tests/test_gpu_plain.py): traveltime within
  1e-5 (max-norm, relative), lambda within 1e-3 (L2, relative), cycle counts within 2;
- a zero residual (picks equal to the model's own traveltimes) gives an exactly zero
  adjoint field, gradient and misfit (src/T_padjoint3d.h:95-124: the boundary values are
  the residuals);
- the adjoint is linear in the residuals: doubling them doubles the gradient and
  quadruples the misfit;
- a source's results do not depend on which sources share its chunk: a shot set of
  sources 4..11 (one chunk) reproduces their fields bit for bit from the 16-source set
  (chunks 0..7 and 8..15);
- the quantities the precision contract bounds, at this size: the summed gradient
  g = -sum_s lambda_s / v^3 of all 16 sources within 1e-3 (L2, relative) and the summed
  misfit within 1e-4 of the same sums over the FP64 form (src/T_PFAST3D.cpp:393-399);
- the FP64 form itself is bit-identical to the FP64 oracle at this size (traveltimes,
  lambda, misfit and cycle counts of one source).
"""
import numpy as np
import pytest

from pfast_amd import synth

pytestmark = pytest.mark.gpu

NSRC = 16


@pytest.fixture(scope="module")
def full():
    import pfast_amd
    cfg = synth.config3d(nsrc=NSRC, nrec=2000)
    ctx = pfast_amd.Context(0)
    m = pfast_amd.Model3D(ctx, cfg["vp"], cfg["h"], cfg["h"], cfg["h"])
    sh = pfast_amd.Shots3D(m, cfg["src"], cfg["srnum"], cfg["rec"], cfg["nxyz"], cfg["T_obs"])
    sh.run(adjoint=True, MAXSwp=100)
    yield pfast_amd, cfg, m, sh
    sh.close()
    m.close()
    ctx.close()


def _shots(pf, cfg, m, sel, T_obs):
    nrec = int(cfg["srnum"][0])
    rows = np.concatenate([np.arange(s * nrec, (s + 1) * nrec) for s in sel])
    sh = pf.Shots3D(m, cfg["src"][sel], cfg["srnum"][sel], cfg["rec"][rows], cfg["nxyz"][rows],
                    T_obs[rows])
    sh.run(adjoint=True, MAXSwp=100)
    return sh


def test_fullsize_fp32_against_fp64(full):
    pf, cfg, m, sh = full
    cyc = sh.cycles()
    for s in (0, 9):
        EE, pc, T64, L64 = sh.plain_source(s, MAXSwp=100, fields=True)
        T32 = sh.field(s)
        reach = T64 < 1e6
        terr = np.abs(T32 - T64)[reach].max() / T64[reach].max()
        L32 = sh.field(s, adjoint=True)
        lerr = np.linalg.norm(L32 - L64) / np.linalg.norm(L64)
        print(f"source {s}: T {terr:.2e}, lambda {lerr:.2e}, cycles FP32 {cyc[s].tolist()} "
              f"FP64 {pc.tolist()}", flush=True)
        assert terr <= 1e-5 and lerr <= 1e-3
        assert np.all(np.abs(cyc[s] - pc) <= 2)


def test_fullsize_zero_residual_gives_zero_gradient(full):
    pf, cfg, m, sh = full
    picks = sh.picks()
    z = _shots(pf, cfg, m, np.arange(NSRC), picks)
    g, E = z.gradient()
    lam = z.field(3, adjoint=True)
    z.close()
    assert E == 0.0 and not np.any(g) and not np.any(lam)


def test_fullsize_adjoint_linear_in_residuals(full):
    pf, cfg, m, sh = full
    picks = sh.picks()
    g1, E1 = sh.gradient()
    d = _shots(pf, cfg, m, np.arange(NSRC), picks + 2.0 * (cfg["T_obs"] - picks))
    g2, E2 = d.gradient()
    d.close()
    rel = np.linalg.norm(g2 - 2.0 * g1) / np.linalg.norm(g1)
    print(f"|g2 - 2 g1| / |g1| = {rel:.2e}, E2 / 4 E1 = {E2 / (4 * E1):.8f}", flush=True)
    assert rel <= 1e-5 and abs(E2 / (4.0 * E1) - 1.0) <= 1e-9


def test_fullsize_result_independent_of_chunk_mates(full):
    pf, cfg, m, sh = full
    sel = np.arange(4, 12)
    part = _shots(pf, cfg, m, sel, cfg["T_obs"])
    for k, s in enumerate(sel):
        assert np.array_equal(part.field(k), sh.field(s)), s
        assert np.array_equal(part.field(k, adjoint=True), sh.field(s, adjoint=True)), s
    part.close()


def test_fullsize_gradient_and_misfit_against_fp64(full):
    pf, cfg, m, sh = full
    g32, E32 = sh.gradient()
    vp = cfg["vp"]
    sumL = np.zeros(vp.shape)
    E64 = 0.0
    for s in range(NSRC):
        EE, _, _, L64 = sh.plain_source(s, MAXSwp=100, fields=True)
        sumL += L64
        E64 += EE
    g64 = -sumL / vp ** 3
    gerr = float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64))
    eerr = abs(E32 - E64) / E64
    print(f"16 sources: gradient {gerr:.2e} (L2), misfit {eerr:.2e}", flush=True)
    assert gerr <= 1e-3 and eerr <= 1e-4


def test_fullsize_plain_fp64_bitexact_vs_oracle(full):
    import oracle_lib as O
    pf, cfg, m, sh = full
    h = cfg["h"]
    s = 5
    EE, pc, T64, L64 = sh.plain_source(s, MAXSwp=100, fields=True)
    To, cf = O.fsm3d(cfg["vp"], cfg["src"], s, h, h, h)
    Lo, EEo, ca = O.adjoint_fsm3d(To, cfg["T_obs"], cfg["rec"], cfg["nxyz"], cfg["src"],
                                  cfg["srnum"], s, h, h, h)
    print(f"source {s}: cycles GPU {pc.tolist()} oracle {[cf, ca]}, EE {EE!r} {EEo!r}", flush=True)
    assert np.array_equal(T64.view(np.int64), To.view(np.int64))
    assert np.array_equal(L64.view(np.int64), Lo.view(np.int64))
    assert EE == EEo and list(pc) == [cf, ca]
