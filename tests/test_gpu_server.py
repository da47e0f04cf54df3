"""The request server (pfast_server_*): each result is the same alone and in the mix.

Bound: bit-identical.  A request's source chunks run exactly the sweep cycles that a
standalone shot-set run gives them, whatever else shares the launches, so the
traveltime fields, gradients, misfits and cycle counts of a request served in a mix
equal those of the same request served alone, and those of Shots3D.run /
Shots2D.run, bit for bit.  Admission by footprint, refusal of requests larger than
the budget, release, and failures reported per request are checked on the way.
"""
import time

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

SHAPE_A, SP_A = (13, 18, 21), (25.0, 25.0, 25.0)
SHAPE_B, SP_B = (9, 24, 17), (10.0, 12.5, 5.0)


def _models():
    a1 = cases.case3d(101, shape=SHAPE_A, spacing=SP_A)["vp"]
    a2 = cases.case3d(102, shape=SHAPE_A, spacing=SP_A)["vp"]
    b = cases.case3d(103, shape=SHAPE_B, spacing=SP_B)["vp"]
    c = cases.case2d(104, shape=(20, 33), NumofR=1)
    d = cases.case2d(105, shape=(26, 37), NumofR=1)
    e = cases.case2d(106, shape=(18, 22), NumofR=1)
    Ze = e["Z"].copy()
    Ze[0, :, 0] = 0.5 * e["dz"]  # above every row: the reference's fatal exit
    return [("3d", a1, SP_A), ("3d", a2, SP_A), ("3d", b, SP_B),
            ("2d", c["vp"], (c["dx"], c["dz"]), None), ("2d", d["vp"], (d["dx"], d["dz"]), d["Z"]),
            ("2d", e["vp"], (e["dx"], e["dz"]), Ze)]


# (model index, kind, sources, seed)
REQS = [(0, "table", 1, 11), (0, "gradient", 5, 12), (0, "gradient", 9, 13), (0, "table", 4, 14),
        (1, "gradient", 3, 15), (1, "table", 7, 16), (2, "gradient", 6, 17), (2, "table", 2, 18),
        (2, "gradient", 1, 19), (3, "table", 3, 20), (3, "gradient", 4, 21), (4, "gradient", 2, 22),
        (4, "table", 5, 23), (5, "gradient", 2, 24)]
FAILING = {13}


def _request_data(mi, nsrc, seed, models):
    m = models[mi]
    if m[0] == "3d":
        return cases.case3d(seed, shape=m[1].shape, spacing=m[2], nsrc=nsrc)
    return cases.case2d(seed, shape=m[1].shape, NumofR=1, nsrc=nsrc)


def _add_models(srv, models):
    ids = []
    for m in models:
        if m[0] == "3d":
            ids.append(srv.add_model3d(m[1], *m[2]))
        else:
            ids.append(srv.add_model2d(m[1], *m[2], Z=m[3]))
    return ids


def _submit(srv, ids, models, k):
    mi, kind, nsrc, seed = REQS[k]
    c = _request_data(mi, nsrc, seed, models)
    if models[mi][0] == "3d":
        if kind == "table":
            return srv.submit3d(ids[mi], kind, c["src"], c["srnum"])
        return srv.submit3d(ids[mi], kind, c["src"], c["srnum"], c["rec"], c["nxyz"], c["T_obs"])
    if kind == "table":
        return srv.submit2d(ids[mi], kind, c["src"], c["srnum"])
    return srv.submit2d(ids[mi], kind, c["src"], c["srnum"], c["rec"], c["nxy"], c["T_obs"],
                        c["TR_obs"])


def _drain(srv, tickets, timeout=120):
    t0 = time.time()
    while any(srv.status(t)[0] in ("queued", "running") for t in tickets):
        srv.progress(2000)
        assert time.time() - t0 < timeout, "server did not finish"


def _results(srv, t, k):
    st, _ = srv.status(t)
    if k in FAILING:
        return st, srv.error(t)
    assert st == "done", srv.error(t)
    nsrc = REQS[k][2]
    out = {"T": np.stack([srv.field(t, s) for s in range(nsrc)]), "cycles": srv.cycles(t)}
    if REQS[k][1] == "gradient":
        out["grad"] = srv.gradient(t)
    return st, out


def _same(a, b):
    if isinstance(a, dict):
        return a.keys() == b.keys() and all(_same(a[k], b[k]) for k in a)
    if isinstance(a, tuple):
        return len(a) == len(b) and all(_same(x, y) for x, y in zip(a, b))
    if isinstance(a, np.ndarray):
        return a.shape == b.shape and np.array_equal(cases.bits(a), cases.bits(b))
    return a == b


def test_same_result_alone_and_in_mix():
    import pfast_amd
    models = _models()
    # alone: one request at a time
    srv = pfast_amd.Server(0, 8 << 30)
    ids = _add_models(srv, models)
    alone = {}
    for k in range(len(REQS)):
        t = _submit(srv, ids, models, k)
        _drain(srv, [t])
        alone[k] = _results(srv, t, k)
        srv.release(t)
    srv.close()
    # mix: three waves, the later ones submitted while earlier requests sweep
    srv = pfast_amd.Server(0, 8 << 30)
    ids = _add_models(srv, models)
    order = [3, 0, 9, 6, 13, 1, 11, 4, 2, 7, 10, 12, 5, 8]
    tickets = {}
    for w, wave in enumerate([order[:5], order[5:10], order[10:]]):
        for k in wave:
            tickets[k] = _submit(srv, ids, models, k)
        for _ in range(3):
            srv.progress(0)
        print(f"wave {w}: {srv.info()}", flush=True)
    _drain(srv, list(tickets.values()))
    info = srv.info()
    print(f"mix done: {info}", flush=True)
    for k, t in tickets.items():
        mixed = _results(srv, t, k)
        assert _same(alone[k], mixed), f"request {k} {REQS[k]} differs in the mix"
        if k in FAILING:
            assert mixed[0] == "failed" and "no k2" in mixed[1]
            with pytest.raises(pfast_amd.FatalError):
                srv.gradient(t)
    for t in tickets.values():
        srv.release(t)
    srv.progress(0)
    assert srv.info()["running"] == 0
    srv.close()
    # the plain shot-set API gives the same bits
    ctx = pfast_amd.Context(0)
    for k, (mi, kind, nsrc, seed) in enumerate(REQS):
        if k in FAILING:
            continue
        c = _request_data(mi, nsrc, seed, models)
        m = models[mi]
        grad = kind == "gradient"
        if m[0] == "3d":
            mod = pfast_amd.Model3D(ctx, m[1], *m[2])
            sh = (pfast_amd.Shots3D(mod, c["src"], c["srnum"], c["rec"], c["nxyz"], c["T_obs"])
                  if grad else pfast_amd.Shots3D(mod, c["src"], c["srnum"]))
            sh.run(adjoint=grad)
            T = np.stack([sh.field(s) for s in range(nsrc)])
            assert _same(T, alone[k][1]["T"]), f"request {k}: traveltimes"
            if grad:
                assert _same(sh.gradient(), alone[k][1]["grad"]), f"request {k}: gradient"
        else:
            mod = pfast_amd.Model2D(ctx, m[1], *m[2], m[3])
            sh = (pfast_amd.Shots2D(mod, c["src"], c["srnum"], c["rec"], c["nxy"], c["T_obs"],
                                    c["TR_obs"]) if grad else pfast_amd.Shots2D(mod, c["src"], c["srnum"]))
            sh.run(adjoint=grad)
            T = np.stack([sh.field(s, "T") for s in range(nsrc)])
            assert _same(T, alone[k][1]["T"]), f"request {k}: traveltimes"
            if grad:
                gT, gR, _, (Et, Er, _) = sh.gradient()
                assert _same((gT, gR, Et, Er), alone[k][1]["grad"]), f"request {k}: gradient"
        sh.close()
        mod.close()
    ctx.close()


def test_chunk_field_bound():
    """The sweep kernels address a chunk's field with 32-bit byte offsets, levels up to 64
    past either end of a column included, as for a shot set: on a 1700 x 200 x 200 model a
    gradient request (the (lambda, T) field: 4.52 GB per chunk with the margin) is refused
    at submission, before anything is queued; the model stays usable."""
    import pfast_amd
    nz, nr, nc = 1700, 200, 200
    srv = pfast_amd.Server(0, 16 << 30)
    mid = srv.add_model3d(np.full((nz, nr, nc), 2000.0), 25.0, 25.0, 25.0)
    src = np.array([[nr // 2, nc // 2, 0]])
    with pytest.raises(pfast_amd.PfastError, match="too large"):
        srv.submit3d(mid, "gradient", src, np.array([1]), np.array([[nr // 2, nc // 2 + 4, 0]]),
                     np.array([[0.0, 0.0, 1.0]]), np.array([0.1]))
    assert srv.info()["queued"] == 0 and srv.info()["running"] == 0
    srv.remove_model(mid)
    srv.close()


def test_admission_by_footprint():
    import pfast_amd
    models = _models()[:1]
    probe = pfast_amd.Server(0, 8 << 30)
    pid = _add_models(probe, models)
    resident = probe.info()["in_use"]
    c = _request_data(0, 5, 31, models)
    t = probe.submit3d(pid[0], "gradient", c["src"], c["srnum"], c["rec"], c["nxyz"], c["T_obs"])
    fp = probe.footprint(t)
    probe.close()
    srv = pfast_amd.Server(0, resident + fp + fp // 2)
    ids = _add_models(srv, models)
    assert srv.info()["in_use"] == resident
    tickets = []
    for seed in (31, 32, 33):
        c = _request_data(0, 5, seed, models)
        tickets.append(srv.submit3d(ids[0], "gradient", c["src"], c["srnum"], c["rec"], c["nxyz"],
                                    c["T_obs"]))
    big = _request_data(0, 40, 34, models)
    with pytest.raises(pfast_amd.PfastError):
        srv.submit3d(ids[0], "gradient", big["src"], big["srnum"], big["rec"], big["nxyz"],
                     big["T_obs"])
    most = 0
    t0 = time.time()
    done = set()
    while len(done) < len(tickets):
        srv.progress(500)
        info = srv.info()
        most = max(most, info["running"])
        assert info["in_use"] <= resident + fp + fp // 2
        for t in tickets:
            if t not in done and srv.status(t)[0] == "done":
                done.add(t)
                srv.release(t)  # frees its memory for the next one
        assert time.time() - t0 < 60
    assert most == 1
    with pytest.raises(pfast_amd.PfastError):  # not while requests reference it
        t = srv.submit3d(ids[0], "table", c["src"], c["srnum"])
        srv.remove_model(ids[0])
    srv.release(t)
    srv.progress(0)
    assert srv.info()["in_use"] == resident
    srv.remove_model(ids[0])
    assert srv.info()["in_use"] == 0
    srv.close()


def test_poisson_stream_completes():
    """A short Poisson stream over the small serving grids: every request finishes,
    latencies are ordered and positive, and the server ends empty."""
    import pfast_amd
    from pfast_amd import loadgen as LG
    grids = LG.SERVING_GRIDS[:3]
    srv = pfast_amd.Server(0, 16 << 30)
    models = LG.load_models(srv, grids=grids, per_grid=2)
    resident = srv.info()["in_use"]
    recs = LG.poisson_run(srv, models, 40.0, 2.0, seed=3)
    res = LG.summarize(recs, grids)
    print(res, flush=True)
    assert res["requests"] == len(recs) > 20
    assert 0 < res["p50_ms"] <= res["p99_ms"]
    assert all(r["queue"] >= 0 and r["latency"] >= r["queue"] for r in recs)
    srv.progress(0)
    info = srv.info()
    assert info["running"] == 0 and info["queued"] == 0 and info["in_use"] == resident
    srv.close()


def test_large_serving_grids_alone_and_in_mix():
    """Requests on the two largest serving grids (3D 129 x 129 x 65 and 321 x 321 x 161,
    h = 25 m, pfast_amd.loadgen models): each result is bit-identical alone, in a mix
    submitted together, and from Shots3D.run."""
    import pfast_amd
    from pfast_amd import loadgen as LG
    grids = [g for g in LG.SERVING_GRIDS if g[0] == "3d" and g[1][0] >= 65 and g[1][1] >= 129]
    assert [g[1] for g in grids] == [(65, 129, 129), (161, 321, 321)]
    rng = np.random.default_rng(77)
    # (grid, kind, sources)
    plan = [(0, "gradient", 8), (0, "table", 3), (1, "gradient", 1), (1, "table", 2),
            (1, "gradient", 5)]
    reqs = [LG.make_request("3d", grids[gi][1], grids[gi][2], kind, n, rng) for gi, kind, n in plan]
    vps = [LG.velocity_model("3d", g[1], g[2], 40 + i) for i, g in enumerate(grids)]

    def serve(order):
        srv = pfast_amd.Server(0, 64 << 30)
        ids = [srv.add_model3d(v, 25.0, 25.0, 25.0) for v in vps]
        out, tickets = {}, {}
        for k in order:
            gi, kind, n = plan[k]
            r = reqs[k]
            args = (r["rec"], r["nxyz"], r["T_obs"]) if kind == "gradient" else ()
            tickets[k] = srv.submit3d(ids[gi], kind, r["src"], r["srnum"], *args)
            if len(order) == 1:
                _drain(srv, [tickets[k]])
        _drain(srv, list(tickets.values()), timeout=300)
        for k, t in tickets.items():
            st, _ = srv.status(t)
            assert st == "done", srv.error(t)
            n = plan[k][2]
            res = {"T": np.stack([srv.field(t, s) for s in range(n)]), "cycles": srv.cycles(t)}
            if plan[k][1] == "gradient":
                res["grad"] = srv.gradient(t)
            out[k] = res
            srv.release(t)
        srv.close()
        return out

    alone = {}
    for k in range(len(plan)):
        print(f"alone: request {k} {plan[k]}", flush=True)
        alone.update(serve([k]))
    print("mix: all requests together", flush=True)
    mixed = serve([4, 0, 2, 1, 3])
    for k in range(len(plan)):
        assert _same(alone[k], mixed[k]), f"request {k} {plan[k]} differs in the mix"
    ctx = pfast_amd.Context(0)
    for k in (0, 4):
        gi, kind, n = plan[k]
        r = reqs[k]
        mod = pfast_amd.Model3D(ctx, vps[gi], 25.0, 25.0, 25.0)
        sh = pfast_amd.Shots3D(mod, r["src"], r["srnum"], r["rec"], r["nxyz"], r["T_obs"])
        sh.run(adjoint=True)
        assert _same(np.stack([sh.field(s) for s in range(n)]), alone[k]["T"]), k
        assert _same(sh.gradient(), alone[k]["grad"]), k
        sh.close()
        mod.close()
    ctx.close()


def test_poisson_stream_completes_on_every_grid():
    """A short Poisson stream over all five serving grids, large ones included."""
    import pfast_amd
    from pfast_amd import loadgen as LG
    srv = pfast_amd.Server(0, 64 << 30)
    models = LG.load_models(srv, per_grid=1)
    resident = srv.info()["in_use"]
    recs = LG.poisson_run(srv, models, 6.0, 3.0, seed=5)
    res = LG.summarize(recs)
    print(res, flush=True)
    assert res["requests"] == len(recs) >= 8
    assert 0 < res["p50_ms"] <= res["p99_ms"]
    srv.progress(0)
    info = srv.info()
    assert info["running"] == 0 and info["queued"] == 0 and info["in_use"] == resident
    srv.close()
