"""Seeded serving workload for the request server (Server): resident velocity
models on a set of grids, and a Poisson stream of traveltime-table and gradient
requests driven against the server in real time.

Models: v = v0 + g z (2D 1500 + 0.8 z, 3D 2000 + 0.6 z) times three seeded
Gaussian anomalies of +-10% (sigma 4 cells).  Requests: a grid and one of its
models, 1-64 surface sources, and a kind (table or gradient), all uniform.
Gradient requests carry 32 surface receivers per source with picks from the
straight-ray time at the surface velocity plus 2 ms noise, normals (0, 0, -1)
in 3D and (-1, 0) in 2D.
"""
import time

import numpy as np

# (dimension, shape in array order, spacing)
SERVING_GRIDS = [
    ("2d", (51, 101), 25.0),        # 2D 101 x 51 (x by z)
    ("2d", (101, 401), 25.0),       # 2D 401 x 101
    ("3d", (65, 65, 65), 25.0),     # 3D 65^3
    ("3d", (65, 129, 129), 25.0),   # 3D 129 x 129 x 65
    ("3d", (161, 321, 321), 25.0),  # 3D 321 x 321 x 161
]


def velocity_model(dim, shape, h, seed):
    rng = np.random.default_rng(seed)
    if dim == "3d":
        nz, nr, nc = shape
        k, i, j = np.meshgrid(np.arange(nz), np.arange(nr), np.arange(nc), indexing="ij",
                              sparse=True)
        v = (2000.0 + 0.6 * k * h) * np.ones(shape)
        for _ in range(3):
            c = [rng.uniform(0, n) for n in shape]
            amp = rng.choice([-0.1, 0.1])
            v *= 1 + amp * np.exp(-((k - c[0]) ** 2 + (i - c[1]) ** 2 + (j - c[2]) ** 2) / (2 * 4.0 ** 2))
        return v
    nr, nc = shape
    i, j = np.meshgrid(np.arange(nr), np.arange(nc), indexing="ij", sparse=True)
    v = (1500.0 + 0.8 * i * h) * np.ones(shape)
    for _ in range(3):
        c = [rng.uniform(0, n) for n in shape]
        amp = rng.choice([-0.1, 0.1])
        v *= 1 + amp * np.exp(-((i - c[0]) ** 2 + (j - c[1]) ** 2) / (2 * 4.0 ** 2))
    return v


def make_request(dim, shape, h, kind, nsrc, rng, nrec_per=32):
    """Source/receiver tables and picks of one request (server argument order)."""
    if dim == "3d":
        nz, nr, nc = shape
        src = np.stack([rng.integers(0, nr, nsrc), rng.integers(0, nc, nsrc),
                        np.zeros(nsrc, np.int64)], 1).astype(np.int32)
        req = dict(src=src, srnum=np.full(nsrc, nrec_per if kind == "gradient" else 0, np.int32))
        if kind == "gradient":
            nrec = nsrc * nrec_per
            rec = np.stack([rng.integers(0, nr, nrec), rng.integers(0, nc, nrec),
                            np.zeros(nrec, np.int64)], 1).astype(np.int32)
            s = np.repeat(src, nrec_per, axis=0)
            dist = h * np.sqrt(((rec[:, :2] - s[:, :2]) ** 2).sum(1))
            req.update(rec=rec, nxyz=np.tile([0.0, 0.0, -1.0], (nrec, 1)),
                       T_obs=dist / 2000.0 + rng.normal(0, 0.002, nrec))
        return req
    nr, nc = shape
    src = np.stack([np.zeros(nsrc, np.int64), rng.integers(0, nc, nsrc)], 1).astype(np.int32)
    req = dict(src=src, srnum=np.full(nsrc, nrec_per if kind == "gradient" else 0, np.int32))
    if kind == "gradient":
        nrec = nsrc * nrec_per
        rec = np.stack([np.zeros(nrec, np.int64), rng.integers(0, nc, nrec)], 1).astype(np.int32)
        s = np.repeat(src, nrec_per, axis=0)
        dist = h * np.abs(rec[:, 1] - s[:, 1])
        req.update(rec=rec, nxy=np.tile([-1.0, 0.0], (nrec, 1)),
                   TT_obs=dist / 1500.0 + rng.normal(0, 0.002, nrec))
    return req


def load_models(server, grids=SERVING_GRIDS, per_grid=2, seed=2026):
    """Add per_grid seeded models of every grid; returns [(dim, shape, h, [model ids])]."""
    out = []
    for g, (dim, shape, h) in enumerate(grids):
        ids = []
        for m in range(per_grid):
            v = velocity_model(dim, shape, h, seed + 100 * g + m)
            ids.append(server.add_model3d(v, h, h, h) if dim == "3d" else server.add_model2d(v, h, h))
            del v
        out.append((dim, shape, h, ids))
    return out


def _submit(server, model, dim, kind, req):
    if dim == "3d":
        if kind == "table":
            return server.submit3d(model, kind, req["src"], req["srnum"])
        return server.submit3d(model, kind, req["src"], req["srnum"], req["rec"], req["nxyz"],
                               req["T_obs"])
    if kind == "table":
        return server.submit2d(model, kind, req["src"], req["srnum"])
    return server.submit2d(model, kind, req["src"], req["srnum"], req["rec"], req["nxy"],
                           req["TT_obs"])


def poisson_run(server, models, rate_hz, duration_s, seed=7, max_src=64, warmup=0, log=None):
    """Submit a Poisson stream (rate_hz, arrivals over duration_s) in real time,
    drive the server until every request finished, release each as it finishes.

    Returns per-request records (grid, kind, nsrc, latency, queue, arrival) and the
    span from the first arrival to the last completion.  The first `warmup`
    requests are served but not recorded.
    """
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_hz, int(rate_hz * duration_s * 2) + 16)
    arrivals = np.cumsum(gaps)
    arrivals = arrivals[arrivals <= duration_s]
    specs = []
    for _ in range(len(arrivals)):
        g = int(rng.integers(len(models)))
        dim, shape, h, ids = models[g]
        kind = "table" if rng.random() < 0.5 else "gradient"
        nsrc = int(rng.integers(1, max_src + 1))
        specs.append((g, int(rng.choice(ids)), kind, nsrc,
                      make_request(dim, shape, h, kind, nsrc, rng)))
    t0 = time.perf_counter()
    live, recs = {}, []
    nxt = 0
    last_log = t0
    while nxt < len(specs) or live:
        now = time.perf_counter() - t0
        while nxt < len(specs) and arrivals[nxt] <= now:
            g, mid, kind, nsrc, req = specs[nxt]
            t = _submit(server, mid, models[g][0], kind, req)
            live[t] = (nxt, g, kind, nsrc)
            nxt += 1
        wait = 2000 if nxt >= len(specs) else max(0, int((arrivals[nxt] - now) * 1e6))
        if server.progress(min(wait, 2000)):
            for t in list(live):
                st, (ts, ta, td) = server.status(t)
                if st in ("done", "failed"):
                    k, g, kind, nsrc = live.pop(t)
                    if st == "failed":
                        raise RuntimeError(f"request {k} failed: {server.error(t)}")
                    if k >= warmup:
                        recs.append(dict(grid=g, kind=kind, nsrc=nsrc, latency=td - ts,
                                         queue=ta - ts, t_submit=ts, t_done=td))
                    server.release(t)
        if log and time.perf_counter() - last_log > 10:
            last_log = time.perf_counter()
            log(f"serving: {nxt}/{len(specs)} submitted, {len(live)} in flight, "
                f"{len(recs)} done")
    return recs


def summarize(recs, grids=SERVING_GRIDS):
    lat = np.array([r["latency"] for r in recs]) * 1e3
    q = np.array([r["queue"] for r in recs]) * 1e3
    span = max(r["t_done"] for r in recs) - min(r["t_submit"] for r in recs)
    src = sum(r["nsrc"] for r in recs)
    per_grid = {}
    for g, (dim, shape, _) in enumerate(grids):
        lg = [r["latency"] * 1e3 for r in recs if r["grid"] == g]
        if lg:
            per_grid["x".join(str(s) for s in shape)] = {
                "n": len(lg), "p50_ms": round(float(np.percentile(lg, 50)), 2),
                "p99_ms": round(float(np.percentile(lg, 99)), 2)}
    return {"requests": len(recs), "p50_ms": round(float(np.percentile(lat, 50)), 2),
            "p99_ms": round(float(np.percentile(lat, 99)), 2),
            "mean_queue_ms": round(float(q.mean()), 2),
            "requests_per_s": round(len(recs) / span, 2), "sources_per_s": round(src / span, 1),
            "per_grid": per_grid}
