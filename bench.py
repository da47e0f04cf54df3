#!/usr/bin/env python3
"""Benchmark: batched FP32 forward+adjoint source-solves/s on the 3D diving-wave
tomography workload (201 x 201 x 101, h = 25 m, 400 sources x 2000 receivers),
against the single-source FP64 GPU form and the CPU path on the host cores.

A step = one forward + adjoint + FP64 gradient/misfit reduction over all sources
(every source of the shot set batched on the GPU), inputs resident in HBM.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes as C
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "batched FP32 vs plain FSM fwd+adj source-solves/s (3D 201²×101); p99 lat"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(cfg, nthreads, nsamp, maxswp):
    """The reference's own FSM/adjoint routines (oracle/_ref, compiled from the
    reference sources) where present, else the FP64 oracle restatement, with the
    reference's rank-per-contiguous-source-block layout over `nthreads` host
    threads, on the first `nsamp` sources of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    vp = cfg["vp"]
    nz, nr, nc = vp.shape
    N = vp.size
    h = cfg["h"]
    src, srnum, rec, nxyz, tobs = cfg["src"], cfg["srnum"], cfg["rec"], cfg["nxyz"], cfg["T_obs"]
    ref = None
    if O.have_ref():
        try:
            ref = O.ref3d()
        except OSError:
            ref = None
    O.oracle()
    kind = "reference" if ref is not None else "port"
    P = max(1, min(nthreads, nsamp))
    per, rmd = divmod(nsamp, P)
    blocks, s0 = [], 0
    for r in range(P):
        cnt = per + (1 if r < rmd else 0)
        blocks.append((s0, s0 + cnt))
        s0 += cnt
    vpf = np.ascontiguousarray(vp.ravel())
    srcf = np.ascontiguousarray(src.ravel())
    recf = np.ascontiguousarray(rec.ravel())
    nf = np.ascontiguousarray(nxyz.ravel())
    grads = [np.zeros(N) for _ in range(P)]
    mis = [0.0] * P

    def rank(r):
        b0, b1 = blocks[r]
        T = np.empty(N)
        lam = np.empty(N)
        for s in range(b0, b1):
            if ref is not None:
                ref.ref_fsm3d(vpf, T, srcf, s, maxswp, h, h, h, nr, nc, nz)
                mis[r] += ref.ref_adjoint_fsm3d(lam, T, tobs, recf, nf, srcf, srnum, s, maxswp,
                                                h, h, h, nr, nc, nz)
            else:
                O.oracle().orc_fsm3d(vpf, T, srcf, s, maxswp, h, h, h, nr, nc, nz)
                cyc = C.c_int(0)
                mis[r] += O.oracle().orc_adjoint_fsm3d(lam, T, tobs, recf, nf, srcf, srnum, s,
                                                       maxswp, h, h, h, nr, nc, nz, C.byref(cyc))
            O.oracle().orc_grad_accum(grads[r], lam, vpf, N)

    th = [threading.Thread(target=rank, args=(r,)) for r in range(P)]
    t0 = time.perf_counter()
    for t in th:
        t.start()
    for t in th:
        t.join()
    wall = time.perf_counter() - t0
    return dict(value=nsamp / wall, unit="source-solves/s", cores=P, kind=kind,
                nproc=os.cpu_count(),
                cores_note="threads = the host's CPU share for one GPU (16 here), not nproc: "
                           "the GPU hosts share their cores between GPUs",
                sample=f"sources 0..{nsamp - 1} of the same 400-source workload, full FP64 "
                       f"forward+adjoint+gradient each, {P} threads (rank per contiguous source "
                       f"block), wall {wall:.1f} s")


def tile_skip_mode(args):
    return "off" if args.no_tile_skip else args.tile_skip


def measured_traffic(kernel, sweep_name, args):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (profiles/traffic.json, copy-calibrated), when they were taken
    on this same workload, sweep variant, tile-skip and pairing setting (a launch
    moves the same bytes whether or not other launches overlap it); else None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    if (t.get("sweep") != sweep_name or t.get("nsrc") != args.nsrc or t.get("nrec") != args.nrec
            or t.get("grid") != [101, 201, 201]
            or t.get("tile_skip", False) != (tile_skip_mode(args) != "off")
            or t.get("start_levels", False) != (tile_skip_mode(args) == "levels")
            or t.get("paired", False) != (args.paired == "on")):
        return None
    return t.get("bytes_per_launch", {}).get(kernel)


def serving(args):
    """p50/p99 latency and throughput of the request server under a Poisson stream
    of table and gradient requests on the five serving grids."""
    import pfast_amd
    from pfast_amd import loadgen as LG
    srv = pfast_amd.Server(0, int(args.serve_budget_gib * (1 << 30)))
    t0 = time.perf_counter()
    models = LG.load_models(srv, per_grid=2)
    log(f"serving: 10 resident models in {time.perf_counter() - t0:.1f} s; warm-up stream ...")
    LG.poisson_run(srv, models, 4.0, 2.0, seed=1)
    log(f"serving: Poisson stream, {args.serve_rate} requests/s for {args.serve_seconds} s ...")
    recs = LG.poisson_run(srv, models, args.serve_rate, args.serve_seconds, seed=7, log=log)
    info = srv.info()
    srv.close()
    res = LG.summarize(recs)
    res.update({"rate_hz": args.serve_rate, "arrival_window_s": args.serve_seconds,
                "workload": "Poisson arrivals; grid uniform over 2D 101x51, 2D 401x101, 3D 65^3, "
                            "3D 129x129x65, 3D 321x321x161 (h=25 m, 2 resident seeded models "
                            "each); 1-64 surface sources; table or gradient (32 receivers per "
                            "source); latency = submission to result resident on the device",
                "sweep_cycles": info["cycles"], "node_sweeps": info["node_sweeps"]})
    log(f"serving: {res['requests']} requests, p50 {res['p50_ms']} ms, p99 {res['p99_ms']} ms, "
        f"{res['requests_per_s']} requests/s")
    return res


def tomo2d(args):
    """The 2D workloads of the survey (101x401 tomography, 60 sources x 200 receivers;
    151x601 joint transmission + reflection, 80 sources x 300 receivers, one reflector):
    one batched FP32 gradient evaluation (every source's forward and adjoints and the
    FP64 gradient reduction) on the starting model, picks from the true model's own
    FP32 forward, against the plain FP64 form of one source."""
    import numpy as np
    import pfast_amd
    from pfast_amd import synth
    res = {}
    ctx = pfast_amd.Context(0)
    for kind, name in ((2, "tomography 101x401, 60 sources x 200 receivers"),
                       (3, "joint T+R 151x601, 80 sources x 300 receivers, 1 reflector")):
        cfg = synth.config2d(kind)
        nsrc, nrec, R = cfg["src"].shape[0], int(cfg["srnum"][0]), cfg["NumofR"]
        rec, h = cfg["rec"], cfg["h"]
        mt = pfast_amd.Model2D(ctx, cfg["vp"], h, h, cfg["Z"])
        st = pfast_amd.Shots2D(mt, cfg["src"], cfg["srnum"], rec, cfg["nxy"])
        st.run(adjoint=False, MAXSwp=500)
        tt, tr = np.empty(nsrc * nrec), np.empty((nsrc * nrec, max(1, R)))
        for s in range(nsrc):
            rows = slice(s * nrec, (s + 1) * nrec)
            tt[rows] = st.field(s, "T")[rec[rows, 0], rec[rows, 1]]
            if R:
                tr[rows, 0] = st.field(s, "TTu", 0)[rec[rows, 0], rec[rows, 1]]
        st.close()
        mt.close()
        m0 = pfast_amd.Model2D(ctx, cfg["vp0"], h, h, cfg["Z"])
        sh = pfast_amd.Shots2D(m0, cfg["src"], cfg["srnum"], rec, cfg["nxy"],
                               *((tt, tr) if R else (tt,)))
        sh.run(adjoint=True, MAXSwp=500)  # warm-up
        best = float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            sh.run(adjoint=True, MAXSwp=500)
            sh.gradient(0.5)
            best = min(best, time.perf_counter() - t0)
        # roofline: algorithmic bytes of the node-sweeps the reference performs (forward:
        # every node per ordering, T read and written and the slowness read; adjoints:
        # interior nodes, lambda read and written and T read: 12 B each) over the time
        cyc = np.asarray(sh.info()[0], np.float64)
        nr_, nc_ = cfg["nrows"], cfg["ncols"]
        fwd_sweeps = cyc[:, :1 + 2 * R].sum() * 4 * nr_ * nc_
        adj_sweeps = cyc[:, 1 + 2 * R:].sum() * 4 * (nr_ - 2) * (nc_ - 2)
        gbps = 12.0 * (fwd_sweeps + adj_sweeps) / best / 1e9
        sh.close()
        m0.close()
        t0 = time.perf_counter()
        r = pfast_amd.RT_fsm2d(cfg["vp0"], cfg["Z"], cfg["src"], 0, R, 500, h, h)
        pfast_amd.Tadjoint_fsm2d(r["T"], tt, rec, cfg["nxy"], cfg["src"], cfg["srnum"], 0, 500,
                                 h, h)
        if R:
            pfast_amd.Radjoint_fsm2d(r["TTu"], tr, rec, cfg["nxy"], cfg["src"], cfg["srnum"], 0,
                                     1, 500, h, h)
        plain = time.perf_counter() - t0
        res[name] = {"gradient_ms": round(best * 1e3, 2),
                     "source_solves_per_s": round(nsrc / best, 1),
                     "plain_fp64_one_source_ms": round(plain * 1e3, 2),
                     "ratio_vs_plain_fp64": round(nsrc / best * plain, 1),
                     "roofline": {"bound": "hbm", "limited_by": "latency: one dependent "
                                  "chain of row and column steps per sweep and source",
                                  "achieved": round(gbps, 1), "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": round(gbps / HBM_PEAK_GBS, 4),
                                  "node_sweeps": fwd_sweeps + adj_sweeps}}
        log(f"2D {name}: {best * 1e3:.1f} ms per gradient ({nsrc / best:.0f} source-solves/s), "
            f"plain FP64 {plain * 1e3:.1f} ms per source")
    ctx.close()
    return res


def driver2d(args):
    """Wall time of the drop-in 2D driver (pfast_amd/bin/pfast_rt2d, built by build()) on
    the reference's shipped joint inversion example (tests/golden/example2d_inv.npz:
    160x600, 56 sources x 600 receivers, two reflectors), RTtag 3, L-BFGS memory 50,
    10 iterations, from process start to exit (file writes included)."""
    import re
    import subprocess
    import tempfile
    import numpy as np
    from pfast_amd import formats as F
    e = np.load(os.path.join(ROOT, "tests", "golden", "example2d_inv.npz"))
    nr, nc = (int(v) for v in e["grid"])
    dx, dz = (float(v) for v in e["spacing"])
    n = len(e["tt"])
    rec = np.concatenate([np.tile(e["geom"], (n // len(e["geom"]), 1)),
                          e["tt"].astype(np.float64)[:, None], np.zeros((n, 1)),
                          e["tr2"].astype(np.float64)[:, None]], 1)
    with tempfile.TemporaryDirectory() as d:
        F.write_bin(os.path.join(d, "m.vp"), e["vp"].astype(np.float64))
        F.write_bin(os.path.join(d, "m.src"), e["src"])
        F.write_bin(os.path.join(d, "m.rec"), rec)
        F.write_bin(os.path.join(d, "m.ref"), e["ref"])
        F.write_bin(os.path.join(d, "m.wr"), np.full((nr, nc), float(e["wr"][0])))
        F.write_inp2d(os.path.join(d, "m.inp"), "m.vp", "m.src", e["src"].shape, "m.rec",
                      rec.shape, "m.ref", e["ref"].shape, nr, nc, dx, dz, vpfile="final.vp",
                      nux=float(e["nu"][0]), nuz=float(e["nu"][1]), twfile="m.wr", schemetag=3,
                      lsrc=2)
        t0 = time.perf_counter()
        r = subprocess.run([os.path.join(ROOT, "pfast_amd", "bin", "pfast_rt2d"), "m.inp", "50",
                            "3", "--max-iter", "10", "--no-shot-gather"], cwd=d,
                           capture_output=True, text=True, timeout=300)
        wall = time.perf_counter() - t0
    if r.returncode != 0:
        raise RuntimeError(r.stdout[-1000:] + r.stderr[-1000:])
    res = [float(v) for v in re.findall(r"Residual \(TR\): ([-0-9.eE+]+) ms", r.stdout)]
    log(f"2D driver: 10 iterations in {wall:.2f} s, residual {res[0]:.4f} -> {res[-1]:.4f} ms")
    return {"workload": "shipped 2D joint inversion example, 160x600, 56 sources x 600 "
                        "receivers, 2 reflectors, RTtag 3, L-BFGS m=50",
            "iterations": 10, "wall_s": round(wall, 3),
            "residual_TR_ms": [res[0], res[-1]]}


def driver3d(cfg, iters=3):
    """Wall time of the drop-in 3D driver (pfast_amd/bin/pfast_t3d) on this workload:
    the model above as the starting model, its 400 x 2000 picks, L-BFGS memory 5,
    `iters` iterations from process start to exit (reading the .inp/.vp/.src/.rec files,
    every gradient evaluation and line-search trial, the FFT regularisation and the
    model writes included)."""
    import re
    import subprocess
    import tempfile
    from pfast_amd import formats as F
    nz, nr, nc = cfg["vp"].shape
    h = cfg["h"]
    with tempfile.TemporaryDirectory() as d:
        F.write_bin(os.path.join(d, "m.vp"), cfg["vp"])
        F.write_bin(os.path.join(d, "m.src"), cfg["src_tab"])
        F.write_bin(os.path.join(d, "m.rec"), cfg["rec_tab"])
        F.write_inp3d(os.path.join(d, "m.inp"), "m.vp", "m.src", cfg["src_tab"].shape, "m.rec",
                      cfg["rec_tab"].shape, nr, nc, nz, h, h, h, vpfile="final.vp")
        t0 = time.perf_counter()
        # --no-autotune: the wall time of the iterations themselves (the driver's own
        # tuning, on by default, times the same candidates the bench tuned above)
        r = subprocess.run([os.path.join(ROOT, "pfast_amd", "bin", "pfast_t3d"), "m.inp", "5",
                            "--max-iter", str(iters), "--no-autotune"], cwd=d,
                           capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
    if r.returncode != 0:
        raise RuntimeError(r.stdout[-1000:] + r.stderr[-1000:])
    trials = len(re.findall(r"\(alpha0,alpha1,phi0,phi1,dphi0,dphi1\)", r.stdout))
    log(f"3D driver: {iters} iterations in {wall:.2f} s ({trials} line-search trial lines)")
    return {"workload": "this 3D workload as the starting model, L-BFGS m=5",
            "iterations": iters, "wall_s": round(wall, 3), "trial_lines": trials}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nsrc", type=int, default=400)
    ap.add_argument("--nrec", type=int, default=2000)
    ap.add_argument("--maxswp", type=int, default=100)
    ap.add_argument("--plain-sources", type=int, default=4,
                    help="sources run through the single-source FP64 GPU form")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="CPU baseline threads (default: the 16-core CPU share of one GPU, "
                         "or nproc if smaller)")
    ap.add_argument("--cpu-sources", type=int, default=0, help="0 = one per thread")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sweep", type=int, default=0,
                    help="batched sweep kernel variant (pfast_amd.sweep_variants())")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain device-to-device copy of MIB MiB (3 times), the "
                         "known-size kernel that scales the HBM byte counters of a profile")
    ap.add_argument("--autotune", dest="autotune", action="store_true", default=True,
                    help="(default) before the warm-up, choose the sweep variant (tile shape "
                         "and batch depths), then the schedule and the convergence-check "
                         "interval, by timing them on this shot set (results are the same "
                         "under each)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep --sweep, --schedule and --check-interval as given")
    ap.add_argument("--check-interval", type=int, default=2,
                    help="sweep cycles between two convergence checks on the host (the "
                         "setting the autotuner starts from; 2 measured fastest on this "
                         "workload, same results under any interval)")
    ap.add_argument("--graphs", choices=["on", "off"], default="on",
                    help="replay the cycles between checks from captured hipGraphs (the "
                         "bench line also carries the step time with the other setting)")
    ap.add_argument("--ab-steps", type=int, default=2,
                    help="steps timed with hipGraphs toggled, for the graph A/B (0: skip)")
    ap.add_argument("--schedule", choices=["auto", "dataflow", "diagonal"], default="auto",
                    help="batched 3D sweep schedule: dataflow launches (two per cycle), one "
                         "launch per tile anti-diagonal, or per cycle the one that suits its "
                         "active chunks (same results)")
    ap.add_argument("--paired", choices=["on", "off"], default="on",
                    help="per-diagonal launches pair orderings 2p and 2p+1, the second "
                         "trailing by ntj + 1 tile diagonals (same results)")
    ap.add_argument("--tile-skip", choices=["off", "tiles", "levels"], default="levels",
                    help="exact tile skipping: off, whole tiles, or whole tiles and depth "
                         "ranges (a sweep starts at the first and ends after the last level "
                         "whose inputs changed); same results")
    ap.add_argument("--no-tile-skip", action="store_true", help="same as --tile-skip off")
    ap.add_argument("--streams", type=int, default=3,
                    help="streams a batched cycle splits its source chunks over in the timed "
                         "steps (1..4); the per-launch roofline step always uses one")
    ap.add_argument("--serve-rate", type=float, default=8.0,
                    help="Poisson arrival rate (requests/s) of the serving measurement")
    ap.add_argument("--serve-seconds", type=float, default=30.0,
                    help="arrival window of the serving measurement")
    ap.add_argument("--serve-budget-gib", type=float, default=200.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-2d", action="store_true",
                    help="skip the 2D tomography and joint T+R gradient timings")
    ap.add_argument("--driver2d", action="store_true",
                    help="also time the drop-in 2D driver on the shipped inversion example")
    ap.add_argument("--driver3d", action="store_true",
                    help="also time 3 iterations of the drop-in 3D driver on this workload")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the batched path keeps every source of a shot set "
              "on a single MI355X and has no multi-GPU mode (across GPUs the sources would be "
              "split, with one gradient all-reduce per evaluation).  Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import pfast_amd
    from pfast_amd import synth

    if args.copy_calibration > 0:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copy calibration: 3 x {args.copy_calibration} MiB device copy "
            f"({n * 4} B read + {n * 4} B written per launch)")
        del a, b

    t0 = time.perf_counter()
    cfg = synth.config3d(nsrc=args.nsrc, nrec=args.nrec)
    vp = cfg["vp"]
    h = cfg["h"]
    N = vp.size
    log(f"workload generated in {time.perf_counter() - t0:.1f}s: grid {vp.shape}, "
        f"{args.nsrc} sources x {args.nrec} receivers")
    ctx = pfast_amd.Context(0)
    ctx.set_streams(args.streams)
    ctx.set_tile_skip({"off": 0, "tiles": 1, "levels": 2}[tile_skip_mode(args)])
    ctx.set_paired(args.paired == "on")
    ctx.set_dataflow({"auto": 2, "dataflow": 1, "diagonal": 0}[args.schedule])
    ctx.set_sweep(args.sweep)
    ctx.set_check_interval(args.check_interval)
    ctx.set_graphs(args.graphs == "on")
    sweep_name = pfast_amd.sweep_variants()[args.sweep]
    log(f"sweep kernels: {sweep_name}")
    model = pfast_amd.Model3D(ctx, vp, h, h, h)
    shots = pfast_amd.Shots3D(model, cfg["src"], cfg["srnum"], cfg["rec"], cfg["nxyz"], cfg["T_obs"])
    ctx.sync()
    log("inputs resident in HBM")
    tune = None
    if args.autotune:
        t = time.perf_counter()
        args.sweep, ms = shots.autotune()
        sweep_name = pfast_amd.sweep_variants()[args.sweep]
        log("autotune variants: " + ", ".join(f"{v} {t:.0f} ms" for v, t in ms.items()) +
            f" -> {sweep_name}")
        (sched, args.check_interval), ms2 = shots.autotune_schedule()
        args.schedule = {0: "diagonal", 1: "dataflow", 2: "auto"}[sched]
        log("autotune (schedule, check interval): " +
            ", ".join(f"{k} {v:.0f} ms" for k, v in ms2.items()) +
            f" -> {args.schedule}, every {args.check_interval} cycle(s)")
        tune = {"variant_ms": {pfast_amd.sweep_variants()[v]: round(t_, 1) for v, t_ in ms.items()},
                "schedule_ms": {f"{k[0]},{k[1]}": round(v, 1) for k, v in ms2.items()},
                "seconds": round(time.perf_counter() - t, 1)}

    for w in range(args.warmup):
        t = time.perf_counter()
        shots.run(adjoint=True, MAXSwp=args.maxswp)
        ctx.sync()
        log(f"warmup {w + 1}/{args.warmup}: {time.perf_counter() - t:.3f}s")

    ctx.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        shots.run(adjoint=True, MAXSwp=args.maxswp)
        log(f"step {k + 1}/{args.steps} done")
    ctx.sync()
    elapsed = time.perf_counter() - t0
    ms_step = 1000.0 * elapsed / args.steps
    value = args.nsrc * args.steps / elapsed
    cyc = shots.cycles()
    grad, misfit = shots.gradient()
    log(f"timed: {ms_step:.1f} ms/step, {value:.1f} source-solves/s; FP32 cycles fwd "
        f"{cyc[:, 0].mean():.2f} adj {cyc[:, 1].mean():.2f}; misfit {misfit:.6g}")

    # A/B: the same steps with hipGraph replay of the cycles toggled (one untimed step
    # first, which captures the graphs when they are turned on)
    graphs_ab = None
    if args.ab_steps > 0:
        on = args.graphs == "on"
        ctx.set_graphs(not on)
        shots.run(adjoint=True, MAXSwp=args.maxswp)
        ctx.sync()
        t = time.perf_counter()
        for _ in range(args.ab_steps):
            shots.run(adjoint=True, MAXSwp=args.maxswp)
        ctx.sync()
        other = 1000.0 * (time.perf_counter() - t) / args.ab_steps
        ctx.set_graphs(on)
        graphs_ab = {"on_ms": round(ms_step if on else other, 2),
                     "off_ms": round(other if on else ms_step, 2),
                     "timed_setting": args.graphs, "ab_steps": args.ab_steps}
        log(f"hipGraph A/B: on {graphs_ab['on_ms']:.1f} ms/step, off {graphs_ab['off_ms']:.1f} ms/step")

    # one more step with per-launch HIP events, launches one at a time on the library's
    # stream, so that a launch's duration is its own (the roofline of the kernel).  Per
    # kernel: device time, launches, the node-sweeps credited (every node of every active
    # chunk) and those actually swept (tiles that exact skipping left out excluded,
    # counted on the device)
    ctx.set_streams(1)
    ctx.timing(True)
    shots.run(adjoint=True, MAXSwp=args.maxswp)
    B = args.nsrc
    per_unit = {"fwd": 8.0 + 4.0 / B,  # T read+write per source, cell time per model
                "adj": 12.0}           # lambda read+write, T read
    kinds = [("fwd3d_col", ctx.FWD_DIAG, "fwd"), ("adj3d_col", ctx.ADJ_DIAG, "adj"),
             ("fwd3d_flow", ctx.FWD_FLOW, "fwd"), ("adj3d_flow", ctx.ADJ_FLOW, "adj")]
    kern = {}
    for name, kind, unit in kinds:
        ms_k, n_k, credited, swept = ctx.timing_read(kind)
        if n_k == 0:
            continue
        kern[name] = {"ms": round(ms_k, 2), "launches": n_k,
                      "bytes_swept": swept * per_unit[unit],
                      "bytes_credited": credited * per_unit[unit],
                      "GBps": round(swept * per_unit[unit] / (ms_k * 1e-3) / 1e9, 1),
                      "GBps_effective": round(credited * per_unit[unit] / (ms_k * 1e-3) / 1e9, 1)}
        log(f"kernel {name}: {ms_k:.1f} ms / {n_k} launches -> {kern[name]['GBps']:.0f} GB/s "
            f"swept, {kern[name]['GBps_effective']:.0f} GB/s effective "
            f"({swept / max(credited, 1):.2f} of the credited node-sweeps swept)")
    ctx.timing(False)
    ctx.set_streams(args.streams)
    dom = max(kern, key=lambda k: kern[k]["ms"])

    # single-source FP64 GPU form (bit-identical to the FP64 oracle) on a few sources
    plain = []
    terr, lerr = [], []
    for s in range(args.plain_sources):
        ctx.sync()
        t = time.perf_counter()
        EE, pc, T64, L64 = shots.plain_source(s, MAXSwp=args.maxswp, fields=True)
        dt = time.perf_counter() - t
        plain.append((dt, pc))
        T32 = shots.field(s)
        reach = T64 < 1e6
        terr.append(float(np.abs(T32 - T64)[reach].max() / T64[reach].max()))
        L32 = shots.field(s, adjoint=True)
        lerr.append(float(np.linalg.norm(L32 - L64) / max(np.linalg.norm(L64), 1e-300)))
        log(f"plain FP64 source {s}: {dt:.2f}s, cycles fwd {pc[0]} adj {pc[1]}; FP32 err T "
            f"{terr[-1]:.2e} lambda {lerr[-1]:.2e}")
    plain_rate = len(plain) / sum(p[0] for p in plain) if plain else None
    If = float(np.mean([p[1][0] for p in plain])) if plain else None
    Ia = float(np.mean([p[1][1] for p in plain])) if plain else None

    # roofline of the dominant kernel: the algorithmic bytes of the node-sweeps its
    # launches performed over their mean duration; the credited ("effective") rate,
    # which also counts the tiles that skipping left out, is reported beside it
    k = kern[dom]
    avg_ms = k["ms"] / k["launches"]
    roofline = {"bound": "hbm", "achieved": k["GBps"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(k["GBps"] / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(dom, sweep_name, args),
                "kernel": dom, "launches": k["launches"], "avg_launch_ms": round(avg_ms, 5),
                "bytes_per_launch": round(k["bytes_swept"] / k["launches"]),
                "effective_GBps": k["GBps_effective"],
                "effective_frac": round(k["GBps_effective"] / HBM_PEAK_GBS, 4), "streams": 1}
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "source-solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded gradient+checkerboard model, analytic v(z) picks + 2 ms noise)",
        "config": {"workload": "3D diving-wave tomography, 201x201x101 nodes, h=25 m, "
                               f"{args.nsrc} sources x {args.nrec} receivers, forward+adjoint+gradient",
                   "nsrc": args.nsrc, "nrec": args.nrec, "grid": [101, 201, 201], "h": h,
                   "parallelism": "all sources batched on 1 GPU", "sweep": sweep_name,
                   "schedule": args.schedule, "check_interval": args.check_interval,
                   "graphs": args.graphs, "autotuned": bool(args.autotune),
                   "streams": args.streams, "paired": args.paired,
                   "tile_skip": tile_skip_mode(args)},
        "roofline": roofline,
        "graphs": graphs_ab,
        "autotune": tune,
        "kernels": {n: {kk: v for kk, v in k.items() if not kk.startswith("bytes")}
                    for n, k in kern.items()},
        "fp32_cycles_mean": [round(float(cyc[:, 0].mean()), 3), round(float(cyc[:, 1].mean()), 3)],
        "misfit": misfit,
    }
    # the whole timed step at chip level: the algorithmic bytes of the node-sweeps that the
    # sweeps performed in a step (the per-launch step above performs the same ones: a
    # chunk's launches do not depend on the streams) over the timed step time
    swept_step = sum(k["bytes_swept"] for k in kern.values())
    out["roofline_step"] = {"bytes_swept_per_step": swept_step,
                            "achieved_GBps": round(swept_step / (ms_step * 1e-3) / 1e9, 1),
                            "frac": round(swept_step / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                            "streams": args.streams}
    if plain:
        bytes_solve = 8.0 * N * (If * (8.0 + 4.0 / B) + Ia * 12.0)
        out["plain_fp64"] = {"value": round(plain_rate, 3), "unit": "source-solves/s",
                             "sources": len(plain), "oracle_cycles": [If, Ia]}
        out["ratio_vs_plain_fp64"] = round(value / plain_rate, 1)
        out["roofline_solve"] = {"bytes_per_source_solve": bytes_solve,
                                 "achieved_GBps": round(value * bytes_solve / 1e9, 1),
                                 "frac": round(value * bytes_solve / 1e9 / HBM_PEAK_GBS, 4)}
        out["accuracy_vs_fp64"] = {"T_maxnorm_rel": max(terr), "lambda_l2_rel": max(lerr)}
    if not args.no_cpu:
        ns = args.cpu_sources or args.cpu_threads
        log(f"CPU baseline: {ns} sources on {args.cpu_threads} threads ...")
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_threads, ns, args.maxswp)
        log(f"CPU baseline: {out['cpu_baseline']['value']:.3f} source-solves/s")
    shots.close()
    model.close()
    ctx.close()
    if not args.no_2d:
        try:
            out["tomo2d"] = tomo2d(args)
        except Exception as exc:  # the headline line must still be printed
            log(f"2D measurement failed: {exc!r}")
            out["tomo2d"] = {"error": repr(exc)}
    if args.driver3d:
        try:
            out["driver3d"] = driver3d(cfg)
        except Exception as exc:  # the headline line must still be printed
            log(f"3D driver measurement failed: {exc!r}")
            out["driver3d"] = {"error": repr(exc)}
    if args.driver2d:
        try:
            out["driver2d"] = driver2d(args)
        except Exception as exc:  # the headline line must still be printed
            log(f"2D driver measurement failed: {exc!r}")
            out["driver2d"] = {"error": repr(exc)}
    if not args.no_serve:
        try:
            out["serving"] = serving(args)
        except Exception as exc:  # the headline line must still be printed
            log(f"serving measurement failed: {exc!r}")
            out["serving"] = {"error": repr(exc)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
