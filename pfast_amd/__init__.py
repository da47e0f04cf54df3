"""pfast_amd -- MI355X-native traveltime-and-gradient engine (FSM + adjoint-state).

Python mirror of the reference's subprogram interface for this path, calling the
HIP kernels of libpfast.so through its C ABI (include/pfast.h):

    fsm3d            <- src/T_pfsm3d.h:131
    adjoint_fsm3d    <- src/T_padjoint3d.h:73
    RT_fsm2d         <- src/RT_pfsm2d.h:116
    Tadjoint_fsm2d   <- src/RT_padjoint2d.h:288 (and the WR overload :67)
    Radjoint_fsm2d   <- src/RT_padjoint2d.h:382 (and the WR overload :159)

plus the batched device-resident shot-set API (Context / Model3D / Shots3D,
Model2D / Shots2D) that replaces the drivers' per-source loops and MPI
reductions, and a request Server (resident models, traveltime-table and gradient
requests admitted by device-memory footprint and batched into shared sweeps).

Arrays use the reference layouts: 3D fields are numpy arrays of shape
(nz, nrows, ncols), 2D fields (nrows, ncols); src/rec are int32 index tables
(3D: row i, col j, depth k; 2D: row i, col j).  There is no CPU fallback.
"""
import ctypes as C
import weakref

import numpy as np

from ._lib import (PFAST_FWD_ADJ, PFAST_FWD_ONLY, PFAST_REQ_GRADIENT, PFAST_REQ_TABLE,  # noqa: F401
                   REQ_STATES, FatalError, PfastError, check, lib, ptr)

INF = 9999999.0

__all__ = ["fsm3d", "adjoint_fsm3d", "RT_fsm2d", "Tadjoint_fsm2d", "Radjoint_fsm2d", "sweep_variants",
           "Context", "Model3D", "Shots3D", "Model2D", "Shots2D", "Server", "device_count", "PfastError", "FatalError", "INF"]


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def device_count():
    n = C.c_int(0)
    check(lib().pfast_device_count(C.byref(n)))
    return n.value


# ----------------------------------------------------------------------- 3D
def fsm3d(vp, src, srcid, MAXSwp, dx, dy, dz, pe=0):
    """First-arrival traveltime of source `srcid` (FP64, bit-identical to the reference).
    Returns (U, ncycles) with U shaped like vp."""
    vp = _f64(vp)
    nz, nrows, ncols = vp.shape
    U = np.empty_like(vp)
    cyc = C.c_int(0)
    check(lib().pfast_fsm3d(vp.ravel(), U.ravel(), _i32(src).ravel(), int(srcid), int(MAXSwp),
                            float(dx), float(dy), float(dz), nrows, ncols, nz, int(pe),
                            C.byref(cyc)))
    return U, cyc.value


def adjoint_fsm3d(T, T_obs, rec, nxyz, src, srnum, srcid, MAXSwp, dx, dy, dz, pe=0):
    """Adjoint state of source `srcid`.  Returns (lambda, EE, ncycles)."""
    T = _f64(T)
    nz, nrows, ncols = T.shape
    U = np.empty_like(T)
    T_obs, rec, nxyz = _f64(T_obs), _i32(rec), _f64(nxyz)
    EE = C.c_double(0.0)
    cyc = C.c_int(0)
    check(lib().pfast_adjoint_fsm3d(U.ravel(), T.ravel(), ptr(T_obs), ptr(rec), ptr(nxyz),
                                    _i32(src).ravel(), _i32(srnum), int(srcid), int(MAXSwp),
                                    float(dx), float(dy), float(dz), nrows, ncols, nz, int(pe),
                                    C.byref(EE), C.byref(cyc)))
    return U, EE.value, cyc.value


# ----------------------------------------------------------------------- 2D
def RT_fsm2d(vp, Z, src, srcid, NumofR, MAXSwp, dx, dz, pe=0):
    """Transmission + reflection traveltimes of source `srcid`.
    Returns dict(T, TTd, TTu, cycles, k2vec).  Raises FatalError where the
    reference exits."""
    vp = _f64(vp)
    nrows, ncols = vp.shape
    R = max(int(NumofR), 1)
    U = np.empty_like(vp)
    TTd = np.zeros((R, nrows, ncols))
    TTu = np.zeros((R, nrows, ncols))
    cyc = np.zeros(1 + 2 * R, np.int32)
    k2 = np.zeros(ncols, np.int32)
    Zf = _f64(Z) if NumofR else np.zeros(3 * ncols)
    check(lib().pfast_rt_fsm2d(vp.ravel(), U.ravel(), ptr(TTd), ptr(TTu), ptr(Zf),
                               _i32(src).ravel(), None, None, int(srcid), int(NumofR),
                               int(MAXSwp), float(dx), float(dz), nrows, ncols, int(pe),
                               ptr(cyc), ptr(k2)))
    return dict(T=U, TTd=TTd[:NumofR], TTu=TTu[:NumofR], cycles=cyc[:1 + 2 * NumofR], k2vec=k2)


def Tadjoint_fsm2d(T, T_obs, rec, nxyz, src, srnum, srcid, MAXSwp, dx, dz, WR=None, pe=0):
    """Transmission adjoint.  WR=None is the plain overload (Cr = 1), else the WR
    overload (Cr = 1 - WR).  Returns (lambda, EE, ncycles)."""
    T = _f64(T)
    nrows, ncols = T.shape
    U = np.empty_like(T)
    Cr = 1.0 if WR is None else 1.0 - float(WR)
    T_obs, rec, nxyz = _f64(T_obs), _i32(rec), _f64(nxyz)
    EE = C.c_double(0.0)
    cyc = np.zeros(1, np.int32)
    check(lib().pfast_tadjoint_fsm2d(U.ravel(), T.ravel(), ptr(T_obs), Cr, ptr(rec), ptr(nxyz),
                                     _i32(src).ravel(), _i32(srnum), int(srcid), int(MAXSwp),
                                     float(dx), float(dz), nrows, ncols, int(pe), C.byref(EE),
                                     ptr(cyc)))
    return U, EE.value, int(cyc[0])


def Radjoint_fsm2d(Tu, TR_obs, rec, nxyz, src, srnum, srcid, NumofR, MAXSwp, dx, dz, WR=None,
                   pe=0):
    """Reflection adjoint summed over reflectors.  WR=None: plain overload (Cr = 1),
    else the WR overload (Cr = WR).  Returns (lambda_R, EE, cycles per reflector)."""
    Tu = _f64(Tu)
    nrows, ncols = Tu.shape[-2:]
    UR = np.empty((nrows, ncols))
    Cr = 1.0 if WR is None else float(WR)
    TR_obs, rec, nxyz = _f64(TR_obs), _i32(rec), _f64(nxyz)
    EE = C.c_double(0.0)
    cyc = np.zeros(max(int(NumofR), 1), np.int32)
    check(lib().pfast_radjoint_fsm2d(UR.ravel(), None, Tu.ravel(), ptr(TR_obs), Cr, ptr(rec),
                                     ptr(nxyz), _i32(src).ravel(), _i32(srnum), int(srcid),
                                     int(NumofR), int(MAXSwp), float(dx), float(dz), nrows, ncols,
                                     int(pe), C.byref(EE), ptr(cyc)))
    return UR, EE.value, cyc[:NumofR]


# ------------------------------------------------------------------ batched
class Context:
    """A device stream plus profiling state (pfast_ctx)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().pfast_ctx_create(int(device), C.byref(h)))
        self.h = h
        self._children = weakref.WeakSet()  # live models (closed before the context)

    def close(self):
        if self.h:
            for m in list(self._children):
                m.close()
            check(lib().pfast_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().pfast_ctx_sync(self.h))

    def timing(self, enable=True):
        check(lib().pfast_ctx_timing_enable(self.h, 1 if enable else 0))
        check(lib().pfast_ctx_timing_reset(self.h))

    def set_sweep(self, variant):
        """Select the batched sweep kernel variant (0 = default); see sweep_variants()."""
        check(lib().pfast_ctx_set_sweep(self.h, int(variant)))

    def set_sweep2d(self, variant):
        """Batched 2D sweep form: 0 (default) row window with a source's independent
        solves side by side, 1 one work-group per source stepping anti-diagonal planes
        (bit-identical results)."""
        check(lib().pfast_ctx_set_sweep2d(self.h, int(variant)))

    def set_tile_skip(self, mode=True):
        """Exact tile skipping of the batched 3D sweeps: False/0 off, 1 whole tiles,
        True/2 (default) whole tiles and depth ranges (a sweep starts at the first and
        ends after the last level whose inputs changed; same results)."""
        m = 2 if mode is True else 0 if mode is False else int(mode)
        check(lib().pfast_ctx_set_tile_skip(self.h, m))

    def set_paired(self, enable=True):
        """Pair orderings 2p and 2p+1 of a batched 3D cycle in shared per-diagonal
        launches, the second trailing by ntj + 1 diagonals (default on; same results)."""
        check(lib().pfast_ctx_set_paired(self.h, 1 if enable else 0))

    def set_dataflow(self, mode=True):
        """Schedule of the batched 3D sweep cycles: False/0 one launch per tile
        anti-diagonal, True/1 dataflow launches (a tile starts as soon as the values it
        reads are final), "auto"/2 (default) dataflow for cycles with few chunks x
        tiles.  Results are the same under every mode."""
        m = 2 if mode == "auto" else int(mode)
        check(lib().pfast_ctx_set_dataflow(self.h, m))

    def set_check_interval(self, cycles):
        """Sweep cycles of a batched 3D phase between two convergence checks on the
        host (1..64; converged chunks retire on the device every cycle either way, so
        results and cycle counts do not change)."""
        check(lib().pfast_ctx_set_check_interval(self.h, int(cycles)))

    def set_source_order(self, mode="clusters"):
        """Internal source order of 3D shot sets created afterwards: "clusters" (default;
        balanced clusters of eight neighbouring sources), "morton" or "caller".  Results
        come back in the caller's order either way."""
        m = {"caller": 0, "morton": 1, "clusters": 2}[mode] if isinstance(mode, str) else int(mode)
        check(lib().pfast_ctx_set_source_order(self.h, m))

    def set_graphs(self, enable=True):
        """Replay the cycles between two checks from captured hipGraphs."""
        check(lib().pfast_ctx_set_graphs(self.h, 1 if enable else 0))

    def set_streams(self, n):
        """Split batched 3D cycles over n streams (1..4); results do not change."""
        check(lib().pfast_ctx_set_streams(self.h, int(n)))

    # timed kinds (pfast_ctx_timing)
    FWD_DIAG, ADJ_DIAG, SOLVE_2D, FWD_FLOW, ADJ_FLOW = 0, 1, 2, 3, 4

    def timing_read(self, kind):
        """(device ms, launches, node-sweeps issued, node-sweeps swept) of a timed kind:
        FWD_DIAG / ADJ_DIAG (per-diagonal launches), FWD_FLOW / ADJ_FLOW (dataflow
        launches), SOLVE_2D."""
        ms, n, w, sw = C.c_double(0), C.c_long(0), C.c_double(0), C.c_double(0)
        check(lib().pfast_ctx_timing(self.h, int(kind), C.byref(ms), C.byref(n)))
        check(lib().pfast_ctx_work(self.h, int(kind), C.byref(w)))
        check(lib().pfast_ctx_work_swept(self.h, int(kind), C.byref(sw)))
        return ms.value, n.value, w.value, sw.value


def source_order(src, spacing=(1.0, 1.0, 1.0), mode="clusters"):
    """The internal order a 3D shot set gives its sources (pfast_source_order): perm[p] =
    the caller's index of internal source p; src rows are (row, column, depth) nodes,
    spacing (dx, dy, dz).  Host only."""
    s = _i32(src).reshape(-1, 3)
    m = {"caller": 0, "morton": 1, "clusters": 2}[mode] if isinstance(mode, str) else int(mode)
    flat = np.ascontiguousarray(s.ravel())
    perm = np.zeros(len(s), np.int32)
    check(lib().pfast_source_order(len(s), ptr(flat), float(spacing[0]), float(spacing[1]),
                                   float(spacing[2]), m, ptr(perm)))
    return perm


def sweep_variants():
    """Names of the batched sweep kernel variants, index = variant number."""
    out = []
    while True:
        n = lib().pfast_sweep_name(len(out))
        if n is None:
            return out
        out.append(n.decode())


class Model3D:
    """A velocity model resident in HBM (vp shaped (nz, nrows, ncols))."""

    def __init__(self, ctx, vp, dx, dy, dz):
        vp = _f64(vp)
        self.ctx = ctx
        self.shape = vp.shape
        self.dx, self.dy, self.dz = float(dx), float(dy), float(dz)
        nz, nrows, ncols = vp.shape
        h = C.c_void_p()
        check(lib().pfast_model3d_create(ctx.h, vp.ravel(), nrows, ncols, nz, self.dx, self.dy,
                                         self.dz, C.byref(h)))
        self.h = h
        self._children = weakref.WeakSet()  # live shot sets (closed before the model)
        ctx._children.add(self)

    def set_vp(self, vp):
        vp = _f64(vp)
        assert vp.shape == self.shape
        check(lib().pfast_model3d_set_vp(self.h, vp.ravel()))

    def close(self):
        if self.h:
            for s in list(self._children):
                s.close()
            check(lib().pfast_model3d_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Shots3D:
    """All sources of a shot set resident on one GPU.

    src: (nsrc, 3) node indices (i, j, k); srnum: receivers per source; rec: (nrec, 3);
    nxyz: (nrec, 3) normals (n_y, n_x, n_z); T_obs: (nrec,) picks (used where < 100).
    """

    def __init__(self, model, src, srnum, rec=None, nxyz=None, T_obs=None):
        self.model = model
        self.src = _i32(src).reshape(-1, 3)
        self.nsrc = self.src.shape[0]
        self.srnum = _i32(srnum)
        self.nrec = int(self.srnum.sum())
        self._rec = None if rec is None else _i32(rec).reshape(-1, 3)
        self._nxyz = None if nxyz is None else _f64(nxyz).reshape(-1, 3)
        self._tobs = None if T_obs is None else _f64(T_obs)
        h = C.c_void_p()
        check(lib().pfast_shots3d_create(model.h, self.nsrc, self.src.ravel(), self.srnum,
                                         ptr(self._rec), ptr(self._nxyz), ptr(self._tobs),
                                         C.byref(h)))
        self.h = h
        model._children.add(self)

    def run(self, adjoint=True, MAXSwp=100):
        check(lib().pfast_shots3d_run(self.h, PFAST_FWD_ADJ if adjoint else PFAST_FWD_ONLY,
                                      int(MAXSwp)))

    def set_source_init(self, halfwidth):
        """Fix the traveltimes of the nodes within `halfwidth` nodes of each source to
        FP64 values of the linear velocity model fitted at the source (0, the default:
        the source node alone, as the reference seeds it).  Applies from the next run."""
        check(lib().pfast_shots3d_set_source_init(self.h, int(halfwidth)))

    def autotune(self, variants=None):
        """Time two forward cycles (and two adjoint cycles, when the shot set has picks)
        of this shot set under each sweep variant (all of them by default), set the
        fastest on the context and return (variant, {variant: ms}).  Results are the
        same under every variant; run the shot set afterwards."""
        cand = None if variants is None else _i32(variants)
        n = len(sweep_variants()) if cand is None else len(cand)
        ms = np.zeros(n)
        chosen = C.c_int(0)
        check(lib().pfast_shots3d_autotune(self.h, ptr(cand), 0 if cand is None else n,
                                           C.byref(chosen), ptr(ms)))
        keys = range(n) if cand is None else cand.tolist()
        return chosen.value, dict(zip(keys, ms.tolist()))

    SCHEDULE_CANDIDATES = ((2, 1), (0, 1), (1, 1), (2, 2), (2, 4))

    def autotune_schedule(self):
        """Time a full solve under each (schedule, check interval) candidate, set the
        fastest on the context and return ((schedule, interval), {candidate: ms}).
        Results are the same under every candidate; run the shot set afterwards."""
        ms = np.zeros(len(self.SCHEDULE_CANDIDATES))
        sc, iv = C.c_int(0), C.c_int(0)
        check(lib().pfast_shots3d_autotune_schedule(self.h, C.byref(sc), C.byref(iv), ptr(ms)))
        return (sc.value, iv.value), dict(zip(self.SCHEDULE_CANDIDATES, ms.tolist()))

    def gradient(self):
        N = int(np.prod(self.model.shape))
        g = np.empty(N)
        m = C.c_double(0)
        check(lib().pfast_shots3d_gradient(self.h, ptr(g), C.byref(m)))
        return g.reshape(self.model.shape), m.value

    def misfit(self):
        m = C.c_double(0)
        check(lib().pfast_shots3d_gradient(self.h, None, C.byref(m)))
        return m.value

    def picks(self):
        out = np.empty(self.nrec)
        check(lib().pfast_shots3d_picks(self.h, out))
        return out

    def field(self, s, adjoint=False):
        out = np.empty(int(np.prod(self.model.shape)))
        check(lib().pfast_shots3d_field(self.h, 1 if adjoint else 0, int(s), out))
        return out.reshape(self.model.shape)

    def cycles(self):
        c = np.zeros(2 * self.nsrc, np.int32)
        check(lib().pfast_shots3d_cycles(self.h, c))
        return c.reshape(-1, 2)

    def plain_source(self, s, MAXSwp=100, fields=False):
        """FP64 single-source forward+adjoint of source s on the resident inputs.
        Returns (EE, cycles[2]) or, with fields=True, (EE, cycles, T, lambda)."""
        EE = C.c_double(0)
        cyc = np.zeros(2, np.int32)
        T = lam = None
        if fields:
            T = np.empty(self.model.shape)
            lam = np.empty(self.model.shape)
        check(lib().pfast_shots3d_plain_source(self.h, int(s), int(MAXSwp), C.byref(EE),
                                               ptr(cyc), ptr(T), ptr(lam)))
        if fields:
            return EE.value, cyc, T, lam
        return EE.value, cyc

    def close(self):
        if self.h:
            lib().pfast_shots3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Model2D:
    """A 2D velocity model resident in HBM (vp shaped (nrows, ncols)) with its
    reflectors Z shaped (NumofR, ncols, 3): depth, n_z, n_x (the .ref table)."""

    def __init__(self, ctx, vp, dx, dz, Z=None):
        vp = _f64(vp)
        self.ctx = ctx
        self.shape = vp.shape
        self.dx, self.dz = float(dx), float(dz)
        nrows, ncols = vp.shape
        self.NumofR = 0 if Z is None else int(np.asarray(Z).shape[0])
        self._Z = None if Z is None else _f64(Z).reshape(self.NumofR, ncols, 3)
        h = C.c_void_p()
        check(lib().pfast_model2d_create(ctx.h, vp.ravel(), nrows, ncols, self.dx, self.dz,
                                         ptr(self._Z), self.NumofR, C.byref(h)))
        self.h = h
        self._children = weakref.WeakSet()
        ctx._children.add(self)

    def set_vp(self, vp):
        vp = _f64(vp)
        assert vp.shape == self.shape
        check(lib().pfast_model2d_set_vp(self.h, vp.ravel()))

    def close(self):
        if self.h:
            for s in list(self._children):
                s.close()
            check(lib().pfast_model2d_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Shots2D:
    """Every source of a 2D survey on the GPU: RT_fsm2d plus both adjoints per source
    in FP32, FP64 gradient and misfit reductions (the 2D driver's per-source loop).

    src: (nsrc, 2) nodes (row, col); srnum: receivers per source; rec: (nrec, 2);
    nxy: (nrec, 2) normals (n_z, n_x); TT_obs: (nrec,) transmission picks (< 100 used);
    TR_obs: (nrec, NumofR) reflection picks (!= 0 used).
    """

    def __init__(self, model, src, srnum, rec=None, nxy=None, TT_obs=None, TR_obs=None):
        self.model = model
        self.src = _i32(src).reshape(-1, 2)
        self.nsrc = self.src.shape[0]
        self.srnum = _i32(srnum)
        self.nrec = int(self.srnum.sum())
        R = model.NumofR
        self._rec = None if rec is None else _i32(rec).reshape(-1, 2)
        self._nxy = None if nxy is None else _f64(nxy).reshape(-1, 2)
        self._tt = None if TT_obs is None else _f64(TT_obs)
        self._tr = None if TR_obs is None else _f64(TR_obs).reshape(-1, max(R, 1))
        h = C.c_void_p()
        check(lib().pfast_shots2d_create(model.h, self.nsrc, self.src.ravel(), self.srnum,
                                         ptr(self._rec), ptr(self._nxy), ptr(self._tt),
                                         ptr(self._tr), C.byref(h)))
        self.h = h
        model._children.add(self)

    def run(self, adjoint=True, MAXSwp=500):
        check(lib().pfast_shots2d_run(self.h, PFAST_FWD_ADJ if adjoint else PFAST_FWD_ONLY,
                                      int(MAXSwp)))

    def gradient(self, WR=None):
        """(gT, gR, gRT, (Et, Er, E)); gRT = gT (1 - WR) + gR WR (WR per node or None)."""
        N = int(np.prod(self.model.shape))
        gT, gR, gRT, E3 = np.empty(N), np.empty(N), np.empty(N), np.empty(3)
        wr = None if WR is None else _f64(np.broadcast_to(WR, self.model.shape))
        check(lib().pfast_shots2d_gradient(self.h, ptr(wr), ptr(gT), ptr(gR), ptr(gRT), ptr(E3)))
        sh = self.model.shape
        return gT.reshape(sh), gR.reshape(sh), gRT.reshape(sh), tuple(E3)

    def field(self, s, which="T", ir=0):
        """which: 'T', 'TTd', 'TTu', 'lambdaT', 'lambdaR'."""
        k = {"T": 0, "TTd": 1, "TTu": 2, "lambdaT": 3, "lambdaR": 4}[which]
        out = np.empty(int(np.prod(self.model.shape)))
        check(lib().pfast_shots2d_field(self.h, k, int(s), int(ir), out))
        return out.reshape(self.model.shape)

    def info(self):
        """(cycles (nsrc, 1+2R+1+R), k2vec (nsrc, ncols))."""
        R = self.model.NumofR
        cyc = np.empty((self.nsrc, 2 + 3 * R), np.int32)
        k2 = np.empty((self.nsrc, self.model.shape[1]), np.int32)
        check(lib().pfast_shots2d_info(self.h, ptr(cyc), ptr(k2)))
        return cyc, k2

    def close(self):
        if self.h:
            lib().pfast_shots2d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request server on one GPU (pfast_server_* of include/pfast.h).

    Models are added once and stay resident; submit3d / submit2d queue a
    traveltime-table ("table") or gradient ("gradient") request for a source set
    on a resident model and return a ticket; progress() admits, launches and
    retires work.  Results stay on the device until release().
    """

    def __init__(self, device=0, budget_bytes=64 << 30):
        h = C.c_void_p()
        check(lib().pfast_server_create(int(device), C.c_size_t(int(budget_bytes)), C.byref(h)))
        self.h = h
        self._models = {}   # id -> (dim, shape, NumofR)
        self._tickets = {}  # ticket -> (model id, nsrc)

    def add_model3d(self, vp, dx, dy, dz):
        vp = _f64(vp)
        nz, nr, nc = vp.shape
        mid = C.c_int(0)
        check(lib().pfast_server_add_model3d(self.h, vp.ravel(), nr, nc, nz, float(dx), float(dy),
                                             float(dz), C.byref(mid)))
        self._models[mid.value] = (3, vp.shape, 0)
        return mid.value

    def add_model2d(self, vp, dx, dz, Z=None):
        vp = _f64(vp)
        nr, nc = vp.shape
        R = 0 if Z is None else int(np.asarray(Z).shape[0])
        Zf = None if Z is None else _f64(Z).ravel()
        mid = C.c_int(0)
        check(lib().pfast_server_add_model2d(self.h, vp.ravel(), nr, nc, float(dx), float(dz),
                                             ptr(Zf), R, C.byref(mid)))
        self._models[mid.value] = (2, vp.shape, R)
        return mid.value

    def remove_model(self, model):
        check(lib().pfast_server_remove_model(self.h, int(model)))
        del self._models[model]

    @staticmethod
    def _kind(kind):
        return {"table": PFAST_REQ_TABLE, "gradient": PFAST_REQ_GRADIENT}[kind]

    def submit3d(self, model, kind, src, srnum, rec=None, nxyz=None, T_obs=None, MAXSwp=100):
        src = _i32(src).reshape(-1, 3)
        srnum = _i32(srnum)
        rec = None if rec is None else _i32(rec).reshape(-1, 3)
        nxyz = None if nxyz is None else _f64(nxyz).reshape(-1, 3)
        T_obs = None if T_obs is None else _f64(T_obs)
        t = C.c_long(0)
        check(lib().pfast_server_submit3d(self.h, int(model), self._kind(kind), src.shape[0],
                                          src.ravel(), srnum, ptr(rec), ptr(nxyz), ptr(T_obs),
                                          int(MAXSwp), C.byref(t)))
        self._tickets[t.value] = (model, src.shape[0])
        return t.value

    def submit2d(self, model, kind, src, srnum, rec=None, nxy=None, TT_obs=None, TR_obs=None,
                 MAXSwp=500):
        src = _i32(src).reshape(-1, 2)
        srnum = _i32(srnum)
        rec = None if rec is None else _i32(rec).reshape(-1, 2)
        nxy = None if nxy is None else _f64(nxy).reshape(-1, 2)
        TT_obs = None if TT_obs is None else _f64(TT_obs)
        TR_obs = None if TR_obs is None else _f64(TR_obs)
        t = C.c_long(0)
        check(lib().pfast_server_submit2d(self.h, int(model), self._kind(kind), src.shape[0],
                                          src.ravel(), srnum, ptr(rec), ptr(nxy), ptr(TT_obs),
                                          ptr(TR_obs), int(MAXSwp), C.byref(t)))
        self._tickets[t.value] = (model, src.shape[0])
        return t.value

    def progress(self, wait_us=0):
        n = C.c_int(0)
        check(lib().pfast_server_progress(self.h, int(wait_us), C.byref(n)))
        return n.value

    def status(self, ticket):
        """(state name, (t_submit, t_admit, t_done) in seconds since creation)."""
        st = C.c_int(0)
        t = np.zeros(3)
        check(lib().pfast_server_status(self.h, int(ticket), C.byref(st), ptr(t)))
        return REQ_STATES[st.value], tuple(t.tolist())

    def error(self, ticket):
        st, _ = self.status(ticket)
        return lib().pfast_last_error().decode(errors="replace") if st == "failed" else None

    def cycles(self, ticket):
        model, nsrc = self._tickets[ticket]
        dim, _, R = self._models[model]
        c = np.zeros(2 * nsrc if dim == 3 else nsrc * (2 + 3 * R), np.int32)
        check(lib().pfast_server_cycles(self.h, int(ticket), ptr(c)))
        return c

    def gradient(self, ticket):
        """3D: (grad, misfit); 2D: (gT, gR, Et, Er)."""
        model, _ = self._tickets[ticket]
        dim, shape, _ = self._models[model]
        g = np.empty(shape)
        m = np.zeros(2)
        if dim == 3:
            check(lib().pfast_server_gradient(self.h, int(ticket), ptr(g), None, ptr(m)))
            return g, float(m[0])
        gr = np.empty(shape)
        check(lib().pfast_server_gradient(self.h, int(ticket), ptr(g), ptr(gr), ptr(m)))
        return g, gr, float(m[0]), float(m[1])

    def field(self, ticket, s):
        model, _ = self._tickets[ticket]
        out = np.empty(self._models[model][1])
        check(lib().pfast_server_field(self.h, int(ticket), int(s), out.ravel()))
        return out

    def footprint(self, ticket):
        b = C.c_size_t(0)
        check(lib().pfast_server_footprint(self.h, int(ticket), C.byref(b)))
        return b.value

    def release(self, ticket):
        check(lib().pfast_server_release(self.h, int(ticket)))
        self._tickets.pop(ticket, None)

    def info(self):
        used, q, r, cyc = C.c_size_t(0), C.c_int(0), C.c_int(0), C.c_long(0)
        ns = C.c_double(0)
        check(lib().pfast_server_info(self.h, C.byref(used), C.byref(q), C.byref(r), C.byref(cyc),
                                      C.byref(ns)))
        return dict(in_use=used.value, queued=q.value, running=r.value, cycles=cyc.value,
                    node_sweeps=ns.value)

    def close(self):
        if self.h:
            lib().pfast_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
