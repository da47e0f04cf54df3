"""ctypes bindings for the FP64 CPU oracle (oracle/liboracle.so) and, where it has
been built, the reference's own routines (oracle/_ref/*.so).

Test infrastructure only: the product never imports this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF3D_SO = os.path.join(ROOT, "oracle", "_ref", "libpfastref3d.so")
REF2D_SO = os.path.join(ROOT, "oracle", "_ref", "libpfastref2d.so")

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
d, i = C.c_double, C.c_int

_orc = None
_ref3 = None
_ref2 = None


def oracle():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(ORACLE_SO)
        L.orc_fsm3d.argtypes = [dp, dp, ip, i, i, d, d, d, i, i, i]
        L.orc_fsm3d.restype = i
        L.orc_adjoint_fsm3d.argtypes = [dp, dp, dp, ip, dp, ip, ip, i, i, d, d, d, i, i, i, C.POINTER(C.c_int)]
        L.orc_adjoint_fsm3d.restype = d
        L.orc_rt_fsm2d.argtypes = [dp, dp, dp, dp, dp, ip, i, i, i, d, d, i, i, ip, ip]
        L.orc_rt_fsm2d.restype = i
        L.orc_tadjoint_fsm2d.argtypes = [dp, dp, dp, d, ip, dp, ip, ip, i, i, d, d, i, i, C.POINTER(C.c_int)]
        L.orc_tadjoint_fsm2d.restype = d
        L.orc_radjoint_fsm2d.argtypes = [dp, dp, dp, d, ip, dp, ip, i, i, i, d, d, i, i, ip]
        L.orc_radjoint_fsm2d.restype = d
        L.orc_map_index.argtypes = [dp, i, i, i, d, ip, i]
        L.orc_map_index.restype = None
        L.orc_grad_accum.argtypes = [dp, dp, dp, C.c_long]
        L.orc_grad_accum.restype = None
        L.orc_fwdadj3d_parallel.argtypes = [i, dp, ip, ip, ip, dp, dp, i, i, i, d, d, d, i, i, i,
                                            dp, C.POINTER(C.c_double), ip, ip]
        L.orc_fwdadj3d_parallel.restype = i
        _orc = L
    return _orc


def have_ref():
    return os.path.exists(REF3D_SO) and os.path.exists(REF2D_SO)


def ref3d():
    global _ref3
    if _ref3 is None:
        L = C.CDLL(REF3D_SO)
        L.ref_fsm3d.argtypes = [dp, dp, ip, i, i, d, d, d, i, i, i]
        L.ref_fsm3d.restype = None
        L.ref_adjoint_fsm3d.argtypes = [dp, dp, dp, ip, dp, ip, ip, i, i, d, d, d, i, i, i]
        L.ref_adjoint_fsm3d.restype = d
        L.ref_regu3d.argtypes = [dp, dp, d, d, d, d, d, d, i, i, i]
        L.ref_regu3d.restype = None
        L.ref_gd2sk3d.argtypes = [dp, dp, dp, d, dp, dp, dp, i, i, i, i, i]
        L.ref_gd2sk3d.restype = d
        L.ref_magnitudeOf3d.argtypes = [dp, i, i, i]
        L.ref_magnitudeOf3d.restype = d
        L.ref_updatevp3d.argtypes = [dp, dp, d, i, i, i]
        L.ref_updatevp3d.restype = None
        _ref3 = L
    return _ref3


def ref2d():
    global _ref2
    if _ref2 is None:
        L = C.CDLL(REF2D_SO)
        L.ref_rt_fsm2d.argtypes = [dp, dp, dp, dp, dp, ip, ip, ip, i, i, i, d, d, i, i]
        L.ref_rt_fsm2d.restype = None
        L.ref_tadjoint_fsm2d.argtypes = [dp, dp, dp, ip, dp, ip, ip, i, i, d, d, i, i]
        L.ref_tadjoint_fsm2d.restype = d
        L.ref_tadjoint_fsm2d_wr.argtypes = [dp, dp, dp, d, ip, dp, ip, ip, i, i, d, d, i, i]
        L.ref_tadjoint_fsm2d_wr.restype = d
        L.ref_radjoint_fsm2d.argtypes = [dp, dp, dp, dp, ip, dp, ip, ip, i, i, i, d, d, i, i]
        L.ref_radjoint_fsm2d.restype = d
        L.ref_radjoint_fsm2d_wr.argtypes = [dp, dp, dp, dp, d, ip, dp, ip, ip, i, i, i, d, d, i, i]
        L.ref_radjoint_fsm2d_wr.restype = d
        L.ref_regu2d.argtypes = [dp, dp, d, d, d, d, i, i]
        L.ref_regu2d.restype = None
        L.ref_beta_HS2d.argtypes = [dp, dp, dp, i, i]
        L.ref_beta_HS2d.restype = d
        L.ref_beta_CGDESCENT2d.argtypes = [dp, dp, dp, i, i]
        L.ref_beta_CGDESCENT2d.restype = d
        L.ref_gd2sk2d.argtypes = [dp, dp, dp, d, dp, dp, dp, i, i, i, i]
        L.ref_gd2sk2d.restype = d
        L.ref_cubicInterp.argtypes = [d, d, d, d, d, d]
        L.ref_cubicInterp.restype = d
        L.ref_magnitudeOf2d.argtypes = [dp, i, i]
        L.ref_magnitudeOf2d.restype = d
        L.ref_updatevp2d.argtypes = [dp, dp, d, i, i]
        L.ref_updatevp2d.restype = None
        _ref2 = L
    return _ref2


# ---------------------------------------------------------------- wrappers
def fsm3d(vp, src, srcid, dx, dy, dz, MAXSwp=100):
    nz, nr, nc = vp.shape
    U = np.empty(vp.size)
    cyc = oracle().orc_fsm3d(np.ascontiguousarray(vp.ravel()), U, np.ascontiguousarray(src.ravel(), np.int32),
                             srcid, MAXSwp, dx, dy, dz, nr, nc, nz)
    return U.reshape(vp.shape), cyc


def adjoint_fsm3d(T, T_obs, rec, nxyz, src, srnum, srcid, dx, dy, dz, MAXSwp=100):
    nz, nr, nc = T.shape
    U = np.empty(T.size)
    cyc = C.c_int(0)
    EE = oracle().orc_adjoint_fsm3d(U, np.ascontiguousarray(T.ravel()), np.ascontiguousarray(T_obs, np.float64),
                                    np.ascontiguousarray(rec.ravel(), np.int32),
                                    np.ascontiguousarray(nxyz.ravel(), np.float64),
                                    np.ascontiguousarray(src.ravel(), np.int32),
                                    np.ascontiguousarray(srnum, np.int32), srcid, MAXSwp, dx, dy, dz,
                                    nr, nc, nz, C.byref(cyc))
    return U.reshape(T.shape), EE, cyc.value


def rt_fsm2d(vp, Z, src, srcid, NumofR, dx, dz, MAXSwp=500):
    nr, nc = vp.shape
    U = np.empty(vp.size)
    TTd = np.zeros(max(NumofR, 1) * vp.size)
    TTu = np.zeros(max(NumofR, 1) * vp.size)
    cyc = np.zeros(1 + 2 * max(NumofR, 1), np.int32)
    k2 = np.zeros(nc, np.int32)
    Zf = np.ascontiguousarray(Z.ravel(), np.float64) if Z is not None and Z.size else np.zeros(3 * nc)
    st = oracle().orc_rt_fsm2d(np.ascontiguousarray(vp.ravel()), U, TTd, TTu, Zf,
                               np.ascontiguousarray(src.ravel(), np.int32), srcid, NumofR, MAXSwp,
                               dx, dz, nr, nc, cyc, k2)
    return dict(status=st, T=U.reshape(vp.shape), TTd=TTd.reshape(-1, nr, nc)[:NumofR],
                TTu=TTu.reshape(-1, nr, nc)[:NumofR], cycles=cyc, k2vec=k2)


def tadjoint_fsm2d(T, T_obs, Cr, rec, nxyz, src, srnum, srcid, dx, dz, MAXSwp=500):
    nr, nc = T.shape
    U = np.empty(T.size)
    cyc = C.c_int(0)
    EE = oracle().orc_tadjoint_fsm2d(U, np.ascontiguousarray(T.ravel()), np.ascontiguousarray(T_obs, np.float64),
                                     Cr, np.ascontiguousarray(rec.ravel(), np.int32),
                                     np.ascontiguousarray(nxyz.ravel(), np.float64),
                                     np.ascontiguousarray(src.ravel(), np.int32),
                                     np.ascontiguousarray(srnum, np.int32), srcid, MAXSwp, dx, dz, nr, nc,
                                     C.byref(cyc))
    return U.reshape(T.shape), EE, cyc.value


def radjoint_fsm2d(Tu, TR_obs, Cr, rec, nxyz, srnum, srcid, NumofR, dx, dz, MAXSwp=500):
    nr, nc = Tu.shape[-2:]
    UR = np.empty(nr * nc)
    cyc = np.zeros(max(NumofR, 1), np.int32)
    EE = oracle().orc_radjoint_fsm2d(UR, np.ascontiguousarray(Tu.ravel()),
                                     np.ascontiguousarray(TR_obs.ravel(), np.float64), Cr,
                                     np.ascontiguousarray(rec.ravel(), np.int32),
                                     np.ascontiguousarray(nxyz.ravel(), np.float64),
                                     np.ascontiguousarray(srnum, np.int32), srcid, NumofR, MAXSwp,
                                     dx, dz, nr, nc, cyc)
    return UR.reshape(nr, nc), EE, cyc


def map_index(table, col, h):
    table = np.ascontiguousarray(table, np.float64)
    n, stride = table.shape
    out = np.empty(n, np.int32)
    oracle().orc_map_index(table.ravel(), n, stride, col, h, out, 1)
    return out


def ref_wavenumbers(n, h):
    """k[j] = -pi/h + 2 pi j / (n h), then the two halves of length n // 2 swapped
    (src/T_padjoint3d.h:218-232, src/RT_padjoint2d.h:520-530); for odd n this is a
    half-bin-shifted permutation that the regularisation reproduces as is."""
    k = -np.pi / h + 2 * np.pi * np.arange(n) / n / h
    m = n // 2
    k[:m], k[m:2 * m] = k[m:2 * m].copy(), k[:m].copy()
    return k


def regu_np(Dj, vp, nus, hs):
    """FFT smoothing of a gradient (src/T_padjoint3d.h:205-262, 2D :506-545 of
    src/RT_padjoint2d.h) restated with numpy's FFT: Dj <- Re IFFT[FFT(Dj) /
    (1 + sum nu_a^2 k_a^2)], air nodes (vp = 330) set to 0.  Axes in array order:
    3D (z, rows, cols) with nus = (nuz, nuy, nux), hs = (dz, dy, dx); 2D (rows, cols)
    with nus = (nuz, nux), hs = (dz, dx)."""
    F = np.fft.fftn(Dj)
    den = np.ones(Dj.shape)
    for ax, (nu, h) in enumerate(zip(nus, hs)):
        k = ref_wavenumbers(Dj.shape[ax], h)
        sh = [1] * Dj.ndim
        sh[ax] = -1
        den = den + (nu * nu * k * k).reshape(sh)
    out = np.real(np.fft.ifftn(F / den))
    out[np.abs(vp - 330) < 1e-9] = 0.0
    return out
