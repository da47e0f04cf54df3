"""GPU check of the allocation-free entry point pfast_fsm3d_noalloc against
pfast_fsm3d (src/T_pfsm3d.h:131) through the C ABI: fields and cycle counts equal
bit for bit on two calls in a row, the second on a smaller grid (the shared scratch
keeps the first call's contents), and a grid above the scratch limit is refused
before anything runs."""
import ctypes as C

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _call(fn, c, srcid):
    nz, nr, nc = c["vp"].shape
    U = np.empty(c["vp"].size)
    cyc = C.c_int(0)
    rc = fn(np.ascontiguousarray(c["vp"].ravel()), U, np.ascontiguousarray(c["src"].ravel()),
            srcid, 100, c["dx"], c["dy"], c["dz"], nr, nc, nz, 0, C.byref(cyc))
    return rc, U, cyc.value


def test_noalloc_matches_fsm3d_two_calls():
    from pfast_amd import _lib
    L = _lib.lib()
    big = cases.case3d(81, shape=(17, 23, 29), spacing=(25.0, 25.0, 25.0), nsrc=2)
    small = cases.case3d(82, shape=(9, 12, 10), spacing=(12.5, 10.0, 8.0), nsrc=2)
    for c, srcid in ((big, 1), (small, 0)):
        print(f"noalloc call: grid {c['vp'].shape} source {srcid}", flush=True)
        rc0, U0, n0 = _call(L.pfast_fsm3d, c, srcid)
        rc1, U1, n1 = _call(L.pfast_fsm3d_noalloc, c, srcid)
        assert rc0 == 0 and rc1 == 0, L.pfast_last_error()
        assert n1 == n0
        assert np.array_equal(U1.view(np.int64), U0.view(np.int64))


def test_noalloc_refuses_grids_above_its_scratch():
    from pfast_amd import _lib
    L = _lib.lib()
    vp = np.full(2, 2000.0)  # never read: the size check comes first
    U = np.empty(2)
    src = np.zeros(3, np.int32)
    rc = L.pfast_fsm3d_noalloc(vp, U, src, 0, 100, 25.0, 25.0, 25.0, 1000, 1000, 6, 0, None)
    assert rc == _lib.PFAST_ERR_ARG
    assert b"too large" in L.pfast_last_error()
