"""Seeded random test cases (small grids the FP64 oracle solves in milliseconds)."""
import numpy as np


def case3d(seed, shape=None, spacing=None, nsrc=3, nrec_per=6, interior_rec=True):
    rng = np.random.default_rng(seed)
    nz, nr, nc = shape if shape is not None else tuple(int(x) for x in rng.integers(5, 13, 3))
    if spacing is None:
        spacing = tuple(float(x) for x in rng.choice([5.0, 10.0, 12.5, 25.0], 3))
    dx, dy, dz = spacing
    z = np.arange(nz)[:, None, None] * dz
    vp = (1800 + 0.8 * z) * (1 + 0.15 * rng.standard_normal((nz, nr, nc)).clip(-2, 2))
    src = np.stack([rng.integers(0, nr, nsrc), rng.integers(0, nc, nsrc),
                    rng.integers(0, nz, nsrc)], 1).astype(np.int32)
    srnum = np.full(nsrc, nrec_per, np.int32)
    nrec = int(srnum.sum())
    kmax = nz if interior_rec else 1
    rec = np.stack([rng.integers(0, nr, nrec), rng.integers(0, nc, nrec),
                    rng.integers(0, kmax, nrec)], 1).astype(np.int32)
    nxyz = rng.standard_normal((nrec, 3))
    nxyz /= np.linalg.norm(nxyz, axis=1, keepdims=True)
    T_obs = rng.random(nrec) * 0.4
    T_obs[rng.random(nrec) < 0.1] = 150.0  # unpicked
    return dict(vp=vp, src=src, srnum=srnum, rec=rec, nxyz=nxyz, T_obs=T_obs,
                dx=dx, dy=dy, dz=dz)


def case2d(seed, shape=None, NumofR=2, nsrc=2, nrec_per=6, dx=12.5, dz=8.0):
    rng = np.random.default_rng(seed)
    nr, nc = shape if shape is not None else (int(rng.integers(14, 30)), int(rng.integers(14, 40)))
    vp = 1500 + 1500 * rng.random((nr, nc))
    x = np.arange(nc)
    Z = np.zeros((NumofR, nc, 3))
    for ir in range(NumofR):
        depth = (0.45 + 0.25 * ir) * nr * dz + 3 * dz * np.sin(2 * np.pi * x / nc * (ir + 1) + rng.random())
        Z[ir, :, 0] = depth
        Z[ir, :, 1] = -1.0 + 0.05 * rng.random(nc)
        Z[ir, :, 2] = 0.1 * rng.standard_normal(nc)
    src = np.stack([rng.integers(0, 3, nsrc), rng.integers(0, nc, nsrc)], 1).astype(np.int32)
    srnum = np.full(nsrc, nrec_per, np.int32)
    nrec = int(srnum.sum())
    rec = np.stack([rng.integers(0, 2, nrec), rng.integers(0, nc, nrec)], 1).astype(np.int32)
    nxy = np.stack([-np.ones(nrec), 0.1 * rng.standard_normal(nrec)], 1)
    T_obs = rng.random(nrec) * 0.3
    T_obs[rng.random(nrec) < 0.1] = 150.0
    TR_obs = rng.random((nrec, NumofR)) * 0.5
    TR_obs[rng.random((nrec, NumofR)) < 0.1] = 0.0
    return dict(vp=vp, Z=Z, src=src, srnum=srnum, rec=rec, nxy=nxy, T_obs=T_obs, TR_obs=TR_obs,
                NumofR=NumofR, dx=dx, dz=dz)


def bits(a):
    return np.ascontiguousarray(a, np.float64).ravel().view(np.int64)
