"""Seeded synthetic inputs for the benchmark and test workloads.

Geometry follows the reference's conventions: node i of an axis sits at coordinate
(i+1)*h and a coordinate c maps to node ceil(c/h)-1 (src/T_PFAST3D.cpp:250-252,
src/RT_PFAST2D.cpp:300-301, 336-337), so generated coordinates lie in (0, n*h].
3D fields are (nz, nrows, ncols) with rows = y, cols = x; 2D fields are
(nrows, ncols) with rows = depth z, cols = x.
"""
import math

import numpy as np

SEED = 20261203


def map_index(coord, h):
    """The reference's coordinate-to-node rule: (int)ceil(coord/h) - 1."""
    return (np.ceil(np.asarray(coord, np.float64) / h) - 1).astype(np.int32)


def vz_traveltime(r, z_s, z_r, v0, g):
    """Analytic traveltime in v(z) = v0 + g z between points at horizontal offset r
    and depths z_s, z_r: T = acosh(1 + g^2 (r^2 + dz^2) / (2 v_s v_r)) / g."""
    vs, vr = v0 + g * z_s, v0 + g * z_r
    d2 = r ** 2 + (z_r - z_s) ** 2
    return np.arccosh(1.0 + g * g * d2 / (2.0 * vs * vr)) / g


# ------------------------------------------------------------------------ 3D
def model3d(nrows=201, ncols=201, nz=101, h=25.0, v0=2000.0, grad=0.6, checker=0.05, cell=500.0):
    """Vp = v0 + grad*z times (1 +- checker) on cells of `cell` metres (x, y, z)."""
    k = np.arange(nz)[:, None, None]
    i = np.arange(nrows)[None, :, None]
    j = np.arange(ncols)[None, None, :]
    z, y, x = k * h, i * h, j * h
    vp = v0 + grad * z + 0.0 * (x + y)
    if checker:
        sgn = (-1.0) ** (np.floor(x / cell) + np.floor(y / cell) + np.floor(z / cell))
        vp = vp * (1.0 + checker * sgn)
    return np.ascontiguousarray(vp, np.float64)


def survey3d(nsrc=400, nrec=2000, nrows=201, ncols=201, h=25.0, v0=2000.0, grad=0.6,
             noise=0.002, seed=SEED):
    """Sources and receivers uniform at random on the surface (z = h, node k = 0);
    every source records every receiver; picks from the analytic traveltime of
    the unperturbed v(z) model plus N(0, noise) seconds.  Returns the reference's
    .src/.rec tables (coordinates) and the mapped index tables."""
    rng = np.random.default_rng(seed)
    Ly, Lx = nrows * h, ncols * h
    sy = Ly * (1.0 - rng.random(nsrc))  # (0, L]
    sx = Lx * (1.0 - rng.random(nsrc))
    ry = Ly * (1.0 - rng.random(nrec))
    rx = Lx * (1.0 - rng.random(nrec))
    src_tab = np.stack([sy, sx, np.full(nsrc, h), np.full(nsrc, float(nrec))], 1)
    src = np.stack([map_index(sy, h), map_index(sx, h), map_index(np.full(nsrc, h), h)], 1)
    ri, rj = map_index(ry, h), map_index(rx, h)
    srnum = np.full(nsrc, nrec, np.int32)
    rec = np.tile(np.stack([ri, rj, np.zeros(nrec, np.int32)], 1), (nsrc, 1)).astype(np.int32)
    # analytic picks between node positions (depth 0 for both ends)
    dy = (src[:, 0][:, None] - ri[None, :]) * h
    dx = (src[:, 1][:, None] - rj[None, :]) * h
    T = vz_traveltime(np.hypot(dx, dy), 0.0, 0.0, v0, grad)
    T_obs = (T + noise * rng.standard_normal(T.shape)).ravel()
    nxyz = np.tile([0.0, 0.0, -1.0], (nsrc * nrec, 1))
    rec_tab = np.concatenate([np.tile(np.stack([ry, rx, np.full(nrec, h)], 1), (nsrc, 1)), nxyz,
                              T_obs[:, None]], 1)
    return dict(src=src.astype(np.int32), srnum=srnum, rec=rec, nxyz=nxyz, T_obs=T_obs,
                src_tab=src_tab, rec_tab=rec_tab)


def config3d(nsrc=400, nrec=2000, seed=SEED, **kw):
    """The 3D diving-wave tomography workload: 201x201x101, h = 25 m,
    Vp = 2000 + 0.6 z with a +-5% checkerboard of 500 m cells."""
    vp = model3d(**kw)
    nz, nrows, ncols = vp.shape
    sv = survey3d(nsrc=nsrc, nrec=nrec, nrows=nrows, ncols=ncols, h=kw.get("h", 25.0), seed=seed)
    sv.update(vp=vp, h=kw.get("h", 25.0))
    return sv


# ------------------------------------------------------------------------ 2D
def model2d(nrows=101, ncols=401, h=10.0, v0=1500.0, grad=0.8, anomalies=3, amp=0.10,
            sigma=80.0, seed=SEED):
    """Vp = v0 + grad*z plus `anomalies` seeded Gaussian anomalies (+-amp, sigma m)."""
    rng = np.random.default_rng(seed)
    z = (np.arange(nrows) * h)[:, None]
    x = (np.arange(ncols) * h)[None, :]
    base = v0 + grad * z + 0.0 * x
    pert = np.zeros_like(base)
    for a in range(anomalies):
        cx = rng.uniform(0.15, 0.85) * ncols * h
        cz = rng.uniform(0.3, 0.8) * nrows * h
        sgn = 1.0 if a % 2 == 0 else -1.0
        pert += sgn * amp * np.exp(-((x - cx) ** 2 + (z - cz) ** 2) / (2 * sigma ** 2))
    return np.ascontiguousarray(base * (1.0 + pert)), np.ascontiguousarray(base)


def survey2d(nsrc, nrec, nrows, ncols, h, NumofR=0):
    """Evenly spaced surface sources and receivers (depth h -> row 0)."""
    L = ncols * h
    sx = L * (np.arange(nsrc) + 0.5) / nsrc
    rx = L * (np.arange(nrec) + 0.5) / nrec
    src_tab = np.stack([np.full(nsrc, h), sx, np.full(nsrc, float(nrec))], 1)
    src = np.stack([map_index(src_tab[:, 0], h), map_index(sx, h)], 1).astype(np.int32)
    ri = np.stack([map_index(np.full(nrec, h), h), map_index(rx, h)], 1)
    rec = np.tile(ri, (nsrc, 1)).astype(np.int32)
    nxy = np.tile([-1.0, 0.0], (nsrc * nrec, 1))  # (n_z, n_x), pointing up
    srnum = np.full(nsrc, nrec, np.int32)
    return dict(src=src, srnum=srnum, rec=rec, nxy=nxy, src_tab=src_tab,
                rec_xz=np.tile(np.stack([np.full(nrec, h), rx], 1), (nsrc, 1)))


def reflector2d(ncols, h, depth=900.0, amp=100.0, seed=SEED):
    """One smooth interface z(x) = depth + sum of 3 seeded sinusoids (|dz| <= amp),
    as the reference's .ref table (depth, n_z, n_x) per column, normal pointing up
    toward the sources (mfiles/RT_SynMod2D.m:42-44)."""
    rng = np.random.default_rng(seed + 7)
    x = (np.arange(ncols) + 1) * h
    L = ncols * h
    w = rng.dirichlet(np.ones(3)) * amp
    zz = np.full(ncols, depth)
    dzdx = np.zeros(ncols)
    for n, wk in enumerate(w):
        kx = 2 * math.pi * (n + 1) / L
        ph = rng.uniform(0, 2 * math.pi)
        zz += wk * np.sin(kx * x + ph)
        dzdx += wk * kx * np.cos(kx * x + ph)
    nrm = np.hypot(1.0, dzdx)
    Z = np.stack([zz, -1.0 / nrm, dzdx / nrm], 1)
    return Z[None, :, :]


def config2d(kind=1, seed=SEED):
    """kind 1: forward, 101x401 h=10, 20 sources / 200 receivers.
    kind 2: tomography, same grid, 60 sources / 200 receivers.
    kind 3: joint T+R, 151x601, 80 sources / 300 receivers, one reflector."""
    if kind in (1, 2):
        nrows, ncols, nsrc, nrec = 101, 401, (20 if kind == 1 else 60), 200
    elif kind == 3:
        nrows, ncols, nsrc, nrec = 151, 601, 80, 300
    else:
        raise ValueError(kind)
    h = 10.0
    vp_true, vp0 = model2d(nrows, ncols, h, seed=seed)
    sv = survey2d(nsrc, nrec, nrows, ncols, h)
    sv.update(vp=vp_true, vp0=vp0, h=h, nrows=nrows, ncols=ncols)
    if kind == 3:
        sv["Z"] = reflector2d(ncols, h, seed=seed)
        sv["NumofR"] = 1
    else:
        sv["Z"] = np.zeros((0, ncols, 3))
        sv["NumofR"] = 0
    return sv
