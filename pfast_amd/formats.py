"""File formats of the PFAST drivers: parameter files (.inp) and the raw FP64
binary tables they point to.

All binary files are headerless little-endian doubles, row-major (the
reference's readmod / writemod, src/T_pfsm3d.h:39-57):
  .vp   model, 2D [nrows][ncols], 3D [nz][nrows][ncols]
  .src  2D: rows (z, x, nchan); 3D: rows (y, x, z, nchan)            coordinates in m
  .rec  2D: rows (z, x, n_z, n_x, TT, TR_1..TR_R); 3D: rows (y, x, z, n_y, n_x, n_z, T)
  .ref  2D reflectors: NumofR*ncols rows (depth, n_z, n_x)
  .wr   2D per-node reflection weights [nrows][ncols]
Node i sits at coordinate (i+1) h; the drivers map a coordinate to
ceil(c/h) - 1 (src/T_PFAST3D.cpp:250-252, src/RT_PFAST2D.cpp:299-302).

Parameter files: every line not starting with '#' is the next parameter line
(src/RT_PFAST2D.cpp:140-177 for 2D, src/T_PFAST3D.cpp:105-136 for 3D).
"""
import math
import os

import numpy as np


def write_bin(path, a):
    np.ascontiguousarray(a, dtype="<f8").tofile(path)


def read_bin(path, shape):
    a = np.fromfile(path, dtype="<f8")
    n = int(np.prod(shape))
    if a.size < n:
        raise ValueError(f"{path}: {a.size} doubles, expected {n}")
    return a[:n].reshape(shape)


def map_index(c, h):
    """Coordinate (m) -> node index, FP64 ceil(c/h) - 1."""
    return np.asarray(np.ceil(np.asarray(c, np.float64) / h) - 1, dtype=np.int64)


def read_inp_lines(path):
    with open(path) as fh:
        return [ln for ln in fh if not ln.startswith("#")]


INP2D_FIELDS = ["vp0file", "srcfile", "src_shape", "recfile", "rec_shape", "refgeom", "ref_shape",
                "grid", "vpfile", "nu", "twfile", "scheme"]
INP3D_FIELDS = ["vp0file", "srcfile", "src_shape", "recfile", "rec_shape", "grid", "vpfile", "nu",
                "scheme"]


def write_inp2d(path, vp0file, srcfile, src_shape, recfile, rec_shape, refgeom, ref_shape,
                nrows, ncols, dx, dz, vpfile="final.vp", nux=60.0, nuz=20.0, twfile="w.wr",
                schemetag=3, lsrc=2):
    """A 2D parameter file in the layout of the reference's par/RT_SynMod2D_Inv.inp."""
    lines = ["#The Input file for 2D travel time tomography (pfast_rt2d)",
             "#The file name of the initial velocity", vp0file,
             "#The file name of the source locations", srcfile, f"{src_shape[0]} {src_shape[1]}",
             "#The file name of the receiver locations for Tran and Reflection", recfile,
             f"{rec_shape[0]} {rec_shape[1]}",
             "#The file name of the reflector locations", refgeom, f"{ref_shape[0]} {ref_shape[1]}",
             "#Size of the model, rows, columns, sample interval",
             f"{nrows:f} {ncols:f} {dx:f} {dz:f}",
             "#===Output the inversed model===", vpfile,
             "#===Regularization Parameter: nux, nuz", f"{nux:g} {nuz:g}",
             "#===Weighting factor of Reflection (WR)", twfile,
             "#===Inversion Scheme tag (schemetag) and Line search tag (lsrc)", f"{schemetag} {lsrc}"]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def write_inp3d(path, vp0file, srcfile, src_shape, recfile, rec_shape, nrows, ncols, nz, dx, dy,
                dz, vpfile="final.vp", nux=10.0, nuy=10.0, nuz=10.0, schemetag=3, lsrc=2):
    """A 3D parameter file in the layout of the reference's par/Field3D.inp."""
    lines = ["#The Input file for 3D travel time tomography (pfast_t3d)",
             "#The file name of the initial velocity", vp0file,
             "#The file name of the source locations", srcfile, f"{src_shape[0]} {src_shape[1]}",
             "#The file name of the receiver locations", recfile, f"{rec_shape[0]} {rec_shape[1]}",
             "#Size of the model, rows, columns, depth, sample intervals",
             f"{nrows} {ncols} {nz} {dx:f} {dy:f} {dz:f}",
             "#===Output the inversed model===", vpfile,
             "#===Regularization Parameters: nux, nuy, nuz", f"{nux:g} {nuy:g} {nuz:g}",
             "#===Inversion Scheme tag and Line search tag", f"{schemetag} {lsrc}"]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def parse_inp2d(path):
    L = read_inp_lines(path)
    t = [ln.split() for ln in L]
    d = os.path.dirname(os.path.abspath(path))
    out = dict(vp0file=t[0][0], srcfile=t[1][0], src_shape=(int(t[2][0]), int(t[2][1])),
               recfile=t[3][0], rec_shape=(int(t[4][0]), int(t[4][1])), refgeom=t[5][0],
               ref_shape=(int(t[6][0]), int(t[6][1])), nrows=int(float(t[7][0])),
               ncols=int(float(t[7][1])), dx=float(t[7][2]), dz=float(t[7][3]), base=d)
    if len(t) >= 12:
        out.update(vpfile=t[8][0], nux=float(t[9][0]), nuz=float(t[9][1]), twfile=t[10][0],
                   schemetag=int(t[11][0]), lsrc=int(t[11][1]))
    out["NumofR"] = int(math.ceil(out["ref_shape"][0] / out["ncols"]))
    return out


def parse_inp3d(path):
    L = read_inp_lines(path)
    t = [ln.split() for ln in L]
    d = os.path.dirname(os.path.abspath(path))
    out = dict(vp0file=t[0][0], srcfile=t[1][0], src_shape=(int(t[2][0]), int(t[2][1])),
               recfile=t[3][0], rec_shape=(int(t[4][0]), int(t[4][1])),
               nrows=int(float(t[5][0])), ncols=int(float(t[5][1])), nz=int(float(t[5][2])),
               dx=float(t[5][3]), dy=float(t[5][4]), dz=float(t[5][5]), base=d)
    if len(t) >= 9:
        out.update(vpfile=t[6][0], nux=float(t[7][0]), nuy=float(t[7][1]), nuz=float(t[7][2]),
                   schemetag=int(t[8][0]), lsrc=int(t[8][1]))
    return out


def load_survey2d(inp_path):
    """Model, node tables and picks of a 2D parameter file, as the driver maps them."""
    p = parse_inp2d(inp_path)
    b = p["base"]
    nr, nc = p["nrows"], p["ncols"]
    vp = read_bin(os.path.join(b, p["vp0file"]), (nr, nc))
    src = read_bin(os.path.join(b, p["srcfile"]), p["src_shape"])
    rec = read_bin(os.path.join(b, p["recfile"]), p["rec_shape"])
    Z = read_bin(os.path.join(b, p["refgeom"]), p["ref_shape"])
    R = p["NumofR"]
    return dict(params=p, vp=vp,
                src=np.stack([map_index(src[:, 0], p["dz"]), map_index(src[:, 1], p["dx"])], 1).astype(np.int32),
                srnum=src[:, 2].astype(np.int32),
                rec=np.stack([map_index(rec[:, 0], p["dz"]), map_index(rec[:, 1], p["dx"])], 1).astype(np.int32),
                nxy=np.ascontiguousarray(rec[:, 2:4]), TT_obs=np.ascontiguousarray(rec[:, 4]),
                TR_obs=np.ascontiguousarray(rec[:, 5:5 + R]), Z=Z.reshape(R, nc, 3))
