"""CPU check of the batched 3D sweep schedule (pfast_amd/csrc/batch3d.hip).

tests/emu/sweep_schedule.cpp replays the kernels' column-window schedule on the CPU
(tile anti-diagonal launches, per-thread level skew, register windows,
parity-buffered exchange of levels m-1 / m+1, halo and INF-slot neighbours,
batched loads and delayed stores) and compares every one of the 8 orderings,
bit for bit, with the plain lexicographic sweep in the reference's loop order
(src/T_pfsm3d.h:131-185 forward, src/T_padjoint3d.h:73-172 adjoint) using the
same per-node arithmetic.  Partial edge tiles, nz not a multiple of the
prefetch depth, sources on faces and fixed receiver nodes are covered.
"""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

HERE = Path(__file__).resolve().parent
SRC = HERE / "emu" / "sweep_schedule.cpp"
INF = np.float32(9999999.0)


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("emu") / "sweep_schedule"
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", str(SRC), "-o", str(exe)],
                   check=True)
    return exe


def _run(exe, tmp, mode, shape, tile, T, U0, src):
    nz, nr, nc = shape
    f = tmp / f"in_{mode}_{nz}_{nr}_{nc}_{tile[0]}_{tile[1]}_{tile[2]}.bin"
    with open(f, "wb") as fh:
        fh.write(np.ascontiguousarray(T, np.float32).tobytes())
        fh.write(np.ascontiguousarray(U0, np.float32).tobytes())
        fh.write(np.int32(src).tobytes())
    r = subprocess.run([str(exe), mode, str(nr), str(nc), str(nz), *map(str, tile), str(f)],
                       capture_output=True, text=True)
    return r.returncode, r.stdout


# the batch depth L is even (static step parity inside a batch)
CASES = [((11, 19, 13), (8, 8, 4)), ((9, 8, 8), (8, 8, 2)), ((13, 21, 17), (4, 4, 4)),
              ((7, 5, 23), (8, 4, 6)), ((12, 16, 9), (8, 8, 4)), ((3, 17, 9), (8, 8, 4)),
              ((10, 14, 11), (4, 4, 6)), ((8, 13, 19), (4, 8, 6)), ((9, 21, 35), (8, 16, 6)),
              ((7, 37, 18), (16, 8, 4)),
              # the default adjoint batch depth with partial edge tiles and depths that are
              # not a multiple of two batches
              ((14, 23, 19), (8, 8, 2)), ((5, 17, 26), (8, 8, 2)), ((21, 9, 30), (8, 8, 2))]


@pytest.mark.parametrize("shape,tile", CASES)
@pytest.mark.parametrize("where", ["interior", "face"])
def test_forward_schedule_equals_lexicographic(emu, tmp_path, shape, tile, where):
    nz, nr, nc = shape
    rng = np.random.default_rng(nz * 1000 + nr * 10 + nc)
    ftime = (25.0 / rng.uniform(1500, 6000, size=shape)).astype(np.float32)
    src = (nz // 2, nr // 3, nc - 2) if where == "interior" else (0, nr - 1, 0)
    sn = int(np.ravel_multi_index(src, shape))
    U0 = np.full(shape, INF, np.float32)
    U0[src] = 0.0
    rc, out = _run(emu, tmp_path, "fwd", shape, tile, ftime, U0, sn)
    assert rc == 0, out
    assert "mismatches 0" in out


@pytest.mark.parametrize("shape,tile", CASES)
def test_adjoint_schedule_equals_lexicographic(emu, tmp_path, shape, tile):
    nz, nr, nc = shape
    rng = np.random.default_rng(7 + nz + nr + nc)
    k, i, j = np.meshgrid(np.arange(nz), np.arange(nr), np.arange(nc), indexing="ij")
    sk, si, sj = 1, nr // 2, nc // 2
    T = (np.sqrt((k - sk) ** 2 + (i - si) ** 2 + (j - sj) ** 2) * 25.0 / 3000.0).astype(np.float32)
    T *= rng.uniform(0.95, 1.05, size=shape).astype(np.float32)
    L0 = np.zeros(shape, np.float32)
    # receivers: fixed nodes (sign bit set) with a boundary value
    nrec = max(3, nr * nc // 10)
    rk = rng.integers(1, nz - 1, nrec)
    ri = rng.integers(1, nr - 1, nrec)
    rj = rng.integers(1, nc - 1, nrec)
    T[rk, ri, rj] = -T[rk, ri, rj]
    L0[rk, ri, rj] = rng.normal(0, 50, nrec).astype(np.float32)
    rc, out = _run(emu, tmp_path, "adj", shape, tile, T, L0, -1)
    assert rc == 0, out
    assert "mismatches 0" in out


@pytest.mark.parametrize("plant,mode", [("PLANT_NO_SKEW", "fwd"), ("PLANT_NO_SKEW", "adj"),
                                        ("PLANT_RING_DEPTH", "adj"),
                                        ("PLANT_SWAP_EXCHANGE", "fwd")])
def test_planted_schedule_mistake_is_caught(tmp_path, plant, mode):
    """The comparison is not vacuous: the emulator built with a deliberate
    schedule mistake (all columns on one level; z-neighbours named by window order
    instead of absolute depth; upstream and downstream exchange slots swapped)
    reports mismatches."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "planted"
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", f"-D{plant}", str(SRC), "-o",
                    str(exe)], check=True)
    shape, tile = (11, 19, 13), (8, 8, 4)
    nz, nr, nc = shape
    rng = np.random.default_rng(3)
    if mode == "fwd":
        T = (25.0 / rng.uniform(1500, 6000, size=shape)).astype(np.float32)
        U0 = np.full(shape, INF, np.float32)
        U0[5, 6, 7] = 0
        src = int(np.ravel_multi_index((5, 6, 7), shape))
    else:
        k, i, j = np.meshgrid(np.arange(nz), np.arange(nr), np.arange(nc), indexing="ij")
        T = (np.sqrt((k - 1.0) ** 2 + (i - 9.0) ** 2 + (j - 6.0) ** 2) * 25 / 3000).astype(np.float32)
        T *= rng.uniform(0.9, 1.1, size=shape).astype(np.float32)
        U0 = np.zeros(shape, np.float32)
        rk, ri, rj = rng.integers(1, nz - 1, 30), rng.integers(1, nr - 1, 30), rng.integers(1, nc - 1, 30)
        T[rk, ri, rj] = -T[rk, ri, rj]
        U0[rk, ri, rj] = rng.normal(0, 50, 30).astype(np.float32)
        src = -1
    rc, out = _run(exe, tmp_path, mode, shape, tile, T, U0, src)
    assert rc == 1 and "mismatches 0" not in out, out
