"""CPU check of the per-diagonal launch plan of the batched 3D sweeps
(pfast_amd/csrc/plan.h): paired launches, in which ordering 2p+1 (the column
direction flipped) trails ordering 2p by pair_lag tile anti-diagonals in shared
launches, and the unpaired one-ordering launches.

tests/emu/pair_schedule.cpp includes plan.h as it is and walks a whole cycle: every
(ordering, tile) is swept once, no two tasks of a launch touch each other's tiles,
and launch by launch, with each launch's tasks in a random order and each tile swept
in its ordering's lexicographic order, the plan gives the bits of the plain
lexicographic sweeps of the 8 orderings in the reference's loop order
(src/T_pfsm3d.h:158-160, src/T_padjoint3d.h:139-141).  A lag one diagonal short must
be caught.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
SRC = HERE / "emu" / "pair_schedule.cpp"

# nrows ncols nz TI TJ
GRIDS = [(19, 21, 5, 8, 8), (9, 30, 4, 4, 4), (1, 20, 3, 8, 8), (20, 1, 3, 8, 8),
         (5, 5, 3, 8, 8), (33, 17, 4, 8, 16), (17, 33, 4, 16, 8), (23, 23, 3, 8, 4),
         (12, 40, 2, 4, 8), (41, 9, 2, 8, 4)]


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("pair") / "pair_schedule"
    subprocess.run([gxx, "-O2", "-std=c++17", str(SRC), "-o", str(exe)], check=True)
    return exe


def _run(exe, grid, paired, seed, lag_delta=0):
    r = subprocess.run([str(exe), *map(str, grid), str(int(paired)), str(seed), str(lag_delta)],
                       capture_output=True, text=True)
    return r.returncode, r.stdout


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("paired", [True, False])
def test_plan_equals_lexicographic_cycle(emu, grid, paired):
    for seed in (1, 2, 3):
        rc, out = _run(emu, grid, paired, seed)
        assert rc == 0 and "conflicts 0 mismatches 0" in out, (grid, paired, seed, out)


def test_paired_plan_saves_launches(emu):
    """201x201 columns in 8x8 tiles: 26 + 2 * 26 = 78 launches per pair of orderings
    against 2 * 51 = 102."""
    rc, on = _run(emu, (201, 201, 1, 8, 8), True, 1)
    rc2, off = _run(emu, (201, 201, 1, 8, 8), False, 1)
    assert rc == 0 and rc2 == 0
    assert on.split()[1] == str(4 * 78) and off.split()[1] == str(8 * 51)


@pytest.mark.parametrize("grid", [g for g in GRIDS if (g[0] + g[3] - 1) // g[3] > 1
                                  or (g[1] + g[4] - 1) // g[4] > 1])
def test_short_lag_is_caught(emu, grid):
    """The check is not vacuous: one diagonal less lag puts neighbouring tiles of the
    two orderings in one launch and reads values the sequential sweep has not yet
    written."""
    rc, out = _run(emu, grid, True, 1, -1)
    assert rc == 1 and "conflicts 0 mismatches 0" not in out, out
