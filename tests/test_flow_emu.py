"""CPU check of the dataflow schedule of the persistent 3D sweep kernels.

tests/emu/flow_schedule.cpp replays the claim/wait/publish protocol of fwd3d_flow /
adj3d_flow (pfast_amd/csrc/batch3d.hip) with the task numbering and dependency rules
of pfast_amd/csrc/flow.h, as a seeded discrete-event simulation in which resident
blocks advance in random order and every load and store lands at a random point
between its issue and the block's next wait.  Each node carries the number of the
last ordering that processed it, and every read must see what the lexicographic
sweep (src/T_pfsm3d.h:152-179, src/T_padjoint3d.h:135-163) shows at that point:
the current ordering's value behind, the previous ordering's ahead.  The schedule
must also never deadlock, whatever the number of resident blocks.  Three planted
protocol mistakes (a smaller margin, no previous-ordering dependencies, progress
published before its stores complete) must be caught.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
SRC = HERE / "emu" / "flow_schedule.cpp"

# nrows ncols nz TI TJ L nchunk nblocks launches
CONFIGS = [
    (19, 21, 13, 8, 8, 4, 3, 16, 2),   # forward default (8x8, batch 4), partial edge tiles
    (19, 21, 13, 8, 8, 2, 2, 7, 2),    # adjoint default (batch 2)
    (19, 21, 13, 8, 8, 4, 3, 1, 1),    # a single resident block
    (19, 21, 13, 8, 8, 4, 1, 64, 2),   # more blocks than tasks on a diagonal
    (33, 17, 21, 8, 16, 6, 2, 24, 1),
    (17, 33, 7, 16, 8, 6, 2, 24, 1),
    (9, 10, 5, 4, 4, 4, 3, 12, 2),
    (21, 45, 11, 4, 16, 6, 2, 30, 1),
    (1, 20, 6, 8, 8, 4, 2, 8, 1),      # one row
    (20, 1, 6, 8, 8, 2, 2, 8, 1),      # one column
    (17, 9, 30, 8, 8, 8, 1, 20, 1),    # deep columns
]


def _build(tmp, define=None):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp / ("flow_" + (define or "ok"))
    cmd = [gxx, "-O2", "-std=c++17", str(SRC), "-o", str(exe)]
    if define:
        cmd.insert(1, "-D" + define)
    subprocess.run(cmd, check=True)
    return exe


@pytest.fixture(scope="module")
def tmp(tmp_path_factory):
    return tmp_path_factory.mktemp("flow")


@pytest.mark.parametrize("cfg", CONFIGS)
def test_flow_schedule_reads_what_the_lexicographic_sweep_reads(tmp, cfg):
    exe = _build(tmp) if not (tmp / "flow_ok").exists() else tmp / "flow_ok"
    for seed in range(1, 6):
        r = subprocess.run([str(exe), *map(str, cfg), str(seed)], capture_output=True, text=True)
        assert r.returncode == 0, (cfg, seed, r.stdout, r.stderr[-500:])
        assert "mismatches 0 deadlock 0" in r.stdout


@pytest.mark.parametrize("plant", ["PLANT_MARGIN", "PLANT_NO_PREV", "PLANT_EARLY_PUBLISH"])
def test_flow_schedule_catches_planted_mistakes(tmp, plant):
    exe = _build(tmp, plant)
    caught = 0
    for seed in range(1, 4):
        for cfg in (CONFIGS[0], CONFIGS[1]):
            r = subprocess.run([str(exe), *map(str, cfg), str(seed)], capture_output=True,
                               text=True)
            caught += r.returncode != 0
    assert caught == 6, plant
