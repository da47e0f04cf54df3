"""The drop-in boundary: the reference's unmodified drivers (src/T_PFAST3D.cpp,
src/RT_PFAST2D.cpp) compile and link against include/dropin/*.h and libpfast with the
command INTEGRATION.md gives (the drop-ins pre-included, the reference headers' include
guards defined so they compile to nothing), the executables import their solver calls
from libpfast, and they start and print the drop-ins' banner.

Runs in the build container only: it reads the reference's sources where they lie and
needs the MPI headers and libraries of /opt/conda; it is skipped elsewhere."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"
MPI = "/opt/conda"
LIB = os.path.join(ROOT, "pfast_amd", "libpfast.so")

CASES = {
    "T_PFAST3D": (["-D_FIM3D_FUNCTIONS", "-D_ADJOINT3D_METHODS"], ["T_pfsm3d.h", "T_padjoint3d.h"],
                  ["pfast_fsm3d", "pfast_adjoint_fsm3d", "pfast_regu3d"]),
    "RT_PFAST2D": (["-D_FIM2D_FUNCTIONS", "-D_ADJOINT2D_METHODS"], ["RT_pfsm2d.h", "RT_padjoint2d.h"],
                   ["pfast_rt_fsm2d", "pfast_tadjoint_fsm2d", "pfast_radjoint_fsm2d", "pfast_regu2d"]),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_reference_driver_builds_against_dropins(tmp_path, name):
    gxx = shutil.which("g++")
    if not (gxx and os.path.exists(os.path.join(REF, name + ".cpp")) and
            os.path.exists(os.path.join(MPI, "include", "mpi.h"))):
        pytest.skip("reference sources or MPI not present (build container only)")
    if not os.path.exists(LIB):
        pytest.skip("libpfast not built")
    defs, heads, syms = CASES[name]
    exe = tmp_path / name
    cmd = [gxx, "-O2", "-w", *defs]
    for h in heads:
        cmd += ["-include", os.path.join(ROOT, "include", "dropin", h)]
    cmd += ["-I" + os.path.join(MPI, "include"), os.path.join(REF, name + ".cpp"),
            "-L" + os.path.dirname(LIB), "-lpfast", "-Wl,-rpath," + os.path.dirname(LIB),
            os.path.join(MPI, "lib", "libmpicxx.so"), os.path.join(MPI, "lib", "libmpi.so"),
            "-Wl,-rpath-link," + os.path.join(MPI, "lib"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    nm = subprocess.run(["nm", "-D", str(exe)], capture_output=True, text=True).stdout
    for s in syms:
        assert f" U {s}" in nm, s
    # MPI's libraries without conda's older libstdc++ beside them
    libdir = tmp_path / "mpilib"
    libdir.mkdir()
    for f in ("libmpi.so.12", "libmpicxx.so.12", "libgfortran.so.4", "libquadmath.so.0"):
        if os.path.exists(os.path.join(MPI, "lib", f)):
            os.symlink(os.path.join(MPI, "lib", f), libdir / f)
    env = dict(os.environ, LD_LIBRARY_PATH=str(libdir))
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=60)
    out = r.stdout + r.stderr
    assert "GPU engine libpfast" in out and "usage" in out, out[-2000:]
