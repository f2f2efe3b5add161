"""The reference's C++ API as a drop-in: callers written against the ocu
headers (tests/cpp/mg_dropin.cpp, tests/cpp/advection_dropin.cpp,
tests/cpp/pcg_sample_dropin.cpp) compile
and link against libocu.so + libocu_hip.so unchanged, and on the GPU solve
their Poisson problem and pass the reference advection tests' symmetry,
conservation and band checks."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "opencurrent_amd" / "lib"


def build(tmp_path, name):
    exe = tmp_path / name
    cmd = ["g++", "-O2", "-std=c++17", "-I", str(ROOT / "include"), "-I",
           str(ROOT / "opencurrent_amd" / "csrc" / "host" / "include"), str(ROOT / "tests" / "cpp" / f"{name}.cpp"),
           "-L", str(LIB), "-locu", "-locu_hip", f"-Wl,-rpath,{LIB}", "-o", str(exe)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    return exe


@pytest.mark.parametrize("name", ["mg_dropin", "advection_dropin", "pcg_sample_dropin"])
def test_dropin_caller_compiles_and_links(tmp_path, name):
    assert build(tmp_path, name).exists()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [32, 64])
def test_dropin_caller_solves_on_the_gpu(gpu, tmp_path, n):
    exe = build(tmp_path, "mg_dropin")
    p = subprocess.run([str(exe), str(n)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ok=1" in p.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", [1, 2, 4])
def test_dropin_scalar_advection_matches_reference_checks(gpu, tmp_path, scheme):
    """advection3ddoubletest.cpp's symmetry test (six flow directions agree
    to 1e-12 and conserve the sum) and band test (exact zeros outside the
    band), through Eqn_ScalarAdvection3DD, for the reference's two schemes
    and second-order upwind"""
    exe = build(tmp_path, "advection_dropin")
    p = subprocess.run([str(exe), str(scheme)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ok=0" not in p.stdout


@pytest.mark.gpu
def test_dropin_pcg_and_sampling(gpu, tmp_path):
    """Sol_PCGPressure3DDeviceD (multigrid preconditioner) returns the
    eigenmode behind its rhs; a two-layer Neumann fill of a linear field
    followed by sample_points_3d returns the field, also within half a cell
    of the x faces."""
    exe = build(tmp_path, "pcg_sample_dropin")
    p = subprocess.run([str(exe), "32"], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.count("ok=1") == 2, p.stdout
