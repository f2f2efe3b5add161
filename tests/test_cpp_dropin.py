"""The reference's C++ API as a drop-in: a caller written against the ocu
headers (tests/cpp/mg_dropin.cpp) compiles and links against libocu.so +
libocu_hip.so unchanged, and on the GPU solves its Poisson problem."""
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


def test_dropin_caller_compiles_and_links(tmp_path):
    assert build(tmp_path, "mg_dropin").exists()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [32, 64])
def test_dropin_caller_solves_on_the_gpu(gpu, tmp_path, n):
    exe = build(tmp_path, "mg_dropin")
    p = subprocess.run([str(exe), str(n)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ok=1" in p.stdout
