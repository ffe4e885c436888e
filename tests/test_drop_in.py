"""A reference-shaped caller of the drop-in host class.

tests/drop_in_caller.cpp follows the reference's main.cpp:78-97: it sets only
the nine fields of the reference's SimulationParameters and constructs
BarnesHutParticleSystem(parameters, n).  The struct must keep the
reference's layout, and the context must get the reference's step
configuration (theta 0.5, eps2 0.025) with the native walk group 64."""
import subprocess
from pathlib import Path

import pytest

from tests.conftest import gpu_available

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "stellarsolver_amd"
LIB = PKG / "libstellar_hip.so"


@pytest.fixture(scope="module")
def caller(tmp_path_factory):
    if not LIB.exists():
        pytest.skip("libstellar_hip.so not built")
    out = tmp_path_factory.mktemp("caller") / "drop_in_caller"
    cmd = ["g++", "-O1", "-std=c++17", "-Wall", "-Werror", f"-I{ROOT / 'include'}",
           f"-I{PKG / 'host'}", str(ROOT / "tests" / "drop_in_caller.cpp"),
           str(PKG / "host" / "BarnesHutParticleSystem.cpp"), "-o", str(out), f"-L{PKG}",
           "-lstellar_hip", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,/opt/rocm/lib"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    return out


def _fields(stdout, tag):
    for ln in stdout.splitlines():
        if ln.startswith(tag + " "):
            return ln.split()[1:]
    raise AssertionError(f"no '{tag}' line in: {stdout}")


def test_reference_struct_layout_and_defaults(caller):
    p = subprocess.run([str(caller)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stdout + p.stderr
    # the reference's struct: enum, four bools, four floats
    assert _fields(p.stdout, "layout") == ["24", "7", "8", "20"]
    theta, eps2, group, flags, dt = _fields(p.stdout, "params")
    assert float(theta) == 0.5 and float(eps2) == pytest.approx(0.025, rel=1e-7)
    assert int(group) == 64 and int(flags) == 0 and float(dt) == pytest.approx(0.001, rel=1e-7)


@pytest.mark.gpu
def test_reference_caller_steps(caller):
    if not gpu_available():
        pytest.skip("no GPU")
    p = subprocess.run([str(caller), "--step"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    theta, eps2, group, flags = _fields(p.stdout, "context")
    assert float(theta) == 0.5 and float(eps2) == pytest.approx(0.025, rel=1e-7)
    assert int(group) == 64 and int(flags) == 0
    n, x0, y0 = _fields(p.stdout, "output")
    assert int(n) == 4096 and float(x0) == float(x0) and float(y0) == float(y0)
