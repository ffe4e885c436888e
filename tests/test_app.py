"""The stellar_app drop-in (reference CLI main.cpp:75-103 over
BarnesHutParticleSystem)."""
import os
import re
import subprocess
from pathlib import Path

import pytest

from tests.conftest import gpu_available

APP = Path(__file__).resolve().parents[1] / "stellarsolver_amd" / "stellar_app"


def run(*args, timeout=120):
    env = dict(os.environ)
    return subprocess.run([str(APP), *args], capture_output=True, text=True, timeout=timeout, env=env)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not APP.exists():
        pytest.skip("stellar_app not built")


def test_fails_cleanly_without_gpu():
    """No device: a clear error and a non-zero exit, never a wait on stdin
    (the reference's gpuErrchk calls getchar(), BarnesHutParticleSystem.cpp:20)."""
    if gpu_available():
        pytest.skip("a GPU is present")
    p = run("-iterations=1", "-numbodies=1024", timeout=60)
    assert p.returncode != 0
    assert "stellar_app:" in p.stderr


@pytest.fixture
def gpu():
    if not gpu_available():
        pytest.skip("no GPU")


@pytest.mark.gpu
def test_benchmark_lines(gpu):
    p = run("-benchmark", "-iterations=5", "-numbodies=65536", "-workload=2", "-eps=0.01")
    assert p.returncode == 0, p.stderr
    lines = [ln for ln in p.stdout.splitlines() if ln.startswith("Timestep:")]
    assert len(lines) == 5
    assert all(re.match(r"Timestep: \d+    Elapsed time: [0-9.e+-]+$", ln) for ln in lines)
    assert "Body-steps/s:" in p.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("model", ["-disk", "-plummer", "-colliding-disks"])
def test_reference_models_and_debug_checks(gpu, model):
    p = run(model, "-debug", "-iterations=2")
    assert p.returncode == 0, p.stderr
    res = [ln for ln in p.stdout.splitlines() if "RESULT" in ln]
    assert len(res) == 16  # 8 checks x 2 steps (debug.cpp:6-178)
    assert all("PASS" in ln for ln in res), res
