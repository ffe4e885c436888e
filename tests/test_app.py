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


@pytest.mark.gpu
def test_debug_checks_in_hilbert_order(gpu):
    """-hilbert (bodies stored along a Hilbert curve of the tree) with the
    quadrupole model: the exported tree is the reference's numbering, so the
    reference's own checks pass on it as in slot order."""
    p = run("-colliding-disks", "-debug", "-iterations=2", "-hilbert", "-quad", "-textbook")
    assert p.returncode == 0, p.stderr
    res = [ln for ln in p.stdout.splitlines() if "RESULT" in ln]
    assert len(res) == 16
    assert all("PASS" in ln for ln in res), res


def _read_pgm(path):
    data = Path(path).read_bytes()
    head = data.split(b"\n", 3)
    assert head[0] == b"P5" and head[2] == b"255"
    w, h = (int(v) for v in head[1].split())
    import numpy as np

    img = np.frombuffer(head[3], np.uint8)
    assert img.size == w * h
    return img.reshape(h, w)


@pytest.mark.gpu
def test_frames_are_the_reference_view_of_each_step(gpu, oracle, tmp_path):
    """-frames=DIR: the headless form of -opengl (StellarSolverVisualizer.cpp:
    208-253) writes the reference's view of every K-th step as a PGM; the
    image equals the oracle's drawing of the same state."""
    import numpy as np

    import stellarsolver_amd as S

    n = 16384
    p = run("-disk", f"-numbodies={n}", "-iterations=3", "-frame-every=2",
            f"-frames={tmp_path}", "-resolution=400")
    assert p.returncode == 0, p.stderr
    assert sorted(f.name for f in tmp_path.iterdir()) == ["frame_00000.pgm", "frame_00002.pgm"]
    # the same simulation through the library: reference disk model, the
    # system's default options
    ic = S.ic_params(0)
    ic.model = S.IC_REF_DISK
    st = S.generate_ic(ic, n)
    prm = S.default_params(theta=0.5, eps2=0.025, group=64, timestep=1e-3)
    view = (-25.0, 25.0, -25.0, 25.0)
    with S.Engine(n, prm) as e:
        e.upload(st["mass"], st["x"], st["y"], st["vx"], st["vy"])
        for k in range(3):
            e.update()
            if k % 2 == 0:
                pos = e.download()
                ref = oracle.render_grey(oracle.render_counts(pos["x"], pos["y"], view, 400, 400))
                assert np.array_equal(_read_pgm(tmp_path / f"frame_{k:05d}.pgm"), ref)
