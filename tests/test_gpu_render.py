"""GPU: headless point drawing (stellar_render, the reference's OpenGL view of
StellarSolverVisualizer.cpp:208-253) against the oracle's per-pixel counts.

Bar: per-pixel counts and 8-bit greys bit-exact (integer work)."""
import numpy as np
import pytest

from tests.helpers import ic

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    from tests.conftest import gpu_available

    if not gpu_available():
        pytest.skip("no GPU")
    import stellarsolver_amd as S

    S.load_library()
    return S


def _engine(S, st, **kw):
    p = S.default_params(theta=0.5, eps2=1e-4, group=64, timestep=1e-3, **kw)
    e = S.Engine(st["x"].size, p)
    e.upload(st["mass"], st["x"], st["y"], st["vx"], st["vy"])
    return e


@pytest.mark.parametrize("workload,n,view,size", [
    (0, 4096, (-25.0, 25.0, -25.0, 25.0), (1000, 1000)),      # the reference's window
    (0, 65536, (-110.0, 110.0, -110.0, 110.0), (640, 480)),   # whole disk, non-square
    (2, 1 << 20, (-520.0, 520.0, -260.0, 260.0), (1024, 512)),  # colliding disks, 2^20
    (1, 262144, (-0.5, 0.5, -0.5, 0.5), (333, 257)),          # Plummer core, dense pixels
])
def test_render_counts_match_oracle(S, oracle, workload, n, view, size):
    st = ic(workload, n)
    w, h = size
    with _engine(S, st) as e:
        # after two steps the storage is in tree order (runs of equal pixels)
        e.run(2)
        pos = e.download()
        counts, grey = e.render(view, w, h)
    ref = oracle.render_counts(pos["x"], pos["y"], view, w, h)
    assert counts.shape == (h, w)
    assert np.array_equal(counts, ref), int((counts != ref).sum())
    assert np.array_equal(grey, oracle.render_grey(ref))
    assert int(counts.sum()) > 0


def test_render_before_any_step_and_clipping(S, oracle):
    st = ic(0, 1000)
    st = {k: v.copy() for k, v in st.items()}
    st["x"][:3] = np.array([25.0, -25.0, np.nan], np.float32)  # edge, corner, NaN
    st["y"][:3] = np.array([0.0, -25.0, 0.0], np.float32)
    view = (-25.0, 25.0, -25.0, 25.0)
    with _engine(S, st) as e:
        counts, grey = e.render(view, 1000, 1000)
        # a second, smaller image reuses the device buffer
        c2, _ = e.render(view, 100, 100)
    assert np.array_equal(counts, oracle.render_counts(st["x"], st["y"], view, 1000, 1000))
    assert np.array_equal(c2, oracle.render_counts(st["x"], st["y"], view, 100, 100))
    assert counts[999, 0] >= 1


def test_render_rejects_bad_views(S):
    st = ic(0, 256)
    with _engine(S, st) as e:
        for view in ((1.0, 1.0, 0.0, 1.0), (0.0, 1.0, 2.0, -2.0)):
            with pytest.raises(S.StellarError):
                e.render(view, 10, 10)
        with pytest.raises(S.StellarError):
            e.render((0.0, 1.0, 0.0, 1.0), 0, 10)
