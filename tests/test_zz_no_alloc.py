"""stellar_bbox_noalloc, the allocation-free bounding box, against stellar_bbox
(the context entry point) through the C ABI: bit-identical boxes on two calls
in a row, the second over fewer bodies (its partial results must not mix
with the first call's, which stay in the library's fixed device array)."""
import ctypes as C

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip():
    if not gpu_available():
        pytest.skip("no GPU")
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


@pytest.fixture(scope="module")
def S():
    import stellarsolver_amd as S

    S.load_library()
    return S


def _to_device(hip, a):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), max(a.nbytes, 16)) == 0
    if a.nbytes:
        assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
    return p


def _noalloc_bbox(S, hip, x, y):
    lib = S.load_library()
    dx, dy = _to_device(hip, x), _to_device(hip, y)
    out = C.c_void_p()
    assert hip.hipMalloc(C.byref(out), 16) == 0
    try:
        rc = lib.stellar_bbox_noalloc(dx, dy, x.size, out, None)
        assert rc == S.STELLAR_OK, S.load_library().stellar_status_string(rc)
        got = np.zeros(4, np.float32)
        assert hip.hipMemcpy(got.ctypes.data, out, 16, D2H) == 0  # synchronises
        return got
    finally:
        for p in (dx, dy, out):
            hip.hipFree(p)


def _context_bbox(S, x, y):
    n = x.size
    z = np.zeros(n, np.float32)
    with S.Engine(n, S.default_params()) as e:
        e.upload(np.ones(n, np.float32), x, y, z, z)
        e.bbox()
        return e.get_bbox()


def test_noalloc_bbox_matches_stellar_bbox(S, hip):
    rng = np.random.default_rng(12)
    cases = [
        # a large ragged input whose box is far from the origin ...
        (rng.uniform(-900.0, -20.0, 1_000_003), rng.uniform(30.0, 700.0, 1_000_003)),
        # ... then a smaller one inside it: stale partials would widen its box
        (rng.uniform(-2.0, 3.0, 5000), rng.uniform(-1.0, 4.0, 5000)),
    ]
    for x, y in cases:
        x, y = x.astype(np.float32), y.astype(np.float32)
        got = _noalloc_bbox(S, hip, x, y)
        ref = _context_bbox(S, x, y)
        assert np.array_equal(got, ref), (x.size, got, ref)
        exp = np.array([min(0.0, x.min()), max(0.0, x.max()), min(0.0, y.min()), max(0.0, y.max())],
                       np.float32)
        assert np.array_equal(got, exp)


def test_noalloc_bbox_limits(S, hip):
    lib = S.load_library()
    x = np.ones(8, np.float32)
    dx = _to_device(hip, x)
    try:
        assert lib.stellar_bbox_noalloc(dx, dx, S.BBOX_NOALLOC_MAX_BODIES + 1, dx, None) == S.STELLAR_ERR_SIZE
        assert lib.stellar_bbox_noalloc(dx, dx, -1, dx, None) == S.STELLAR_ERR_ARG
        assert lib.stellar_bbox_noalloc(None, None, 4, dx, None) == S.STELLAR_ERR_ARG
    finally:
        hip.hipFree(dx)
    # no bodies: the origin alone, as a context of zero bodies reports
    assert np.array_equal(_noalloc_bbox(S, hip, np.zeros(0, np.float32), np.zeros(0, np.float32)),
                          np.zeros(4, np.float32))
