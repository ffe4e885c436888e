"""C-ABI library checks that need no GPU: it loads, exports every symbol the
header declares, and its host-side initial-condition generators behave."""
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "stellar.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(stellar_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(stellar):
    lib = stellar.load_library()
    syms = declared_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(lib, s), s
    # and the Python binding covers all of them
    assert set(syms) == set(stellar.SIGNATURES)


def test_status_strings(stellar):
    lib = stellar.load_library()
    for code in (0, -1, -2, -3, -4, -5, -6, -7, -8):
        assert lib.stellar_status_string(code)


def test_default_params_match_reference(stellar):
    p = stellar.default_params()
    # kernels.cu:8-9 (eps2 0.025, theta 0.5); main.cpp:95-97
    assert p.theta == 0.5 and np.float32(p.eps2) == np.float32(0.025)
    assert np.float32(p.timestep) == np.float32(0.001) and p.dampening == 1.0 and p.gravity == 1.0


def test_ic_deterministic_and_index_local(stellar):
    a = stellar.generate_ic(stellar.ic_params(1), 5000)
    b = stellar.generate_ic(stellar.ic_params(1), 5000)
    for k in a:
        assert np.array_equal(a[k], b[k])
    c = stellar.generate_ic(stellar.ic_params(1, seed=99), 5000)
    assert not np.array_equal(a["x"], c["x"])


def test_plummer_profile(stellar):
    n = 200000
    st = stellar.generate_ic(stellar.ic_params(1), n)
    r = np.hypot(st["x"], st["y"]).astype(np.float64)
    assert np.isclose(st["mass"].sum(dtype=np.float64), 1.0, rtol=1e-5)
    assert r.max() < 20.0
    # cumulative mass fraction of a Plummer sphere truncated at 20a
    m = lambda q: q ** 3 / (q ** 2 + 1) ** 1.5
    for q in (0.5, 1.0, 2.0, 5.0):
        frac = (r < q).mean()
        assert abs(frac - m(q) / m(20.0)) < 0.01, (q, frac)
    # speeds below the escape speed sqrt(2)(r^2+1)^-1/4
    v = np.hypot(st["vx"], st["vy"])
    assert np.all(v <= np.sqrt(2.0) * (r ** 2 + 1) ** -0.25 * (1 + 1e-5))


def test_disk_circular_speed(stellar):
    n = 4096
    p = stellar.ic_params(0)
    st = stellar.generate_ic(p, n)
    assert st["mass"][0] == 1000.0 and st["x"][0] == 0.0 and st["y"][0] == 0.0
    r = np.hypot(st["x"][1:], st["y"][1:]).astype(np.float64)
    assert r.min() >= 2.0 - 1e-4 and r.max() <= 100.0 + 1e-3
    inside = np.argsort(np.argsort(r))  # rank = bodies strictly inside (no ties)
    vc = np.sqrt((1000.0 + inside) / r)
    # tangential (counter-clockwise) component equals the circular speed
    vt = (st["x"][1:] * st["vy"][1:] - st["y"][1:] * st["vx"][1:]) / r
    assert np.sqrt(np.mean((vt - vc) ** 2)) < 0.02
    vr = (st["x"][1:] * st["vx"][1:] + st["y"][1:] * st["vy"][1:]) / r
    assert abs(np.std(vr) - 0.01) < 0.002


def test_colliding_disks_layout(stellar):
    n = 20000
    st = stellar.generate_ic(stellar.ic_params(2), n)
    h = n // 2
    assert st["mass"][0] == 1e4 and st["mass"][h] == 1e4
    assert st["x"][0] == 300.0 and st["x"][h] == -300.0
    assert (st["vx"][0], st["vy"][0]) == (-5.0, 1.0)
    assert (st["vx"][h], st["vy"][h]) == (5.0, -1.0)
    assert np.all(np.abs(st["x"][1:h] - 300.0) <= 200.0 + 1e-3)


def test_reference_models(stellar):
    # BarnesHutParticleSystem.cpp:439-483 (disk) and :494-560 (colliding)
    n = 4096
    d = stellar.generate_ic(stellar.ic_params(None, model=stellar.IC_REF_DISK, gravity=1.0), n)
    assert d["mass"][0] == 200000 and np.all(d["mass"][1:] == 1.0)
    r = np.hypot(d["x"][1:], d["y"][1:])
    assert r.min() >= 1.5 - 1e-5 and r.max() <= 12.0 + 1e-5
    v = np.hypot(d["vx"][1:], d["vy"][1:])
    assert np.allclose(v, np.sqrt(200000.0 / r), rtol=1e-5)
    c = stellar.generate_ic(stellar.ic_params(None, model=stellar.IC_REF_COLLIDING, gravity=1.0), n)
    assert c["mass"][0] == 100000 and c["mass"][1] == 25000
    p = stellar.generate_ic(stellar.ic_params(None, model=stellar.IC_REF_PLUMMER, gravity=1.0), n)
    assert np.all(p["mass"] == 1.0)
    # the reference's default-seeded engine: deterministic across calls
    p2 = stellar.generate_ic(stellar.ic_params(None, model=stellar.IC_REF_PLUMMER, gravity=1.0), n)
    assert np.array_equal(p["x"], p2["x"])


def test_bad_arguments_rejected_without_gpu(stellar):
    import ctypes as C

    lib = stellar.load_library()
    p = stellar.StellarParams()
    lib.stellar_default_params(C.byref(p))
    h = C.c_void_p()
    p.group = 48
    assert lib.stellar_create(C.byref(p), 10, 0, C.byref(h)) == stellar.STELLAR_ERR_ARG
    p.group = 64
    p.eps2 = 0.0
    assert lib.stellar_create(C.byref(p), 10, 0, C.byref(h)) == stellar.STELLAR_ERR_ARG
    p.eps2 = 0.01
    assert lib.stellar_create(C.byref(p), -1, 0, C.byref(h)) == stellar.STELLAR_ERR_ARG
    p.precision = stellar.PRECISION_MIXED  # reserved, not implemented
    assert lib.stellar_create(C.byref(p), 10, 0, C.byref(h)) == stellar.STELLAR_ERR_ARG
    p.precision = stellar.PRECISION_FP32
    counts = (C.c_int32 * 3)(10, 0, 5)  # an empty instance
    assert lib.stellar_batch_create(C.byref(p), 3, counts, 0, C.byref(h)) == stellar.STELLAR_ERR_ARG
    assert lib.stellar_batch_create(C.byref(p), 0, counts, 0, C.byref(h)) == stellar.STELLAR_ERR_ARG
    tune = stellar.StellarTuning()
    assert lib.stellar_autotune(None, 0, 1, C.byref(tune)) == stellar.STELLAR_ERR_ARG
    bad = stellar.ic_params(None, model=77)
    with pytest.raises(stellar.StellarError):
        stellar.generate_ic(bad, 10)
