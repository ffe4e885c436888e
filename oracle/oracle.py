"""ctypes binding of liboracle.so, the CPU restatement of the reference step.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import io
import os
import subprocess
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_DEBUG_PATH = HERE / "_ref" / "libref_debug.so"

OR_OK = 0
OR_ERR_ARG = -1
OR_ERR_POOL = -3
OR_ERR_DEPTH = -5
OR_ERR_BUCKET = -8
MAX_LEVEL = 64

_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int32)
_D = C.POINTER(C.c_double)
_U64 = C.POINTER(C.c_uint64)
_I64 = C.POINTER(C.c_int64)


class WalkParams(C.Structure):
    _fields_ = [("theta", C.c_float), ("eps2", C.c_float), ("group", C.c_int32), ("threads", C.c_int32),
                ("opening", C.c_int32), ("mixed", C.c_int32), ("cq", _F), ("cthr", _F), ("cqh", _F),
                ("clevel", _I), ("ref_root", C.c_int32), ("pad2", C.c_int32), ("order", _I)]


class StepParams(C.Structure):
    _fields_ = [("theta", C.c_float), ("eps2", C.c_float), ("dt", C.c_float), ("dampening", C.c_float),
                ("group", C.c_int32), ("threads", C.c_int32), ("m", C.c_int32)]


_SIG = {
    "or_bbox": (C.c_int, [C.c_int32, _F, _F, _F]),
    "or_path_keys": (None, [C.c_int32, _F, _F, _F, _U64, _U64]),
    "or_build_insert": (C.c_int, [C.c_int32, C.c_int32, _F, _F, _F, _F, _I, _I, _I, _F, _F, _F, _I]),
    "or_com_ref": (None, [C.c_int32, C.c_int32, _F, _F, _F]),
    "or_sort_ref": (None, [C.c_int32, C.c_int32, _I, _I, _I, _I]),
    "or_canonicalize": (C.c_int, [C.c_int32, C.c_int32, _I, _I, _I, _F, _F, _F, _I, _I, _I, _I, _I, _F, _F, _F]),
    "or_build_keys": (C.c_int, [C.c_int32, _F, _F, _F, C.c_int64, _I, _I, _I, _I, _I, _I, _I64]),
    "or_com_exact": (None, [C.c_int32, C.c_int64, _I, _F, _F, _F, _F, _F, _F]),
    "or_bmax_exact": (None, [C.c_int32, C.c_int64, _I, _F, _F, _F, C.c_float, C.c_float, _F]),
    "or_hilbert_order": (C.c_int, [C.c_int32, _I, _I, C.c_int32, _I]),
    "or_quad_exact": (None, [C.c_int32, C.c_int64, _I, _F, _F, _F, _F]),
    "or_quad_mixed": (None, [C.c_int32, C.c_int64, _I, _I, _F, _F, _F, _F, _F]),
    "or_f16_round": (C.c_double, [C.c_double, C.c_int32]),
    "or_walk": (C.c_int, [C.c_int32, _F, _F, _F, _F, _I, _I, _I, C.c_int64, _F, _F, _F,
                          C.POINTER(WalkParams), _F, _F, _I]),
    "or_walk_recursive": (C.c_int, [C.c_int32, _F, _F, _F, _F, _I, _I, _I, C.c_int64, _F, _F, _F,
                                    C.POINTER(WalkParams), _D, _D, _I]),
    "or_direct_fp64": (None, [C.c_int32, _F, _F, _F, C.c_double, C.c_int32, _I, _D, _D, C.c_int32]),
    "or_update": (None, [C.c_int32, _F, _F, _F, _F, _F, _F, C.c_float, C.c_float]),
    "or_invariants": (C.c_int32, [C.c_int32, C.c_int32, _I, _I, _I, _F]),
    "or_energy": (None, [C.c_int32, _F, _F, _F, _F, _F, C.c_double, C.c_int32, _D, _D, _D, _D]),
    "or_diagnostics": (None, [C.c_int32, _F, _F, _F, _F, _F, _F, _F, C.c_double, C.c_double,
                              C.c_int32, _D]),
    "or_step": (C.c_int, [C.c_int32, _F, _F, _F, _F, _F, _F, _F, C.POINTER(StepParams), _D]),
    "or_num_threads": (C.c_int, []),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            subprocess.run(["make", "-C", str(HERE), "liboracle.so"], check=True,
                           capture_output=True)
        _lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIG.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def _f(a):
    return None if a is None else a.ctypes.data_as(_F)


def _i(a):
    return None if a is None else a.ctypes.data_as(_I)


def _d(a):
    return None if a is None else a.ctypes.data_as(_D)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class OracleError(RuntimeError):
    def __init__(self, rc, what):
        super().__init__(f"oracle status {rc} in {what}")
        self.status = rc


def num_threads() -> int:
    return int(lib().or_num_threads())


def bbox(x, y) -> np.ndarray:
    x, y = f32(x), f32(y)
    b = np.zeros(4, np.float32)
    lib().or_bbox(x.size, _f(x), _f(y), _f(b))
    return b


def path_keys(x, y, bb) -> tuple[np.ndarray, np.ndarray]:
    x, y, bb = f32(x), f32(y), f32(bb)
    k1 = np.zeros(x.size, np.uint64)
    k2 = np.zeros(x.size, np.uint64)
    lib().or_path_keys(x.size, _f(x), _f(y), _f(bb), k1.ctypes.data_as(_U64), k2.ctypes.data_as(_U64))
    return k1, k2


def build_insert(x, y, mass, bb, m: int | None = None) -> dict:
    """Reference-layout tree from the sequential insertion restatement."""
    x, y, mass, bb = f32(x), f32(y), f32(mass), f32(bb)
    n = x.size
    m = 2 * n + 12000 if m is None else int(m)
    child = np.zeros(4 * m, np.int32)
    count = np.zeros(m, np.int32)
    start = np.zeros(m, np.int32)
    cx = np.zeros(m, np.float32)
    cy = np.zeros(m, np.float32)
    cm = np.zeros(m, np.float32)
    nc = C.c_int32()
    rc = lib().or_build_insert(n, m, _f(x), _f(y), _f(mass), _f(bb), _i(child), _i(count), _i(start),
                               _f(cx), _f(cy), _f(cm), C.byref(nc))
    if rc:
        raise OracleError(rc, "build_insert")
    return {"n": n, "m": m, "child": child, "count": count, "start": start, "cx": cx, "cy": cy,
            "cm": cm, "ncells": nc.value}


def com_ref(t: dict) -> None:
    lib().or_com_ref(t["n"], t["ncells"], _f(t["cx"]), _f(t["cy"]), _f(t["cm"]))


def sort_ref(t: dict) -> None:
    t["sorted"] = np.zeros(t["m"], np.int32)
    lib().or_sort_ref(t["n"], t["ncells"], _i(t["child"]), _i(t["count"]), _i(t["start"]),
                      _i(t["sorted"]))


def canonicalize(t: dict) -> dict:
    n, c = t["n"], t["ncells"]
    cc = max(c, 1)
    out = {"root": np.zeros(4, np.int32), "child": np.zeros(4 * cc, np.int32),
           "count": np.zeros(cc, np.int32), "start": np.zeros(cc, np.int32),
           "level": np.zeros(cc, np.int32), "com_x": np.zeros(cc, np.float32),
           "com_y": np.zeros(cc, np.float32), "com_m": np.zeros(cc, np.float32)}
    rc = lib().or_canonicalize(n, c, _i(t["child"]), _i(t["count"]), _i(t["start"]), _f(t["cx"]),
                               _f(t["cy"]), _f(t["cm"]), _i(out["root"]), _i(out["child"]),
                               _i(out["count"]), _i(out["start"]), _i(out["level"]),
                               _f(out["com_x"]), _f(out["com_y"]), _f(out["com_m"]))
    if rc:
        raise OracleError(rc, "canonicalize")
    for k in ("count", "start", "level", "com_x", "com_y", "com_m"):
        out[k] = out[k][:c]
    out["child"] = out["child"][: 4 * c]
    out["sorted"] = t["sorted"][:n].copy() if "sorted" in t else None
    out["ncells"] = c
    return out


def build_keys(x, y, bb, capacity: int | None = None) -> dict:
    """Canonical tree from the independent path-key construction."""
    x, y, bb = f32(x), f32(y), f32(bb)
    n = x.size
    cap = 2 * n + 12000 if capacity is None else int(capacity)
    t = {"sorted": np.zeros(max(n, 1), np.int32), "root": np.zeros(4, np.int32),
         "child": np.zeros(4 * max(cap, 1), np.int32), "count": np.zeros(max(cap, 1), np.int32),
         "start": np.zeros(max(cap, 1), np.int32), "level": np.zeros(max(cap, 1), np.int32)}
    nc = C.c_int64()
    rc = lib().or_build_keys(n, _f(x), _f(y), _f(bb), cap, _i(t["sorted"]), _i(t["root"]),
                             _i(t["child"]), _i(t["count"]), _i(t["start"]), _i(t["level"]),
                             C.byref(nc))
    if rc:
        raise OracleError(rc, "build_keys")
    c = nc.value
    return {"root": t["root"], "child": t["child"][: 4 * c], "count": t["count"][:c],
            "start": t["start"][:c], "level": t["level"][:c], "sorted": t["sorted"][:n],
            "ncells": c}


def com_exact(tree: dict, x, y, mass) -> None:
    x, y, mass = f32(x), f32(y), f32(mass)
    c = tree["ncells"]
    for k in ("com_x", "com_y", "com_m"):
        tree[k] = np.zeros(max(c, 1), np.float32)
    lib().or_com_exact(x.size, c, _i(tree["child"]), _f(x), _f(y), _f(mass), _f(tree["com_x"]),
                       _f(tree["com_y"]), _f(tree["com_m"]))
    for k in ("com_x", "com_y", "com_m"):
        tree[k] = tree[k][:c]


def quad_exact(tree: dict, x, y, mass) -> None:
    """Quadrupole moments (3 per cell) into tree["quad"]."""
    x, y, mass = f32(x), f32(y), f32(mass)
    c = tree["ncells"]
    q = np.zeros(3 * max(c, 1), np.float32)
    lib().or_quad_exact(x.size, c, _i(tree["child"]), _f(x), _f(y), _f(mass), _f(q))
    tree["quad"] = q[: 3 * c]


def _wp(tree, theta, eps2, group, threads, opening, quad, bmax=False, mixed=False, ref_root=False,
        order=None):
    cq = None
    if quad:
        cq = tree["quad"] if tree["quad"].size else np.zeros(3, np.float32)
    th = None
    if bmax:
        th = tree["thr"] if tree["thr"].size else np.zeros(1, np.float32)
    qh = lv = None
    if mixed:
        qh = tree["qh"] if tree["qh"].size else np.zeros(3, np.float32)
        lv = np.ascontiguousarray(tree["level"], dtype=np.int32)
        lv = lv if lv.size else np.zeros(1, np.int32)
    od = None if order is None else np.ascontiguousarray(order, dtype=np.int32)
    p = WalkParams(theta, eps2, group, threads, opening, 1 if mixed else 0, _f(cq), _f(th), _f(qh),
                   _i(lv), 1 if ref_root else 0, 0, _i(od))
    return p, (cq, th, qh, lv, od)


def quad_mixed(tree: dict, x, y, mass, bb) -> None:
    """Qh = Q / (m ws^2) per cell, rounded to FP16, into tree["qh"]."""
    x, y, mass, bb = f32(x), f32(y), f32(mass), f32(bb)
    c = tree["ncells"]
    q = np.zeros(3 * max(c, 1), np.float32)
    lv = np.ascontiguousarray(tree["level"], dtype=np.int32)
    lib().or_quad_mixed(x.size, c, _i(tree["child"]), _i(lv if lv.size else np.zeros(1, np.int32)),
                        _f(x), _f(y), _f(mass), _f(bb), _f(q))
    tree["qh"] = q[: 3 * c]


def bmax_exact(tree: dict, x, y, mass, theta: float, eps2: float) -> None:
    """Per-cell opening thresholds bmax^2/theta^2 + eps2 into tree["thr"]."""
    x, y, mass = f32(x), f32(y), f32(mass)
    c = tree["ncells"]
    t = np.zeros(max(c, 1), np.float32)
    lib().or_bmax_exact(x.size, c, _i(tree["child"]), _f(x), _f(y), _f(mass), float(theta),
                        float(eps2), _f(t))
    tree["thr"] = t[:c]


def walk(tree: dict, x, y, mass, bb, theta: float, eps2: float, group: int, threads: int = 0,
         opening: int = 0, quad: bool = False, bmax: bool = False, mixed: bool = False,
         ref_root: bool = False, order=None):
    """Group-vote walk (kernels.cu:340-435) over a canonical tree; by body id.
    bmax: per-cell thresholds tree["thr"] (bmax_exact) instead of levels.
    mixed: accepted cells' terms in FP16 from tree["qh"] (quad_mixed).
    order: body order of the walk groups (default tree["sorted"])."""
    x, y, mass, bb = f32(x), f32(y), f32(mass), f32(bb)
    n = x.size
    p, _keep = _wp(tree, theta, eps2, group, threads, opening, quad, bmax, mixed, ref_root, order)
    ax = np.zeros(n, np.float32)
    ay = np.zeros(n, np.float32)
    it = np.zeros(n, np.int32)
    ch = tree["child"] if tree["child"].size else np.zeros(4, np.int32)
    cx = tree["com_x"] if tree["com_x"].size else np.zeros(1, np.float32)
    cy = tree["com_y"] if tree["com_y"].size else np.zeros(1, np.float32)
    cm = tree["com_m"] if tree["com_m"].size else np.zeros(1, np.float32)
    rc = lib().or_walk(n, _f(x), _f(y), _f(mass), _f(bb), _i(tree["sorted"]), _i(tree["root"]),
                       _i(ch), tree["ncells"], _f(cx), _f(cy), _f(cm), C.byref(p), _f(ax), _f(ay),
                       _i(it))
    if rc:
        raise OracleError(rc, "walk")
    return ax, ay, it


def walk_recursive(tree: dict, x, y, mass, bb, theta: float, eps2: float, group: int, threads: int = 0,
                   opening: int = 0, quad: bool = False, bmax: bool = False, ref_root: bool = False,
                   order=None):
    x, y, mass, bb = f32(x), f32(y), f32(mass), f32(bb)
    n = x.size
    p, _keep = _wp(tree, theta, eps2, group, threads, opening, quad, bmax, ref_root=ref_root,
                   order=order)
    ax = np.zeros(n, np.float64)
    ay = np.zeros(n, np.float64)
    it = np.zeros(n, np.int32)
    ch = tree["child"] if tree["child"].size else np.zeros(4, np.int32)
    cx = tree["com_x"] if tree["com_x"].size else np.zeros(1, np.float32)
    cy = tree["com_y"] if tree["com_y"].size else np.zeros(1, np.float32)
    cm = tree["com_m"] if tree["com_m"].size else np.zeros(1, np.float32)
    rc = lib().or_walk_recursive(n, _f(x), _f(y), _f(mass), _f(bb), _i(tree["sorted"]),
                                 _i(tree["root"]), _i(ch), tree["ncells"], _f(cx), _f(cy), _f(cm),
                                 C.byref(p), _d(ax), _d(ay), _i(it))
    if rc:
        raise OracleError(rc, "walk_recursive")
    return ax, ay, it


def hilbert_order(tree: dict, n: int, l1: int = 16) -> np.ndarray:
    """Body ids in Hilbert order of a canonical tree (Hilbert children at
    levels 1..l1, slot order below; or_hilbert_order)."""
    out = np.zeros(max(n, 1), np.int32)
    ch = tree["child"] if tree["child"].size else np.zeros(4, np.int32)
    rc = lib().or_hilbert_order(n, _i(tree["root"]), _i(ch), int(l1), _i(out))
    if rc:
        raise OracleError(rc, "hilbert_order")
    return out[:n]


def direct_fp64(x, y, mass, eps2: float, targets=None, threads: int = 0):
    x, y, mass = f32(x), f32(y), f32(mass)
    if targets is None:
        nt, tp = x.size, None
    else:
        targets = np.ascontiguousarray(targets, dtype=np.int32)
        nt, tp = targets.size, _i(targets)
    ax = np.zeros(nt, np.float64)
    ay = np.zeros(nt, np.float64)
    lib().or_direct_fp64(x.size, _f(x), _f(y), _f(mass), float(eps2), nt, tp, _d(ax), _d(ay), threads)
    return ax, ay


def update(x, y, vx, vy, ax, ay, dt: float, d: float):
    x, y, vx, vy = (f32(a).copy() for a in (x, y, vx, vy))
    ax, ay = f32(ax), f32(ay)
    lib().or_update(x.size, _f(x), _f(y), _f(vx), _f(vy), _f(ax), _f(ay), dt, d)
    return x, y, vx, vy


def invariants(n: int, m: int, count, sorted_, child, bb) -> int:
    count, sorted_, child = (np.ascontiguousarray(a, dtype=np.int32) for a in (count, sorted_, child))
    return int(lib().or_invariants(n, m, _i(count), _i(sorted_), _i(child), _f(f32(bb))))


def energy(x, y, vx, vy, mass, eps2: float, threads: int = 0) -> dict:
    v = [C.c_double() for _ in range(4)]
    a = [f32(q) for q in (x, y, vx, vy, mass)]
    lib().or_energy(a[0].size, *(_f(q) for q in a), float(eps2), threads, *(C.byref(q) for q in v))
    return {"kinetic": v[0].value, "potential": v[1].value, "px": v[2].value, "py": v[3].value}


def diagnostics(x, y, vx, vy, mass, eps2: float, ax=None, ay=None, half_dt: float = 0.0,
                threads: int = 0) -> dict:
    """Kinetic and potential energy, momentum and angular momentum about the
    origin in FP64; with ax, ay given the velocities are v + half_dt * a."""
    a = [f32(q) for q in (x, y, vx, vy, mass)]
    acc = [f32(q) for q in (ax, ay)] if ax is not None else None
    out = np.zeros(5, np.float64)
    lib().or_diagnostics(a[0].size, *(_f(q) for q in a),
                         _f(acc[0]) if acc else None, _f(acc[1]) if acc else None,
                         float(half_dt), float(eps2), threads, _d(out))
    return dict(zip(("kinetic", "potential", "px", "py", "lz"), (float(v) for v in out)))


def step(state: dict, theta: float, eps2: float, dt: float, d: float, group: int = 32,
         threads: int = 0, m: int = 0):
    """One reference step on the CPU, in place on state (x, y, vx, vy, mass)."""
    for k in ("x", "y", "vx", "vy", "mass"):
        state[k] = f32(state[k])
    n = state["x"].size
    ax = np.zeros(n, np.float32)
    ay = np.zeros(n, np.float32)
    secs = np.zeros(6, np.float64)
    p = StepParams(theta, eps2, dt, d, group, threads, m)
    rc = lib().or_step(n, _f(state["x"]), _f(state["y"]), _f(state["vx"]), _f(state["vy"]),
                       _f(state["mass"]), _f(ax), _f(ay), C.byref(p), _d(secs))
    if rc:
        raise OracleError(rc, "step")
    state["ax"], state["ay"] = ax, ay
    return secs


def render_counts(x, y, view, width: int, height: int) -> np.ndarray:
    """Bodies per pixel of the reference's point drawing, headless
    (StellarSolverVisualizer.cpp:13-19, 208-253: glm::ortho onto the window,
    one pixel per point).  The body at (x, y) falls in column
    floor((x - left) * sx) and row height - 1 - floor((y - bottom) * sy),
    sx = width / (right - left), sy = height / (top - bottom), FP32 as
    written; bodies outside the image are clipped.  Returns height x width
    uint32, row 0 at the top."""
    x, y = f32(x), f32(y)
    left, right, bottom, top = (np.float32(v) for v in view)
    sx = np.float32(width) / (right - left)
    sy = np.float32(height) / (top - bottom)
    with np.errstate(invalid="ignore"):
        fx = (x - left) * sx
        fy = (y - bottom) * sy
        ok = (fx >= 0) & (fx < np.float32(width)) & (fy >= 0) & (fy < np.float32(height))
    col = fx[ok].astype(np.int64)
    row = height - 1 - fy[ok].astype(np.int64)
    counts = np.bincount(row * width + col, minlength=width * height)
    return counts.astype(np.uint32).reshape(height, width)


def render_grey(counts) -> np.ndarray:
    """The reference's blend of white at alpha 0.1 per body over black
    (glBlendFunc(GL_SRC_ALPHA, GL_ONE)), clamped, quantised once to 8 bits:
    min(255, round(25.5 count))."""
    c = np.asarray(counts, dtype=np.uint64)
    return np.where(c >= 10, 255, (51 * c + 1) // 2).astype(np.uint8)


def to_reference_layout(tree: dict, n: int, m: int | None = None) -> dict:
    """Canonical tree -> reference pool arrays (cells at n + c, root in child[0..3])."""
    c = tree["ncells"]
    m = 2 * n + 12000 if m is None else m
    child = np.full(4 * m, -1, np.int32)
    count = np.zeros(m, np.int32)
    count[:n] = 1
    if c:
        child[4 * n: 4 * (n + c)] = tree["child"]
        count[n: n + c] = tree["count"]
    child[0:4] = tree["root"]
    sorted_ = np.zeros(m, np.int32)
    sorted_[:n] = tree["sorted"]
    return {"child": child, "count": count, "sorted": sorted_, "m": m}


# --------------------------------------------------------------------------
# The reference's own debug.cpp (DEBUG_RUN_TESTS), compiled from its source
# into _ref/ by `make -C oracle ref`; it prints PASS/FAIL lines.
_REF_SYM = "_Z15DEBUG_RUN_TESTSPfS_S_PiS0_S0_S0_S_S_S_S_ii"


def ref_debug_available() -> bool:
    return REF_DEBUG_PATH.exists()


def ref_debug_run(x, y, mass, count, start, sorted_, child, bb, n: int, m: int) -> list[str]:
    """Run DEBUG_RUN_TESTS; returns its result lines (captured from fd 1)."""
    lib_ref = C.CDLL(str(REF_DEBUG_PATH))
    fn = getattr(lib_ref, _REF_SYM)
    fn.restype = None
    arrs_f = [np.ascontiguousarray(a, dtype=np.float32) for a in (x, y, mass)]
    arrs_i = [np.ascontiguousarray(a, dtype=np.int32) for a in (count, start, sorted_, child)]
    b = [np.array([v], np.float32) for v in bb]
    libc = C.CDLL(None)
    with tempfile.TemporaryFile(mode="w+b") as tmp:
        libc.fflush(None)
        saved = os.dup(1)
        os.dup2(tmp.fileno(), 1)
        try:
            fn(*(_f(a) for a in arrs_f), *(_i(a) for a in arrs_i), *(_f(q) for q in b),
               C.c_int(n), C.c_int(m))
            libc.fflush(None)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        tmp.seek(0)
        text = tmp.read().decode()
    return [ln.strip() for ln in io.StringIO(text) if "RESULT" in ln]
