"""stellarsolver_amd: MI355X (gfx950) Barnes-Hut step behind the libstellar_hip C ABI.

This module is a thin ctypes binding of ``include/stellar.h``.  Every compute
call goes to the HIP library; there is no CPU fallback.  If the library is
missing the import of :class:`Engine` raises, naming the build command.

Mirrors the reference's host object ``BarnesHutParticleSystem``
(BarnesHutParticleSystem.h:61-71): ``update()``, ``reset()``,
``get_output_buffer()``, ``get_num_particles()``, plus the stage-level calls
of the launch wrappers (Particle_cuda.cuh:5-13).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libstellar_hip.so"

STELLAR_OK = 0
STELLAR_ERR_ARG = -1
STELLAR_ERR_OOM = -2
STELLAR_ERR_POOL = -3
STELLAR_ERR_STACK = -4
STELLAR_ERR_DEPTH = -5
STELLAR_ERR_DEVICE = -6
STELLAR_ERR_STATE = -7
STELLAR_ERR_BUCKET = -8
STELLAR_ERR_SIZE = -9
WALK_STATIC, WALK_QUEUE, WALK_QUEUE_WAVE = 0, 1, 2
BBOX_NOALLOC_MAX_BODIES = 4096 * 4096
STELLAR_MAX_LEVEL = 64

PRECISION_FP32 = 0
PRECISION_MIXED = 1
FLAG_STAGE_TIMING = 1
FLAG_QUADRUPOLE = 2
FLAG_TEXTBOOK_OPENING = 4
FLAG_GRAPH = 8
FLAG_BMAX_OPENING = 16
FLAG_FORCES_TIMING = 32
FLAG_REFERENCE_ROOT = 64
FLAG_HILBERT_ORDER = 128

IC_DISK = 0
IC_PLUMMER = 1
IC_COLLIDING = 2
IC_REF_DISK = 10
IC_REF_PLUMMER = 11
IC_REF_COLLIDING = 12

STAGES = ("bbox", "build", "com", "sort", "forces", "update", "output")


class StellarParams(C.Structure):
    _fields_ = [
        ("theta", C.c_float),
        ("eps2", C.c_float),
        ("timestep", C.c_float),
        ("dampening", C.c_float),
        ("gravity", C.c_float),
        ("group", C.c_int32),
        ("precision", C.c_int32),
        ("flags", C.c_int32),
    ]


class StellarICParams(C.Structure):
    _fields_ = [
        ("model", C.c_int32),
        ("pad", C.c_int32),
        ("seed", C.c_uint64),
        ("gravity", C.c_double),
        ("central_mass", C.c_double),
        ("rmin", C.c_double),
        ("rmax", C.c_double),
        ("jitter", C.c_double),
        ("plummer_a", C.c_double),
        ("plummer_rcut", C.c_double),
        ("offset_x", C.c_double),
        ("bulk_vx", C.c_double),
        ("bulk_vy", C.c_double),
    ]


class StellarTuning(C.Structure):
    _fields_ = [
        ("group", C.c_int32),
        ("nodes_per_pop", C.c_int32),
        ("block_threads", C.c_int32),
        ("walk_ms", C.c_float),
        ("schedule", C.c_int32),
    ]


class StellarDiagnostics(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("kinetic", "potential", "px", "py", "lz")]


# exported symbol -> (restype, argtypes)
_P = C.c_void_p
_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int32)
_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
SIGNATURES = {
    "stellar_default_params": (None, [C.POINTER(StellarParams)]),
    "stellar_create": (C.c_int, [C.POINTER(StellarParams), C.c_int32, C.c_int64, C.POINTER(_P)]),
    "stellar_batch_create": (C.c_int, [C.POINTER(StellarParams), C.c_int32, _I, C.c_int64, C.POINTER(_P)]),
    "stellar_num_instances": (C.c_int32, [_P]),
    "stellar_instance_offsets": (C.c_int, [_P, _I]),
    "stellar_destroy": (C.c_int, [_P]),
    "stellar_set_params": (C.c_int, [_P, C.POINTER(StellarParams)]),
    "stellar_get_params": (C.c_int, [_P, C.POINTER(StellarParams)]),
    "stellar_num_bodies": (C.c_int32, [_P]),
    "stellar_upload": (C.c_int, [_P, _F, _F, _F, _F, _F]),
    "stellar_download": (C.c_int, [_P, _F, _F, _F, _F, _F, _F, _F]),
    "stellar_download_positions": (C.c_int, [_P, _F]),
    "stellar_render": (C.c_int, [_P, _F, C.c_int32, C.c_int32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint8)]),
    "stellar_bbox": (C.c_int, [_P]),
    "stellar_bbox_noalloc": (C.c_int, [_P, _P, C.c_int32, _P, _P]),
    "stellar_build": (C.c_int, [_P]),
    "stellar_com": (C.c_int, [_P]),
    "stellar_sort": (C.c_int, [_P]),
    "stellar_forces": (C.c_int, [_P]),
    "stellar_update": (C.c_int, [_P]),
    "stellar_step": (C.c_int, [_P]),
    "stellar_run": (C.c_int, [_P, C.c_int32]),
    "stellar_sync": (C.c_int, [_P]),
    "stellar_get_bbox": (C.c_int, [_P, _F]),
    "stellar_get_bbox_instance": (C.c_int, [_P, C.c_int32, _F]),
    "stellar_num_cells": (C.c_int, [_P, _I64]),
    "stellar_export_tree": (C.c_int, [_P, _I, _I, _I, _I, _I, _I, _F, _F, _F, C.c_int64, _I64]),
    "stellar_export_tree_instance": (C.c_int, [_P, C.c_int32, _I, _I, _I, _I, _I, _I, _F, _F, _F, C.c_int64, _I64]),
    "stellar_export_forces": (C.c_int, [_P, _F, _F, _I]),
    "stellar_last_interactions": (C.c_int, [_P, _I64]),
    "stellar_last_quad_interactions": (C.c_int, [_P, _I64]),
    "stellar_stage_times": (C.c_int, [_P, _F]),
    "stellar_last_error": (C.c_char_p, [_P]),
    "stellar_status_string": (C.c_char_p, [C.c_int]),
    "stellar_direct_fp64": (C.c_int, [_P, C.c_int32, _I, C.c_double, _D, _D]),
    "stellar_energy_fp64": (C.c_int, [_P, C.c_double, _D, _D, _D, _D]),
    "stellar_energy_fp64_instances": (C.c_int, [_P, C.c_double, _D, _D, _D, _D]),
    "stellar_diagnostics_fp64": (C.c_int, [_P, C.c_double, C.c_double, C.POINTER(StellarDiagnostics)]),
    "stellar_autotune": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(StellarTuning)]),
    "stellar_set_walk_shape": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "stellar_set_walk_schedule": (C.c_int, [_P, C.c_int32]),
    "stellar_debug_sort_cap": (C.c_int, [_P, C.c_int32]),
    "stellar_ic_defaults": (C.c_int, [C.c_int32, C.POINTER(StellarICParams)]),
    "stellar_generate_ic": (C.c_int, [C.POINTER(StellarICParams), C.c_int32, _F, _F, _F, _F, _F]),
}

_lib = None


class StellarError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"stellar status {status}: {msg}")
        self.status = status


def load_library(path: os.PathLike | str | None = None) -> C.CDLL:
    """Load libstellar_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build it with `python -m stellarsolver_amd.build` "
            "(the engine has no CPU fallback)"
        )
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _fp(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_F)


def _ip(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_I)


def _dp(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_D)


def default_params(**kw) -> StellarParams:
    lib = load_library()
    p = StellarParams()
    lib.stellar_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def ic_params(workload: int | None = None, **kw) -> StellarICParams:
    lib = load_library()
    p = StellarICParams()
    if workload is not None:
        rc = lib.stellar_ic_defaults(workload, C.byref(p))
        if rc:
            raise StellarError(rc, f"no workload {workload}")
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def generate_ic(p: StellarICParams, n: int) -> dict[str, np.ndarray]:
    """Initial conditions on the host: dict of mass, x, y, vx, vy (float32)."""
    lib = load_library()
    out = {k: np.zeros(n, np.float32) for k in ("mass", "x", "y", "vx", "vy")}
    rc = lib.stellar_generate_ic(C.byref(p), n, *(_fp(out[k]) for k in ("mass", "x", "y", "vx", "vy")))
    if rc:
        raise StellarError(rc, "stellar_generate_ic")
    return out


class Engine:
    """One simulation, or an ensemble of independent simulations (``counts``),
    on the current HIP device (a stellar_ctx)."""

    def __init__(self, n: int | None = None, params: StellarParams | None = None,
                 node_capacity: int = 0, counts=None):
        self._lib = load_library()
        self.params = params if params is not None else default_params()
        h = C.c_void_p()
        if counts is None:
            self.n = int(n)
            self.offsets = np.array([0, self.n], np.int64)
            rc = self._lib.stellar_create(C.byref(self.params), self.n, int(node_capacity), C.byref(h))
        else:
            cnt = np.ascontiguousarray(counts, dtype=np.int32)
            self.offsets = np.concatenate([[0], np.cumsum(cnt, dtype=np.int64)])
            self.n = int(self.offsets[-1])
            if n is not None and int(n) != self.n:
                raise ValueError("n must equal sum(counts)")
            rc = self._lib.stellar_batch_create(C.byref(self.params), cnt.size, _ip(cnt),
                                                int(node_capacity), C.byref(h))
        if rc:
            raise StellarError(rc, self._lib.stellar_status_string(rc).decode())
        self._h = h
        self.ninst = len(self.offsets) - 1

    # -- lifetime -------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.stellar_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc: int, what: str) -> int:
        if rc:
            msg = self._lib.stellar_last_error(self._h)
            raise StellarError(rc, f"{what}: {msg.decode() if msg else ''}")
        return rc

    # -- configuration --------------------------------------------------
    def set_params(self, **kw) -> None:
        for k, v in kw.items():
            setattr(self.params, k, v)
        self._check(self._lib.stellar_set_params(self._h, C.byref(self.params)), "set_params")

    # -- data movement --------------------------------------------------
    def upload(self, mass, x, y, vx, vy) -> None:
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (mass, x, y, vx, vy)]
        for a in arrs:
            if a.shape != (self.n,):
                raise ValueError(f"expected shape ({self.n},), got {a.shape}")
        self._check(self._lib.stellar_upload(self._h, *(_fp(a) for a in arrs)), "upload")

    def download(self) -> dict[str, np.ndarray]:
        out = {k: np.zeros(self.n, np.float32) for k in ("mass", "x", "y", "vx", "vy", "ax", "ay")}
        self._check(self._lib.stellar_download(self._h, *(_fp(out[k]) for k in ("mass", "x", "y", "vx", "vy", "ax", "ay"))), "download")
        return out

    def get_output_buffer(self) -> np.ndarray:
        """Interleaved (x0, y0, x1, y1, ...) in body-id order (kernels.cu:541-553)."""
        out = np.zeros(2 * self.n, np.float32)
        self._check(self._lib.stellar_download_positions(self._h, _fp(out)), "download_positions")
        return out

    def get_num_particles(self) -> int:
        return self.n

    def render(self, view=(-25.0, 25.0, -25.0, 25.0), width: int = 1000,
               height: int = 1000) -> tuple[np.ndarray, np.ndarray]:
        """Headless form of the reference's point drawing
        (StellarSolverVisualizer.cpp:208-253): (counts, grey), each
        height x width with row 0 at the top."""
        v = np.ascontiguousarray(view, dtype=np.float32)
        counts = np.zeros((height, width), np.uint32)
        grey = np.zeros((height, width), np.uint8)
        self._check(self._lib.stellar_render(
            self._h, _fp(v), int(width), int(height),
            counts.ctypes.data_as(C.POINTER(C.c_uint32)),
            grey.ctypes.data_as(C.POINTER(C.c_uint8))), "render")
        return counts, grey

    # -- stages ---------------------------------------------------------
    def bbox(self) -> None:
        self._check(self._lib.stellar_bbox(self._h), "bbox")

    def build(self) -> None:
        self._check(self._lib.stellar_build(self._h), "build")

    def com(self) -> None:
        self._check(self._lib.stellar_com(self._h), "com")

    def sort(self) -> None:
        self._check(self._lib.stellar_sort(self._h), "sort")

    def forces(self) -> None:
        self._check(self._lib.stellar_forces(self._h), "forces")

    def update_stage(self) -> None:
        self._check(self._lib.stellar_update(self._h), "update")

    def update(self) -> None:
        """One whole synchronous step (BarnesHutParticleSystem::update)."""
        self._check(self._lib.stellar_step(self._h), "step")

    step = update

    def run(self, steps: int) -> None:
        self._check(self._lib.stellar_run(self._h, int(steps)), "run")

    def sync(self) -> None:
        self._check(self._lib.stellar_sync(self._h), "sync")

    # -- inspection -----------------------------------------------------
    def get_bbox(self, instance: int = 0) -> np.ndarray:
        b = np.zeros(4, np.float32)
        self._check(self._lib.stellar_get_bbox_instance(self._h, int(instance), _fp(b)), "get_bbox")
        return b

    def num_cells(self) -> int:
        c = C.c_int64()
        self._check(self._lib.stellar_num_cells(self._h, C.byref(c)), "num_cells")
        return c.value

    def export_tree(self, with_com: bool = True, instance: int = 0) -> dict[str, np.ndarray]:
        ncell = self.num_cells()
        cap = max(ncell, 1)
        ni = int(self.offsets[instance + 1] - self.offsets[instance])
        t = {
            "root": np.zeros(4, np.int32),
            "child": np.zeros(4 * cap, np.int32),
            "count": np.zeros(cap, np.int32),
            "start": np.zeros(cap, np.int32),
            "level": np.zeros(cap, np.int32),
            "sorted": np.zeros(max(ni, 1), np.int32),
        }
        if with_com:
            for k in ("com_x", "com_y", "com_m"):
                t[k] = np.zeros(cap, np.float32)
        got = C.c_int64()
        self._check(
            self._lib.stellar_export_tree_instance(
                self._h, int(instance), _ip(t["root"]), _ip(t["child"]), _ip(t["count"]), _ip(t["start"]),
                _ip(t["level"]), _ip(t["sorted"]), _fp(t.get("com_x")), _fp(t.get("com_y")),
                _fp(t.get("com_m")), cap, C.byref(got)),
            "export_tree",
        )
        c = got.value
        out = {"root": t["root"], "child": t["child"][: 4 * c], "count": t["count"][:c],
               "start": t["start"][:c], "level": t["level"][:c], "sorted": t["sorted"][:ni],
               "ncells": c}
        if with_com:
            for k in ("com_x", "com_y", "com_m"):
                out[k] = t[k][:c]
        return out

    def export_forces(self) -> dict[str, np.ndarray]:
        ax = np.zeros(self.n, np.float32)
        ay = np.zeros(self.n, np.float32)
        it = np.zeros(self.n, np.int32)
        self._check(self._lib.stellar_export_forces(self._h, _fp(ax), _fp(ay), _ip(it)), "export_forces")
        return {"ax": ax, "ay": ay, "inter": it}

    def last_interactions(self) -> int:
        v = C.c_int64()
        self._check(self._lib.stellar_last_interactions(self._h, C.byref(v)), "last_interactions")
        return v.value

    def last_quad_interactions(self) -> int:
        v = C.c_int64()
        self._check(self._lib.stellar_last_quad_interactions(self._h, C.byref(v)), "last_quad_interactions")
        return v.value

    def stage_times(self) -> dict[str, float]:
        ms = np.zeros(7, np.float32)
        self._check(self._lib.stellar_stage_times(self._h, _fp(ms)), "stage_times")
        return dict(zip(STAGES, map(float, ms)))

    def direct_fp64(self, targets: np.ndarray | None = None, eps2: float | None = None):
        e2 = float(self.params.eps2) if eps2 is None else float(eps2)
        if targets is None:
            nt = self.n
            tp = None
        else:
            targets = np.ascontiguousarray(targets, dtype=np.int32)
            nt = targets.size
            tp = _ip(targets)
        ax = np.zeros(nt, np.float64)
        ay = np.zeros(nt, np.float64)
        self._check(self._lib.stellar_direct_fp64(self._h, nt, tp, e2, _dp(ax), _dp(ay)), "direct_fp64")
        return ax, ay

    def energy_fp64(self, eps2: float | None = None) -> dict[str, float]:
        e2 = float(self.params.eps2) if eps2 is None else float(eps2)
        v = [C.c_double() for _ in range(4)]
        self._check(self._lib.stellar_energy_fp64(self._h, e2, *(C.byref(q) for q in v)), "energy_fp64")
        return {"kinetic": v[0].value, "potential": v[1].value, "px": v[2].value, "py": v[3].value}

    def autotune(self, group_mask: int = 0, reps: int = 3) -> dict:
        t = StellarTuning()
        self._check(self._lib.stellar_autotune(self._h, int(group_mask), int(reps), C.byref(t)), "autotune")
        self.params.group = t.group
        return {"group": t.group, "nodes_per_pop": t.nodes_per_pop,
                "block_threads": t.block_threads, "walk_ms": t.walk_ms, "schedule": t.schedule}

    def set_walk_schedule(self, schedule: int) -> None:
        """WALK_STATIC, WALK_QUEUE or WALK_QUEUE_WAVE (stellar_set_walk_schedule)."""
        self._check(self._lib.stellar_set_walk_schedule(self._h, int(schedule)), "set_walk_schedule")

    def set_walk_shape(self, nodes_per_pop: int, block_threads: int) -> None:
        self._check(self._lib.stellar_set_walk_shape(self._h, int(nodes_per_pop), int(block_threads)),
                    "set_walk_shape")

    def diagnostics_fp64(self, eps2: float | None = None, half_dt: float = 0.0) -> dict[str, np.ndarray]:
        """Per-instance FP64 kinetic and potential energy, momentum and angular
        momentum about the origin; half_dt != 0 takes velocities as
        v + half_dt * a with the accelerations of the last forces stage."""
        e2 = float(self.params.eps2) if eps2 is None else float(eps2)
        out = (StellarDiagnostics * self.ninst)()
        self._check(self._lib.stellar_diagnostics_fp64(self._h, e2, float(half_dt), out), "diagnostics_fp64")
        return {k: np.array([getattr(d, k) for d in out], np.float64)
                for k, _ in StellarDiagnostics._fields_}

    def energy_fp64_instances(self, eps2: float | None = None) -> dict[str, np.ndarray]:
        e2 = float(self.params.eps2) if eps2 is None else float(eps2)
        v = {k: np.zeros(self.ninst, np.float64) for k in ("kinetic", "potential", "px", "py")}
        self._check(self._lib.stellar_energy_fp64_instances(
            self._h, e2, *(_dp(v[k]) for k in ("kinetic", "potential", "px", "py"))), "energy_fp64_instances")
        return v


__all__ = [
    "Engine", "StellarParams", "StellarICParams", "StellarTuning", "StellarDiagnostics", "StellarError",
    "load_library",
    "default_params", "ic_params", "generate_ic", "LIB_PATH", "SIGNATURES", "STAGES",
]
