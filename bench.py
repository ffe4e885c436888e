#!/usr/bin/env python3
"""Barnes-Hut step throughput on one MI355X (libstellar_hip).

Workload (default): colliding disks, N = 2^20 (two disks of 524,288 bodies,
central masses 1e4, r ~ U[2,200], centres (+-300, 0), bulk velocities
(-+5, +-1)), theta = 0.5, eps = 0.01, dt = 1e-3, seed 2, FP32 state.
Force model (default): cell quadrupoles with the textbook opening test
w/d <= theta, the configuration that holds the rms force error under 1e-3;
the reference's own criterion (monopoles, w^2/theta <= d^2 + eps^2) is run in
the same invocation and reported beside it ("reference_opening").

One "step" = the six stages (bbox, build, centre of mass, tree-order sort,
tree walk, update) plus the output buffer, over bodies already resident in
HBM.  Prints one JSON line: body-steps/s, per-stage times, the walk's
roofline position (interaction rate against FP32 VALU issue peak), the rms
relative force error against an FP64 direct sum on a seeded 8192-body
subsample, FP64 energy, momentum and angular-momentum drift (velocities
synchronised with the positions), and the reference algorithm timed on
the host cores (the CPU oracle) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# MI355X: 256 CUs x 4 SIMD16 x 2.4 GHz = 39.3e12 FP32 VALU lane-ops/s (non-packed)
VALU_PEAK_LANE_OPS = 256 * 4 * 16 * 2.4e9
# algorithmic VALU issue slots of one body-node interaction: 2 sub, 2 FMA for
# d^2+eps^2, 1 rsqrt at quarter rate (4 slots), 3 mul, 2 FMA accumulate
OPS_PER_INTERACTION = 13
# a cell's quadrupole term, in the combined form the walk evaluates
# (rs^5 [(m r + (2.5 (d.Q.d) + T) rs^2) d - Q.d], T the softening term read
# with the moments), issues 11 more than the monopole's 3 multiplies (already
# in the 13): rs^2, rs^5 (2), Q.d (2 mul + 2 FMA), d.Q.d (mul + FMA), the
# radial factor (mul + 2 FMA), one FMA per component (2), less the
# monopole's 3 multiplies
OPS_PER_QUAD_TERM = 11
HBM_PEAK_GBS = 8000.0
# Counter figures of the default workload (colliding disks, N = 2^20, theta 0.5,
# quadrupoles + textbook opening, group 64, Hilbert order) from rocprofv3 runs of this script
# (profiles/hbm_traffic_colliding_disks_1M.md, profiles/walk_counters_colliding_disks_1M.md).
DEFAULT_KEY = (2, 1 << 20, 0.5, 64, "quad", "textbook", "hilbert")
# HBM bytes per step of each stage's kernels: FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0
# (KiB per dispatch), the factors being what the same counters read for a 1 GiB
# device copy; "forces" is the walk launch (with the update inside it).
STAGE_TRAFFIC_BYTES = {DEFAULT_KEY: {
    "bbox": (4130 * 2 + 8) * 1024,
    "build": (89755 * 2 + 193908) * 1024,
    "com": (32028 * 2 + 39696) * 1024,
    "forces": (161850 * 2 + 27876) * 1024,
}}
# The walk's VALU busy fraction: SQ_ACTIVE_INST_VALU / (256 CUs x GRBM_GUI_ACTIVE
# per XCD) per dispatch, 413.9e6 / (256 x 15.62e6 / 8).
WALK_VALU_BUSY = {DEFAULT_KEY: 413.9e6 / (256 * 15.62e6 / 8)}


def progress(msg: str) -> None:
    """One progress line on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--workload", type=int, default=2, help="initial-condition workload (0,1,2,4)")
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--eps", type=float, default=0.01)
    ap.add_argument("--dt", type=float, default=1e-3)
    ap.add_argument("--group", type=int, default=64)
    ap.add_argument("--multipole", choices=("quad", "mono"), default="quad",
                    help="cell terms: monopole+quadrupole (default) or monopole only")
    ap.add_argument("--opening", choices=("textbook", "reference", "bmax"), default="textbook",
                    help="opening test: textbook w/d <= theta (default), the reference's "
                         "w^2/theta <= d^2+eps^2, or bmax/d <= theta with a per-cell bound")
    ap.add_argument("--order", choices=("hilbert", "slot"), default="hilbert",
                    help="storage order of the bodies, and so the walk groups: along a Hilbert "
                         "curve of the tree (default) or the reference's depth-first slot order")
    ap.add_argument("--no-reference-opening", action="store_true",
                    help="skip the side run with the reference's criterion")
    ap.add_argument("--error-sample", type=int, default=8192)
    ap.add_argument("--no-energy", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=3, help="timed oracle steps after 1 warm-up (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--ensemble", action="store_true",
                    help="batched ensemble workload: 256 instances, N_i log-uniform in "
                         "[1024, 65536] (multiples of 64), models cycling disk / Plummer / "
                         "colliding disks, theta 0.5, eps 0.01, dt 1e-3, graph-captured steps, "
                         "autotuned walk shape; compared with the same instances run one "
                         "context at a time")
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device three times (a plain copy of known "
                         "size that scales rocprofv3's FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--graph", action="store_true", help="replay captured two-step graphs")
    ap.add_argument("--accuracy-sweep", type=str, default="",
                    help="comma-separated thetas: after --warmup steps, one JSON line per theta "
                         "with the force error against the FP64 direct sum (rms, p50, p99, max, "
                         "rms by decile of |a| and of radius from the median position) and the "
                         "walk's time, all on the same positions (no step line)")
    ap.add_argument("--precision", choices=("fp32", "mixed"), default="fp32",
                    help="force path: FP32 (default) or mixed (accepted cells' terms in packed "
                         "FP16 with Kahan-compensated FP32 sums; needs --multipole quad)")
    return ap.parse_args()


def rms_rel(ax, ay, rx, ry):
    num = (ax.astype(np.float64) - rx) ** 2 + (ay.astype(np.float64) - ry) ** 2
    den = rx ** 2 + ry ** 2
    q = num / den
    return float(np.sqrt(q.mean())), float(np.sqrt(np.percentile(q, 99))), float(np.sqrt(q.max()))


def cpu_baseline(st, args, eps2):
    """The reference algorithm (oracle restatement: sequential insertion build,
    FP32 centre of mass, DFS sort, group-of-32 vote walk, kick-drift) on the
    host cores, same workload; median per-step time after one warm-up.  Each
    timed step starts from the initial state: within a few steps two bodies of
    this workload land on the same FP32 position, which the reference's
    insertion build cannot separate (it splits until its pool overflows)."""
    from oracle import oracle as O

    threads = args.cpu_threads if args.cpu_threads > 0 else host_cores()
    times = []
    for k in range(args.cpu_steps + 1):
        state = {q: st[q].copy() for q in ("x", "y", "vx", "vy", "mass")}
        t0 = time.perf_counter()
        O.step(state, args.theta, eps2, args.dt, 1.0, group=32, threads=threads)
        t1 = time.perf_counter()
        if k > 0:
            times.append(t1 - t0)
        progress(f"cpu baseline step {k} ({'timed' if k else 'warm-up'}): {(t1 - t0) * 1e3:.0f} ms")
    med = statistics.median(times)
    n = st["x"].size
    return {
        "value": n / med,
        "unit": "body-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"N={n} {WORKLOAD_NAMES.get(args.workload, args.workload)}, {args.cpu_steps} "
                  f"steps from the initial state after "
                  f"1 warm-up, median {med * 1e3:.1f} ms/step, reference criterion, monopoles, "
                  "group 32, OpenMP walk, sequential insertion build",
    }


def host_cores() -> int:
    """Cores this process may run on: its CPU affinity, bounded by
    OMP_NUM_THREADS when that is set (a host that shares its cores between
    GPUs sets it to one GPU's share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


WORKLOAD_NAMES = {0: "disk", 1: "Plummer a=1", 2: "colliding disks", 4: "deep Plummer a=0.01"}


def model_flags(S, args):
    f = S.FLAG_QUADRUPOLE if args.multipole == "quad" else 0
    if args.opening == "textbook":
        f |= S.FLAG_TEXTBOOK_OPENING
    elif args.opening == "bmax":
        f |= S.FLAG_BMAX_OPENING
    if args.order == "hilbert":
        f |= S.FLAG_HILBERT_ORDER
    return f


def multipoles(args):
    return "monopole+quadrupole" if args.multipole == "quad" else "monopole"


def opening_text(args):
    return {"textbook": "textbook (w/d <= theta, unsoftened d",
            "reference": "reference (w^2/theta <= d^2+eps^2",
            "bmax": "bmax (bmax/d <= theta"}[args.opening] + ", group vote)"


def force_error(eng, n, sample):
    """One untimed force pass on the current state against the FP64 direct
    sum on a seeded subsample (rms, p99, max relative error)."""
    eng.bbox()
    eng.build()
    eng.com()
    eng.sort()
    eng.forces()
    f = eng.export_forces()
    rng = np.random.default_rng(2)
    tg = np.sort(rng.choice(n, min(n, sample), replace=False)).astype(np.int32)
    dx, dy = eng.direct_fp64(tg)
    return rms_rel(f["ax"][tg], f["ay"][tg], dx, dy), int(tg.size)


def stage_bytes(n, cells):
    """Algorithmic HBM bytes of each memory-bound stage per step (N bodies,
    C cells): bbox 8N (x, y read); build 8N + 20C + 48N (x, y read; per cell
    its four 4-byte child links and a 4-byte first-body index written; the
    six body arrays moved into tree order, 24 B read and 24 B written per
    body, which the sort stage then swaps in); centre of mass 12N + 12C
    (bodies' x, y, m read; a cell's x, y, m written); update 40N (x, y, vx,
    vy, ax, ay read, x, y, vx, vy written) when it runs as its own kernel
    (walks that pop one node at a time do it inside the walk).  What a stage
    moves at least, not what it fetches: the walk's second 64-byte block per
    node (quadrupole moments and opening thresholds) is a term of this
    layout on top of these."""
    return {"bbox": 8 * n, "build": 8 * n + 20 * cells + 48 * n, "com": 12 * n + 12 * cells,
            "update": 40 * n}


def stage_roofline(stages, n, cells, key):
    """Per memory-bound stage: algorithmic bytes over the measured time
    against HBM peak, and, for a workload with committed counter runs, the
    counter bytes and their ratio to the algorithmic bytes."""
    traffic = STAGE_TRAFFIC_BYTES.get(key, {})
    out = {}
    for k, b in stage_bytes(n, cells).items():
        ms = stages.get(k, 0.0)
        if ms <= 1e-3:  # a stage under 1 us ran inside another
            continue
        gbs = b / (ms * 1e-3) / 1e9
        t = traffic.get(k)
        out[k] = {"bound": "hbm", "ms": ms, "algorithmic_bytes": b, "achieved": gbs,
                  "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                  "traffic": t, "traffic_over_algorithmic": t / b if t else None}
    return out


def synced_diagnostics(eng, dt, forces):
    """FP64 energy, momentum and angular momentum with velocities brought
    level with the positions (v + dt/2 a), summed over instances; forces=True
    first evaluates the accelerations at the current positions."""
    if forces:
        for stage in ("bbox", "build", "com", "sort", "forces"):
            getattr(eng, stage)()
    d = eng.diagnostics_fp64(half_dt=0.5 * dt)
    return {k: float(np.sum(v)) for k, v in d.items()}


def reference_opening_run(S, args, st, eps2):
    """The reference's force model (monopoles, w^2/theta <= d^2+eps^2, walk
    groups of 64 consecutive bodies of its depth-first slot order) on the same
    workload: speed and error."""
    n = st["x"].size
    p = S.default_params(theta=args.theta, eps2=eps2, timestep=args.dt, group=args.group,
                         flags=S.FLAG_STAGE_TIMING)
    with S.Engine(n, p) as e:
        e.upload(st["mass"], st["x"], st["y"], st["vx"], st["vy"])
        e.run(args.warmup)
        e.sync()
        stages = e.stage_times() if args.warmup > 0 else {}  # split from the warm-up
        e.set_params(flags=S.FLAG_FORCES_TIMING)  # timed steps: walk events only
        k = max(2, min(args.steps, 20))
        t0 = time.perf_counter()
        e.run(k)
        t1 = time.perf_counter()
        stages = dict(stages, forces=e.stage_times()["forces"])
        (err, p99, emax), ns = force_error(e, n, args.error_sample)
    return {"value": n * k / (t1 - t0), "unit": "body-steps/s", "steps": k,
            "ms_per_step": (t1 - t0) / k * 1e3, "rms_rel_force_err": err,
            "force_err_p99": p99, "force_err_max": emax, "stage_ms": stages,
            "opening": "reference (w^2/theta <= d^2+eps^2, group vote)",
            "multipoles": "monopole", "order": "slot"}


def ensemble_members(S, ninst, seed=3):
    """N_i log-uniform in [1024, 65536] rounded to multiples of 64; models
    cycle disk (workload 0), Plummer (1), colliding disks (2); instance i uses
    seed 1000*seed + i."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(1024), np.log(65536), ninst))
    counts = (np.round(counts / 64) * 64).astype(np.int64).clip(1024, 65536)
    sts = []
    for i, n in enumerate(counts):
        p = S.ic_params(i % 3, seed=1000 * seed + i)
        sts.append(S.generate_ic(p, int(n)))
    return [int(c) for c in counts], sts


def precision_of(S, args):
    return S.PRECISION_MIXED if args.precision == "mixed" else S.PRECISION_FP32


def dtype_of(args):
    # arithmetic of the path: FP32 throughout, or packed FP16 cell terms in
    # FP32 sums
    return "f16+f32" if args.precision == "mixed" else "f32"


def run_ensemble(args, S):
    eps2 = float(np.float32(args.eps * args.eps))
    prec = precision_of(S, args)
    counts, sts = ensemble_members(S, args.instances)
    ntot = int(sum(counts))
    cat = {k: np.concatenate([s[k] for s in sts]) for k in ("mass", "x", "y", "vx", "vy")}
    base_flags = model_flags(S, args)

    # batched form: one context, graph-captured steps, autotuned walk shape
    p = S.default_params(theta=args.theta, eps2=eps2, timestep=args.dt, group=args.group,
                         precision=prec, flags=base_flags | S.FLAG_STAGE_TIMING)
    eng = S.Engine(params=p, counts=counts)
    eng.upload(cat["mass"], cat["x"], cat["y"], cat["vx"], cat["vy"])
    tuning = eng.autotune(group_mask=0, reps=2)
    progress(f"ensemble of {len(counts)} instances ({ntot} bodies) tuned: {tuning}")
    eng.run(2)  # eager step timing for the stage split
    stages = eng.stage_times()
    inter = eng.last_interactions()
    qinter = eng.last_quad_interactions()
    eng.set_params(flags=base_flags | S.FLAG_GRAPH)
    eng.run(args.warmup + (args.warmup % 2))
    eng.sync()
    t0 = time.perf_counter()
    eng.run(args.steps)
    t1 = time.perf_counter()
    wall = t1 - t0
    batched = ntot * args.steps / wall
    progress(f"batched: {wall / args.steps * 1e3:.3f} ms/step")
    # accuracy on a sample of instances
    eng.bbox(); eng.build(); eng.com(); eng.sort(); eng.forces()
    f = eng.export_forces()
    rng = np.random.default_rng(4)
    off = np.concatenate([[0], np.cumsum(counts)])
    tg = []
    for i in rng.choice(len(counts), min(16, len(counts)), replace=False):
        tg.append(off[i] + rng.choice(counts[i], min(512, counts[i]), replace=False))
    tg = np.sort(np.concatenate(tg)).astype(np.int32)
    dx, dy = eng.direct_fp64(tg)
    err_rms, err_p99, err_max = rms_rel(f["ax"][tg], f["ay"][tg], dx, dy)
    eng.close()

    # plain form: the same instances, one context each (the unbatched API),
    # in the three ways a caller could drive them: one after another with
    # direct steps, one after another with graph-replayed steps, and from
    # eight host threads at once (every context has its own stream; the
    # library's calls release the interpreter lock) with graph-replayed steps
    plain_steps = max(2, min(args.steps, 10))
    engines = []
    for st in sts:
        q = S.default_params(theta=args.theta, eps2=eps2, timestep=args.dt, group=args.group,
                             precision=prec, flags=base_flags)
        e = S.Engine(st["x"].size, q)
        e.upload(st["mass"], st["x"], st["y"], st["vx"], st["vy"])
        engines.append(e)

    def timed(fn):
        t0 = time.perf_counter()
        fn()
        return ntot * plain_steps / (time.perf_counter() - t0)

    def sequential():
        for e in engines:
            e.run(plain_steps)

    for e in engines:
        e.run(1)
    plain_forms = {"sequential_direct": timed(sequential)}
    for e in engines:
        e.set_params(flags=base_flags | S.FLAG_GRAPH)
        e.run(2)  # captures the graphs
    plain_forms["sequential_graph"] = timed(sequential)

    def threaded():
        import threading

        def worker(k):
            for e in engines[k::8]:
                e.run(plain_steps)

        ts = [threading.Thread(target=worker, args=(k,)) for k in range(8)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()

    plain_forms["threads8_graph"] = timed(threaded)
    plain = max(plain_forms.values())
    progress("plain forms (body-steps/s): " +
             ", ".join(f"{k} {v:.3e}" for k, v in plain_forms.items()))
    for e in engines:
        e.close()

    walk_ms = stages["forces"]
    achieved = (inter * OPS_PER_INTERACTION + qinter * OPS_PER_QUAD_TERM) / (walk_ms * 1e-3) / 1e12
    peak = VALU_PEAK_LANE_OPS / 1e12
    out = {
        "metric": "BH body-steps/s, batched ensemble of 256 instances (N_i 1024..65536)",
        "value": batched,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype_of(args),
        "data": "synthetic",
        "config": {"workload": f"ensemble of {len(counts)} instances, N_i log-uniform "
                               "[1024, 65536] (x64), disk/Plummer/colliding disks",
                   "n_bodies": ntot, "theta": args.theta, "eps": args.eps, "dt": args.dt,
                   "precision": args.precision, "graph": True, "tuning": tuning,
                   "multipoles": multipoles(args), "opening": opening_text(args),
                   "order": args.order},
        "plain_form_body_steps_per_s": plain,
        "plain_form": f"{len(counts)} single-instance contexts, {plain_steps} steps each; "
                      "the fastest of: one after another with direct steps, one after "
                      "another with graph-replayed steps, eight host threads with "
                      "graph-replayed steps",
        "plain_forms_body_steps_per_s": plain_forms,
        "speedup_over_plain_form": batched / plain,
        "same_error_as_plain_form": "every instance of the batched form evolves bit for bit "
                                    "as its single context (same forces, same error)",
        "rms_rel_force_err": err_rms,
        "force_err_p99": err_p99,
        "force_err_max": err_max,
        "force_err_sample": int(tg.size),
        "stage_ms": stages,
        "interactions_per_step": inter,
        "roofline": {"bound": "valu", "kernel": "walk64_kernel", "achieved": achieved,
                     "peak": peak, "unit": "Tlane-op/s", "frac": achieved / peak,
                     "traffic": None},
    }
    print(json.dumps(out))
    return 0


def decile_rms(q, key):
    """rms of the squared relative errors q within each decile of key."""
    edges = np.percentile(key, np.linspace(0, 100, 11))
    b = np.clip(np.searchsorted(edges, key, side="right") - 1, 0, 9)
    return [float(np.sqrt(q[b == k].mean())) if np.any(b == k) else None for k in range(10)]


def accuracy_sweep(args, S):
    """Where the force error sits and what a smaller theta buys: the state
    after --warmup steps, the FP64 direct sum on a seeded subsample, then
    for each theta one tree and walk (the configured multipoles and opening
    test) with the per-body relative errors binned by |a| and by radius,
    and the walk's time per theta."""
    n = args.n
    eps2 = float(np.float32(args.eps * args.eps))
    st = S.generate_ic(S.ic_params(args.workload), n)
    p = S.default_params(theta=args.theta, eps2=eps2, timestep=args.dt, group=args.group,
                         precision=precision_of(S, args), flags=model_flags(S, args))
    eng = S.Engine(n, p)
    eng.upload(st["mass"], st["x"], st["y"], st["vx"], st["vy"])
    eng.run(args.warmup)
    eng.sync()
    rng = np.random.default_rng(2)
    tg = np.sort(rng.choice(n, min(n, args.error_sample), replace=False)).astype(np.int32)
    rx, ry = eng.direct_fp64(tg)
    xy = eng.get_output_buffer()
    x, y = xy[0::2][tg].astype(np.float64), xy[1::2][tg].astype(np.float64)
    radius = np.hypot(x - np.median(x), y - np.median(y))
    aref = np.hypot(rx, ry)
    thetas = [float(t) for t in args.accuracy_sweep.split(",")]
    rows = []
    for theta in thetas:
        eng.set_params(theta=theta, flags=model_flags(S, args))
        eng.bbox(); eng.build(); eng.com(); eng.sort(); eng.forces()
        f = eng.export_forces()
        q = ((f["ax"][tg] - rx) ** 2 + (f["ay"][tg] - ry) ** 2) / (rx ** 2 + ry ** 2)
        dabs = (f["ax"][tg] - rx) ** 2 + (f["ay"][tg] - ry) ** 2
        rows.append({
            "workload": WORKLOAD_NAMES.get(args.workload, f"workload {args.workload}"),
            "n_bodies": n, "eps": args.eps, "steps_before": args.warmup, "theta": theta,
            "multipoles": multipoles(args), "opening": opening_text(args), "order": args.order,
            "precision": args.precision, "sample": int(tg.size),
            "rms_rel_force_err": float(np.sqrt(q.mean())),
            "p50": float(np.sqrt(np.median(q))), "p99": float(np.sqrt(np.percentile(q, 99))),
            "max": float(np.sqrt(q.max())),
            "rms_abs_err_over_rms_a": float(np.sqrt(dabs.mean() / (aref ** 2).mean())),
            "rms_by_abs_a_decile": decile_rms(q, aref),
            "rms_by_radius_decile": decile_rms(q, radius),
            "interactions_per_body": eng.last_interactions() / n})
    # walk time per theta (HIP events around the walk over three steps; the
    # steps move the bodies, so this comes after every error above)
    for theta, row in zip(thetas, rows):
        eng.set_params(theta=theta, flags=model_flags(S, args) | S.FLAG_FORCES_TIMING)
        eng.run(3)
        eng.sync()
        row["walk_ms"] = eng.stage_times()["forces"]
        print(json.dumps(row), flush=True)
    eng.close()
    return 0


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU figures are not "
                          "produced (the walk would shard by contiguous ranges of the "
                          "tree order with one position all-gather per step)"}))
        return 2
    if args.copy_calibration:
        import torch

        a = torch.ones(256 << 20, device="cuda", dtype=torch.float32)  # 1 GiB
        torch.cuda.synchronize()
        for _ in range(3):
            b = a.clone()
            torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": b.numel() * 4, "copies": 3}))
        return 0
    import stellarsolver_amd as S

    S.load_library()
    if args.ensemble:
        return run_ensemble(args, S)
    if args.accuracy_sweep:
        return accuracy_sweep(args, S)
    n = args.n
    eps2 = float(np.float32(args.eps * args.eps))
    st = S.generate_ic(S.ic_params(args.workload), n)
    prec = precision_of(S, args)
    p = S.default_params(theta=args.theta, eps2=eps2, timestep=args.dt, dampening=1.0,
                         group=args.group, precision=prec,
                         flags=S.FLAG_STAGE_TIMING | model_flags(S, args)
                         | (S.FLAG_GRAPH if args.graph else 0))
    eng = S.Engine(n, p)
    eng.upload(st["mass"], st["x"], st["y"], st["vx"], st["vy"])
    progress(f"N={n} uploaded")
    e0 = None if args.no_energy else synced_diagnostics(eng, args.dt, forces=True)

    if args.graph:
        # stage times come from direct steps (graph replays record none)
        eng.set_params(flags=p.flags & ~S.FLAG_GRAPH)
        eng.run(2)
        stages = eng.stage_times()
        eng.set_params(flags=p.flags | S.FLAG_GRAPH)
    if not args.graph and args.warmup >= 2:
        eng.run(1)  # the first sort after the upload starts from an arbitrary order
        eng.run(args.warmup - 1)
    else:
        eng.run(args.warmup)
    eng.sync()
    if not args.graph:
        # the warm-up steps after the first give the per-stage split (eight
        # HIP events per step); the timed steps record only the two events
        # around the walk, whose time enters the roofline (the other six cost
        # ~35 us a step)
        stages = eng.stage_times() if args.warmup > 0 else {}
        eng.set_params(flags=(p.flags & ~S.FLAG_STAGE_TIMING) | S.FLAG_FORCES_TIMING)
    progress("warm-up done")
    t0 = time.perf_counter()
    eng.run(args.steps)  # synchronises once at the end
    t1 = time.perf_counter()
    progress(f"{args.steps} timed steps: {(t1 - t0) / args.steps * 1e3:.3f} ms/step")
    if not args.graph:
        # mean device ms per timed step, HIP events on the engine stream
        stages = dict(stages, forces=eng.stage_times()["forces"])
        eng.set_params(flags=p.flags)
    inter = eng.last_interactions()
    qinter = eng.last_quad_interactions()
    ncells = eng.num_cells()
    wall = t1 - t0
    ms_per_step = wall / args.steps * 1e3

    # accuracy of the walk on the final state: one more untimed force pass
    (err_rms, err_p99, err_max), nsample = force_error(eng, n, args.error_sample)
    progress(f"force error {err_rms:.2e}")
    drift = None
    if e0 is not None:
        # the force pass of the error check left forces at the final positions
        e1 = synced_diagnostics(eng, args.dt, forces=False)
        E0 = e0["kinetic"] + e0["potential"]
        E1 = e1["kinetic"] + e1["potential"]
        msum = float(st["mass"].astype(np.float64).sum())
        vsum = float((st["mass"].astype(np.float64) * np.hypot(st["vx"], st["vy"])).sum())
        drift = {"dE_over_E": (E1 - E0) / abs(E0),
                 "dP_over_sum_m_v": float(np.hypot(e1["px"] - e0["px"], e1["py"] - e0["py"]) / vsum),
                 "dLz_over_Lz": (e1["lz"] - e0["lz"]) / abs(e0["lz"]),
                 "velocities": "synchronised, v + dt/2 a(x) (the update stores v half a step behind x)",
                 "steps": args.warmup + args.steps,
                 "total_mass": msum}
    graph_replay = None
    if not args.graph and args.steps >= 2:
        # side figure: the same number of steps replayed from captured
        # two-step graphs (`value` stays the direct launches, whose walk the
        # HIP events time)
        try:
            eng.set_params(flags=p.flags | S.FLAG_GRAPH)
            eng.run(2)
            eng.sync()
            k = args.steps - args.steps % 2
            tg0 = time.perf_counter()
            eng.run(k)
            tg1 = time.perf_counter()
            graph_replay = {"value": n * k / (tg1 - tg0), "unit": "body-steps/s",
                            "ms_per_step": (tg1 - tg0) / k * 1e3, "steps": k}
        except S.StellarError as err:  # a side figure: report it, keep the line
            graph_replay = {"error": str(err)}
    eng.close()

    walk_ms = stages["forces"]
    achieved = (inter * OPS_PER_INTERACTION + qinter * OPS_PER_QUAD_TERM) / (walk_ms * 1e-3) / 1e12
    peak = VALU_PEAK_LANE_OPS / 1e12
    model_key = (args.workload, n, args.theta, args.group, args.multipole, args.opening,
                 args.order)
    if args.precision != "fp32" or args.graph:
        model_key = None  # counter runs exist for the default FP32 direct launches only
    out = {
        "metric": "BH body-steps/s at N=1M, theta=0.5 + rms rel force err vs FP64 direct sum",
        "value": n * args.steps / wall,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype_of(args),
        "data": "synthetic",
        "config": {"workload": WORKLOAD_NAMES.get(args.workload, f"workload {args.workload}"),
                   "n_bodies": n, "theta": args.theta, "eps": args.eps, "dt": args.dt,
                   "group": args.group, "precision": args.precision, "seed": args.workload,
                   "graph": bool(args.graph),
                   "opening": opening_text(args), "multipoles": multipoles(args),
                   "order": args.order},
        "rms_rel_force_err": err_rms,
        "force_err_p99": err_p99,
        "force_err_max": err_max,
        "force_err_sample": nsample,
        "stage_ms": stages,
        "stage_ms_source": ("HIP events: forces over the timed steps, the other stages over the "
                            "warm-up steps after the first" if not args.graph else
                            "HIP events over two direct steps before the graph replays"),
        "stage_roofline": stage_roofline(stages, n, ncells, model_key),
        "cells": ncells,
        "interactions_per_step": inter,
        "interactions_per_body": inter / n,
        "roofline": {
            "bound": "valu",
            "kernel": "walk64_kernel",
            "achieved": achieved,
            "peak": peak,
            "unit": "Tlane-op/s",
            "frac": achieved / peak,
            "traffic": STAGE_TRAFFIC_BYTES.get(model_key, {}).get("forces"),
            "valu_busy_counters": WALK_VALU_BUSY.get(model_key),
            "interactions_per_s": inter / (walk_ms * 1e-3),
            "ops_per_interaction": OPS_PER_INTERACTION,
            "ops_per_quadrupole_term": OPS_PER_QUAD_TERM,
            "quadrupole_interactions": qinter,
        },
        "drift": drift,
        "graph_replay": graph_replay,
        "output_buffer": "filled when read (stellar_download_positions), outside the timed "
                         "steps; the reference fills it every step (about 5 us at N = 2^20)",
    }
    if not args.no_reference_opening and args.opening != "reference":
        out["reference_opening"] = reference_opening_run(S, args, st, eps2)
        progress("reference-opening side run done")
    if args.cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(st, args, eps2)
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
        if "reference_opening" in out:
            # the same force model on both sides (the CPU port runs the
            # reference's criterion with monopoles)
            out["gpu_over_cpu_same_force_model"] = (out["reference_opening"]["value"]
                                                    / out["cpu_baseline"]["value"])
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
