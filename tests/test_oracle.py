"""Pins for the CPU oracle (oracle/oracle.c) -- CPU only.

Every oracle function is checked against something that does not share its
code: hand-computed known answers, closed forms, an independent construction
of the same object, brute force, or the reference's own debug.cpp checker
(oracle/_ref, built from the reference source where it is available).
"""
from fractions import Fraction

import numpy as np
import pytest

from tests.helpers import ic, tree_equal


# ---------------------------------------------------------------- bbox ----
def test_bbox_closed_form_includes_origin(oracle):
    rng = np.random.default_rng(3)
    for lo, hi in ((1.0, 5.0), (-5.0, -1.0), (-3.0, 4.0)):
        x = rng.uniform(lo, hi, 1000).astype(np.float32)
        y = rng.uniform(lo, hi, 1000).astype(np.float32)
        b = oracle.bbox(x, y)
        exp = [min(0.0, x.min()), max(0.0, x.max()), min(0.0, y.min()), max(0.0, y.max())]
        assert np.array_equal(b, np.array(exp, np.float32))


def test_bbox_negative_zero_does_not_replace_origin(oracle):
    b = oracle.bbox(np.array([-0.0, 1.0], np.float32), np.array([2.0, -0.0], np.float32))
    assert not np.signbit(b[0]) and not np.signbit(b[2])


# ----------------------------------------------------- build: known answers
def _full_build(oracle, x, y, m=None):
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    m = np.ones_like(x) if m is None else np.asarray(m, np.float32)
    bb = oracle.bbox(x, y)
    t = oracle.build_insert(x, y, m, bb)
    oracle.com_ref(t)
    oracle.sort_ref(t)
    return bb, t, oracle.canonicalize(t), oracle.build_keys(x, y, bb)


def test_two_bodies_known_answer(oracle):
    # bbox [-1,1]^2, midpoints 0: (1,1) is east/north -> slot 0, (-1,-1) slot 3
    bb, t, can, kt = _full_build(oracle, [1, -1], [1, -1])
    assert list(bb) == [-1, 1, -1, 1]
    assert list(can["root"]) == [0, -1, -1, 1] == list(kt["root"])
    assert can["ncells"] == 0 == kt["ncells"]
    assert list(kt["sorted"]) == [0, 1]


def test_three_bodies_known_answer(oracle):
    # bbox [0,3]^2, mid 1.5: (1,1) -> slot 3; (2,2),(3,3) share slot 0 and are
    # split by a level-1 cell over [1.5,3]^2 with mid 2.25: (3,3) slot 0,
    # (2,2) slot 3.  DFS order: cell{body2, body1}, body0.
    n = 3
    bb, t, can, kt = _full_build(oracle, [1, 2, 3], [1, 2, 3])
    assert list(bb) == [0, 3, 0, 3]
    assert list(kt["root"]) == [n + 0, -1, -1, 0]
    assert list(kt["child"]) == [2, -1, -1, 1]
    assert list(kt["count"]) == [2] and list(kt["start"]) == [0] and list(kt["level"]) == [1]
    assert list(kt["sorted"]) == [2, 1, 0]
    assert tree_equal(can, kt)
    # reference pool layout: the cell took id n (kernels.cu:199)
    assert t["ncells"] == 1 and list(t["child"][4 * n: 4 * n + 4]) == [2, -1, -1, 1]
    # centre of mass of the cell: ((2+3)/2, (2+3)/2)
    assert t["cx"][n] == 2.5 and t["cy"][n] == 2.5 and t["cm"][n] == 2.0


def test_tie_goes_east_north(oracle):
    # bbox [0,2]^2, mid 1: (1,1) sits on both midpoints -> slot 0 (x<mid false)
    bb, t, can, kt = _full_build(oracle, [0, 2, 1], [0, 2, 1])
    assert list(kt["root"])[3] == 0
    cell = kt["root"][0] - 3
    # level-1 cell [1,2]^2 (mid 1.5): (2,2) slot 0, (1,1) slot 3
    assert list(kt["child"][4 * cell: 4 * cell + 4]) == [1, -1, -1, 2]


def test_midpoint_is_fp32_sum_halved(oracle):
    """kernels.cu:144: mid = 0.5*(l+r) with l+r rounded to FP32 first.
    When the FP32 sum rounds down, a body sitting exactly on fl(l+r)/2 is
    below the exact midpoint but not below the FP32 one: the reference sends
    it east (slot bit 0 clear); an exact-midpoint rule would send it west."""
    rng = np.random.default_rng(11)
    found = 0
    for _ in range(200000):
        r = np.float32(rng.uniform(1.0, 2.0))
        l = np.float32(r * np.float32(rng.uniform(0.3, 0.7)))
        s = np.float32(l + r)
        fmid = Fraction(float(s)) / 2
        emid = (Fraction(float(l)) + Fraction(float(r))) / 2
        if not fmid < emid:
            continue
        cand = np.float32(float(fmid))
        assert Fraction(float(cand)) == fmid
        bb = np.array([l, r, 0.0, 1.0], np.float32)
        k1, _ = oracle.path_keys(np.array([cand], np.float32), np.array([0.9], np.float32), bb)
        west = int(k1[0] >> np.uint64(62)) & 1
        assert west == 0
        found += 1
        if found >= 20:
            break
    assert found >= 5


# ----------------------------------------- build: two constructions agree --
@pytest.mark.parametrize("workload,n", [(0, 4096), (1, 20000), (2, 30000)])
def test_insertion_and_key_builds_agree(oracle, workload, n):
    st = ic(workload, n)
    bb, t, can, kt = _full_build(oracle, st["x"], st["y"], st["mass"])
    assert tree_equal(can, kt)
    # sorted = stable argsort of the path keys (the DFS-by-slot order)
    k1, k2 = oracle.path_keys(st["x"], st["y"], bb)
    order = np.lexsort((np.arange(n), k2, k1))
    assert np.array_equal(order, kt["sorted"])


def test_brute_force_cells_tiny(oracle):
    """A cell at level L exists iff >= 2 bodies share L slot codes."""
    rng = np.random.default_rng(5)
    for trial in range(30):
        n = int(rng.integers(2, 40))
        x = rng.normal(0, 1, n).astype(np.float32)
        y = rng.normal(0, 1, n).astype(np.float32)
        bb = oracle.bbox(x, y)
        k1, _ = oracle.path_keys(x, y, bb)
        kt = oracle.build_keys(x, y, bb)
        expect = 0
        for L in range(1, 33):
            pref = k1 >> np.uint64(64 - 2 * L)
            _, cnt = np.unique(pref, return_counts=True)
            expect += int((cnt >= 2).sum())
        assert kt["ncells"] == expect
        # every cell's count equals the number of bodies sharing its prefix
        for c in range(kt["ncells"]):
            L = kt["level"][c]
            b0 = kt["sorted"][kt["start"][c]]
            pref = k1 >> np.uint64(64 - 2 * L)
            assert kt["count"][c] == int((pref == pref[b0]).sum())


def test_reference_debug_checks_pass(oracle):
    if not oracle.ref_debug_available():
        pytest.skip("reference debug.cpp not built (oracle/_ref)")
    st = ic(2, 8192)
    bb, t, can, kt = _full_build(oracle, st["x"], st["y"], st["mass"])
    ref = oracle.to_reference_layout(kt, 8192)
    m = ref["m"]
    lines = oracle.ref_debug_run(np.zeros(m), np.zeros(m), np.zeros(m), ref["count"],
                                 np.zeros(m, np.int32), ref["sorted"], ref["child"], bb, 8192, m)
    assert len(lines) == 8 and all("PASS" in ln for ln in lines), lines
    assert oracle.invariants(8192, m, ref["count"], ref["sorted"], ref["child"], bb) == 0


def test_invariants_catch_planted_mistakes(oracle):
    """Plant mistakes in an oracle tree; the restated checks and the
    reference's own debug.cpp must both flag them."""
    st = ic(0, 4096)
    bb, t, can, kt = _full_build(oracle, st["x"], st["y"], st["mass"])
    base = oracle.to_reference_layout(kt, 4096)
    m = base["m"]
    cases = []
    bad = {k: v.copy() for k, v in base.items() if k != "m"}
    bad["count"][4096] += 1
    cases.append((bad, 1))
    bad = {k: v.copy() for k, v in base.items() if k != "m"}
    bad["sorted"][5] = bad["sorted"][6]
    cases.append((bad, 32 | 64))
    bad = {k: v.copy() for k, v in base.items() if k != "m"}
    i = int(np.flatnonzero(bad["child"][4 * 4096:] >= 0)[0]) + 4 * 4096
    bad["child"][i] = -2
    cases.append((bad, 8))
    for arrays, bits in cases:
        got = oracle.invariants(4096, m, arrays["count"], arrays["sorted"], arrays["child"], bb)
        assert got & bits, (got, bits)
        if oracle.ref_debug_available():
            lines = oracle.ref_debug_run(np.zeros(m), np.zeros(m), np.zeros(m), arrays["count"],
                                         np.zeros(m, np.int32), arrays["sorted"], arrays["child"],
                                         bb, 4096, m)
            assert any("FAIL" in ln for ln in lines)


def test_coincident_bodies(oracle):
    x = np.array([0.5, 0.5, 0.5, -1.0], np.float32)
    y = np.array([0.25, 0.25, 0.25, 2.0], np.float32)
    bb = oracle.bbox(x, y)
    with pytest.raises(oracle.OracleError) as e:
        oracle.build_insert(x, y, np.ones(4, np.float32), bb)
    assert e.value.status == oracle.OR_ERR_DEPTH
    kt = oracle.build_keys(x, y, bb)
    deep = [c for c in range(kt["ncells"]) if kt["level"][c] == oracle.MAX_LEVEL]
    assert len(deep) == 1
    c = deep[0]
    assert list(kt["child"][4 * c: 4 * c + 4]) == [0, 1, 2, -1]  # by rank (id)
    x5 = np.full(5, 0.5, np.float32)
    with pytest.raises(oracle.OracleError) as e:
        oracle.build_keys(x5, x5, oracle.bbox(x5, x5))
    assert e.value.status == oracle.OR_ERR_BUCKET


def test_pool_overflow(oracle):
    st = ic(0, 2000)
    bb = oracle.bbox(st["x"], st["y"])
    with pytest.raises(oracle.OracleError) as e:
        oracle.build_insert(st["x"], st["y"], st["mass"], bb, m=2000 + 10)
    assert e.value.status == oracle.OR_ERR_POOL


# ------------------------------------------------------ centre of mass ----
def test_com_matches_range_sums(oracle):
    st = ic(1, 20000)
    x, y, m = st["x"], st["y"], st["mass"]
    bb = oracle.bbox(x, y)
    kt = oracle.build_keys(x, y, bb)
    oracle.com_exact(kt, x, y, m)
    s = kt["sorted"]
    cm = np.cumsum(np.concatenate([[0.0], m[s].astype(np.float64)]))
    cx = np.cumsum(np.concatenate([[0.0], (m[s].astype(np.float64) * x[s])]))
    cy = np.cumsum(np.concatenate([[0.0], (m[s].astype(np.float64) * y[s])]))
    a, c = kt["start"], kt["count"]
    M = cm[a + c] - cm[a]
    X = (cx[a + c] - cx[a]) / M
    Y = (cy[a + c] - cy[a]) / M
    ulp = lambda v: np.spacing(np.abs(v).astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(kt["com_x"] - X) <= ulp(X) + 1e-12 * np.abs(bb).max())
    assert np.all(np.abs(kt["com_y"] - Y) <= ulp(Y) + 1e-12 * np.abs(bb).max())
    assert np.all(np.abs(kt["com_m"] - M) <= ulp(M))


def test_reference_fp32_com_close_to_exact(oracle):
    st = ic(2, 20000)
    bb, t, can, kt = _full_build(oracle, st["x"], st["y"], st["mass"])
    oracle.com_exact(kt, st["x"], st["y"], st["mass"])
    rel = np.abs(can["com_x"] - kt["com_x"]) / (np.abs(kt["com_x"]) + 1.0)
    assert rel.max() < 1e-4
    assert np.array_equal(can["com_m"], kt["com_m"])  # integer masses sum exactly


# ----------------------------------------------------------------- walk ---
def _tree(oracle, st):
    bb = oracle.bbox(st["x"], st["y"])
    kt = oracle.build_keys(st["x"], st["y"], bb)
    oracle.com_exact(kt, st["x"], st["y"], st["mass"])
    return bb, kt


def test_walk_theta_to_zero_is_direct_sum(oracle):
    st = ic(1, 3000)
    bb, kt = _tree(oracle, st)
    eps2 = 1e-3
    ax, ay, it = oracle.walk(kt, st["x"], st["y"], st["mass"], bb, 1e-30, eps2, 32)
    dx, dy = oracle.direct_fp64(st["x"], st["y"], st["mass"], eps2)
    assert np.all(it == 3000)
    rel = np.sqrt(((ax - dx) ** 2 + (ay - dy) ** 2) / (dx ** 2 + dy ** 2))
    assert rel.max() < 2e-4


def test_two_body_closed_form(oracle):
    x = np.array([0.0, 3.0], np.float32)
    y = np.array([0.0, 4.0], np.float32)
    m = np.array([2.0, 5.0], np.float32)
    eps2 = 0.25
    st = {"x": x, "y": y, "mass": m}
    bb, kt = _tree(oracle, st)
    ax, ay, it = oracle.walk(kt, x, y, m, bb, 0.5, eps2, 32)
    r3 = (25.0 + eps2) ** 1.5
    assert np.allclose([ax[0], ay[0]], [5.0 * 3 / r3, 5.0 * 4 / r3], rtol=1e-6)
    assert np.allclose([ax[1], ay[1]], [-2.0 * 3 / r3, -2.0 * 4 / r3], rtol=1e-6)
    dx, dy = oracle.direct_fp64(x, y, m, eps2)
    assert np.allclose(dx, [5.0 * 3 / r3, -2.0 * 3 / r3], rtol=1e-12)


def test_symmetric_ring_cancels(oracle):
    k = 16
    ang = 2 * np.pi * np.arange(k) / k
    x = np.concatenate([[0.0], 10 * np.cos(ang)]).astype(np.float32)
    y = np.concatenate([[0.0], 10 * np.sin(ang)]).astype(np.float32)
    m = np.ones(k + 1, np.float32)
    dx, dy = oracle.direct_fp64(x, y, m, 0.01, targets=np.array([0], np.int32))
    assert abs(dx[0]) < 1e-7 and abs(dy[0]) < 1e-7
    st = {"x": x, "y": y, "mass": m}
    bb, kt = _tree(oracle, st)
    ax, ay, _ = oracle.walk(kt, x, y, m, bb, 1e-30, 0.01, 32)
    assert abs(ax[0]) < 1e-6 and abs(ay[0]) < 1e-6


@pytest.mark.parametrize("group", [1, 16, 32, 64])
def test_walk_matches_recursive_statement(oracle, group):
    st = ic(2, 20000)
    bb, kt = _tree(oracle, st)
    ax, ay, it = oracle.walk(kt, st["x"], st["y"], st["mass"], bb, 0.5, 1e-4, group)
    rx, ry, rit = oracle.walk_recursive(kt, st["x"], st["y"], st["mass"], bb, 0.5, 1e-4, group)
    assert np.array_equal(it, rit)
    rel = np.sqrt(((ax - rx) ** 2 + (ay - ry) ** 2) / (rx ** 2 + ry ** 2))
    assert np.sqrt(np.mean(rel ** 2)) < 5e-6


# ---- the opening threshold, stated from the cell widths ----------------------
def _anisotropic_state(seed=11, n=400):
    """Bodies in [-1, 7] x [-0.5, 1.5] with the four extremes present: the box
    holds the origin, its x extent (8) is four times its y extent (2)."""
    rng = np.random.default_rng(seed)
    x = np.concatenate([[-1.0, 7.0, 0.5, 3.0], rng.uniform(-1.0, 7.0, n - 4)]).astype(np.float32)
    y = np.concatenate([[0.2, -0.1, -0.5, 1.5], rng.uniform(-0.5, 1.5, n - 4)]).astype(np.float32)
    m = rng.uniform(0.5, 2.0, n).astype(np.float32)
    return {"x": x, "y": y, "mass": m}


def _brute_force_counts(kt, x, y, bb, theta, eps2, group, textbook):
    """Interactions per walk group from the acceptance rule stated with exact
    rationals (kernels.cu:349,376,394-406): a cell at level L is W/2^L wide,
    W = right - left (the box's x extent); a group accepts it when every body
    has (W/2^L)^2/theta <= d^2 + eps2 (textbook: (W/2^L)^2/theta^2 <= d^2,
    unsoftened); level-1 cells are always opened, bodies always accepted.
    Also returns how many decisions lie within 1e-5 (relative) of the
    threshold, where FP32 rounding could decide instead."""
    n = x.size
    W = Fraction(float(bb[1])) - Fraction(float(bb[0]))
    th = Fraction(float(np.float32(theta)))
    e2 = Fraction(float(np.float32(eps2)))
    child, root, srt = kt["child"], kt["root"], kt["sorted"]
    cx = [Fraction(float(v)) for v in kt["com_x"]]
    cy = [Fraction(float(v)) for v in kt["com_y"]]
    counts = np.zeros(n, np.int32)
    borderline = 0
    for s0 in range(0, n, group):
        ids = srt[s0:s0 + group]
        P = [(Fraction(float(x[b])), Fraction(float(y[b]))) for b in ids]
        cnt = 0
        todo = [(int(ch) - n, 1) for ch in root if ch >= n]  # (cell, level), opened
        cnt += sum(1 for ch in root if 0 <= ch < n)
        while todo:
            c, lev = todo.pop()
            for k in range(4):
                ch = int(child[4 * c + k])
                if ch < 0:
                    continue
                if ch < n:
                    cnt += 1
                    continue
                c2 = ch - n
                w = W / 2 ** (lev + 1)
                thr = w * w / (th * th if textbook else th)
                d2 = min((cx[c2] - px) ** 2 + (cy[c2] - py) ** 2 for px, py in P)
                rhs = d2 if textbook else d2 + e2
                if abs(rhs - thr) <= thr * Fraction(1, 100000):
                    borderline += 1
                if thr <= rhs:
                    cnt += 1
                else:
                    todo.append((c2, lev + 1))
        counts[ids] = cnt
    return counts, borderline


@pytest.mark.parametrize("textbook", [0, 1])
@pytest.mark.parametrize("group", [1, 32])
def test_opening_threshold_from_cell_widths(oracle, group, textbook):
    """The walk's per-level thresholds (root depth from half the box's x
    extent, a quarter per level) against the widths of the cells themselves,
    on a box whose x and y extents differ: taking the y extent, or halving
    instead of quartering per level, changes the counts."""
    st = _anisotropic_state()
    x, y, m = st["x"], st["y"], st["mass"]
    bb, kt = _tree(oracle, st)
    assert list(bb) == [-1.0, 7.0, -0.5, 1.5]
    theta, eps2 = 0.5, 1e-4
    _, _, it = oracle.walk(kt, x, y, m, bb, theta, eps2, group, opening=textbook)
    bf, borderline = _brute_force_counts(kt, x, y, bb, theta, eps2, group, textbook)
    assert borderline == 0
    assert np.array_equal(it, bf)
    # the rule is not vacuous here: the x-extent threshold accepts cells a
    # y-extent threshold (16 times smaller) would not
    assert 1 < it.min() < it.max() < x.size


def test_step_is_the_composed_stages(oracle):
    """or_step (the CPU baseline's timed step) equals bbox -> insertion build
    -> FP32 CoM -> sort -> group-of-32 walk -> update, composed bit for bit,
    and does not depend on the thread count."""
    st = ic(1, 3000)
    theta, eps2, dt, d = 0.5, 1e-3, 1e-3, 0.9
    x, y, vx, vy, m = (np.asarray(st[k], np.float32).copy() for k in ("x", "y", "vx", "vy", "mass"))
    bb = oracle.bbox(x, y)
    t = oracle.build_insert(x, y, m, bb)
    oracle.com_ref(t)
    oracle.sort_ref(t)
    ct = oracle.canonicalize(t)
    ax, ay, _ = oracle.walk(ct, x, y, m, bb, theta, eps2, 32)
    ex = oracle.update(x, y, vx, vy, ax, ay, dt, d)
    for threads in (1, 4):
        s = {"x": x.copy(), "y": y.copy(), "vx": vx.copy(), "vy": vy.copy(), "mass": m.copy()}
        oracle.step(s, theta, eps2, dt, d, group=32, threads=threads)
        assert np.array_equal(s["ax"], ax) and np.array_equal(s["ay"], ay)
        for k, v in zip(("x", "y", "vx", "vy"), ex):
            assert np.array_equal(s[k], v), k


def test_walk_error_falls_with_theta_and_group_is_conservative(oracle):
    st = ic(2, 20000)
    bb, kt = _tree(oracle, st)
    x, y, m = st["x"], st["y"], st["mass"]
    dx, dy = oracle.direct_fp64(x, y, m, 1e-4)
    errs, counts = [], []
    for th in (1.0, 0.5, 0.25):
        ax, ay, it = oracle.walk(kt, x, y, m, bb, th, 1e-4, 32)
        errs.append(np.sqrt(np.mean(((ax - dx) ** 2 + (ay - dy) ** 2) / (dx ** 2 + dy ** 2))))
        counts.append(it.mean())
    assert errs[0] > errs[1] > errs[2] and counts[0] < counts[1] < counts[2]
    _, _, i1 = oracle.walk(kt, x, y, m, bb, 0.5, 1e-4, 1)
    _, _, i64 = oracle.walk(kt, x, y, m, bb, 0.5, 1e-4, 64)
    assert i1.mean() < counts[1] < i64.mean()


# --------------------------------------------------------------- update ---
def test_update_closed_form(oracle):
    rng = np.random.default_rng(9)
    a = [rng.normal(0, 1, 1000).astype(np.float32) for _ in range(6)]
    dt, d = np.float32(1e-3), np.float32(0.9)
    x, y, vx, vy = oracle.update(*a, dt, d)
    ex_vx = a[2].astype(np.float64) + float(dt) * a[4]
    ex_x = a[0].astype(np.float64) + float(np.float32(d * dt)) * vx.astype(np.float64)
    assert np.all(np.abs(vx - ex_vx) <= np.spacing(np.abs(vx)))
    assert np.all(np.abs(x - ex_x) <= np.spacing(np.abs(x)))


def test_energy_two_body(oracle):
    x = np.array([0.0, 3.0], np.float32)
    y = np.array([0.0, 4.0], np.float32)
    m = np.array([2.0, 5.0], np.float32)
    vx = np.array([1.0, 0.0], np.float32)
    vy = np.array([0.0, 2.0], np.float32)
    e = oracle.energy(x, y, vx, vy, m, 0.0)
    assert np.isclose(e["kinetic"], 0.5 * 2 * 1 + 0.5 * 5 * 4)
    assert np.isclose(e["potential"], -10.0 / 5.0)
    assert np.isclose(e["px"], 2.0) and np.isclose(e["py"], 10.0)


def test_diagnostics_two_body_closed_form(oracle):
    x = np.array([0.0, 3.0], np.float32)
    y = np.array([0.0, 4.0], np.float32)
    m = np.array([2.0, 5.0], np.float32)
    vx = np.array([1.0, 0.0], np.float32)
    vy = np.array([0.0, 2.0], np.float32)
    d = oracle.diagnostics(x, y, vx, vy, m, 0.0)
    assert np.isclose(d["kinetic"], 11.0) and np.isclose(d["potential"], -2.0)
    assert np.isclose(d["px"], 2.0) and np.isclose(d["py"], 10.0)
    assert np.isclose(d["lz"], 5.0 * (3.0 * 2.0 - 4.0 * 0.0))
    # velocities advanced by half_dt * a before the sums
    ax = np.array([2.0, -1.0], np.float32)
    ay = np.array([4.0, 0.5], np.float32)
    s = oracle.diagnostics(x, y, vx, vy, m, 0.0, ax=ax, ay=ay, half_dt=0.5)
    u = vx + 0.5 * ax
    v = vy + 0.5 * ay
    assert np.isclose(s["kinetic"], float(np.sum(0.5 * m * (u * u + v * v))))
    assert np.isclose(s["px"], float(np.sum(m * u))) and np.isclose(s["py"], float(np.sum(m * v)))
    assert np.isclose(s["lz"], float(np.sum(m * (x * v - y * u))))
    assert s["potential"] == d["potential"]


def test_synchronised_energy_of_kick_drift_orbit(oracle):
    """Two equal masses on an eccentric orbit (half the circular speed, the
    pair falls to 0.43 of its separation) stepped with the reference update
    (kick then drift) under direct forces: energy taken with v + dt/2 a(x)
    is conserved to O(dt^2), the raw half-step velocities, or v - dt/2 a,
    only to O(dt), which pins the sign and size of the synchronisation;
    angular momentum is conserved to rounding (central pair forces)."""
    x = np.array([-0.5, 0.5], np.float32)
    y = np.zeros(2, np.float32)
    m = np.ones(2, np.float32)
    vc = np.float32(0.5 * np.sqrt(2.0) / 2.0)
    vx = np.zeros(2, np.float32)
    vy = np.array([-vc, vc], np.float32)
    dt, eps2 = 1e-3, 1e-10  # the direct sum's self term needs eps2 > 0
    sync, raw, wrong, lz = [], [], [], []
    for _ in range(800):
        ax, ay = (a.astype(np.float32) for a in oracle.direct_fp64(x, y, m, eps2))
        s = oracle.diagnostics(x, y, vx, vy, m, eps2, ax=ax, ay=ay, half_dt=dt / 2)
        r = oracle.diagnostics(x, y, vx, vy, m, eps2)
        w = oracle.diagnostics(x, y, vx, vy, m, eps2, ax=ax, ay=ay, half_dt=-dt / 2)
        sync.append(s["kinetic"] + s["potential"])
        raw.append(r["kinetic"] + r["potential"])
        wrong.append(w["kinetic"] + w["potential"])
        lz.append(s["lz"])
        x, y, vx, vy = oracle.update(x, y, vx, vy, ax, ay, dt, 1.0)
    sync, raw, wrong, lz = (np.array(q) for q in (sync, raw, wrong, lz))
    e0 = abs(sync[0])
    assert np.ptp(sync) / e0 < 5e-5
    assert np.ptp(raw) / e0 > 1e-3 and np.ptp(wrong) / e0 > 1e-3
    assert np.ptp(lz) / abs(lz[0]) < 1e-5


# -------------------------------------------------------- Hilbert order ---
def test_hilbert_order_known_answer(oracle):
    """One body per quadrant (ids NE, NW, SE, SW): the root's curve visits
    SW, NW, NE, SE."""
    x = np.array([1, -1, 1, -1], np.float32)
    y = np.array([1, 1, -1, -1], np.float32)
    bb = oracle.bbox(x, y)
    kt = oracle.build_keys(x, y, bb)
    assert list(oracle.hilbert_order(kt, 4)) == [3, 1, 0, 2]


def test_hilbert_order_is_a_continuous_curve(oracle):
    """Bodies at the centres of a 16 x 16 grid (the tree's midpoints fall
    between grid lines): consecutive bodies of the order are grid neighbours
    (a Morton order jumps), the order starts in the south-west corner and ends
    in the south-east one, and every cell's bodies are one contiguous run."""
    k = 16
    i, j = np.meshgrid(np.arange(k), np.arange(k), indexing="ij")
    perm = np.random.default_rng(1).permutation(k * k)
    gi, gj = i.ravel()[perm], j.ravel()[perm]
    x = (gi + 0.5).astype(np.float32)
    y = (gj + 0.5).astype(np.float32)
    bb = oracle.bbox(x, y)
    kt = oracle.build_keys(x, y, bb)
    h = oracle.hilbert_order(kt, k * k)
    assert np.array_equal(np.sort(h), np.arange(k * k))
    step = np.abs(np.diff(gi[h])) + np.abs(np.diff(gj[h]))
    assert np.all(step == 1)
    assert (gi[h[0]], gj[h[0]]) == (0, 0) and (gi[h[-1]], gj[h[-1]]) == (k - 1, 0)
    ms = kt["sorted"]
    assert np.any(np.abs(np.diff(gi[ms])) + np.abs(np.diff(gj[ms])) > 1)
    rank = np.empty(k * k, np.int64)
    rank[h] = np.arange(k * k)
    for c in range(kt["ncells"]):
        r = np.sort(rank[ms[kt["start"][c]: kt["start"][c] + kt["count"][c]]])
        assert r[-1] - r[0] == r.size - 1, c


def test_hilbert_order_below_l1_is_slot_order(oracle):
    """Below level l1 the order follows the slot codes: with l1 = 0 it is the
    reference's sort order."""
    st = ic(1, 3000)
    bb, kt = _tree(oracle, st)
    assert np.array_equal(oracle.hilbert_order(kt, 3000, l1=0), kt["sorted"])
    assert not np.array_equal(oracle.hilbert_order(kt, 3000), kt["sorted"])


def test_hilbert_groups_walk(oracle):
    """Walk groups taken along the Hilbert order: the stack walk equals the
    recursive statement with the same groups; against Morton groups fewer
    interactions per body on the colliding disks, forces as accurate."""
    st = ic(2, 20000)
    bb, kt = _tree(oracle, st)
    oracle.quad_exact(kt, st["x"], st["y"], st["mass"])
    h = oracle.hilbert_order(kt, 20000)
    args = (kt, st["x"], st["y"], st["mass"], bb, 0.5, 1e-4, 64)
    ax, ay, it = oracle.walk(*args, opening=1, quad=True, order=h)
    rx, ry, rit = oracle.walk_recursive(*args, opening=1, quad=True, order=h)
    assert np.array_equal(it, rit)
    assert np.sqrt(np.mean(((ax - rx) ** 2 + (ay - ry) ** 2) / (rx ** 2 + ry ** 2))) < 5e-6
    mx, my, mit = oracle.walk(*args, opening=1, quad=True)
    assert it.mean() < mit.mean()
    tg = np.arange(0, 20000, 10, dtype=np.int32)
    dx, dy = oracle.direct_fp64(st["x"], st["y"], st["mass"], 1e-4, targets=tg)
    e = lambda gx, gy: np.sqrt(np.mean(((gx[tg] - dx) ** 2 + (gy[tg] - dy) ** 2) / (dx ** 2 + dy ** 2)))
    assert e(ax, ay) < 1.5 * e(mx, my) and e(ax, ay) < 1e-3


# ------------------------------------------------------------ quadrupole ---
def test_quadrupole_matches_body_range_moments(oracle):
    """Q of each cell = sum over its bodies of m (3 d d^T - |d|^2 I), d from
    the cell's centre of mass, computed directly from the body range."""
    st = ic(1, 5000)
    x, y, m = st["x"], st["y"], st["mass"]
    bb = oracle.bbox(x, y)
    kt = oracle.build_keys(x, y, bb)
    oracle.com_exact(kt, x, y, m)
    oracle.quad_exact(kt, x, y, m)
    s = kt["sorted"]
    X, Y, M = x[s].astype(np.float64), y[s].astype(np.float64), m[s].astype(np.float64)
    for c in range(0, kt["ncells"], 7):
        a, k = kt["start"][c], kt["count"][c]
        mm, xx, yy = M[a:a + k], X[a:a + k], Y[a:a + k]
        ox, oy = (mm * xx).sum() / mm.sum(), (mm * yy).sum() / mm.sum()
        dx, dy = xx - ox, yy - oy
        q = np.array([(mm * (2 * dx * dx - dy * dy)).sum(), (mm * 3 * dx * dy).sum(),
                      (mm * (2 * dy * dy - dx * dx)).sum()])
        scale = (mm * (dx * dx + dy * dy)).sum() + 1e-300
        assert np.all(np.abs(kt["quad"][3 * c: 3 * c + 3] - q) <= 1e-6 * scale), c


def test_quadrupole_improves_far_field(oracle):
    """A tight cluster seen from far away: the quadrupole-corrected cell term
    is much closer to the direct sum than the monopole alone."""
    rng = np.random.default_rng(4)
    k = 200
    x = np.concatenate([rng.normal(0, 1.0, k) * [1.0], [40.0]]).astype(np.float32)
    y = np.concatenate([rng.normal(0, 0.3, k), [25.0]]).astype(np.float32)
    m = np.concatenate([rng.uniform(0.5, 1.5, k), [1.0]]).astype(np.float32)
    st = {"x": x, "y": y, "mass": m}
    bb, kt = _tree(oracle, st)
    oracle.quad_exact(kt, x, y, m)
    tg = np.array([k], np.int32)
    dx, dy = oracle.direct_fp64(x, y, m, 1e-4, targets=tg)
    errs = []
    for quad in (False, True):
        ax, ay, it = oracle.walk(kt, x, y, m, bb, 1.0, 1e-4, 1, quad=quad)
        errs.append(np.hypot(ax[k] - dx[0], ay[k] - dy[0]) / np.hypot(dx[0], dy[0]))
        assert it[k] < 20  # the cluster is taken as a few far cells
    assert errs[1] < 0.2 * errs[0]


def test_softened_quadrupole_is_the_second_order_term(oracle):
    """A symmetric pair (no octupole) seen by a lone body at a distance
    comparable to the softening length: the walk's monopole + quadrupole term
    equals the softened direct sum m_i / sqrt(|d + s_i|^2 + eps2) up to the
    fourth-order residual, in FP32 and in the FP64 recursive statement.  The
    unsoftened quadrupole form (-Q.d/r^5 + 2.5 (d.Q.d) d/r^7 with softened r,
    from the closed-form Q of the pair) misses by the eps2 tr Q term, 400
    times more here, so a dropped, mis-signed or mis-scaled softening term
    fails."""
    half, ang = 0.5, 0.7
    c = np.array([8.0, 6.0])
    u = np.array([np.cos(ang), np.sin(ang)])
    pts = np.array([[0.0, 0.0], c + half * u, c - half * u], np.float32)
    x, y = pts[:, 0].copy(), pts[:, 1].copy()
    m = np.ones(3, np.float32)
    eps2 = 25.0
    bb, kt = _tree(oracle, {"x": x, "y": y, "mass": m})
    oracle.quad_exact(kt, x, y, m)
    ex, ey = oracle.direct_fp64(x, y, m, eps2, targets=np.array([0], np.int32))
    ref = np.array([ex[0], ey[0]])
    # closed form of the pair's expansion about its centre of mass
    s2 = 2 * half * half  # sum m |s|^2 = tr Q
    q = 3 * s2 * np.outer(u, u) - s2 * np.eye(2)
    r2 = c @ c + eps2
    unsoft = 2 * c / r2 ** 1.5 - q @ c / r2 ** 2.5 + 2.5 * (c @ q @ c) * c / r2 ** 3.5
    soft = unsoft - 2.5 * s2 * eps2 * c / r2 ** 3.5
    rel = lambda a: float(np.linalg.norm(a - ref) / np.linalg.norm(ref))
    assert rel(unsoft) > 5e-4 and rel(soft) < 5e-6
    args = (kt, x, y, m, bb, 1.0, eps2, 1)
    ax, ay, it = oracle.walk(*args, opening=1, quad=True)
    assert it[0] == 2  # itself (zero force) and the pair's cell
    assert rel(np.array([ax[0], ay[0]])) < 1e-5
    rx, ry, rit = oracle.walk_recursive(*args, opening=1, quad=True)
    assert rit[0] == 2 and rel(np.array([rx[0], ry[0]])) < 5e-6


@pytest.mark.parametrize("opening", [0, 1])
def test_quad_walk_matches_recursive_statement(oracle, opening):
    st = ic(2, 20000)
    bb, kt = _tree(oracle, st)
    oracle.quad_exact(kt, st["x"], st["y"], st["mass"])
    args = (kt, st["x"], st["y"], st["mass"], bb, 0.5, 1e-4, 64)
    ax, ay, it = oracle.walk(*args, opening=opening, quad=True)
    rx, ry, rit = oracle.walk_recursive(*args, opening=opening, quad=True)
    assert np.array_equal(it, rit)
    assert np.sqrt(np.mean(((ax - rx) ** 2 + (ay - ry) ** 2) / (rx ** 2 + ry ** 2))) < 5e-6


def test_textbook_opening_is_theta_squared(oracle):
    """w/d <= theta (textbook, unsoftened d) equals the reference test at
    theta^2 when the softening vanishes (eps2 = 2^-126, below every
    threshold's rounding)."""
    st = ic(2, 20000)
    bb, kt = _tree(oracle, st)
    x, y, m = st["x"], st["y"], st["mass"]
    tiny = float(np.finfo(np.float32).tiny)
    _, _, a = oracle.walk(kt, x, y, m, bb, 0.5, tiny, 64, opening=1)
    _, _, b = oracle.walk(kt, x, y, m, bb, 0.25, tiny, 64, opening=0)
    assert np.array_equal(a, b)


def test_textbook_opening_never_accepts_own_cell(oracle):
    """A tight cluster of three bodies (well inside the softening length) in
    a field of 300: the reference's softened test accepts every cell
    narrower than about theta*eps, including the one that holds the target
    itself, so its error stays ~1e-3 however small theta gets; the textbook
    test on the unsoftened distance opens it and converges."""
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(-50, 50, 300), [10.0, 10.0001, 10.0]]).astype(np.float32)
    y = np.concatenate([rng.uniform(-50, 50, 300), [10.0, 10.0, 10.00013]]).astype(np.float32)
    m = np.ones(x.size, np.float32)
    m[-1] = 5.0
    st = {"x": x, "y": y, "mass": m}
    bb, kt = _tree(oracle, st)
    eps2 = 1e-2
    tg = np.array([300, 301, 302], np.int32)
    dx, dy = oracle.direct_fp64(x, y, m, eps2, targets=tg)
    angle = 0.01
    ax, ay, _ = oracle.walk(kt, x, y, m, bb, angle * angle, eps2, 1, opening=0)
    ref = np.hypot(ax[tg] - dx, ay[tg] - dy) / np.hypot(dx, dy)
    ax, ay, _ = oracle.walk(kt, x, y, m, bb, angle, eps2, 1, opening=1)
    tb = np.hypot(ax[tg] - dx, ay[tg] - dy) / np.hypot(dx, dy)
    assert ref.max() > 1e-3
    assert tb.max() < 1e-5


# ---- bmax opening (per-cell thresholds) ------------------------------------
EPS2_BMAX = 1e-4
def test_bmax_two_bodies_closed_form(oracle):
    """One cell of two bodies: bmax is the larger body distance from the
    centre of mass, threshold bmax^2/theta^2."""
    x = np.array([1.0, 1.5], np.float32)
    y = np.array([1.0, 1.25], np.float32)
    m = np.array([3.0, 1.0], np.float32)
    st = {"x": x, "y": y, "mass": m}
    bb, kt = _tree(oracle, st)
    oracle.bmax_exact(kt, x, y, m, 0.5, EPS2_BMAX)
    ox = (3.0 * 1.0 + 1.0 * 1.5) / 4.0
    oy = (3.0 * 1.0 + 1.0 * 1.25) / 4.0
    bm = max(np.hypot(1.0 - ox, 1.0 - oy), np.hypot(1.5 - ox, 1.25 - oy))
    # every cell on the chain holds both bodies: same bound everywhere
    assert kt["ncells"] >= 1
    np.testing.assert_allclose(kt["thr"], np.float32(bm * bm / 0.25 + EPS2_BMAX), rtol=1e-6)


def test_bmax_bounds_the_body_distances(oracle):
    """bmax of each cell is at least the largest distance from its centre of
    mass to any of its bodies, exactly that when every child is a body, and
    at most the cell diagonal plus the largest child bound."""
    st = ic(1, 5000)
    x, y, m = st["x"], st["y"], st["mass"]
    theta = 0.7
    bb, kt = _tree(oracle, st)
    tiny = float(np.finfo(np.float32).tiny)  # keeps bmax recoverable from the threshold
    oracle.bmax_exact(kt, x, y, m, theta, tiny)
    s = kt["sorted"]
    X, Y, M = x[s].astype(np.float64), y[s].astype(np.float64), m[s].astype(np.float64)
    n = x.size
    bmax = np.sqrt(kt["thr"].astype(np.float64)) * theta
    for c in range(0, kt["ncells"], 5):
        a, k = kt["start"][c], kt["count"][c]
        mm, xx, yy = M[a:a + k], X[a:a + k], Y[a:a + k]
        ox, oy = (mm * xx).sum() / mm.sum(), (mm * yy).sum() / mm.sum()
        true = np.max(np.hypot(xx - ox, yy - oy))
        assert bmax[c] >= true * (1 - 1e-5) - 1e-9, c
        ch = kt["child"][4 * c:4 * c + 4]
        if np.all(ch < n):
            assert bmax[c] == pytest.approx(true, rel=1e-5, abs=1e-9), c


def test_bmax_walk_theta_to_zero_is_direct_sum(oracle):
    st = ic(0, 2000)
    x, y, m = st["x"], st["y"], st["mass"]
    bb, kt = _tree(oracle, st)
    oracle.bmax_exact(kt, x, y, m, 1e-6, EPS2_BMAX)
    ax, ay, it = oracle.walk(kt, x, y, m, bb, 1e-6, 1e-4, 32, bmax=True)
    dx, dy = oracle.direct_fp64(x, y, m, 1e-4)
    assert np.all(it == x.size)
    assert np.sqrt(np.mean(((ax - dx) ** 2 + (ay - dy) ** 2) / (dx ** 2 + dy ** 2))) < 1e-5


@pytest.mark.parametrize("quad", [False, True])
def test_bmax_walk_matches_recursive_statement(oracle, quad):
    st = ic(2, 20000)
    bb, kt = _tree(oracle, st)
    x, y, m = st["x"], st["y"], st["mass"]
    oracle.bmax_exact(kt, x, y, m, 0.5, EPS2_BMAX)
    if quad:
        oracle.quad_exact(kt, x, y, m)
    args = (kt, x, y, m, bb, 0.5, 1e-4, 64)
    ax, ay, it = oracle.walk(*args, quad=quad, bmax=True)
    rx, ry, rit = oracle.walk_recursive(*args, quad=quad, bmax=True)
    assert np.array_equal(it, rit)
    assert np.sqrt(np.mean(((ax - rx) ** 2 + (ay - ry) ** 2) / (rx ** 2 + ry ** 2))) < 5e-6


def test_bmax_error_falls_with_theta(oracle):
    st = ic(2, 20000)
    x, y, m = st["x"], st["y"], st["mass"]
    bb, kt = _tree(oracle, st)
    tg = np.arange(0, x.size, 37, dtype=np.int32)
    dx, dy = oracle.direct_fp64(x, y, m, 1e-4, targets=tg)
    errs = []
    for th in (0.8, 0.5, 0.3):
        oracle.bmax_exact(kt, x, y, m, th, EPS2_BMAX)
        ax, ay, _ = oracle.walk(kt, x, y, m, bb, th, 1e-4, 64, bmax=True)
        errs.append(np.sqrt(np.mean(((ax[tg] - dx) ** 2 + (ay[tg] - dy) ** 2) / (dx ** 2 + dy ** 2))))
    assert errs[0] > errs[1] > errs[2]


# ---- mixed precision (FP16 far field) ---------------------------------------
def test_f16_rounding_against_numpy_and_brute_force(oracle):
    """The oracle's FP16 rounding: nearest-even against numpy's float16
    conversion, toward-zero against the largest-magnitude FP16 value not
    above |x| (found by search over every finite FP16 value)."""
    lib = oracle.lib()
    rng = np.random.default_rng(21)
    xs = np.concatenate([rng.normal(0, 1, 4000), rng.normal(0, 1e-5, 2000),
                         rng.uniform(-7e4, 7e4, 2000), 2.0 ** rng.integers(-26, 17, 500),
                         [0.0, -0.0, 65504.0, 65519.99, 65520.0, 6e-8, 2.98e-8, 5.96e-8]])
    finite = np.arange(65536, dtype=np.uint16).view(np.float16)
    finite = np.unique(finite[np.isfinite(finite)].astype(np.float64))
    for x in xs:
        got = lib.or_f16_round(float(x), 0)
        with np.errstate(over="ignore"):
            want = float(np.float64(x).astype(np.float16))
        assert got == want or (np.isinf(got) and np.isinf(want)), (x, got, want)
        rz = lib.or_f16_round(float(x), 1)
        ax = min(abs(x), 65504.0)
        want_z = finite[np.searchsorted(finite, ax, side="right") - 1]
        assert rz == np.copysign(want_z, x) or (x == 0 and rz == 0), (x, rz, want_z)


def test_mixed_far_field_matches_fp32_quadrupoles(oracle):
    """The FP16 bracket k B with B = e + t (-Qh.e + (5/2)(e.Qh.e) e) is the
    monopole + quadrupole term: the mixed walk makes the FP32 walk's decisions
    (same counts) and its forces differ from the FP32 quadrupole walk's by no
    more than FP16 rounding allows; a wrong sign or factor in the quadrupole
    bracket (about 1e-2 of a term) or a level off by one between the stored
    Qh and the walk's ws^2 (a factor 4) does not fit in that bound."""
    st = ic(2, 20000)
    x, y, m = st["x"], st["y"], st["mass"]
    bb, kt = _tree(oracle, st)
    oracle.quad_exact(kt, x, y, m)
    oracle.quad_mixed(kt, x, y, m, bb)
    assert np.abs(kt["qh"]).max() < 4.0
    eps2 = 1e-4
    fx, fy, fi = oracle.walk(kt, x, y, m, bb, 0.5, eps2, 64, opening=1, quad=True)
    mx, my, mi = oracle.walk(kt, x, y, m, bb, 0.5, eps2, 64, opening=1, quad=True, mixed=True)
    assert np.array_equal(fi, mi)
    d = np.sqrt(np.mean(((mx - fx) ** 2 + (my - fy) ** 2)
                        / (fx.astype(np.float64) ** 2 + fy.astype(np.float64) ** 2)))
    assert 1e-6 < d < 6e-4, d
    dx, dy = oracle.direct_fp64(x, y, m, eps2)
    err = np.sqrt(np.mean(((mx - dx) ** 2 + (my - dy) ** 2) / (dx ** 2 + dy ** 2)))
    assert err < 1e-3, err
    # the same walk with the quadrupole bracket dropped (Qh = 0) is the
    # monopole walk, an order of magnitude less accurate
    kq = dict(kt, qh=np.zeros_like(kt["qh"]))
    zx, zy, _ = oracle.walk(kq, x, y, m, bb, 0.5, eps2, 64, opening=1, quad=True, mixed=True)
    zerr = np.sqrt(np.mean(((zx - dx) ** 2 + (zy - dy) ** 2) / (dx ** 2 + dy ** 2)))
    assert zerr > 5 * err


# ---- the reference's handling of a body alone in a root slot ------------------
def test_reference_lone_root_body_known_answers(oracle):
    """kernels.cu:373-387,397-398: a lone body in a root slot is pushed like a
    cell and its child slots read.  Body 0's slots are the root's, so the root's
    children are walked again one level down (the close pair's cells are
    opened again and counted twice); any other lone body exerts no force.
    Bodies: one alone in the north-east quadrant, a close pair in the
    south-west (a chain of cells down to the level that separates them)."""
    eps2 = 1e-4
    x = np.array([1.0, -1.0, -1.1], np.float32)
    y = np.array([1.0, -1.0, -1.05], np.float32)
    m = np.array([3.0, 1.0, 2.0], np.float32)
    st = {"x": x, "y": y, "mass": m}
    bb, kt = _tree(oracle, st)
    assert list(kt["root"]) == [0, -1, -1, 3]  # body 0 alone, the pair's cell
    _, _, plain = oracle.walk(kt, x, y, m, bb, 0.5, eps2, 1)
    # body 0: itself, the pair's level-2 cell; bodies 1, 2: body 0, the pair
    assert list(plain) == [2, 3, 3]
    ax, ay, it = oracle.walk(kt, x, y, m, bb, 0.5, eps2, 1, ref_root=True)
    # body 0: the pair's level-2 cell, then the root's slots again: itself and
    # the level-1 cell (accepted at the level-2 threshold); bodies 1, 2: the
    # pair, then the root's slots again: body 0 and the pair once more
    assert list(it) == [3, 5, 5]

    def f(i, j_list, xs=x, ys=y):
        gx = gy = 0.0
        for (px, py, pm) in j_list:
            dx, dy = px - float(xs[i]), py - float(ys[i])
            r = (dx * dx + dy * dy + eps2) ** 1.5
            gx += pm * dx / r
            gy += pm * dy / r
        return gx, gy

    cm = (float(x[1]) + 2.0 * float(x[2])) / 3.0, (float(y[1]) + 2.0 * float(y[2])) / 3.0
    pair = [(float(x[1]), float(y[1]), 1.0), (float(x[2]), float(y[2]), 2.0)]
    e0 = f(0, [(cm[0], cm[1], 3.0), (cm[0], cm[1], 3.0)])
    e1 = f(1, pair + pair + [(1.0, 1.0, 3.0)])
    assert np.allclose([ax[0], ay[0]], e0, rtol=1e-5)
    assert np.allclose([ax[1], ay[1]], e1, rtol=1e-5)
    # a lone body other than 0 exerts no force
    x2 = np.array([-1.0, 1.0, -1.1], np.float32)
    y2 = np.array([-1.0, 1.0, -1.05], np.float32)
    st2 = {"x": x2, "y": y2, "mass": m}
    bb2, kt2 = _tree(oracle, st2)
    assert list(kt2["root"]) == [1, -1, -1, 3]
    bx, by, it2 = oracle.walk(kt2, x2, y2, m, bb2, 0.5, eps2, 1, ref_root=True)
    # body 1 (lone): the pair's level-2 cell only; the pair: itself, no body 1
    assert list(it2) == [2, 1, 2]
    gx, gy = f(0, [(float(x2[0]), float(y2[0]), 3.0), (float(x2[2]), float(y2[2]), 2.0)], x2, y2)
    assert np.allclose([bx[0], by[0]], [gx, gy], rtol=1e-5)  # no pull from body 1
    c2 = ((3.0 * float(x2[0]) + 2.0 * float(x2[2])) / 5.0, (3.0 * float(y2[0]) + 2.0 * float(y2[2])) / 5.0)
    assert np.allclose([bx[1], by[1]], f(1, [(c2[0], c2[1], 5.0)], x2, y2), rtol=1e-5)


def _lone_body_disk(n=4096):
    """Workload 0 (a disk about a central mass) with body 0 moved far to the
    south-west, alone in its root quadrant."""
    st = ic(0, n)
    st = {k: v.copy() for k, v in st.items()}
    st["x"][0], st["y"][0] = np.float32(-500.0), np.float32(-500.0)
    return st


@pytest.mark.parametrize("group", [1, 32, 64])
def test_reference_lone_root_body_matches_recursive_statement(oracle, group):
    st = _lone_body_disk()
    x, y, m = st["x"], st["y"], st["mass"]
    bb, kt = _tree(oracle, st)
    assert sorted(kt["root"].tolist())[:3] == [-1, -1, 0]
    ax, ay, it = oracle.walk(kt, x, y, m, bb, 0.5, 1e-4, group, ref_root=True)
    rx, ry, rit = oracle.walk_recursive(kt, x, y, m, bb, 0.5, 1e-4, group, ref_root=True)
    assert np.array_equal(it, rit)
    assert np.sqrt(np.mean(((ax - rx) ** 2 + (ay - ry) ** 2) / (rx ** 2 + ry ** 2))) < 5e-6
    _, _, plain = oracle.walk(kt, x, y, m, bb, 0.5, 1e-4, group)
    assert (it > plain).all()  # every group walks the root's children twice


def test_render_known_answers(oracle):
    """Headless point drawing (StellarSolverVisualizer.cpp:208-253): the
    reference's glm::ortho(-25, 25, -25, 25) view of a 1000 x 1000 window,
    20 pixels per unit; pixels are counted from the top-left corner."""
    view = (-25.0, 25.0, -25.0, 25.0)
    x = np.array([0.0, -25.0, 24.99, 25.0, -25.001, 0.0, 0.0, 0.0, np.nan, 0.025], np.float32)
    y = np.array([0.0, -25.0, 24.99, 0.0, 0.0, 30.0, 0.0, 0.0, 0.0, -0.025], np.float32)
    c = oracle.render_counts(x, y, view, 1000, 1000)
    assert c.shape == (1000, 1000) and c.dtype == np.uint32
    # the origin: window (500, 500) -> column 500, row 999 - 500; three bodies
    assert c[499, 500] == 3
    assert c[999, 0] == 1          # bottom-left corner
    assert c[0, 999] == 1          # just inside the top-right corner
    assert c[500, 500] == 1        # (0.025, -0.025): half a pixel right of and below the origin
    # right edge, left of the view, above it, NaN: clipped
    assert int(c.sum()) == 6
    # the blend: white at alpha 0.1 per body, clamped, quantised once
    g = oracle.render_grey(np.array([0, 1, 2, 3, 9, 10, 11, 1000], np.uint32))
    assert g.tolist() == [0, 26, 51, 77, 230, 255, 255, 255]


def test_render_invariants(oracle):
    """Sum of counts = bodies inside the view; a view shifted by exactly one
    pixel shifts the image by one column; the mirrored view of mirrored
    bodies is the mirrored image (pixel-exact coordinates only)."""
    rng = np.random.default_rng(7)
    # coordinates on a 1/64 grid, so every product below is exact
    x = (rng.integers(-2000, 2000, 20000) / 64.0).astype(np.float32)
    y = (rng.integers(-2000, 2000, 20000) / 64.0).astype(np.float32)
    view = (-16.0, 16.0, -8.0, 8.0)
    c = oracle.render_counts(x, y, view, 256, 128)  # 8 pixels per unit
    inside = (x >= -16) & (x < 16) & (y >= -8) & (y < 8)
    assert int(c.sum()) == int(inside.sum())
    c1 = oracle.render_counts(x, y, (-16.125, 15.875, -8.0, 8.0), 256, 128)
    assert np.array_equal(c1[:, 1:], c[:, :-1])
    # on the 1/64 grid, x -> -x - 1/64 maps pixel column k to 255 - k
    cm = oracle.render_counts((-x - np.float32(1 / 64)).astype(np.float32), y, view, 256, 128)
    assert np.array_equal(cm, c[:, ::-1])
