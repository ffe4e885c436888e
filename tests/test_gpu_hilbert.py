"""GPU parity of STELLAR_FLAG_HILBERT_ORDER: bodies stored, cells numbered and
walk groups formed along a Hilbert curve of the same tree.

The tree itself does not change: exported in the reference's numbering it is
bit-exact against the oracle's, centres of mass included.  The walk groups are
consecutive bodies of the Hilbert order, so the interaction counts equal the
oracle's walk over the groups of oracle.hilbert_order (the oracle's Hilbert
restatement, pinned by its own known-answer and continuity tests), and the
forces agree within the rsqrt tolerance.
"""
import numpy as np
import pytest

from tests.helpers import first_tree_difference, ic, rms_rel, tree_equal

pytestmark = pytest.mark.gpu

EPS = {0: 0.01, 1: 0.005, 2: 0.01, 4: 1e-4}
FORCE_RMS_TOL = 2e-6
MIXED_RMS_TOL = 2e-5


@pytest.fixture(scope="module")
def S():
    from tests.conftest import gpu_available

    if not gpu_available():
        pytest.skip("no GPU")
    import stellarsolver_amd as S

    S.load_library()
    return S


def make_engine(S, st, eps2, theta=0.5, group=64, **kw):
    p = S.default_params(theta=theta, eps2=eps2, group=group, **kw)
    e = S.Engine(st["x"].size, p)
    e.upload(st["mass"], st["x"], st["y"], st["vx"], st["vy"])
    return e


def oracle_tree(oracle, st):
    bb = oracle.bbox(st["x"], st["y"])
    kt = oracle.build_keys(st["x"], st["y"], bb)
    oracle.com_exact(kt, st["x"], st["y"], st["mass"])
    return bb, kt


def gpu_tree_and_forces(e):
    e.bbox()
    bb = e.get_bbox()
    e.build()
    e.com()
    t = e.export_tree()
    e.sort()
    e.forces()
    return bb, t, e.export_forces()


# model: (flags, precision, group, theta, oracle walk keywords)
MODELS = {
    "textbook+quad": ("QUADRUPOLE|TEXTBOOK_OPENING", 0, 64, 0.5, {"opening": 1, "quad": True}),
    "reference": ("", 0, 64, 0.5, {}),
    "reference_g32": ("", 0, 32, 0.5, {}),
    "bmax+quad_g16": ("BMAX_OPENING|QUADRUPOLE", 0, 16, 0.6, {"bmax": True, "quad": True}),
    "mixed": ("QUADRUPOLE|TEXTBOOK_OPENING", 1, 64, 0.5,
              {"opening": 1, "quad": True, "mixed": True}),
    "reference_root": ("REFERENCE_ROOT", 0, 64, 0.5, {"ref_root": True}),
}


def model_flags(S, spec):
    f = S.FLAG_HILBERT_ORDER
    for name in filter(None, spec.split("|")):
        f |= getattr(S, "FLAG_" + name)
    return f


def oracle_walk(oracle, ot, st, obb, theta, eps2, group, kw, order):
    x, y, m = st["x"], st["y"], st["mass"]
    if kw.get("quad"):
        oracle.quad_exact(ot, x, y, m)
    if kw.get("bmax"):
        oracle.bmax_exact(ot, x, y, m, theta, eps2)
    if kw.get("mixed"):
        oracle.quad_mixed(ot, x, y, m, obb)
    return oracle.walk(ot, x, y, m, obb, theta, eps2, group, order=order, **kw)


@pytest.mark.parametrize("model", list(MODELS))
@pytest.mark.parametrize("workload,n", [(2, 65536), (1, 65536), (0, 4096), (4, 65536)])
def test_hilbert_tree_and_walk(S, oracle, workload, n, model):
    print(f"hilbert parity workload={workload} n={n} model={model}", flush=True)
    spec, prec, group, theta, kw = MODELS[model]
    st = ic(workload, n)
    eps2 = float(np.float32(EPS[workload] ** 2))
    with make_engine(S, st, eps2, theta=theta, group=group, flags=model_flags(S, spec),
                     precision=prec) as e:
        gbb, gt, f = gpu_tree_and_forces(e)
    obb, ot = oracle_tree(oracle, st)
    assert np.array_equal(gbb, obb)
    assert tree_equal(gt, ot), first_tree_difference(gt, ot)
    for k in ("com_x", "com_y", "com_m"):
        assert np.array_equal(gt[k], ot[k]), k
    order = oracle.hilbert_order(ot, n)
    ax, ay, it = oracle_walk(oracle, ot, st, obb, theta, eps2, group, kw, order)
    assert np.array_equal(f["inter"], it)
    tol = MIXED_RMS_TOL if kw.get("mixed") else FORCE_RMS_TOL
    assert rms_rel(f["ax"], f["ay"], ax, ay) < tol
    # the groups of the Hilbert order are more compact: fewer interactions
    _, _, it_slot = oracle_walk(oracle, ot, st, obb, theta, eps2, group, kw, None)
    assert it.sum() < it_slot.sum()


def test_hilbert_steps_match_oracle(S, oracle):
    """Each step = the oracle's Hilbert-group walk on the GPU's own pre-step
    state, then the oracle update applied to the GPU forces (bit-exact); the
    sample sort is seeded by the previous step's Hilbert order."""
    n, workload = 65536, 2
    st = ic(workload, n)
    eps2 = float(np.float32(EPS[workload] ** 2))
    dt, d = np.float32(1e-3), np.float32(1.0)
    flags = S.FLAG_HILBERT_ORDER | S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING
    with make_engine(S, st, eps2, flags=flags, timestep=float(dt), dampening=float(d)) as e:
        e.run(2)
        for k in range(3):
            print(f"hilbert step {k}", flush=True)
            pre = e.download()
            e.update()
            post = e.download()
            obb, ot = oracle_tree(oracle, pre)
            gt = e.export_tree()
            assert tree_equal(gt, ot), (k, first_tree_difference(gt, ot))
            order = oracle.hilbert_order(ot, n)
            oracle.quad_exact(ot, pre["x"], pre["y"], pre["mass"])
            ax, ay, it = oracle.walk(ot, pre["x"], pre["y"], pre["mass"], obb, 0.5, eps2, 64,
                                     opening=1, quad=True, order=order)
            assert np.array_equal(e.export_forces()["inter"], it), k
            assert rms_rel(post["ax"], post["ay"], ax, ay) < FORCE_RMS_TOL
            x, y, vx, vy = oracle.update(pre["x"], pre["y"], pre["vx"], pre["vy"], post["ax"],
                                         post["ay"], dt, d)
            for a, b in ((x, post["x"]), (y, post["y"]), (vx, post["vx"]), (vy, post["vy"])):
                assert np.array_equal(a, b)


def test_hilbert_flag_switch(S, oracle):
    """Switching the order on and off between steps: each tree and walk is the
    oracle's for the order in force (the sort's splitters of the other order
    are not reused)."""
    n, workload = 65536, 1
    st = ic(workload, n)
    eps2 = float(np.float32(EPS[workload] ** 2))
    base = S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING
    with make_engine(S, st, eps2, flags=base) as e:
        for k, hil in enumerate((False, True, True, False, True)):
            print(f"hilbert switch {k} hilbert={hil}", flush=True)
            e.set_params(flags=base | (S.FLAG_HILBERT_ORDER if hil else 0))
            e.run(1)
            cur = e.download()
            gbb, gt, f = gpu_tree_and_forces(e)
            obb, ot = oracle_tree(oracle, cur)
            assert tree_equal(gt, ot), (k, first_tree_difference(gt, ot))
            order = oracle.hilbert_order(ot, n) if hil else None
            oracle.quad_exact(ot, cur["x"], cur["y"], cur["mass"])
            _, _, it = oracle.walk(ot, cur["x"], cur["y"], cur["mass"], obb, 0.5, eps2, 64,
                                   opening=1, quad=True, order=order)
            assert np.array_equal(f["inter"], it), k


COUNTS = [1000, 64, 5000, 1, 2, 333, 4096, 777, 65, 20000]


@pytest.mark.parametrize("group", [64, 16])
def test_hilbert_ensemble(S, oracle, group):
    """Ensembles: each instance's tree bit-exact, its walk groups consecutive
    bodies of its own Hilbert order (the same order as alone: Hilbert children
    at levels 1..16, which every key holds), and its forces, counts and steps
    equal to a single context's on the same bodies bit for bit."""
    sts = []
    for i, c in enumerate(COUNTS):
        sts.append(S.generate_ic(S.ic_params([0, 1, 2][i % 3], seed=5 + i), c))
    eps2 = float(np.float32(1e-4))
    flags = S.FLAG_HILBERT_ORDER | S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING
    p = S.default_params(eps2=eps2, group=group, flags=flags)
    cat = {k: np.concatenate([s[k] for s in sts]) for k in ("mass", "x", "y", "vx", "vy")}
    with S.Engine(params=p, counts=COUNTS) as e:
        e.upload(cat["mass"], cat["x"], cat["y"], cat["vx"], cat["vy"])
        e.bbox()
        e.build()
        e.com()
        trees = [e.export_tree(instance=i) for i in range(len(COUNTS))]
        e.sort()
        e.forces()
        f = e.export_forces()
        e.run(3)
        after = e.download()
    off = np.concatenate([[0], np.cumsum(COUNTS)])
    for i, st in enumerate(sts):
        print(f"hilbert ensemble instance {i} n={COUNTS[i]} group={group}", flush=True)
        obb, ot = oracle_tree(oracle, st)
        assert tree_equal(trees[i], ot), (i, first_tree_difference(trees[i], ot))
        order = oracle.hilbert_order(ot, COUNTS[i])
        oracle.quad_exact(ot, st["x"], st["y"], st["mass"])
        ax, ay, it = oracle.walk(ot, st["x"], st["y"], st["mass"], obb, 0.5, eps2, group,
                                 opening=1, quad=True, order=order)
        sl = slice(off[i], off[i + 1])
        assert np.array_equal(f["inter"][sl], it), i
        if COUNTS[i] > 1:  # a lone body feels no force
            assert rms_rel(f["ax"][sl], f["ay"][sl], ax, ay) < FORCE_RMS_TOL, i
        s1 = S.default_params(eps2=eps2, group=group, flags=flags)
        with S.Engine(COUNTS[i], s1) as one:
            one.upload(st["mass"], st["x"], st["y"], st["vx"], st["vy"])
            one.bbox()
            one.build()
            one.com()
            one.sort()
            one.forces()
            fo = one.export_forces()
            one.run(3)
            ao = one.download()
        for k in fo:
            assert np.array_equal(f[k][sl], fo[k]), (i, k)
        for k in ao:
            assert np.array_equal(after[k][sl], ao[k]), (i, k)


@pytest.mark.parametrize("ensemble", [False, True])
def test_hilbert_graph_replay(S, ensemble):
    """Graph-replayed steps in Hilbert order equal direct steps bit for bit
    (replays interleaved with direct steps and downloads)."""
    flags = S.FLAG_HILBERT_ORDER | S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING
    if ensemble:
        sts = [S.generate_ic(S.ic_params(i % 3, seed=40 + i), c) for i, c in enumerate(COUNTS)]
        st = {k: np.concatenate([s[k] for s in sts]) for k in ("mass", "x", "y", "vx", "vy")}
        counts = COUNTS
    else:
        st = ic(2, 65536)
        counts = None
    eps2 = float(np.float32(1e-4))
    out = []
    for graph in (False, True):
        p = S.default_params(eps2=eps2, flags=flags | (S.FLAG_GRAPH if graph else 0))
        e = S.Engine(params=p, counts=counts) if counts else S.Engine(st["x"].size, p)
        with e:
            e.upload(st["mass"], st["x"], st["y"], st["vx"], st["vy"])
            got = []
            for k in (2, 4, 1, 2):
                print(f"hilbert graph={graph} ensemble={ensemble} run {k}", flush=True)
                e.run(k)
                got.append(e.download())
            out.append(got)
    for a, b in zip(*out):
        for k in a:
            assert np.array_equal(a[k], b[k]), k


def test_hilbert_headline_size(S, oracle):
    """The bench default at its real size (N = 2^20 colliding disks, theta 0.5,
    quadrupoles, textbook opening, Hilbert order): bit-exact tree and
    interaction counts, forces within the rsqrt tolerance of the oracle's
    walk, and inside the 1e-3 bound against the FP64 direct sum."""
    n = 1 << 20
    print(f"hilbert headline parity n={n}", flush=True)
    st = ic(2, n)
    eps2 = float(np.float32(1e-4))
    flags = S.FLAG_HILBERT_ORDER | S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING
    with make_engine(S, st, eps2, flags=flags) as e:
        gbb, gt, f = gpu_tree_and_forces(e)
        tg = np.sort(np.random.default_rng(2).choice(n, 8192, replace=False)).astype(np.int32)
        dx, dy = e.direct_fp64(tg)
    obb, ot = oracle_tree(oracle, st)
    assert tree_equal(gt, ot), first_tree_difference(gt, ot)
    oracle.quad_exact(ot, st["x"], st["y"], st["mass"])
    ax, ay, it = oracle.walk(ot, st["x"], st["y"], st["mass"], obb, 0.5, eps2, 64, opening=1,
                             quad=True, order=oracle.hilbert_order(ot, n))
    assert np.array_equal(f["inter"], it)
    assert rms_rel(f["ax"], f["ay"], ax, ay) < 1e-5
    err = rms_rel(f["ax"][tg], f["ay"][tg], dx, dy)
    print(f"rms relative error vs FP64 direct sum: {err:.3e}", flush=True)
    assert err < 1e-3


def test_hilbert_below_level_16(S, oracle):
    """A tree deeper than the curve's 16 levels (deep Plummer core, N = 2^20,
    cells to level 23): the children below level 16 follow the slot order on
    both sides; tree bit-exact, counts equal to the oracle's."""
    n = 1 << 20
    st = ic(4, n)
    eps2 = float(np.float32(EPS[4] ** 2))
    flags = S.FLAG_HILBERT_ORDER | S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING
    with make_engine(S, st, eps2, flags=flags) as e:
        gbb, gt, f = gpu_tree_and_forces(e)
    obb, ot = oracle_tree(oracle, st)
    assert int(ot["level"][: ot["ncells"]].max()) > 16
    assert tree_equal(gt, ot), first_tree_difference(gt, ot)
    oracle.quad_exact(ot, st["x"], st["y"], st["mass"])
    order = oracle.hilbert_order(ot, n)
    _, _, it = oracle.walk(ot, st["x"], st["y"], st["mass"], obb, 0.5, eps2, 64, opening=1,
                           quad=True, order=order)
    assert np.array_equal(f["inter"], it)
    # the curve to a deeper level is a different order: its groups differ
    deeper = oracle.hilbert_order(ot, n, 32)
    assert not np.array_equal(order, deeper)
