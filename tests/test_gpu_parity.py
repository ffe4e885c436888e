"""GPU parity: libstellar_hip (through its C ABI) against the CPU oracle.

Bar: bit-exact for the bounding box, the tree (canonical child/count/start/
level/sorted), the centres of mass, the per-body interaction counts and the
update; forces within the rsqrt tolerance stated in each test (the GPU uses
the hardware v_rsq_f32, the oracle a correctly rounded 1/sqrt).
"""
import numpy as np
import pytest

from tests.helpers import first_tree_difference, ic, rms_rel, tree_equal

pytestmark = pytest.mark.gpu

EPS = {0: 0.01, 1: 0.005, 2: 0.01, 4: 1e-4}

# Forces: the walk makes identical accept/open decisions (bit-exact counts);
# the accumulated values differ only through rsqrt rounding (<= 2 ulp per
# term) and so agree to a few parts in 1e7.
FORCE_RMS_TOL = 2e-6
FORCE_MAX_TOL = 1e-4


@pytest.fixture(scope="module")
def S():
    from tests.conftest import gpu_available

    if not gpu_available():
        pytest.skip("no GPU")
    import stellarsolver_amd as S

    S.load_library()
    return S


def make_engine(S, st, eps2, theta=0.5, group=64, cap=0, **kw):
    n = st["x"].size
    p = S.default_params(theta=theta, eps2=eps2, group=group, **kw)
    e = S.Engine(n, p, node_capacity=cap)
    e.upload(st["mass"], st["x"], st["y"], st["vx"], st["vy"])
    return e


def oracle_tree(oracle, st):
    bb = oracle.bbox(st["x"], st["y"])
    kt = oracle.build_keys(st["x"], st["y"], bb)
    oracle.com_exact(kt, st["x"], st["y"], st["mass"])
    return bb, kt


def gpu_tree(e):
    e.bbox()
    bb = e.get_bbox()
    e.build()
    e.com()
    t = e.export_tree()
    return bb, t


@pytest.mark.parametrize("workload,n", [(0, 4096), (1, 65536), (2, 65536), (1, 262144),
                                        (4, 1 << 20)])
def test_tree_bit_exact(S, oracle, workload, n):
    print(f"tree parity workload={workload} n={n}", flush=True)
    st = ic(workload, n)
    eps2 = np.float32(EPS[workload] ** 2)
    with make_engine(S, st, eps2) as e:
        gbb, gt = gpu_tree(e)
    obb, ot = oracle_tree(oracle, st)
    assert np.array_equal(gbb, obb)
    assert tree_equal(gt, ot), first_tree_difference(gt, ot)
    for k in ("com_x", "com_y", "com_m"):
        assert np.array_equal(gt[k], ot[k]), k


@pytest.mark.parametrize("quad", [False, True])
def test_tree_over_stale_blocks(S, oracle, quad):
    """The blocks a build fills are cleared by that build, not only when the
    context is created: a tree built where earlier steps and a different body
    set left their trees is bit-exact against the oracle."""
    print(f"tree over stale blocks quad={quad}", flush=True)
    n = 65536
    flags = (S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING) if quad else 0
    eps2 = float(np.float32(EPS[2] ** 2))
    with make_engine(S, ic(2, n), eps2, flags=flags) as e:
        e.run(3)
        for w in (1, 0, 2):
            st = ic(w, n)
            e.upload(st["mass"], st["x"], st["y"], st["vx"], st["vy"])
            e.run(2)
            d = e.download()
            gbb, gt = gpu_tree(e)
            cur = {"x": d["x"], "y": d["y"], "mass": d["mass"]}
            obb, ot = oracle_tree(oracle, cur)
            assert np.array_equal(gbb, obb)
            assert tree_equal(gt, ot), (w, first_tree_difference(gt, ot))
            for k in ("com_x", "com_y", "com_m"):
                assert np.array_equal(gt[k], ot[k]), (w, k)
            e.sort()
            e.forces()
            f = e.export_forces()
            kw = {"opening": 1, "quad": True} if quad else {}
            if quad:
                oracle.quad_exact(ot, cur["x"], cur["y"], cur["mass"])
            _, _, it = oracle.walk(ot, cur["x"], cur["y"], cur["mass"], obb, 0.5, eps2, 64, **kw)
            assert np.array_equal(f["inter"], it), w


# Mixed precision: the GPU and the oracle compute identical FP16 inputs except
# where the hardware rsqrt's last bit moves a value across an FP16 rounding
# boundary (rare), so the forces agree to FP32-walk precision in rms.
MIXED_RMS_TOL = 2e-5


@pytest.mark.parametrize("schedule", [0, 1])
@pytest.mark.parametrize("workload,n", [(2, 65536), (1, 65536), (0, 4096), (4, 65536)])
def test_mixed_precision_parity(S, oracle, workload, n, schedule):
    """STELLAR_PRECISION_MIXED: interaction counts equal the oracle's (the vote
    is the FP32 walk's), forces within MIXED_RMS_TOL of the oracle's FP16
    restatement, and against the FP64 direct sum at most 4e-4 (rms relative)
    less accurate than the FP32 quadrupole walk of the same tree."""
    print(f"mixed parity workload={workload} n={n} schedule={schedule}", flush=True)
    st = ic(workload, n)
    eps2 = float(np.float32(EPS[workload] ** 2))
    flags = S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING
    with make_engine(S, st, eps2, flags=flags, precision=S.PRECISION_MIXED) as e:
        e.set_walk_schedule(schedule)
        gpu_tree(e)
        e.sort()
        e.forces()
        f = e.export_forces()
        tg = np.arange(0, n, max(1, n // 2048), dtype=np.int32)
        dx, dy = e.direct_fp64(tg, eps2=eps2)
    x, y, m = st["x"], st["y"], st["mass"]
    obb, ot = oracle_tree(oracle, st)
    oracle.quad_exact(ot, x, y, m)
    oracle.quad_mixed(ot, x, y, m, obb)
    ax, ay, it = oracle.walk(ot, x, y, m, obb, 0.5, eps2, 64, opening=1, quad=True, mixed=True)
    assert np.array_equal(f["inter"], it)
    assert rms_rel(f["ax"], f["ay"], ax, ay) < MIXED_RMS_TOL
    qx, qy, _ = oracle.walk(ot, x, y, m, obb, 0.5, eps2, 64, opening=1, quad=True)
    fp32_err = rms_rel(qx[tg], qy[tg], dx, dy)
    assert rms_rel(f["ax"][tg], f["ay"][tg], dx, dy) < fp32_err + 4e-4


def test_mixed_precision_needs_quadrupoles_and_group_64(S):
    st = ic(2, 4096)
    for kw in ({"flags": S.FLAG_TEXTBOOK_OPENING}, {"flags": S.FLAG_QUADRUPOLE, "group": 32}):
        with pytest.raises(S.StellarError):
            make_engine(S, st, 1e-4, precision=S.PRECISION_MIXED, **kw)


@pytest.mark.parametrize("group,workload,n",
                         [(g, w, n) for g in (64, 32, 16) for w, n in ((0, 4096), (2, 65536), (1, 65536))]
                         + [(64, 1, 262144)])
def test_forces_parity(S, oracle, workload, n, group):
    print(f"force parity workload={workload} n={n} group={group}", flush=True)
    st = ic(workload, n)
    eps2 = float(np.float32(EPS[workload] ** 2))
    with make_engine(S, st, eps2, group=group) as e:
        gbb, gt = gpu_tree(e)
        e.sort()
        e.forces()
        f = e.export_forces()
        total = e.last_interactions()
    obb, ot = oracle_tree(oracle, st)
    ax, ay, it = oracle.walk(ot, st["x"], st["y"], st["mass"], obb, 0.5, eps2, group)
    assert np.array_equal(f["inter"], it)
    assert total == int(it.astype(np.int64).sum())
    mag = np.hypot(ax.astype(np.float64), ay.astype(np.float64))
    diff = np.hypot(f["ax"] - ax.astype(np.float64), f["ay"] - ay.astype(np.float64))
    # at 2^18 Plummer bodies the dense core's bodies sum ~2x the terms, with
    # more cancellation: the same per-term rsqrt rounding shows as ~2e-6
    tol = FORCE_RMS_TOL if n <= 65536 else 2.5 * FORCE_RMS_TOL
    assert np.sqrt(np.mean((diff / mag) ** 2)) < tol
    # per body, against a floor of the 10th-percentile magnitude: a body whose
    # terms nearly cancel has a tiny |a| and no meaningful relative error
    assert (diff / np.maximum(mag, np.percentile(mag, 10))).max() < FORCE_MAX_TOL


@pytest.mark.parametrize("cap", [64, 1000])
def test_sort_multipass_buckets(S, oracle, cap):
    """The key sort's buckets taken through the multi-pass path (sorted in
    runs of `cap` keys, then merged): the tree stays bit-exact, and three
    steps equal those of the default sort bit for bit."""
    print(f"sort multipass cap={cap}", flush=True)
    st = ic(2, 65536)
    eps2 = float(np.float32(EPS[2] ** 2))
    with make_engine(S, st, eps2) as e:
        assert S.load_library().stellar_debug_sort_cap(e._h, cap) == 0
        gbb, gt = gpu_tree(e)
        e.sort()
        e.forces()
        e.update()
        e.run(2)
        got = e.download()
    obb, ot = oracle_tree(oracle, st)
    assert np.array_equal(gbb, obb)
    assert tree_equal(gt, ot), first_tree_difference(gt, ot)
    with make_engine(S, st, eps2) as e:
        e.run(3)
        want = e.download()
    for k in want:
        assert np.array_equal(got[k], want[k]), k


def test_sort_dense_clusters(S, oracle):
    """Bodies much closer together than their sort bucket's 32-bit key window
    resolves: a 12 x 12 grid (a run of equal windows longer than the
    per-thread fix-up takes, so the bucket falls back to the full-key sort),
    a 4 x 2 grid and pairs (runs fixed up by insertion).  The tree stays
    bit-exact."""
    print("sort dense clusters", flush=True)
    st = ic(0, 65536)
    x, y = st["x"].copy(), st["y"].copy()
    h = np.float32(5e-7)
    gi, gj = np.meshgrid(np.arange(12), np.arange(12), indexing="ij")
    big = slice(65536 - 144, 65536)
    x[big] = np.float32(1.0) + h * gi.ravel().astype(np.float32)
    y[big] = np.float32(1.0) + h * gj.ravel().astype(np.float32)
    gi, gj = np.meshgrid(np.arange(4), np.arange(2), indexing="ij")
    small = slice(65536 - 152, 65536 - 144)
    x[small] = np.float32(-1.0) + h * gi.ravel().astype(np.float32)
    y[small] = np.float32(0.5) + h * gj.ravel().astype(np.float32)
    for k in range(16):  # pairs
        x[1000 + k] = x[2000 + k] + np.float32(4.0) * np.spacing(x[2000 + k])
        y[1000 + k] = y[2000 + k]
    st = dict(st, x=x, y=y)
    for q in ("vx", "vy"):
        st[q] = st[q].copy()
        st[q][65536 - 152:] = 0.0
    eps2 = float(np.float32(EPS[0] ** 2))
    with make_engine(S, st, eps2) as e:
        gbb, gt = gpu_tree(e)
    obb, ot = oracle_tree(oracle, st)
    assert np.array_equal(gbb, obb)
    assert tree_equal(gt, ot), first_tree_difference(gt, ot)
    for k in ("com_x", "com_y", "com_m"):
        assert np.array_equal(gt[k], ot[k]), k


@pytest.mark.parametrize("graph", [False, True])
def test_radix_sort_path(S, oracle, graph):
    """The device-wide radix sort that contexts above 2^24 bodies use (forced
    here at 2^16): the tree is bit-exact, and four steps -- direct, or two
    replays of a captured two-step graph (rocPRIM's merge-sort variant) --
    equal those of the sample sort bit for bit."""
    print(f"radix sort path graph={graph}", flush=True)
    st = ic(2, 65536)
    eps2 = float(np.float32(EPS[2] ** 2))
    lib = S.load_library()
    with make_engine(S, st, eps2) as e:
        assert lib.stellar_debug_sort_cap(e._h, -1) == 0
        gbb, gt = gpu_tree(e)
    obb, ot = oracle_tree(oracle, st)
    assert np.array_equal(gbb, obb)
    assert tree_equal(gt, ot), first_tree_difference(gt, ot)
    flags = S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING
    with make_engine(S, st, eps2, flags=flags | (S.FLAG_GRAPH if graph else 0)) as e:
        assert lib.stellar_debug_sort_cap(e._h, -1) == 0
        e.run(4)
        got = e.download()
    with make_engine(S, st, eps2, flags=flags) as e:
        e.run(4)
        want = e.download()
    for k in want:
        assert np.array_equal(got[k], want[k]), k


def test_stage_timing_modes(S):
    """Eight HIP events per step time every stage (zero for a stage that
    launched nothing: the sort stage only swaps buffers, the output buffer is
    filled when read, and the update runs inside walks that pop one node at
    a time); the forces-only mode times the walk alone and reports zeros for
    the rest."""
    st = ic(2, 65536)
    eps2 = float(np.float32(EPS[2] ** 2))
    with make_engine(S, st, eps2, flags=S.FLAG_STAGE_TIMING) as e:  # monopoles: 2 nodes per pop
        e.run(3)
        full = e.stage_times()
    quad = S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING
    with make_engine(S, st, eps2, flags=S.FLAG_STAGE_TIMING | quad) as e:
        e.run(3)
        fused = e.stage_times()
    with make_engine(S, st, eps2, flags=S.FLAG_FORCES_TIMING) as e:
        e.run(3)
        walk = e.stage_times()
    assert all(full[k] > 0 for k in ("bbox", "build", "com", "forces", "update"))
    assert full["sort"] == 0 and full["output"] == 0
    assert all(fused[k] > 0 for k in ("bbox", "build", "com", "forces"))
    assert fused["sort"] == 0 and fused["update"] == 0
    assert walk["forces"] > 0
    assert all(walk[k] == 0 for k in walk if k != "forces")


def test_sort_cap_rejected(S):
    st = ic(0, 4096)
    with make_engine(S, st, 1e-4) as e:
        lib = S.load_library()
        assert lib.stellar_debug_sort_cap(e._h, 0) == S.STELLAR_ERR_ARG
        assert lib.stellar_debug_sort_cap(e._h, 2049) == S.STELLAR_ERR_ARG


def test_steps_match_oracle(S, oracle):
    """Each GPU step = oracle walk on the GPU's own pre-step state, then the
    oracle update applied to the GPU forces (bit-exact)."""
    n, workload = 20000, 2
    st = ic(workload, n)
    eps2 = float(np.float32(1e-4))
    dt, d = np.float32(1e-3), np.float32(1.0)
    with make_engine(S, st, eps2, timestep=float(dt), dampening=float(d)) as e:
        for k in range(3):
            print(f"step {k}", flush=True)
            pre = e.download()
            e.update()
            post = e.download()
            obb, ot = oracle_tree(oracle, pre)
            ax, ay, it = oracle.walk(ot, pre["x"], pre["y"], pre["mass"], obb, 0.5, eps2, 64)
            assert rms_rel(post["ax"], post["ay"], ax, ay) < FORCE_RMS_TOL
            x, y, vx, vy = oracle.update(pre["x"], pre["y"], pre["vx"], pre["vy"], post["ax"],
                                         post["ay"], dt, d)
            for a, b in ((x, post["x"]), (y, post["y"]), (vx, post["vx"]), (vy, post["vy"])):
                assert np.array_equal(a, b)
            out = e.get_output_buffer()
            assert np.array_equal(out[0::2], post["x"]) and np.array_equal(out[1::2], post["y"])


@pytest.mark.parametrize("model", ["quad", "quad_queue", "bmax", "mixed", "mono_kpop1"])
def test_fused_update_matches_separate_update(S, oracle, model):
    """Walks that pop one node at a time apply the step's update to their own
    bodies right after the forces: a step (stellar_step) equals the stage
    calls with the separate update kernel bit for bit, and equals the oracle
    update applied to its forces."""
    n = 20000
    st = ic(2, n)
    eps2 = float(np.float32(1e-4))
    dt, d = np.float32(1e-3), np.float32(0.999)
    flags = {"quad": S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING,
             "quad_queue": S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING,
             "bmax": S.FLAG_QUADRUPOLE | S.FLAG_BMAX_OPENING,
             "mixed": S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING,
             "mono_kpop1": 0}[model]
    prec = S.PRECISION_MIXED if model == "mixed" else S.PRECISION_FP32
    kw = dict(timestep=float(dt), dampening=float(d), flags=flags, precision=prec)
    with make_engine(S, st, eps2, **kw) as a, make_engine(S, st, eps2, **kw) as b:
        for e in (a, b):
            if model == "quad_queue":
                e.set_walk_schedule(S.WALK_QUEUE)
            if model == "mono_kpop1":
                e.set_walk_shape(1, 512)
        for k in range(3):
            print(f"{model} step {k}", flush=True)
            pre = a.download()
            a.update()  # one step: the update inside the walk
            b.bbox(); b.build(); b.com(); b.sort(); b.forces(); b.update_stage()
            pa, pb = a.download(), b.download()
            for key in ("x", "y", "vx", "vy", "ax", "ay"):
                assert np.array_equal(pa[key], pb[key]), (model, k, key)
            x, y, vx, vy = oracle.update(pre["x"], pre["y"], pre["vx"], pre["vy"], pa["ax"],
                                         pa["ay"], dt, d)
            for u, v in ((x, pa["x"]), (y, pa["y"]), (vx, pa["vx"]), (vy, pa["vy"])):
                assert np.array_equal(u, v), (model, k)


def test_run_is_deterministic(S):
    st = ic(1, 65536)
    eps2 = float(np.float32(0.005 ** 2))
    res = []
    for _ in range(2):
        with make_engine(S, st, eps2, timestep=5e-4) as e:
            e.run(5)
            res.append(e.download())
    for k in res[0]:
        assert np.array_equal(res[0][k], res[1][k]), k


@pytest.mark.parametrize("n", [1, 2, 3, 5, 63, 65, 100, 1000])
def test_small_sizes(S, oracle, n):
    rng = np.random.default_rng(n)
    st = {"x": rng.normal(0, 3, n).astype(np.float32), "y": rng.normal(1, 2, n).astype(np.float32),
          "vx": np.zeros(n, np.float32), "vy": np.zeros(n, np.float32),
          "mass": rng.uniform(0.5, 2, n).astype(np.float32)}
    with make_engine(S, st, 0.01) as e:
        gbb, gt = gpu_tree(e)
        e.sort()
        e.forces()
        f = e.export_forces()
    obb, ot = oracle_tree(oracle, st)
    assert tree_equal(gt, ot), first_tree_difference(gt, ot)
    ax, ay, it = oracle.walk(ot, st["x"], st["y"], st["mass"], obb, 0.5, 0.01, 64)
    assert np.array_equal(f["inter"], it)
    if n > 1:
        assert rms_rel(f["ax"], f["ay"], ax, ay) < FORCE_RMS_TOL


def test_three_body_known_answer(S):
    st = {"x": np.array([1, 2, 3], np.float32), "y": np.array([1, 2, 3], np.float32),
          "vx": np.zeros(3, np.float32), "vy": np.zeros(3, np.float32),
          "mass": np.ones(3, np.float32)}
    with make_engine(S, st, 0.01) as e:
        bb, t = gpu_tree(e)
    assert list(bb) == [0, 3, 0, 3]
    assert list(t["root"]) == [3, -1, -1, 0] and list(t["child"]) == [2, -1, -1, 1]
    assert list(t["sorted"]) == [2, 1, 0] and list(t["count"]) == [2]
    assert t["com_x"][0] == 2.5 and t["com_m"][0] == 2.0


def test_coincident_bodies_bucket(S, oracle):
    rng = np.random.default_rng(1)
    n = 200
    x = rng.normal(0, 1, n).astype(np.float32)
    y = rng.normal(0, 1, n).astype(np.float32)
    x[10:13] = x[5]
    y[10:13] = y[5]  # four bodies at one point -> one level-64 cell
    st = {"x": x, "y": y, "vx": np.zeros(n, np.float32), "vy": np.zeros(n, np.float32),
          "mass": np.ones(n, np.float32)}
    with make_engine(S, st, 0.01) as e:
        gbb, gt = gpu_tree(e)
        e.sort()
        e.forces()
        f = e.export_forces()
    obb, ot = oracle_tree(oracle, st)
    assert tree_equal(gt, ot), first_tree_difference(gt, ot)
    assert (gt["level"] == 64).sum() == 1
    ax, ay, it = oracle.walk(ot, x, y, st["mass"], obb, 0.5, 0.01, 64)
    assert np.array_equal(f["inter"], it)


def test_error_codes(S):
    n = 100
    x = np.zeros(n, np.float32)
    st = {"x": x, "y": x, "vx": x, "vy": x, "mass": np.ones(n, np.float32)}
    with make_engine(S, st, 0.01) as e:
        e.bbox()
        with pytest.raises(S.StellarError) as err:
            e.build()
        assert err.value.status == S.STELLAR_ERR_BUCKET
    st = ic(1, 5000)
    with make_engine(S, st, 0.01, cap=100) as e:
        e.bbox()
        with pytest.raises(S.StellarError) as err:
            e.build()
        assert err.value.status == S.STELLAR_ERR_POOL
    with make_engine(S, st, 0.01) as e:
        with pytest.raises(S.StellarError) as err:
            e.forces()
        assert err.value.status == S.STELLAR_ERR_STATE
        e.update()  # a full step still works after a refused stage


@pytest.mark.parametrize("how", ["step", "run"])
def test_failed_step_leaves_state_untouched(S, how):
    """A step whose build overflows the node pool reports STELLAR_ERR_POOL and
    moves no body: the state downloaded afterwards is the one uploaded (the
    sort stage copies it unchanged, the update is skipped)."""
    st = ic(1, 5000)
    with make_engine(S, st, 0.01, cap=100) as e:
        with pytest.raises(S.StellarError) as err:
            e.update() if how == "step" else e.run(3)
        assert err.value.status == S.STELLAR_ERR_POOL
        d = e.download()
    for k in ("x", "y", "vx", "vy", "mass"):
        assert np.array_equal(d[k], st[k]), k


def test_changed_force_model_needs_a_new_tree(S):
    """The build and centre-of-mass stages store per-child thresholds and
    moments for the force model: after theta, eps2 or the opening/multipole
    flags change, stellar_forces refuses the old tree (STELLAR_ERR_STATE);
    timing flags keep it."""
    st = ic(2, 8192)
    with make_engine(S, st, 0.01) as e:
        for change in ({"theta": 0.4}, {"eps2": 2e-4}, {"flags": S.FLAG_QUADRUPOLE},
                       {"flags": S.FLAG_TEXTBOOK_OPENING}):
            e.set_params(theta=0.5, eps2=1e-4, flags=0)
            e.bbox()
            e.build()
            e.com()
            e.sort()
            e.set_params(flags=S.FLAG_FORCES_TIMING)
            e.set_params(flags=0)
            e.forces()  # timing flags do not touch the tree
            e.set_params(**change)
            with pytest.raises(S.StellarError) as err:
                e.forces()
            assert err.value.status == S.STELLAR_ERR_STATE, change


def test_direct_and_energy_fp64(S, oracle):
    st = ic(0, 4096)
    with make_engine(S, st, 1e-4) as e:
        tg = np.arange(0, 4096, 7, dtype=np.int32)
        gx, gy = e.direct_fp64(tg, eps2=1e-4)
        en = e.energy_fp64(eps2=1e-4)
    ox, oy = oracle.direct_fp64(st["x"], st["y"], st["mass"], 1e-4, targets=tg)
    assert np.allclose(gx, ox, rtol=1e-11, atol=1e-13) and np.allclose(gy, oy, rtol=1e-11, atol=1e-13)
    oe = oracle.energy(st["x"], st["y"], st["vx"], st["vy"], st["mass"], 1e-4)
    for k in oe:
        assert np.isclose(en[k], oe[k], rtol=1e-11, atol=1e-9), k


def test_diagnostics_fp64_with_synchronised_velocities(S, oracle):
    """Energy, momentum and angular momentum on the device equal the oracle's
    FP64 sums, with raw velocities and with v + dt/2 a from the forces stage
    (the accelerations exported by the engine); the synchronised form is
    refused unless the forces stage was the last one run."""
    st = ic(1, 6000)
    eps2 = 1e-4
    half = 0.5e-3
    with make_engine(S, st, eps2) as e:
        e.run(2)
        raw = e.diagnostics_fp64(eps2=eps2)
        s0 = e.download()
        with pytest.raises(S.StellarError) as err:
            e.diagnostics_fp64(eps2=eps2, half_dt=half)
        assert err.value.status == S.STELLAR_ERR_STATE
        for stage in ("bbox", "build", "com", "sort", "forces"):
            getattr(e, stage)()
        f = e.export_forces()
        syn = e.diagnostics_fp64(eps2=eps2, half_dt=half)
    o = oracle.diagnostics(s0["x"], s0["y"], s0["vx"], s0["vy"], s0["mass"], eps2)
    os_ = oracle.diagnostics(s0["x"], s0["y"], s0["vx"], s0["vy"], s0["mass"], eps2,
                             ax=f["ax"], ay=f["ay"], half_dt=half)
    for k in o:
        print(k, raw[k][0], o[k], syn[k][0], os_[k], flush=True)
        assert raw[k][0] == pytest.approx(o[k], rel=1e-11, abs=1e-9), k
        assert syn[k][0] == pytest.approx(os_[k], rel=1e-11, abs=1e-9), k
    assert syn["kinetic"][0] != raw["kinetic"][0]
    assert syn["potential"][0] == raw["potential"][0]


def test_headline_size_tree_and_counts(S, oracle):
    """N = 2^20 colliding disks: bit-exact tree and interaction counts; force
    error against the FP64 direct sum on a seeded 8192-body subsample."""
    n = 1 << 20
    print(f"headline parity n={n}", flush=True)
    st = ic(2, n)
    eps2 = float(np.float32(1e-4))
    with make_engine(S, st, eps2) as e:
        gbb, gt = gpu_tree(e)
        e.sort()
        e.forces()
        f = e.export_forces()
        tg = np.sort(np.random.default_rng(2).choice(n, 8192, replace=False)).astype(np.int32)
        dx, dy = e.direct_fp64(tg)
    obb, ot = oracle_tree(oracle, st)
    assert tree_equal(gt, ot), first_tree_difference(gt, ot)
    ax, ay, it = oracle.walk(ot, st["x"], st["y"], st["mass"], obb, 0.5, eps2, 64)
    assert np.array_equal(f["inter"], it)
    err = rms_rel(f["ax"][tg], f["ay"][tg], dx, dy)
    print(f"rms relative error vs FP64 direct sum: {err:.3e}")
    # the reference opening test at theta = 0.5 is a monopole walk with an
    # effective opening angle sqrt(0.5); measured 5.4e-2 here
    assert err < 0.1


@pytest.mark.parametrize("textbook", [False, True])
@pytest.mark.parametrize("workload,n", [(2, 65536), (1, 65536), (4, 65536)])
def test_quadrupole_parity(S, oracle, workload, n, textbook):
    """Quadrupole moments bit-exact against the oracle's FP64 restatement;
    interaction counts exact; forces within the rsqrt tolerance (the deep
    Plummer core, eps = 1e-4, is where the softening part of the quadrupole
    term is largest)."""
    print(f"quadrupole parity workload={workload} n={n} textbook={textbook}", flush=True)
    st = ic(workload, n)
    eps2 = float(np.float32(EPS[workload] ** 2))
    flags = S.FLAG_QUADRUPOLE | (S.FLAG_TEXTBOOK_OPENING if textbook else 0)
    with make_engine(S, st, eps2, flags=flags) as e:
        gbb, gt = gpu_tree(e)
        e.sort()
        e.forces()
        f = e.export_forces()
    obb, ot = oracle_tree(oracle, st)
    oracle.quad_exact(ot, st["x"], st["y"], st["mass"])
    for k in ("com_x", "com_y", "com_m"):
        assert np.array_equal(gt[k], ot[k]), k
    ax, ay, it = oracle.walk(ot, st["x"], st["y"], st["mass"], obb, 0.5, eps2, 64,
                             opening=int(textbook), quad=True)
    assert np.array_equal(f["inter"], it)
    assert rms_rel(f["ax"], f["ay"], ax, ay) < FORCE_RMS_TOL


@pytest.mark.parametrize("group", [32, 16])
@pytest.mark.parametrize("workload", [2, 4])
def test_quadrupole_parity_small_groups(S, oracle, workload, group):
    """Quadrupole walks of 32- and 16-body groups (several groups per
    wavefront, the softening term read from the node's second block):
    counts exact, forces within the rsqrt tolerance."""
    print(f"quadrupole parity small groups workload={workload} group={group}", flush=True)
    st = ic(workload, 65536)
    eps2 = float(np.float32(EPS[workload] ** 2))
    flags = S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING
    with make_engine(S, st, eps2, group=group, flags=flags) as e:
        gbb, gt = gpu_tree(e)
        e.sort()
        e.forces()
        f = e.export_forces()
    obb, ot = oracle_tree(oracle, st)
    oracle.quad_exact(ot, st["x"], st["y"], st["mass"])
    ax, ay, it = oracle.walk(ot, st["x"], st["y"], st["mass"], obb, 0.5, eps2, group,
                             opening=1, quad=True)
    assert np.array_equal(f["inter"], it)
    assert rms_rel(f["ax"], f["ay"], ax, ay) < FORCE_RMS_TOL


@pytest.mark.parametrize("kpop,threads,schedule", [(1, 512, 0), (2, 512, 0), (4, 256, 0), (1, 1024, 0),
                                                   (2, 64, 0), (1, 512, 1), (2, 256, 1),
                                                   (1, 512, 2), (2, 256, 2)])
@pytest.mark.parametrize("model", ["reference", "textbook+quad", "bmax+quad"])
def test_walk_shapes_parity(S, oracle, model, kpop, threads, schedule):
    """Every walk shape and schedule the tuner may pick (nodes per pop,
    workgroup size, static launch or persistent work queues) makes the
    oracle's decisions: interaction counts exact, forces within the rsqrt
    tolerance."""
    print(f"walk shape parity model={model} kpop={kpop} threads={threads} schedule={schedule}",
          flush=True)
    st = ic(2, 65536)
    eps2 = float(np.float32(EPS[2] ** 2))
    flags = {"reference": 0,
             "textbook+quad": S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING,
             "bmax+quad": S.FLAG_QUADRUPOLE | S.FLAG_BMAX_OPENING}[model]
    with make_engine(S, st, eps2, flags=flags) as e:
        e.set_walk_shape(kpop, threads)
        e.set_walk_schedule(schedule)
        gbb, gt = gpu_tree(e)
        e.sort()
        e.forces()
        f = e.export_forces()
        total = e.last_interactions()
    obb, ot = oracle_tree(oracle, st)
    x, y, m = st["x"], st["y"], st["mass"]
    kw = {}
    if model == "bmax+quad":
        oracle.bmax_exact(ot, x, y, m, 0.5, eps2)
        kw = {"quad": True, "bmax": True}
    if model != "reference":
        oracle.quad_exact(ot, x, y, m)
    if model == "textbook+quad":
        kw = {"opening": 1, "quad": True}
    ax, ay, it = oracle.walk(ot, x, y, m, obb, 0.5, eps2, 64, **kw)
    assert np.array_equal(f["inter"], it)
    assert total == int(it.astype(np.int64).sum())  # every group walked once
    assert rms_rel(f["ax"], f["ay"], ax, ay) < FORCE_RMS_TOL


@pytest.mark.parametrize("theta", [0.3, 0.8])
@pytest.mark.parametrize("model", ["reference", "textbook+quad"])
def test_theta_parity(S, oracle, theta, model):
    """Opening angles 0.3 and 0.8 (the colliding-disks runs at theta in
    {0.3, 0.5, 0.8}): exact interaction counts and forces within the rsqrt
    tolerance, for the reference's monopole walk and the quadrupole walk with
    the textbook test (whose thresholds the build stores per child)."""
    print(f"theta parity theta={theta} model={model}", flush=True)
    st = ic(2, 65536)
    eps2 = float(np.float32(EPS[2] ** 2))
    quad = model != "reference"
    flags = (S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING) if quad else 0
    with make_engine(S, st, eps2, theta=theta, flags=flags) as e:
        gpu_tree(e)
        e.sort()
        e.forces()
        f = e.export_forces()
    obb, ot = oracle_tree(oracle, st)
    if quad:
        oracle.quad_exact(ot, st["x"], st["y"], st["mass"])
    ax, ay, it = oracle.walk(ot, st["x"], st["y"], st["mass"], obb, theta, eps2, 64,
                             opening=int(quad), quad=quad)
    assert np.array_equal(f["inter"], it)
    assert rms_rel(f["ax"], f["ay"], ax, ay) < FORCE_RMS_TOL


def test_quadrupole_headline_error(S):
    """N = 2^20 colliding disks with quadrupoles: the force error against the
    FP64 direct sum on a seeded 8192-body subsample."""
    n = 1 << 20
    st = ic(2, n)
    eps2 = float(np.float32(1e-4))
    res = {}
    for name, flags in (("ref", S.FLAG_QUADRUPOLE),
                        ("textbook", S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING)):
        with make_engine(S, st, eps2, flags=flags) as e:
            gpu_tree(e)
            e.sort()
            e.forces()
            f = e.export_forces()
            tg = np.sort(np.random.default_rng(2).choice(n, 8192, replace=False)).astype(np.int32)
            dx, dy = e.direct_fp64(tg)
        res[name] = rms_rel(f["ax"][tg], f["ay"][tg], dx, dy)
        print(f"quadrupole {name} opening: rms relative error {res[name]:.3e}, "
              f"interactions/body {f['inter'].mean():.1f}", flush=True)
    assert res["textbook"] < 1e-3


@pytest.mark.parametrize("quad", [False, True])
@pytest.mark.parametrize("group", [64, 32, 16])
@pytest.mark.parametrize("workload,n", [(2, 65536), (1, 65536)])
def test_bmax_opening_parity(S, oracle, workload, n, group, quad):
    """Per-cell bmax thresholds: the interaction set equals the oracle's (so
    the FP64 bound and its FP32 threshold agree bit for bit); forces within
    the rsqrt tolerance."""
    print(f"bmax parity workload={workload} n={n} group={group} quad={quad}", flush=True)
    st = ic(workload, n)
    eps2 = float(np.float32(EPS[workload] ** 2))
    theta = 0.6
    flags = S.FLAG_BMAX_OPENING | (S.FLAG_QUADRUPOLE if quad else 0)
    with make_engine(S, st, eps2, theta=theta, group=group, flags=flags) as e:
        gpu_tree(e)
        e.sort()
        e.forces()
        f = e.export_forces()
    obb, ot = oracle_tree(oracle, st)
    x, y, m = st["x"], st["y"], st["mass"]
    oracle.bmax_exact(ot, x, y, m, theta, eps2)
    if quad:
        oracle.quad_exact(ot, x, y, m)
    ax, ay, it = oracle.walk(ot, x, y, m, obb, theta, eps2, group, quad=quad, bmax=True)
    assert np.array_equal(f["inter"], it)
    assert rms_rel(f["ax"], f["ay"], ax, ay) < FORCE_RMS_TOL


def test_deep_plummer_tree_and_forces(S, oracle):
    """Clustered core (Plummer a = 0.01, eps = 1e-4, the deep-tree workload at
    2^18 bodies, about 20 levels): the tree is bit-exact and the
    quadrupole/textbook walk agrees with the oracle's interaction set."""
    n = 1 << 18
    st = ic(4, n)
    eps2 = float(np.float32(1e-4 ** 2))
    flags = S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING
    with make_engine(S, st, eps2, flags=flags) as e:
        gbb, gt = gpu_tree(e)
        e.sort()
        e.forces()
        f = e.export_forces()
    obb, ot = oracle_tree(oracle, st)
    assert tree_equal(gt, ot), first_tree_difference(gt, ot)
    print(f"deep plummer: max level {int(gt['level'].max())}, cells {gt['ncells']}", flush=True)
    assert gt["level"].max() >= 20
    oracle.quad_exact(ot, st["x"], st["y"], st["mass"])
    ax, ay, it = oracle.walk(ot, st["x"], st["y"], st["mass"], obb, 0.5, eps2, 64,
                             opening=1, quad=True)
    assert np.array_equal(f["inter"], it)
    # ~1700 FP32 terms per body whose pulls largely cancel in the dense core:
    # summation order and rsqrt rounding leave a few parts in 1e6
    assert rms_rel(f["ax"], f["ay"], ax, ay) < 1e-5


@pytest.mark.parametrize("group", [64, 32, 16])
@pytest.mark.parametrize("case", ["disk", "body0", "body1"])
def test_reference_lone_root_body(S, oracle, case, group):
    """STELLAR_FLAG_REFERENCE_ROOT: a body alone in a root slot handled as the
    reference does (body 0 revisits the root's children, any other lone body
    exerts no force): counts equal the oracle's, forces within the rsqrt
    tolerance; the plain walk differs."""
    print(f"reference root case={case} group={group}", flush=True)
    if case == "disk":
        st = ic(0, 4096)
        st = {k: v.copy() for k, v in st.items()}
        st["x"][0], st["y"][0] = np.float32(-500.0), np.float32(-500.0)
    else:
        a, b = (0, 1) if case == "body0" else (1, 0)
        x = np.zeros(3, np.float32)
        y = np.zeros(3, np.float32)
        x[a], y[a] = 1.0, 1.0
        x[b], y[b] = -1.0, -1.0
        x[2], y[2] = -1.1, -1.05
        z = np.zeros(3, np.float32)
        st = {"x": x, "y": y, "vx": z, "vy": z, "mass": np.array([3.0, 1.0, 2.0], np.float32)}
    eps2 = 1e-4
    flags = S.FLAG_REFERENCE_ROOT
    with make_engine(S, st, eps2, group=group, flags=flags) as e:
        gpu_tree(e)
        e.sort()
        e.forces()
        f = e.export_forces()
    obb, ot = oracle_tree(oracle, st)
    ax, ay, it = oracle.walk(ot, st["x"], st["y"], st["mass"], obb, 0.5, eps2, group, ref_root=True)
    assert np.array_equal(f["inter"], it)
    assert rms_rel(f["ax"], f["ay"], ax, ay) < FORCE_RMS_TOL
    _, _, plain = oracle.walk(ot, st["x"], st["y"], st["mass"], obb, 0.5, eps2, group)
    assert not np.array_equal(it, plain)


def test_reference_root_needs_level_thresholds(S):
    st = ic(2, 4096)
    with pytest.raises(S.StellarError):
        make_engine(S, st, 1e-4, flags=S.FLAG_REFERENCE_ROOT | S.FLAG_QUADRUPOLE)
