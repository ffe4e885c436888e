"""GPU ensembles: many independent simulations advanced by one context.

Each instance of an ensemble must evolve exactly as it would alone: its
tree equals the oracle's tree of its own bodies (bit-exact), and its forces,
interaction counts and updated state equal those of a single-instance
context on the same bodies bit for bit (same tree, same walk groups, same
traversal).  Instances of 1 and 2 bodies and sizes that are not multiples of
the walk group are included.
"""
import numpy as np
import pytest

from tests.helpers import first_tree_difference, ic, tree_equal

pytestmark = pytest.mark.gpu

COUNTS = [1000, 64, 5000, 1, 2, 333, 4096, 777, 65, 20000]
EPS2 = float(np.float32(1e-4))


@pytest.fixture(scope="module")
def S():
    from tests.conftest import gpu_available

    if not gpu_available():
        pytest.skip("no GPU")
    import stellarsolver_amd as S

    S.load_library()
    return S


def members(S, counts, base_seed=7):
    """Per-instance initial conditions: models cycle disk, Plummer, colliding
    disks; seed base_seed + i."""
    out = []
    for i, n in enumerate(counts):
        p = S.ic_params([0, 1, 2][i % 3], seed=base_seed + i)
        out.append(S.generate_ic(p, n))
    return out


def concat(sts):
    return {k: np.concatenate([s[k] for s in sts]) for k in ("mass", "x", "y", "vx", "vy")}


def batch_engine(S, sts, **kw):
    p = S.default_params(eps2=EPS2, **kw)
    e = S.Engine(params=p, counts=[s["x"].size for s in sts])
    st = concat(sts)
    e.upload(st["mass"], st["x"], st["y"], st["vx"], st["vy"])
    return e


def single_engine(S, st, **kw):
    p = S.default_params(eps2=EPS2, **kw)
    e = S.Engine(st["x"].size, p)
    e.upload(st["mass"], st["x"], st["y"], st["vx"], st["vy"])
    return e


def test_batch_layout(S):
    sts = members(S, COUNTS)
    with batch_engine(S, sts) as e:
        assert e.ninst == len(COUNTS)
        assert e.n == sum(COUNTS)
        off = np.zeros(len(COUNTS) + 1, np.int32)
        assert S.load_library().stellar_instance_offsets(e._h, S._ip(off)) == 0
        assert np.array_equal(off, np.concatenate([[0], np.cumsum(COUNTS)]))


def test_batch_trees_bit_exact(S, oracle):
    sts = members(S, COUNTS)
    with batch_engine(S, sts) as e:
        e.bbox()
        e.build()
        e.com()
        for i, st in enumerate(sts):
            print(f"instance {i} n={st['x'].size}", flush=True)
            bb = oracle.bbox(st["x"], st["y"])
            assert np.array_equal(e.get_bbox(i), np.asarray(bb, np.float32))
            kt = oracle.build_keys(st["x"], st["y"], bb)
            oracle.com_exact(kt, st["x"], st["y"], st["mass"])
            t = e.export_tree(instance=i)
            assert tree_equal(t, kt), first_tree_difference(t, kt)
            for k in ("com_x", "com_y", "com_m"):
                assert np.array_equal(t[k], kt[k]), k


@pytest.mark.parametrize("group,mixed", [(64, False), (32, False), (16, False), (64, True)])
def test_batch_steps_match_single_contexts(S, group, mixed):
    sts = members(S, COUNTS, base_seed=11)
    steps = 3
    kw = {"group": group}
    if mixed:  # the mixed-precision far field (quadrupoles required)
        kw.update(precision=S.PRECISION_MIXED, flags=S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING)
    with batch_engine(S, sts, **kw) as e:
        e.run(steps)
        got = e.download()
        e.bbox()
        e.build()
        e.com()
        e.sort()
        e.forces()
        fb = e.export_forces()
        total = e.last_interactions()
    off = np.concatenate([[0], np.cumsum(COUNTS)])
    tot = 0
    for i, st in enumerate(sts):
        print(f"instance {i} n={st['x'].size}", flush=True)
        with single_engine(S, st, **kw) as s:
            s.run(steps)
            want = s.download()
            s.bbox()
            s.build()
            s.com()
            s.sort()
            s.forces()
            fs = s.export_forces()
            tot += s.last_interactions()
        sl = slice(off[i], off[i + 1])
        for k in want:
            assert np.array_equal(got[k][sl], want[k]), (i, k)
        for k in fs:
            assert np.array_equal(fb[k][sl], fs[k]), (i, k)
    assert total == tot


def test_batch_fp64_diagnostics(S):
    sts = members(S, COUNTS[:6], base_seed=3)
    with batch_engine(S, sts) as e:
        e.run(2)
        ens = e.energy_fp64_instances()
        dia = e.diagnostics_fp64()
        tot = e.energy_fp64()
        ids = np.array([0, 1000, 1063, 1064, 6064, 6065, 6066], np.int32)
        bx, by = e.direct_fp64(ids)
        pos = e.download()
    off = np.concatenate([[0], np.cumsum(COUNTS[:6])])
    for i, st in enumerate(sts):
        with single_engine(S, st) as s:
            s.run(2)
            en = s.energy_fp64()
            ds = s.diagnostics_fp64()
            for k in en:
                assert ens[k][i] == pytest.approx(en[k], rel=1e-12, abs=1e-12), (i, k)
            for k in ds:
                assert dia[k][i] == pytest.approx(ds[k][0], rel=1e-12, abs=1e-12), (i, k)
            loc = ids[(ids >= off[i]) & (ids < off[i + 1])]
            if loc.size:
                sx, sy = s.direct_fp64(loc - off[i])
                sel = np.isin(ids, loc)
                np.testing.assert_allclose(bx[sel], sx, rtol=1e-12, atol=0)
                np.testing.assert_allclose(by[sel], sy, rtol=1e-12, atol=0)
    for k in ("kinetic", "potential", "px", "py"):
        assert tot[k] == pytest.approx(float(np.sum(ens[k])), rel=1e-12, abs=1e-12)
    assert np.all(np.isfinite(pos["x"]))


def test_batch_rejects_empty_instance(S):
    p = S.default_params()
    with pytest.raises(S.StellarError):
        S.Engine(params=p, counts=[10, 0, 5])


def run_states(S, make, steps_plan, graph):
    """Run the step counts of steps_plan one stellar_run call each (graph or
    direct), downloading the state after each call."""
    out = []
    with make(graph) as e:
        for k in steps_plan:
            e.run(k)
            out.append(e.download())
    return out


@pytest.mark.parametrize("kind", ["single", "ensemble"])
def test_graph_replays_match_direct_steps(S, kind):
    """Captured two-step graphs, replayed several times with direct steps and
    host work in between, give the direct steps' results bit for bit."""
    if kind == "single":
        st = ic(2, 20000)

        def make(graph):
            flags = S.FLAG_GRAPH if graph else 0
            p = S.default_params(eps2=EPS2, flags=flags)
            e = S.Engine(st["x"].size, p)
            e.upload(st["mass"], st["x"], st["y"], st["vx"], st["vy"])
            return e
    else:
        sts = members(S, COUNTS, base_seed=5)

        def make(graph):
            flags = S.FLAG_GRAPH if graph else 0
            return batch_engine(S, sts, flags=flags)
    plan = [4, 3, 2, 5]
    want = run_states(S, make, plan, graph=False)
    got = run_states(S, make, plan, graph=True)
    for i, (a, b) in enumerate(zip(want, got)):
        for k in a:
            assert np.array_equal(a[k], b[k]), (i, k)


def test_autotune_keeps_interactions(S):
    st = ic(2, 65536)
    with single_engine(S, st) as ref:
        ref.bbox()
        ref.build()
        ref.com()
        ref.sort()
        ref.forces()
        f0 = ref.export_forces()
    with single_engine(S, st) as e:
        t = e.autotune(group_mask=0, reps=2)
        print("tuning", t, flush=True)
        assert t["group"] == 64 and t["nodes_per_pop"] in (1, 2, 4)
        assert t["block_threads"] in (64, 128, 256, 512, 1024) and t["walk_ms"] > 0
        e.bbox()
        e.build()
        e.com()
        e.sort()
        e.forces()
        f1 = e.export_forces()
        # a second context of the same size class takes the cached choice
        with single_engine(S, st) as e2:
            t2 = e2.autotune(group_mask=0, reps=2)
            assert t2 == t
    assert np.array_equal(f0["inter"], f1["inter"])
    num = (f1["ax"].astype(np.float64) - f0["ax"]) ** 2 + (f1["ay"].astype(np.float64) - f0["ay"]) ** 2
    den = f0["ax"].astype(np.float64) ** 2 + f0["ay"].astype(np.float64) ** 2
    assert np.sqrt(np.mean(num / den)) < 1e-6


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_ensemble_256_full_scale(S, precision):
    """The batched-ensemble workload at its real scale: 256 instances, N_i
    log-uniform in [1024, 65536] (multiples of 64, about 3.8M bodies), models
    cycling disk / Plummer / colliding disks, quadrupoles + textbook opening,
    graph-captured steps.  Sampled instances evolve bit for bit as single
    contexts do, and the forces of sampled bodies are within 2e-3 (rms
    relative) of the FP64 direct sum over their own instance."""
    rng = np.random.default_rng(3)
    counts = np.exp(rng.uniform(np.log(1024), np.log(65536), 256))
    counts = (np.round(counts / 64) * 64).astype(np.int64).clip(1024, 65536).tolist()
    sts = [S.generate_ic(S.ic_params(i % 3, seed=3000 + i), int(c)) for i, c in enumerate(counts)]
    kw = {"flags": S.FLAG_QUADRUPOLE | S.FLAG_TEXTBOOK_OPENING}
    if precision == "mixed":
        kw["precision"] = S.PRECISION_MIXED
    print(f"ensemble 256 precision={precision} bodies={sum(counts)}", flush=True)
    with batch_engine(S, sts, **kw) as e:
        e.set_params(flags=kw["flags"] | S.FLAG_GRAPH)
        e.run(4)
        got = e.download()
        e.set_params(flags=kw["flags"])
        for stage in ("bbox", "build", "com", "sort", "forces"):
            getattr(e, stage)()
        f = e.export_forces()
        off = np.concatenate([[0], np.cumsum(counts)])
        pick = [0, 1, 97, 255, int(np.argmax(counts))]
        tg = np.concatenate([off[i] + np.arange(0, counts[i], 61) for i in pick]).astype(np.int32)
        dx, dy = e.direct_fp64(tg)
    num = (f["ax"][tg].astype(np.float64) - dx) ** 2 + (f["ay"][tg].astype(np.float64) - dy) ** 2
    err = float(np.sqrt(np.mean(num / (dx ** 2 + dy ** 2))))
    print(f"ensemble rms force error {err:.2e}", flush=True)
    assert err < 2e-3
    for i in pick:
        with single_engine(S, sts[i], **kw) as s:
            s.run(4)
            want = s.download()
        sl = slice(off[i], off[i + 1])
        for k in want:
            assert np.array_equal(got[k][sl], want[k]), (i, k)
