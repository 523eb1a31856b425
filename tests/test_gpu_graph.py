"""The Henon step replayed from HIP graphs, alone and for an ensemble of clusters
(merged into one segmented launch per kernel, or each cluster's graph on its own
stream), against the same step enqueued eagerly: every field and scalar
bit-identical after every step (each replay is checked, with host reads and eager
work between replays)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("r", "m", "phi", "vr", "vt", "E", "J", "id", "rOld", "r_peri", "r_apo", "X", "Y", "sigma")
SCAL = ("n_max", "Mtotal", "Rtidal", "Dt", "TotalTime", "TidalMassLoss", "Eescaped", "Jescaped", "Etot", "Ekin",
        "Epot")


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _ctx(cm, cl):
    c = cm.Context(cl.n, cm.default_params(cl.n, orbit_r=cl.rtid, seed=7))
    c.upload(cl)
    c.init_cluster()
    return c


def _same(a, b, tag, fields=FIELDS):
    sa, sb = a.scalars(), b.scalars()
    for k in SCAL:
        assert sa[k] == sb[k], (tag, k, sa[k], sb[k])
    da, db = a.download(), b.download()
    n = sa["n_max"]
    for k in fields:
        assert np.array_equal(da[k][1:n + 1], db[k][1:n + 1]), (tag, k)


def test_gpu_step_graph_matches_eager(gpu):
    cm = gpu
    say("step graph vs eager, Plummer 5e4, 6 steps")
    cl = cm.make_plummer(50000, seed=9)
    eager, graph = _ctx(cm, cl), _ctx(cm, cl)
    try:
        for k in range(6):
            eager.step()
            graph.step_graph()
            _same(eager, graph, f"step {k}")
            say("step", k, "identical, N", graph.scalars()["n_max"])
    finally:
        eager.close()
        graph.close()


def test_gpu_ensemble_step_matches_eager(gpu):
    """Merged ensemble steps (both buffer states, then replays of the kept plans), a
    step of one cluster alone in between (its buffer state then differs from the
    others', so a plan for the mixed state is built), and a concurrent-streams step."""
    cm = gpu
    ns = [10000, 30000, 20000, 50000]
    say("ensemble of", len(ns), "clusters, 5 steps")
    cls = [cm.make_plummer(n, seed=3 + i) if i % 2 == 0 else cm.make_king(n, 5.0, seed=3 + i)
           for i, n in enumerate(ns)]
    eager = [_ctx(cm, cl) for cl in cls]
    ens = [_ctx(cm, cl) for cl in cls]
    try:
        for k in range(5):
            if k == 2:
                eager[0].step()
                ens[0].step_graph()
                _same(eager[0], ens[0], "cluster 0 alone")
            for c in eager:
                c.step()
            stats, errs = cm.ensemble_step(ens, concurrent=(k == 4))
            assert errs == [0] * len(ns)
            for i in range(len(ns)):
                _same(eager[i], ens[i], f"cluster {i} step {k}")
            if k < 4:
                launches, recorded = cm.ensemble_stats()
                say("step", k, "merged launches", launches, "recorded", recorded)
                assert launches * len(ns) == recorded
            say("step", k, "all clusters identical")
        # one cluster's context closed: the plans that stepped it are dropped, and the
        # other three step on together
        ens.pop(1).close()
        eager.pop(1).close()
        for k in range(2):
            for c in eager:
                c.step()
            stats, errs = cm.ensemble_step(ens)
            assert errs == [0] * len(ens)
            for i in range(len(ens)):
                _same(eager[i], ens[i], f"after close: cluster {i} step {k}")
        say("after closing one context: identical")
    finally:
        for c in eager + ens:
            c.close()


def test_gpu_ensemble_64_clusters_merged(gpu):
    """64 clusters of uneven N (log-uniform on [2e3, 3e4]), half Plummer and half King
    W0 = 3, 5, 7, 9, Kroupa masses: three merged steps, every cluster bit-identical to
    its own eager step after each; each kernel of the step is one launch for all 64.
    Then the clusters take three different orbit-sampler launch shapes: two more
    steps, still identical, with the orbit kernel three launches and every other
    kernel still one."""
    cm = gpu
    rng = np.random.default_rng(11)
    ns = np.exp(rng.uniform(np.log(2e3), np.log(3e4), 64)).astype(np.int64)
    say("ensemble of 64 clusters,", int(ns.sum()), "stars")
    eager, ens = [], []
    try:
        for i, n in enumerate(ns):
            n = int(n)
            if i < 32:
                cl = cm.make_plummer(n, seed=200 + i, kroupa=True)
            else:
                cl = cm.make_king(n, (3.0, 5.0, 7.0, 9.0)[i % 4], seed=200 + i, kroupa=True, segregate=False)
            for lst in (eager, ens):
                c = cm.Context(n, cm.default_params(n, orbit_r=cl.rtid, seed=200 + i))
                c.upload(cl)
                c.init_cluster()
                lst.append(c)
        base = None
        for k in range(5):
            if k == 3:
                for i in range(64):
                    for c in (eager[i], ens[i]):
                        c.test_hook(cm.HOOK_ORBIT_VARIANT, (0, 7, 10)[i % 3])
            for c in eager:
                c.step()
            stats, errs = cm.ensemble_step(ens)
            assert errs == [0] * 64
            launches, recorded = cm.ensemble_stats()
            say("step", k, "merged launches", launches, "recorded", recorded)
            if k < 3:
                assert launches * 64 == recorded
                base = launches
            else:
                assert launches == base + 2
            for i in range(64):
                _same(eager[i], ens[i], f"cluster {i} step {k}")
            say("step", k, "all 64 clusters identical")
    finally:
        for c in eager + ens:
            c.close()


def test_gpu_ensemble_learned_brackets_identical(gpu):
    """The merged ensemble with the learned brackets on (one network, trained on
    another cluster's harvested rows, set on every context): each cluster stays
    bit-identical to its own plain eager step, and most proposals are accepted."""
    cm = gpu
    say("ensemble with learned brackets, 4 clusters, 3 steps")
    tr = cm.make_plummer(50000, seed=21, kroupa=True)
    t = _ctx(cm, tr)
    try:
        t.mlp_init(seed=1, max_rows=1 << 20)
        t.mlp_harvest(True)
        for _ in range(2):
            t.step()
        t.mlp_harvest(False)
        rows = t.mlp_rows()
        t.mlp_fit_norm()
        t.mlp_train(400, batch=4096, lr=2e-3, seed=3)
        val = rows[::97][:8000]
        err = float(np.sqrt(np.mean((t.mlp_predict(val[:, :6]) - val[:, 6:]) ** 2)))
        theta, norm = t.mlp_params()
    finally:
        t.close()
    say("trained on", rows.shape[0], "rows, rms log r error", err)
    ns = [20000, 40000, 30000, 60000]
    cls = [cm.make_plummer(n, seed=30 + i, kroupa=True) if i % 2 == 0 else cm.make_king(n, 7.0, seed=30 + i)
           for i, n in enumerate(ns)]
    eager = [_ctx(cm, cl) for cl in cls]
    ens = [_ctx(cm, cl) for cl in cls]
    try:
        for c in ens:
            c.mlp_init(seed=1, max_rows=1 << 12)
            c.mlp_set_params(theta, norm)
            c.set_learned_brackets(True, margin=max(3.0 * err, 0.01))
        hit = miss = 0
        for k in range(3):
            for c in eager:
                c.step()
            stats, errs = cm.ensemble_step(ens)
            assert errs == [0] * len(ns)
            hit += sum(x["n_bracket_hit"] for x in stats)
            miss += sum(x["n_bracket_fallback"] for x in stats)
            for i in range(len(ns)):
                _same(eager[i], ens[i], f"cluster {i} step {k}")
            say("step", k, "all clusters identical")
        say("learned brackets accepted", hit, "rejected", miss)
        assert hit > 0.5 * (hit + miss)
    finally:
        for c in eager + ens:
            c.close()


def test_gpu_autotuned_shape_gives_identical_stars(gpu):
    """The autotuner picks one of the orbit sampler's launch shapes; whichever it
    picks, the stars after a step are bit-identical to the default shape's (only the
    block order of the escape sums may round differently)."""
    cm = gpu
    say("autotune, King W0=5 N=2e5")
    cl = cm.make_king(200000, 5.0, seed=4)
    a, b = _ctx(cm, cl), _ctx(cm, cl)
    try:
        t = b.autotune(reps=2)
        say("autotune", t)
        assert t["ms_best"] <= t["ms_default"] * 1.0001
        # the trials rewrite only the orbit phase's outputs, which the next step recomputes
        _same(a, b, "after tuning (state untouched)", ("r", "m", "phi", "vr", "vt", "E", "J", "id", "sigma"))
        for k in range(2):
            a.step()
            b.step_graph()
        sa, sb = a.scalars(), b.scalars()
        assert sa["n_max"] == sb["n_max"]
        for k in ("Eescaped", "Jescaped", "TidalMassLoss", "Etot"):
            assert sa[k] == pytest.approx(sb[k], rel=1e-12, abs=1e-300), k
        da, db = a.download(), b.download()
        n = sa["n_max"]
        for k in ("r", "vr", "vt", "E", "J", "id", "phi", "m"):
            assert np.array_equal(da[k][1:n + 1], db[k][1:n + 1]), k
    finally:
        a.close()
        b.close()


def test_gpu_every_relax_shape_gives_identical_stars(gpu):
    """Every block size of the relaxation pass, from the same state, gives a
    bit-identical step (eager and graph); the relaxation autotuner restores the state it
    times on, and the step after it is identical too."""
    cm = gpu
    say("relaxation shapes, King W0=7 N=2e5")
    cl = cm.make_king(200000, 7.0, seed=8)
    ref = _ctx(cm, cl)
    try:
        ref.step()
        for th in (128, 512, 1024):
            c = _ctx(cm, cl)
            try:
                c.test_hook(cm.HOOK_RELAX_SHAPE, th)
                c.step()
                _same(ref, c, f"relax shape {th}, eager")
            finally:
                c.close()
        c = _ctx(cm, cl)
        try:
            t = c.autotune_relax(reps=2)
            say("relax autotune", t)
            assert t["threads"] in (128, 256, 512, 1024)
            assert t["ms_best"] <= t["ms_default"] * 1.0001
            c.step_graph()
            _same(ref, c, "after relax autotune, graph")
        finally:
            c.close()
        with pytest.raises(cm.CmcgError):
            ref.test_hook(cm.HOOK_RELAX_SHAPE, 96)
    finally:
        ref.close()


def test_gpu_every_orbit_variant_gives_identical_stars(gpu):
    """Every launch variant of the orbit sampler (threads per block, LDS samples, tries
    cap), run from the same state, writes bit-identical new radii,
    velocities and apsides."""
    cm = gpu
    say("all orbit variants, King W0=7 N=3e5")
    cl = cm.make_king(300000, 7.0, seed=6)
    with cm.Context(cl.n, cm.default_params(cl.n, orbit_r=cl.rtid, seed=2)) as ctx:
        ctx.upload(cl)
        ctx.init_cluster()
        ctx.energy_conservation1()
        state = ctx.download()
        sc = ctx.scalars()
        nvar = len(ctx.orbit_variants())
        ref = None
        for v in range(nvar):
            ctx.upload(state)
            ctx.set_scalars(**{k: sc[k] for k in ("n_max", "n_max_new", "Mtotal", "Rtidal", "Dt", "TotalTime")})
            ctx.test_hook(cm.HOOK_ORBIT_VARIANT, v)
            st = ctx.orbits(step=3)
            d = ctx.download()
            say("variant", v, "tries", st["tries_total"], "removed", st["n_removed_E"])
            if ref is None:
                ref = (d, st)
                continue
            n = sc["n_max"]
            for k in ("rnew", "vrnew", "vtnew", "r_peri", "r_apo", "ntry", "flags"):
                assert np.array_equal(d[k][1:n + 1], ref[0][k][1:n + 1]), (v, k)
        with pytest.raises(cm.CmcgError):
            ctx.test_hook(cm.HOOK_ORBIT_VARIANT, nvar)


def test_gpu_long_stripping_cascade_graph_matches_eager(gpu):
    """A tidal radius inside the cluster (Plummer 1e5, orbit_r = 1.5): the stripping
    loop of the first step runs about 8 passes (the orbit phase removes ~30% of the
    stars, and the mass loss keeps shrinking Rtidal).  The captured step does 3 grid
    passes and finishes in its one-block tail; the eager step loops on the host.
    The tail plays the grid's blocks one after another and writes the same partial
    sums, so every star and every scalar is bit-identical."""
    cm = gpu
    say("stripping cascade: graph tail vs eager loop")
    cl = cm.make_plummer(100000, seed=5)
    ctxs = []
    try:
        for _ in range(2):
            c = cm.Context(cl.n, cm.default_params(cl.n, orbit_r=1.5, seed=7))
            c.upload(cl)
            c.init_cluster()
            ctxs.append(c)
        eager, graph = ctxs
        eager.step()
        graph.step_graph()
        sa, sb = eager.scalars(), graph.scalars()
        say("n_max", sa["n_max"], sb["n_max"], "TidalMassLoss", sa["TidalMassLoss"], sb["TidalMassLoss"])
        assert sa["n_max"] == sb["n_max"] and sa["n_max"] < 0.7 * cl.n
        for k in ("TidalMassLoss", "Eescaped", "Jescaped", "Etidal", "Rtidal", "Mtotal", "Etot"):
            assert sa[k] == sb[k], k
        da, db = eager.download(), graph.download()
        n = sa["n_max"]
        for k in ("r", "m", "vr", "vt", "id", "phi", "E", "J"):
            assert np.array_equal(da[k][1:n + 1], db[k][1:n + 1]), k
    finally:
        for c in ctxs:
            c.close()
