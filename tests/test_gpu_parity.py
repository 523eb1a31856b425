"""Parity of the HIP hot path (through the C-ABI) with the CPU oracle on the same
seeded inputs.

Bars (tolerances written in each test):
  - RNG draws, sort permutation, time step, apsides indices: bit-exact;
  - potential: <= 1e-12 relative;
  - relaxation: bit-exact against the oracle with the device's correctly rounded
    sin/cos; <= 1e-10 relative on the speed scale against the C library's sin/cos
    that the reference calls;
  - orbit sampling from identical inputs: bit-exact against the oracle's FP64-Q
    mode; against the reference's long-double Q, r / v_r / v_t within 1e-10 except
    for reported branch flips;
  - whole step from the same state: bit-exact against the oracle with the device's
    sin/cos and FP64 Q; against the reference's libm and long-double Q, r / v_r / v_t
    within 1e-10 except reported flips;
  - energy drift over 100 steps within 1e-6 of the oracle's with the device's
    sin/cos; against the C library's sin/cos within the drift's own sensitivity to
    last-bit libm differences (tests/golden/drift_golden.json).
"""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def oracle_for(O, cl, **kw):
    kw.setdefault("orbit_r", cl.rtid)
    return O.Oracle(cl.a, cl.n, **kw)


def ctx_for(cm, cl, cap=None, seed=1, **kw):
    p = cm.default_params(cl.n, orbit_r=cl.rtid, seed=seed, **kw)
    ctx = cm.Context(cap or cl.n, p)
    return ctx


def upload_state(ctx, cl, o, extra=()):
    """Upload the cluster with the oracle's current phi/E/J and scalars."""
    h = cl.copy()
    for k in ("phi", "E", "J", "sigma") + tuple(extra):
        h[k] = o[k]
    ctx.upload(h)
    ctx.set_scalars(n_max=o.N, n_max_new=o.N, Mtotal=o.sc("Mtotal"), Rtidal=o.sc("Rtidal"),
                    TidalMassLoss=0.0, TidalMassLoss_old=0.0, OldTidalMassLoss=0.0, Eescaped=0.0,
                    Jescaped=0.0, Etidal=0.0, Dt=0.0, Prev_Dt=0.0, TotalTime=0.0, step=0)


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    d = np.abs(a - b)
    s = np.maximum(np.abs(a), np.abs(b))
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(d == 0, 0.0, d / s)


# ------------------------------------------------------------------ RNG
def test_gpu_rng_streams_bit_exact(gpu, O):
    cm = gpu
    say("rng streams seed 1234")
    cl = cm.make_plummer(1000, seed=1)
    with ctx_for(cm, cl, seed=1234) as ctx:
        seeded = O.seed_state(1234)
        for phase, step, first in [(0, 0, 0), (0, 3, 100000), (1, 0, 1), (1, 7, 3999000)]:
            got = ctx.rng_draws(phase, step, first, 64, 5)
            for i in range(0, 64, 9):
                fn = O.lib.orc_stream_relax if phase == 0 else O.lib.orc_stream_orbit
                st = fn(C.byref(seeded), step, first + i)
                want = [O.lib.orc_t113_dbl(C.byref(st)) for _ in range(5)]
                assert list(got[i]) == want, (phase, step, first + i)


# ------------------------------------------------------------------ potential
@pytest.mark.parametrize("case", ["plummer_1e5", "plummer_kroupa_1e6", "king_4e6"])
def test_gpu_potential_parity(gpu, O, case):
    cm = gpu
    say("potential", case)
    if case == "plummer_1e5":
        cl = cm.make_plummer(100000, seed=1)
    elif case == "plummer_kroupa_1e6":
        cl = cm.make_plummer(1000000, seed=1, kroupa=True)
    else:
        cl = cm.make_king(4000000, 7.0, seed=1)
    o = oracle_for(O, cl)
    n = o.potential_calculate()
    with ctx_for(cm, cl) as ctx:
        ctx.upload(cl)
        ng, mt, rt = ctx.potential()
        d = ctx.download()
    assert ng == n
    assert mt == pytest.approx(o.sc("Mtotal"), rel=1e-15)
    assert rt == pytest.approx(o.sc("Rtidal"), rel=1e-14)
    err = rel(d["phi"][: n + 2], o["phi"][: n + 2])
    neq = int(np.sum(d["phi"][: n + 2] != o["phi"][: n + 2]))
    say(case, "max rel phi err", err.max(), "unequal", neq)
    # the device reproduces the reference's sequential rounding: bit-exact
    # (the stated bar is 1e-12 relative)
    assert neq == 0
    assert mt == o.sc("Mtotal")
    assert d["r"][n + 1] == 1e10 and d["phi"][n + 1] == 0.0


@pytest.mark.parametrize("drop", [30, 60])
def test_gpu_potential_fallback_paths(gpu, O, drop):
    """A test hook widens the chains' classification bound by 2^drop, so that most
    elements are classified sensitive: the appended list overflows (ordered
    compaction), the compacted list (65536 entries) overflows too, and the
    sequential repair finishes each chain from the first element after the list.
    The potential must still be the reference's recurrence bit for bit."""
    cm = gpu
    say("potential fallback paths, scale drop", drop)
    cl = cm.make_plummer(100000, seed=3, kroupa=True)
    o = oracle_for(O, cl)
    n = o.potential_calculate()
    with ctx_for(cm, cl) as ctx:
        ctx.upload(cl)
        ctx.test_hook(cm.HOOK_CHAIN_K_DROP, drop)
        ng, mt, rt = ctx.potential()
        d = ctx.download()
        ctx.test_hook(cm.HOOK_CHAIN_K_DROP, 0)
        ng2, mt2, _ = ctx.potential()
        d2 = ctx.download()
    assert ng == n == ng2
    assert mt == o.sc("Mtotal") == mt2
    neq = int(np.sum(d["phi"][: n + 2] != o["phi"][: n + 2]))
    say("drop", drop, "unequal", neq)
    assert neq == 0
    assert np.array_equal(d2["phi"][: n + 2], d["phi"][: n + 2])


def test_gpu_potential_tail_and_count(gpu, O):
    cm = gpu
    say("potential with stripped tail")
    cl = cm.make_plummer(50001, seed=2)
    cl["r"][49000:50002] = 1e10
    o = oracle_for(O, cl)
    n = o.potential_calculate()
    with ctx_for(cm, cl) as ctx:
        ctx.upload(cl)
        ng, mt, rt = ctx.potential()
        d = ctx.download(n_hint=cl.n)
    assert ng == n == 48999
    assert np.array_equal(d["phi"][: n + 2], o["phi"][: n + 2])


# ------------------------------------------------------------------ sort
@pytest.mark.parametrize("n,ties", [(100003, "short"), (300007, "long"), (600011, "huge"), (4000000, None)])
def test_gpu_sort_bit_exact(gpu, O, n, ties):
    """Bit-exact (r, id) order against the oracle.  The device sorts on a 31-bit
    range-normalised key and orders runs of equal keys by (r, id): "short" has many
    runs of equal radii; "long" has runs of 3000 and 500 equal radii and a run of 1200
    radii that agree in the top 40 bits but differ below; "huge" has runs of 20000 and
    70000 equal radii (beyond one block's LDS sort: merge-sorted) and 6000 runs of 60
    to 100 equal radii."""
    cm = gpu
    say("sort n", n, "ties", ties)
    rng = np.random.default_rng(n)
    cl = cm.Cluster(n)
    if ties == "short":
        r = rng.choice(rng.random(n // 3) * 5 + 1e-3, n)
    elif ties == "long":
        r = np.exp(rng.normal(size=n))
        r[rng.choice(n, 3000, replace=False)] = 0.75
        r[rng.choice(n, 500, replace=False)] = 2.5
        near = rng.choice(n, 1200, replace=False)
        r[near] = 1.3 * (1.0 + rng.integers(0, 4000, 1200) * 2.0 ** -52)
    elif ties == "huge":
        r = np.exp(rng.normal(size=n))
        perm = rng.permutation(n)
        r[perm[:20000]] = 0.9
        r[perm[20000:90000]] = 1.7
        pos = 90000
        for q in range(6000):
            ln = int(rng.integers(60, 101))
            r[perm[pos:pos + ln]] = 2.0 + q * 1e-4
            pos += ln
    else:
        r = np.exp(rng.normal(size=n) * 2)
    r[rng.choice(n, n // 200, replace=False)] = 1e10
    ids = rng.permutation(n).astype(np.int64) + 1
    cl["r"][1:n + 1] = r
    cl["r"][0], cl["r"][n + 1] = 1e-20, 1e10
    cl["id"][1:n + 1] = ids
    for k in ("m", "vr", "vt", "E", "J", "rOld", "Uoldrold", "Uoldrnew", "vrold", "vtold", "r_apo", "X"):
        cl[k][1:n + 1] = rng.random(n)
    cl.rtid = 100.0
    o = oracle_for(O, cl)
    nn = o.sort()
    with ctx_for(cm, cl) as ctx:
        ctx.upload(cl)
        ng = ctx.sort_by_r()
        d = ctx.download()
    assert ng == nn
    for k in ("r", "id", "m", "vr", "vt", "E", "J", "rOld", "Uoldrold", "Uoldrnew", "vrold", "vtold", "r_apo", "X"):
        assert np.array_equal(d[k][1:nn + 1], o[k][1:nn + 1]), k
    assert d["r"][nn + 1] == 1e10


# ------------------------------------------------------------------ relaxation
@pytest.mark.parametrize("case", ["plummer_1e5", "plummer_kroupa_1e6"])
def test_gpu_relaxation_parity(gpu, O, case):
    cm = gpu
    say("relaxation", case)
    cl = cm.make_plummer(100000, seed=1) if case == "plummer_1e5" else cm.make_plummer(1000000, seed=2, kroupa=True)
    o = oracle_for(O, cl, rng_mode=O.RNG_PER_STAR, libm_mode=O.LIBM_DD)
    o.init()
    osys = oracle_for(O, cl, rng_mode=O.RNG_PER_STAR, libm_mode=O.LIBM_SYSTEM)
    osys.init()
    dt = O.lib.orc_simul_relax_new(C.byref(o.s))
    with ctx_for(cm, cl) as ctx:
        upload_state(ctx, cl, o)
        dt_gpu = ctx.simul_relax()
        assert dt_gpu == dt                      # bit-exact window sums and minimum
        ctx.relax(dt, step=0)
        d = ctx.download()
    o.relax(dt)
    osys.relax(dt)
    N = o.N
    assert np.array_equal(d["Y"][1:N], o["Y"][1:N])
    # same sin/cos: bit-exact
    for k in ("vr", "vt", "E", "J"):
        same = int(np.sum(d[k][1:N + 1] == o[k][1:N + 1]))
        say(case, k, "bit-equal to the oracle with the device sin/cos:", same, "of", N)
        assert same == N, k
    # the C library's sin/cos (the reference's): last-bit differences only
    scale = np.sqrt(osys["vr"][1:N + 1] ** 2 + osys["vt"][1:N + 1] ** 2) + 1e-300
    for k in ("vr", "vt"):
        err = np.abs(d[k][1:N + 1] - osys[k][1:N + 1]) / scale
        say(case, k, "vs system libm: max err / |v|", err.max(), "stars differing", int(np.sum(err > 0)))
        assert err.max() <= 1e-10
    for k in ("E", "J"):
        err = rel(d[k][1:N + 1], osys[k][1:N + 1])
        assert err.max() <= 1e-10, k
    # sigma (direct window sums on the device, sliding sums in the reference)
    err = rel(d["sigma"][1:N + 1], o["sigma"][1:N + 1])
    say(case, "sigma max rel", err.max())
    assert err.max() <= 1e-10


# ------------------------------------------------------------------ orbits
def _orbit_case(cm, O, cl, qmode):
    o = oracle_for(O, cl, rng_mode=O.RNG_PER_STAR, q_mode=qmode)
    o.init()
    with ctx_for(cm, cl) as ctx:
        upload_state(ctx, cl, o)
        st = ctx.orbits(step=0)
        d = ctx.download()
    o.s.step = 0
    o.orbits()
    return o, d, st


@pytest.mark.parametrize("case", ["plummer_1e5", "plummer_kroupa_1e6", "king_1e6"])
def test_gpu_orbits_bit_exact_vs_fp64_oracle(gpu, O, case):
    cm = gpu
    say("orbits fp64-Q parity", case)
    if case == "plummer_1e5":
        cl = cm.make_plummer(100000, seed=1)
    elif case == "plummer_kroupa_1e6":
        cl = cm.make_plummer(1000000, seed=3, kroupa=True)
    else:
        cl = cm.make_king(1000000, 7.0, seed=1)
    o, d, st = _orbit_case(cm, O, cl, O.Q_FP64)
    N = o.N
    say(case, "stats", st)
    assert st["error"] == 0 and o.s.error == 0
    for k in ("rnew", "vrnew", "vtnew", "r_peri", "r_apo", "X"):
        same = np.sum(d[k][1:N + 1] == o[k][1:N + 1])
        say(case, k, "bit-equal", same, "of", N)
        assert same == N, k
    assert np.array_equal(d["ntry"][1:N + 1], o["ntry"][1:N + 1])
    assert st["n_circular"] == o.sc("n_circular")
    assert st["max_r"] == o.sc("max_r")


@pytest.mark.parametrize("case", ["plummer_1e5", "king_1e6"])
def test_gpu_orbits_vs_reference_long_double(gpu, O, case):
    """Against the reference's x87 long-double Q: r, v_r, v_t within 1e-10 except
    for stars whose apsides bracket or acceptance branch flips (reported)."""
    cm = gpu
    say("orbits vs long-double oracle", case)
    cl = cm.make_plummer(100000, seed=1) if case == "plummer_1e5" else cm.make_king(1000000, 7.0, seed=2)
    o, d, st = _orbit_case(cm, O, cl, O.Q_LONG_DOUBLE)
    N = o.N
    bad = np.zeros(N, bool)
    for k in ("rnew", "vrnew", "vtnew"):
        bad |= rel(d[k][1:N + 1], o[k][1:N + 1]) > 1e-10
    flips = int(bad.sum())
    say(case, "flipped stars", flips, "fraction", flips / N)
    assert flips / N < 1e-3


def test_gpu_apsides_and_find_zero_Q(gpu, O):
    cm = gpu
    say("apsides / find_zero_Q batch")
    cl = cm.make_king(50000, 7.0, seed=5)
    o = oracle_for(O, cl, q_mode=O.Q_FP64)
    o.init()
    N = o.N
    js = np.arange(1, N + 1, 37)
    E = o["E"][js] + o["m"][js] / o["r"][js] / o.s.N_STAR
    J = o["J"][js]
    with ctx_for(cm, cl) as ctx:
        upload_state(ctx, cl, o)
        res = {k: [] for k in ("rp", "ra", "kmin", "kmax", "circular")}
        for j, e, jj in zip(js, E, J):
            r1 = ctx.apsides(int(j), [e], [jj])
            for k in res:
                res[k].append(r1[k][0])
        kq = ctx.find_zero_Q(js, np.zeros_like(js), js, E, J)
    for i, j in enumerate(js):
        w = o.apsides(int(j), E[i], J[i])
        for k in res:
            assert res[k][i] == w[k], (j, k)
        assert kq[i] == O.lib.orc_find_zero_Q(C.byref(o.s), int(j), 0, int(j), E[i], J[i])


def test_gpu_find_root_vr_batch(gpu, O):
    """find_root_vr (cmc_nr.c:500-603) on the apsis intervals of ordinary stars and
    of near-tangent stars: the device's Brent iteration returns the oracle's root bit
    for bit (the oracle's root is pinned against the exact root of the interval's
    quadratic in tests/test_oracle_hotpath.py)."""
    from cases import near_tangent_stars
    cm = gpu
    say("find_root_vr batch")
    cl = cm.make_plummer(20000, seed=21, kroupa=True)
    o = oracle_for(O, cl, q_mode=O.Q_FP64)
    o.init()
    near_tangent_stars(o, list(range(5, o.N - 1, 53)))
    js, ks, Es, Js = [], [], [], []
    for j in range(2, o.N, 29):
        E = o["E"][j] + o["m"][j] / o["r"][j] / o.s.N_STAR
        J = o["J"][j]
        res = o.apsides(j, E, J)
        if res["circular"]:
            continue
        for k in (res["kmin"], res["kmax"]):
            js.append(j); ks.append(k); Es.append(E); Js.append(J)
    want = np.array([O.lib.orc_find_root_vr(C.byref(o.s), j, k, E, J) for j, k, E, J in zip(js, ks, Es, Js)])
    assert o.s.error == 0
    with ctx_for(cm, cl) as ctx:
        upload_state(ctx, cl, o, extra=("vr", "vt"))
        got = ctx.find_root_vr(js, ks, Es, Js)
    say("intervals", len(js), "bit-equal", int(np.sum(got == want)))
    assert len(js) > 500
    assert np.array_equal(got, want)


def test_gpu_brent_near_tangent_orbits(gpu, O):
    """The orbit phase with near-tangent stars (cases.near_tangent_stars), where the
    reference's calc_orbit_new falls back to find_root_vr (cmc_orbit.c:459-474,
    485-500): the device takes the same path for the same stars (n_brent equal and
    > 0) and every star's sampled orbit is bit-identical to the oracle's, including
    where its narrowed searches meet Q within rounding of zero.  On some of these
    degenerate orbits the reference's rejection loop runs out of tries and stops
    (N_TRY, cmc_evolution_thr.c:890-893); the device reports the same error there."""
    from cases import near_tangent_sampleable, near_tangent_stars
    cm = gpu
    say("orbits with near-tangent stars (Brent path)")
    cl = cm.make_plummer(100000, seed=22)
    js0 = list(range(3, cl.n - 1, 41))
    o, js, dropped = near_tangent_sampleable(O, cl, js0)
    say("near-tangent stars", len(js), "left out (the reference stops on them)", dropped)
    assert dropped
    with ctx_for(cm, cl) as ctx:
        upload_state(ctx, cl, o, extra=("vr", "vt"))
        st = ctx.orbits(step=0)
        d = ctx.download()
    o.s.step = 0
    o.orbits()
    N = o.N
    brent_stars = int(np.sum((o["flags"][1:N + 1] & 16) != 0))      # stars with a find_root_vr apsis
    say("stars through find_root_vr: gpu", st["n_brent"], "oracle", brent_stars, "(oracle calls", o.s.n_brent,
        ") circular", st["n_circular"], o.sc("n_circular"))
    assert st["error"] == 0 and o.s.error == 0
    assert brent_stars > 50
    assert st["n_brent"] == brent_stars
    assert st["n_circular"] == o.sc("n_circular")
    for k in ("rnew", "vrnew", "vtnew", "r_peri", "r_apo", "X"):
        same = int(np.sum(d[k][1:N + 1] == o[k][1:N + 1]))
        assert same == N, (k, same, N)
    assert np.array_equal(d["flags"][1:N + 1] & 16, o["flags"][1:N + 1] & 16)
    # with the first of the left-out stars back in, both stop with N_TRY exceeded
    o2 = oracle_for(O, cl, rng_mode=O.RNG_PER_STAR, q_mode=O.Q_FP64)
    o2.init()
    near_tangent_stars(o2, [j for j in js0 if j not in dropped[1:]])
    with ctx_for(cm, cl) as ctx:
        upload_state(ctx, cl, o2, extra=("vr", "vt"))
        with pytest.raises(cm.CmcgError) as ei:
            ctx.orbits(step=0)
    o2.s.step = 0
    o2.orbits()
    assert ei.value.code == cm.E_NTRY == o2.s.error

# ------------------------------------------------------------------ glue
def test_gpu_glue_parity(gpu, O):
    cm = gpu
    say("glue: energy conservation, strip, intermediate energy, set_velocities3, energy")
    cl = cm.make_king(200000, 7.0, seed=4)
    o = oracle_for(O, cl, rng_mode=O.RNG_PER_STAR, q_mode=O.Q_FP64)
    o.init()
    with ctx_for(cm, cl) as ctx:
        upload_state(ctx, cl, o)
        ctx.energy_conservation1()
        O.lib.orc_energy_conservation1(C.byref(o.s))
        ctx.orbits(step=0)
        o.s.step = 0
        o.s.OldTidalMassLoss = 0.0
        o.orbits()
        ctx.energy_conservation2()
        O.lib.orc_energy_conservation2(C.byref(o.s))
        d = ctx.download()
        N = o.N
        for k in ("Uoldrold", "Uoldrnew", "vrold", "vtold", "rnew"):
            assert np.array_equal(d[k][1:N + 1], o[k][1:N + 1]), k
        ctx.tidally_strip_stars()
        O.lib.orc_tidally_strip_stars(C.byref(o.s))
        ctx.compute_intermediate_energy()
        O.lib.orc_compute_intermediate_energy(C.byref(o.s))
        d = ctx.download()
        for k in ("r", "rOld", "vr", "vt", "m"):
            assert np.array_equal(d[k][1:N + 1], o[k][1:N + 1]), k
        live = o["rnew"][1:N + 1] < 1e6
        assert rel(d["EI"][1:N + 1][live], o["EI"][1:N + 1][live]).max() <= 1e-12
        ng = ctx.sort_by_r()
        nn = o.sort()
        assert ng == nn
        ctx.potential()
        o.potential_calculate()
        # continue from the oracle's potential so set_velocities3 sees identical inputs
        d = ctx.download()
        assert rel(d["phi"][1:nn + 1], o["phi"][1:nn + 1]).max() <= 1e-12
        h = cm.Cluster(nn)
        for k in h.a:
            if k in o.arr:
                h[k] = o[k][: nn + 2]
        ctx.upload(h)
        ctx.set_velocities3()
        O.lib.orc_set_velocities3(C.byref(o.s))
        ctx.compute_energy()
        O.lib.orc_compute_energy(C.byref(o.s))
        d = ctx.download()
        sc = ctx.scalars()
    for k in ("vr", "vt", "E", "J"):
        assert np.array_equal(d[k][1:nn + 1], o[k][1:nn + 1]), k
    assert sc["Ekin"] == pytest.approx(o.sc("Ekin"), rel=1e-12)
    assert sc["Epot"] == pytest.approx(o.sc("Epot"), rel=1e-12)


@pytest.mark.parametrize("case", ["sparse", "dense", "huge_excess"])
def test_gpu_set_velocities3_excess_chain(gpu, O, case):
    """Excess-energy chain of set_velocities3 (cmc_utils.c:268-310) with planted
    events (stars whose new kinetic energy would be negative): isolated, merging
    and very long episodes must give bit-identical velocities."""
    cm = gpu
    say("set_velocities3 excess chain", case)
    n = 300000
    cl = cm.make_plummer(n, seed=11)
    o = oracle_for(O, cl)
    o.potential_calculate()
    rng = np.random.default_rng(5)
    N = o.N
    p_ev, scale = {"sparse": (0.002, 2.0), "dense": (0.3, 1.5), "huge_excess": (2e-5, 2e5)}[case]
    v2 = o["vr"][1:N + 1] ** 2 + o["vt"][1:N + 1] ** 2
    ev = rng.random(N) < p_ev
    delta = np.where(ev, v2 * (1.0 + scale * rng.random(N)), v2 * 0.2 * (rng.random(N) - 0.5))
    o["rOld"][1:N + 1] = o["r"][1:N + 1]
    o["vrold"][1:N + 1] = o["vr"][1:N + 1]
    o["vtold"][1:N + 1] = o["vt"][1:N + 1]
    o["Uoldrold"][1:N + 1] = 0.0
    o["Uoldrnew"][1:N + 1] = delta
    h = cm.Cluster(N)
    for k in h.a:
        if k in o.arr:
            h[k] = o[k][: N + 2]
    with ctx_for(cm, cl) as ctx:
        ctx.upload(h)
        ctx.set_scalars(n_max=N, n_max_new=N, Mtotal=o.sc("Mtotal"))
        ctx.set_velocities3()
        d = ctx.download()
        assert ctx.scalars()["error"] == 0
    O.lib.orc_set_velocities3(C.byref(o.s))
    say(case, "events", o.s.n_excess_events, "dumps", o.s.n_excess_dumps)
    assert o.s.n_excess_events > 0
    for k in ("vr", "vt", "vrnew", "vtnew"):
        assert np.array_equal(d[k][1:N + 1], o[k][1:N + 1]), k


# ------------------------------------------------------------------ whole step
def _id_order(d, o, N):
    gi = {int(i): k for k, i in enumerate(d["id"][1:N + 1])}
    return np.array([gi[int(i)] for i in o["id"][1:N + 1]]) + 1


@pytest.mark.parametrize("case", ["plummer_1e5", "plummer_kroupa_1e6"])
def test_gpu_step_vs_oracle(gpu, O, case):
    """One Henon step from the same state (Plummer 1e5 equal masses, seed 1; Plummer
    1e6 Kroupa).  With the device's sin/cos and FP64 Q in the oracle every star is
    bit-identical; against the reference's C-library sin/cos and long-double Q, a
    few acceptance branches flip and the rest agree within 1e-10."""
    cm = gpu
    say("whole step", case)
    cl = cm.make_plummer(100000, seed=1) if case == "plummer_1e5" else cm.make_plummer(1000000, seed=1, kroupa=True)
    o = oracle_for(O, cl, rng_mode=O.RNG_PER_STAR, q_mode=O.Q_FP64, libm_mode=O.LIBM_DD)
    o.init()
    oref = oracle_for(O, cl, rng_mode=O.RNG_PER_STAR, q_mode=O.Q_LONG_DOUBLE, libm_mode=O.LIBM_SYSTEM)
    oref.init()
    with ctx_for(cm, cl) as ctx:
        ctx.upload(cl)
        ctx.init_cluster()
        st = ctx.step()
        d = ctx.download()
        sc = ctx.scalars()
    assert o.step() == 0
    assert oref.step() == 0
    say(case, "gpu stats", st)
    assert sc["n_max"] == o.N
    assert sc["Dt"] == o.sc("Dt")
    N = o.N
    idx = _id_order(d, o, N)
    for k in ("r", "vr", "vt", "E", "J", "m", "phi"):
        same = int(np.sum(d[k][idx] == o[k][1:N + 1]))
        say(case, k, "bit-equal to the oracle with the device's libm and FP64 Q:", same, "of", N)
        assert same == N, k
    assert sc["Etot"] == pytest.approx(o.sc("Etot"), rel=1e-12)
    # the reference's libm and long-double Q
    Nr = oref.N
    gi = {int(i): k for k, i in enumerate(d["id"][1:sc["n_max"] + 1])}
    bad = 0
    for j in range(1, Nr + 1):
        k = gi.get(int(oref["id"][j]))
        if k is None:
            bad += 1
            continue
        for f in ("r", "vr", "vt"):
            if rel(d[f][k + 1], oref[f][j]) > 1e-10:
                bad += 1
                break
    say(case, "vs reference libm / long-double Q: stars outside 1e-10:", bad, "fraction", bad / Nr)
    assert bad / Nr < 2e-3
    assert sc["Etot"] == pytest.approx(oref.sc("Etot"), rel=1e-8)


def test_gpu_energy_drift_100_steps(gpu, O):
    """Energy drift over 100 steps (Plummer 2e4, seed 1).  The oracle with the
    device's sin/cos follows the same trajectory, so the drifts agree to rounding of
    the energy sums (bar 1e-6).  The reference's own drift moves by ~1e-5 when only
    the last bit of sin/cos changes: tests/golden/drift_golden.json holds it for the
    C library's sin()/cos() (unoptimised build), sincos() (optimised build) and long
    double; against the C-library run the bar is the spread of those variants."""
    import json
    import os
    cm = gpu
    say("energy drift, 100 steps, Plummer 2e4")
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "drift_golden.json")))
    cl = cm.make_plummer(gold["n"], seed=gold["seed"])
    o = oracle_for(O, cl, rng_mode=O.RNG_PER_STAR, q_mode=O.Q_FP64, libm_mode=O.LIBM_DD)
    o.init()
    e0 = o.sc("Etot")
    with ctx_for(cm, cl) as ctx:
        ctx.upload(cl)
        ctx.init_cluster()
        g0 = ctx.scalars()["Etot"]
        for _ in range(gold["steps"]):
            ctx.step()
        sc = ctx.scalars()
    for _ in range(gold["steps"]):
        assert o.step() == 0
    drift_gpu = (sc["Etot"] + sc["Eoops"] - g0) / abs(g0)
    drift_ref = (o.sc("Etot") + o.sc("Eoops") - e0) / abs(e0)
    say("drift gpu", drift_gpu, "oracle (device libm)", drift_ref, "N", sc["n_max"], o.N)
    assert sc["n_max"] == o.N
    assert abs(drift_gpu - drift_ref) < 1e-6
    sys_d, floor = gold["system"]["drift"], gold["libm_spread"]
    say("drift reference (C library sin/cos)", sys_d, "spread over the reference's libm variants", floor,
        "gpu - reference", drift_gpu - sys_d)
    assert abs(drift_gpu - sys_d) <= max(1e-6, floor)


# ------------------------------------------------------------------ edge cases
def test_gpu_edge_cases(gpu, O):
    cm = gpu
    say("edge cases: odd N, unbound, destroyed, small N, capacity")
    cl = cm.make_plummer(2001, seed=9)
    o0 = oracle_for(O, cl)
    o0.init()
    # make some stars unbound and one massless
    cl2 = cl.copy()
    cl2["vr"][10:15] *= 20.0
    cl2["m"][30] = 1e-30
    cl2["vr"][[40, 700, 1500]] = 0.0          # own-index Q <= 0: the linear-scan path
    o = oracle_for(O, cl2, rng_mode=O.RNG_PER_STAR, q_mode=O.Q_FP64)
    o.init()
    with ctx_for(cm, cl2) as ctx:
        upload_state(ctx, cl2, o)
        st = ctx.orbits(step=0)
        d = ctx.download()
    o.orbits()
    N = o.N
    assert st["n_removed_E"] == o.sc("n_removed_E") >= 5
    assert st["n_destroyed"] == 1
    say("linear-scan stars gpu", st["n_linscan"], "oracle", o.sc("n_linscan"))
    assert st["n_linscan"] == o.sc("n_linscan")
    for k in ("rnew", "vrnew", "vtnew", "r_apo", "r_peri"):
        assert np.array_equal(d[k][1:N + 1], o[k][1:N + 1]), k
    # capacity and argument errors are reported, not raised inside the library
    with pytest.raises(cm.CmcgError) as ei:
        with ctx_for(cm, cl, cap=100) as ctx:
            ctx.upload(cl)
    assert ei.value.code == cm.E_CAPACITY


def test_gpu_many_linear_scan_stars(gpu, O):
    """8000 stars with v_r = 0 (Q = 0 at their own radius: about 60% of them take the
    reference's linear-scan case, cmc_orbit.c:40-56), more than the one-thread sort of
    the deferred list takes:
    the list is put in index order by a bitonic sort, so the phase's sums do not depend
    on the order of appending.  Bit-exact against the oracle, twice in a row."""
    cm = gpu
    say("8000 stars with v_r = 0: linear-scan stars")
    cl = cm.make_plummer(20000, seed=11)
    rng = np.random.default_rng(11)
    cl["vr"][1 + rng.choice(20000, 8000, replace=False)] = 0.0
    o = oracle_for(O, cl, rng_mode=O.RNG_PER_STAR, q_mode=O.Q_FP64)
    o.init()
    runs = []
    for _ in range(2):
        with ctx_for(cm, cl) as ctx:
            upload_state(ctx, cl, o)
            st = ctx.orbits(step=0)
            runs.append((st, ctx.download()))
    o.orbits()
    N = o.N
    say("linear-scan stars gpu", runs[0][0]["n_linscan"], "oracle", o.sc("n_linscan"))
    assert runs[0][0]["n_linscan"] == o.sc("n_linscan") > 4096
    for st, d in runs:
        for k in ("rnew", "vrnew", "vtnew", "r_apo", "r_peri", "X"):
            assert np.array_equal(d[k][1:N + 1], o[k][1:N + 1]), k
    assert runs[0][0] == runs[1][0]


def test_gpu_linear_scan_stars_beyond_65536(gpu, O):
    """120000 stars all with v_r = 0: more linear-scan stars in one step than 65536
    (the deferred list holds as many entries as the context holds stars, so any
    count is finished).  Bit-exact against the oracle, and the phase's sums too."""
    cm = gpu
    say("120000 stars with v_r = 0: linear-scan stars beyond 65536")
    cl = cm.make_plummer(120000, seed=12)
    cl["vr"][1:120001] = 0.0
    o = oracle_for(O, cl, rng_mode=O.RNG_PER_STAR, q_mode=O.Q_FP64)
    o.init()
    with ctx_for(cm, cl) as ctx:
        upload_state(ctx, cl, o)
        st = ctx.orbits(step=0)
        d = ctx.download()
    o.orbits()
    N = o.N
    say("linear-scan stars gpu", st["n_linscan"], "oracle", o.sc("n_linscan"))
    assert st["n_linscan"] == o.sc("n_linscan") > 65536
    for k in ("rnew", "vrnew", "vtnew", "r_apo", "r_peri", "X"):
        assert np.array_equal(d[k][1:N + 1], o[k][1:N + 1]), k


def test_gpu_host_shim_reference_names(gpu, O):
    """The reference-named C entry points drive the same kernels."""
    cm = gpu
    say("host shim")
    cl = cm.make_plummer(30000, seed=4)
    h = cm.hostlib
    p = cm.default_params(cl.n, orbit_r=cl.rtid, seed=1)
    assert h.cmc_hotpath_init(C.byref(p), cl.n, C.byref(cl.soa()), 0) == 0
    h.calc_potential_new.restype = None
    h.calc_potential_new()
    h.potential.restype = C.c_double
    h.potential.argtypes = [C.c_double]
    o = oracle_for(O, cl)
    o.potential_calculate()
    for x in [0.0, 1e-3, 0.5, 1.0, 3.0]:
        assert h.potential(x) == pytest.approx(o.potential(x), rel=1e-12)
    h.cmc_hotpath_finalize()


_GLOBALS_SCRIPT = r"""
import ctypes as C, json, os, sys
repo = sys.argv[1]
g = C.CDLL(os.path.join(repo, "tests", "libref_globals.so"), mode=C.RTLD_GLOBAL)
sys.path.insert(0, repo)
import importlib
cm = importlib.import_module("cmc-cosmic_amd")
h = cm.hostlib
cl = cm.make_king(40000, 7.0, seed=5)
p = cm.default_params(cl.n, orbit_r=cl.rtid, seed=3)
assert h.cmc_hotpath_init(C.byref(p), cl.n, C.byref(cl.soa()), 0) == 0
h.cmc_hotpath_ctx.restype = C.c_void_p
shim = cm.Context.__new__(cm.Context)
shim.h = C.c_void_p(h.cmc_hotpath_ctx())
shim.init_cluster()
h.calc_potential_new.restype = None
h.calc_potential_new()
dbl = lambda name: C.c_double.in_dll(g, name).value
out = {"after_init": {"Mtotal": dbl("Mtotal"), "N_MAX": C.c_long.in_dll(g, "clus").value,
                      "dev_Mtotal": shim.scalars()["Mtotal"], "dev_n_max": shim.scalars()["n_max"]}}
h.dynamics_apply.argtypes = [C.c_double, C.c_void_p]
h.calc_timestep.argtypes = [C.c_void_p]
h.calc_timestep(None)
out["Dt_after_calc_timestep"] = [dbl("Dt"), shim.scalars()["Dt"]]
h.dynamics_apply(dbl("Dt"), None)
for name in ("energy_conservation1", "new_orbits_calculate", "energy_conservation2", "tidally_strip_stars",
             "ComputeIntermediateEnergy", "qsorts_new", "post_sort_comm", "calc_potential_new",
             "energy_conservation3", "compute_energy_new"):
    getattr(h, name).restype = None
    getattr(h, name)()
sc = shim.scalars()
class Clus(C.Structure):
    _fields_ = [("N_MAX", C.c_long), ("N_MAX_NEW", C.c_long), ("N_STAR", C.c_long), ("N_STAR_NEW", C.c_long),
                ("N_BINARY", C.c_long)]
class Etot(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("tot", "New", "ini", "K", "P", "Eint", "Eb")]
cs, et = Clus.in_dll(g, "clus"), Etot.in_dll(g, "Etotal")
out["ref"] = {"N_MAX": cs.N_MAX, "N_MAX_NEW": cs.N_MAX_NEW, "Etot": et.tot, "Ekin": et.K, "Epot": et.P,
              **{k: dbl(k) for k in ("Mtotal", "Rtidal", "max_r", "Dt", "TotalTime", "TidalMassLoss", "Eescaped",
                                     "Jescaped", "Etidal", "Eoops")}}
out["dev"] = {"N_MAX": sc["n_max"], "N_MAX_NEW": sc["n_max_new"], "Etot": sc["Etot"], "Ekin": sc["Ekin"],
              "Epot": sc["Epot"], **{k: sc[k] for k in ("Mtotal", "Rtidal", "max_r", "Dt", "TotalTime",
                                                       "TidalMassLoss", "Eescaped", "Jescaped", "Etidal", "Eoops")}}
shim.h = C.c_void_p()
h.cmc_hotpath_finalize()
print("GLOBALS " + json.dumps(out), flush=True)
"""


def test_gpu_shim_writes_reference_globals(gpu):
    """Linked beside the reference's own globals (a fixture library with the names and
    layouts of include/cmc/cmc_vars.h, loaded first as the reference executable's
    definitions would be), the shim's calls leave clus.N_MAX, Dt, Mtotal, Rtidal,
    Etotal ... equal to the device scalars after each call, so the reference's main
    loop reads them unchanged.  Runs in a fresh process (the fixture must be loaded
    before the shim)."""
    import json
    import os
    import subprocess
    import sys
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    say("shim writes the reference's globals (subprocess)")
    r = subprocess.run([sys.executable, "-c", _GLOBALS_SCRIPT, repo], capture_output=True, text=True, timeout=300)
    line = [l for l in r.stdout.splitlines() if l.startswith("GLOBALS ")]
    assert r.returncode == 0 and line, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    out = json.loads(line[0][len("GLOBALS "):])
    ai = out["after_init"]
    assert ai["N_MAX"] == ai["dev_n_max"] > 0 and ai["Mtotal"] == ai["dev_Mtotal"] > 0, ai
    dt = out["Dt_after_calc_timestep"]
    assert dt[0] == dt[1] > 0, dt
    assert out["ref"] == out["dev"], (out["ref"], out["dev"])


STEP_FIELDS = ("r", "m", "phi", "vr", "vt", "E", "J", "id", "r_peri", "r_apo")
STEP_SCALARS = ("n_max", "Mtotal", "Rtidal", "max_r", "Dt", "TotalTime", "TidalMassLoss", "Eescaped", "Jescaped",
                "Etidal", "Etot", "Ekin", "Epot")


def test_gpu_shim_step_in_reference_order_equals_cmcg_step(gpu):
    """One step driven through the reference-named entry points in the order of the
    reference's main loop (src/cmc/cmc.c:268-383: calc_timestep, dynamics_apply,
    energy_conservation1, new_orbits_calculate, energy_conservation2,
    tidally_strip_stars, ComputeIntermediateEnergy, qsorts_new, post_sort_comm,
    calc_potential_new, energy_conservation3, compute_energy_new) leaves the same
    cluster as cmcg_step, which fuses several of them: every field carried across
    steps and the scalars, bit for bit."""
    cm = gpu
    say("reference-ordered step through the host shim vs cmcg_step")
    cl = cm.make_king(60000, 7.0, seed=3)
    h = cm.hostlib
    errors = []
    handler_t = C.CFUNCTYPE(None, C.c_int, C.c_int64, C.c_char_p)
    handler = handler_t(lambda code, idx, where: errors.append((code, idx, where)))
    h.cmc_hotpath_set_error_handler(handler)
    p = cm.default_params(cl.n, orbit_r=cl.rtid, seed=7)
    assert h.cmc_hotpath_init(C.byref(p), cl.n, C.byref(cl.soa()), 0) == 0
    h.cmc_hotpath_ctx.restype = C.c_void_p
    shim = cm.Context.__new__(cm.Context)
    shim.h = C.c_void_p(h.cmc_hotpath_ctx())
    shim.init_cluster()
    h.dynamics_apply.argtypes = [C.c_double, C.c_void_p]
    h.calc_timestep.argtypes = [C.c_void_p]
    h.calc_timestep(None)
    h.dynamics_apply(shim.scalars()["Dt"], None)
    for name in ("energy_conservation1", "new_orbits_calculate", "energy_conservation2", "tidally_strip_stars",
                 "ComputeIntermediateEnergy", "qsorts_new", "post_sort_comm", "calc_potential_new",
                 "energy_conservation3", "compute_energy_new"):
        getattr(h, name).restype = None
        getattr(h, name)()
    assert not errors, errors
    a, sa = shim.download(), shim.scalars()
    shim.h = C.c_void_p()               # the shim owns that context
    h.cmc_hotpath_set_error_handler(None)
    h.cmc_hotpath_finalize()
    with cm.Context(cl.n, p) as ctx:
        ctx.upload(cl)
        ctx.init_cluster()
        ctx.step()
        b, sb = ctx.download(), ctx.scalars()
    n = sb["n_max"]
    say("stars after the step", n, "shim", sa["n_max"])
    for k in STEP_SCALARS:
        assert sa[k] == sb[k], (k, sa[k], sb[k])
    for k in STEP_FIELDS:
        assert np.array_equal(a[k][1:n + 1], b[k][1:n + 1]), k
