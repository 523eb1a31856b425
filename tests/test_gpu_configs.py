"""Whole Henon steps at the benchmark's own workloads, against the oracle.

  - King W0 = 7, N = 4e6, Kroupa IMF, heaviest 1% in the inner 10%, seed 1 (the
    benchmark's workload): one step from the same state, every star bit-identical to
    the oracle with the device's sin/cos and FP64 Q; the same step with learned
    brackets on gives the identical state; against the reference's C-library sin/cos
    and long-double Q the stars outside 1e-10 (branch flips) are counted and bounded.
  - Plummer N = 1e6 with the Kroupa IMF, seed 1: ten steps, bit-identical.
  - Energy drift over 100 steps at two more seeds (Plummer 2e4).
"""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("r", "vr", "vt", "E", "J", "m", "phi")


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    d = np.abs(a - b)
    s = np.maximum(np.abs(a), np.abs(b))
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(d == 0, 0.0, d / s)


def by_id(d, ids, n):
    """Slots of the device state d holding the stars with the given ids."""
    dev_ids = d["id"][1:n + 1]
    order = np.argsort(dev_ids, kind="stable")
    pos = np.searchsorted(dev_ids[order], ids)
    pos = np.minimum(pos, n - 1)
    found = dev_ids[order][pos] == ids
    return order[pos] + 1, found


def ctx_for(cm, cl, seed=1):
    return cm.Context(cl.n, cm.default_params(cl.n, orbit_r=cl.rtid, seed=seed))


def gpu_step(cm, cl, learned=None):
    with ctx_for(cm, cl) as ctx:
        ctx.upload(cl)
        ctx.init_cluster()
        if learned is not None:
            th, nr, margin = learned
            ctx.mlp_init(seed=1, max_rows=1024)
            ctx.mlp_set_params(th, nr)
            ctx.set_learned_brackets(True, margin=margin)
        st = ctx.step()
        d = ctx.download()
        sc = ctx.scalars()
    return d, sc, st


def assert_state_equal(d, sc, o, tag):
    N = o.N
    assert sc["n_max"] == N, (tag, sc["n_max"], N)
    slots, found = by_id(d, o["id"][1:N + 1], N)
    assert found.all(), tag
    for k in FIELDS:
        same = int(np.sum(d[k][slots] == o[k][1:N + 1]))
        say(tag, k, "bit-equal", same, "of", N)
        assert same == N, (tag, k)


@pytest.fixture(scope="module")
def king4e6(gpu, O):
    cm = gpu
    say("King W0=7 N=4e6 seed 1: generator + oracle step (device sin/cos, FP64 Q)")
    cl = cm.make_king(4000000, 7.0, seed=1)
    o = O.Oracle(cl.a, cl.n, orbit_r=cl.rtid, rng_mode=O.RNG_PER_STAR, q_mode=O.Q_FP64, libm_mode=O.LIBM_DD)
    o.init()
    assert o.step() == 0
    say("oracle step done: N", o.N, "Dt", o.sc("Dt"), "n_brent", o.s.n_brent)
    return cl, o


def test_gpu_king4e6_step_bit_exact(gpu, king4e6):
    cm = gpu
    cl, o = king4e6
    d, sc, st = gpu_step(cm, cl)
    say("gpu stats", st)
    assert st["error"] == 0
    assert sc["Dt"] == o.sc("Dt")
    assert st["n_circular"] == o.sc("n_circular")
    assert st["n_brent"] == int(np.sum((o["flags"] & 16) != 0))      # stars with a find_root_vr apsis
    assert_state_equal(d, sc, o, "king4e6")
    assert sc["Etot"] == pytest.approx(o.sc("Etot"), rel=1e-12)


def test_gpu_king4e6_learned_step_identical(gpu, king4e6):
    """Learned brackets on (a network trained on the same cluster's first step):
    the step's stars are the oracle's bit for bit, and most proposals are used."""
    cm = gpu
    cl, o = king4e6
    say("training the bracket network on one harvested step")
    with ctx_for(cm, cl) as ctx:
        ctx.upload(cl)
        ctx.init_cluster()
        ctx.mlp_init(seed=1, max_rows=1 << 22)
        ctx.mlp_harvest(True)
        ctx.step()
        ctx.mlp_harvest(False)
        rows = ctx.mlp_rows()
        ctx.mlp_fit_norm()
        ctx.mlp_train(800, batch=4096, lr=2e-3, seed=3)
        val = rows[::211][:20000]
        err = float(np.sqrt(np.mean((ctx.mlp_predict(val[:, :6]) - val[:, 6:]) ** 2)))
        th, nr = ctx.mlp_params()
    say("rows", rows.shape[0], "rms log r error", err)
    d, sc, st = gpu_step(cm, cl, learned=(th, nr, max(3.0 * err, 0.01)))
    hit, miss = st["n_bracket_hit"], st["n_bracket_fallback"]
    say("learned brackets accepted", hit, "rejected", miss)
    assert st["error"] == 0
    assert_state_equal(d, sc, o, "king4e6 learned")
    assert hit > 0.8 * (hit + miss)


def test_gpu_king4e6_vs_reference_libm_and_long_double(gpu, O, king4e6):
    """The reference's own arithmetic (C-library sin/cos, x87 long-double Q): stars
    whose r, v_r or v_t differ by more than 1e-10 are counted (acceptance and bracket
    flips) and must stay rare; the total energy agrees to 1e-8."""
    cm = gpu
    cl, _ = king4e6
    say("oracle step with the reference's libm and long-double Q")
    oref = O.Oracle(cl.a, cl.n, orbit_r=cl.rtid, rng_mode=O.RNG_PER_STAR, q_mode=O.Q_LONG_DOUBLE,
                    libm_mode=O.LIBM_SYSTEM)
    oref.init()
    assert oref.step() == 0
    d, sc, st = gpu_step(cm, cl)
    Nr = oref.N
    n = sc["n_max"]
    slots, found = by_id(d, oref["id"][1:Nr + 1], n)
    bad = ~found
    for f in ("r", "vr", "vt"):
        bad |= rel(d[f][slots], oref[f][1:Nr + 1]) > 1e-10
    say("stars outside 1e-10 of the reference arithmetic", int(bad.sum()), "of", Nr, "N gpu", n)
    assert bad.sum() / Nr < 2e-3
    assert sc["Etot"] == pytest.approx(oref.sc("Etot"), rel=1e-8)


def test_gpu_plummer_kroupa_1e6_ten_steps(gpu, O):
    cm = gpu
    say("Plummer N=1e6 Kroupa seed 1: ten steps")
    cl = cm.make_plummer(1000000, seed=1, kroupa=True)
    o = O.Oracle(cl.a, cl.n, orbit_r=cl.rtid, rng_mode=O.RNG_PER_STAR, q_mode=O.Q_FP64, libm_mode=O.LIBM_DD)
    o.init()
    etot = []
    for i in range(10):
        assert o.step() == 0
        etot.append(o.sc("Etot"))
    say("oracle done: N", o.N)
    with ctx_for(cm, cl) as ctx:
        ctx.upload(cl)
        ctx.init_cluster()
        for i in range(10):
            st = ctx.step()
            assert st["error"] == 0
            assert ctx.scalars()["Etot"] == pytest.approx(etot[i], rel=1e-12), i
        d = ctx.download()
        sc = ctx.scalars()
    assert sc["step"] == o.s.step
    assert_state_equal(d, sc, o, "plummer-kroupa 1e6, step 10")


@pytest.mark.parametrize("seed", [2, 3])
def test_gpu_energy_drift_100_steps_more_seeds(gpu, O, seed):
    """100 steps of Plummer 2e4 at further seeds: the drift agrees with the oracle's
    (device sin/cos) to 1e-6, and with the oracle run on the reference's C-library
    sin/cos within the spread of the reference's own drift over the math libraries it
    may be built with (C-library sin/cos, sincos from an optimising compiler, long
    double): last-bit differences of sin/cos move the drift by ~1e-5."""
    cm = gpu
    steps, n = 100, 20000
    say("energy drift, 100 steps, Plummer 2e4, seed", seed)
    cl = cm.make_plummer(n, seed=seed)
    drift = {}
    for lm in (O.LIBM_DD, O.LIBM_SYSTEM, O.LIBM_SINCOS, O.LIBM_LONG_DOUBLE):
        o = O.Oracle(cl.a, cl.n, orbit_r=cl.rtid, rng_mode=O.RNG_PER_STAR, q_mode=O.Q_FP64, libm_mode=lm)
        o.init()
        e0 = o.sc("Etot")
        for _ in range(steps):
            assert o.step() == 0
        drift[lm] = (o.sc("Etot") + o.sc("Eoops") - e0) / abs(e0)
    with ctx_for(cm, cl) as ctx:
        ctx.upload(cl)
        ctx.init_cluster()
        g0 = ctx.scalars()["Etot"]
        for _ in range(steps):
            assert ctx.step()["error"] == 0
        sc = ctx.scalars()
    dg = (sc["Etot"] + sc["Eoops"] - g0) / abs(g0)
    ref = [drift[O.LIBM_SYSTEM], drift[O.LIBM_SINCOS], drift[O.LIBM_LONG_DOUBLE]]
    spread = max(ref) - min(ref)
    say("seed", seed, "drift gpu", dg, "oracle dd", drift[O.LIBM_DD], "reference libm variants", ref,
        "spread", spread)
    assert abs(dg - drift[O.LIBM_DD]) < 1e-6
    assert abs(dg - drift[O.LIBM_SYSTEM]) <= max(1e-6, spread)
