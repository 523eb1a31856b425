"""Learned apsides brackets on the device: the forward and backward kernels against
the CPU oracle (oracle/mlp_oracle.py), the harvested training rows against the
state they come from, training, and the orbit phase with learned brackets, which
must give exactly the plain path's orbits."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def ctx_for(cm, cl, seed=1):
    return cm.Context(cl.n, cm.default_params(cl.n, orbit_r=cl.rtid, seed=seed))


def _norm(rng):
    return np.concatenate([rng.normal(size=6), rng.uniform(0.5, 2.0, 6), rng.normal(size=2),
                           rng.uniform(0.5, 2.0, 2)]).astype(np.float32)


def _standardise(rows, nr):
    x = (rows[:, :6].astype(np.float64) - nr[:6]) / nr[6:12]
    t = (rows[:, 6:].astype(np.float64) - nr[12:14]) / nr[14:16]
    return x, t


@pytest.mark.parametrize("n", [1, 300, 4096])
def test_gpu_mlp_gradient_matches_oracle(gpu, n):
    """FP32 device gradients against the float64 oracle: every parameter block
    within 2e-4 of the block's largest gradient (FP32 sums over the batch)."""
    cm = gpu
    say("mlp gradient, batch", n)
    rng = np.random.default_rng(n)
    cl = cm.make_plummer(1000, seed=1)
    rows = rng.normal(size=(n, 8)).astype(np.float32)
    nr = _norm(rng)
    with ctx_for(cm, cl) as ctx:
        ctx.mlp_init(seed=11, max_rows=1024)
        th, _ = ctx.mlp_params()
        ctx.mlp_set_params(th, nr)
        loss, g = ctx.mlp_loss_grad(rows)
    x, t = _standardise(rows, nr)
    lo, go = M.loss_and_grad(th.astype(np.float64), x, t)
    say("loss gpu", loss, "oracle", lo)
    assert abs(loss - lo) <= 1e-5 * abs(lo) + 1e-7
    bounds = [M.OFF_W1, M.OFF_B1, M.OFF_W2, M.OFF_B2, M.OFF_W3, M.OFF_B3, M.NPARAM]
    for a, b in zip(bounds[:-1], bounds[1:]):
        scale = np.max(np.abs(go[a:b])) + 1e-12
        err = np.max(np.abs(g[a:b] - go[a:b])) / scale
        say("block", a, b, "max err / max |g|", err)
        assert err <= 2e-4, (a, b)


def test_gpu_mlp_predict_matches_oracle(gpu):
    cm = gpu
    say("mlp predict")
    rng = np.random.default_rng(4)
    cl = cm.make_plummer(1000, seed=1)
    xr = rng.normal(size=(777, 6)).astype(np.float32)
    nr = _norm(rng)
    with ctx_for(cm, cl) as ctx:
        ctx.mlp_init(seed=5, max_rows=1024)
        th, _ = ctx.mlp_params()
        ctx.mlp_set_params(th, nr)
        y = ctx.mlp_predict(xr)
    xs = (xr.astype(np.float64) - nr[:6]) / nr[6:12]
    yo, _ = M.forward(th.astype(np.float64), xs)
    yo = yo * nr[14:16] + nr[12:14]
    err = np.max(np.abs(y - yo))
    say("max |y - oracle|", err)
    assert err <= 1e-4 * (1 + np.max(np.abs(yo)))


def test_gpu_mlp_mfma_inference_matches_fp32(gpu):
    """The matrix-core (f16 operands, f32 accumulation) forward pass that proposes the
    orbit phase's learned brackets against the per-thread FP32 pass and the float64
    oracle: within 1e-2 of each output's standard deviation, on ragged row counts."""
    cm = gpu
    say("mlp inference on the matrix cores")
    rng = np.random.default_rng(6)
    cl = cm.make_plummer(1000, seed=1)
    nr = _norm(rng)
    with ctx_for(cm, cl) as ctx:
        ctx.mlp_init(seed=9, max_rows=1024)
        th, _ = ctx.mlp_params()
        ctx.mlp_set_params(th, nr)
        for n in (1, 15, 16, 17, 1000, 4099):
            xr = (nr[:6] + nr[6:12] * rng.normal(size=(n, 6))).astype(np.float32)
            ctx.test_hook(cm.HOOK_MLP_MFMA, 0)
            y32 = ctx.mlp_predict(xr)
            ctx.test_hook(cm.HOOK_MLP_MFMA, 1)
            y16 = ctx.mlp_predict(xr)
            xs = (xr.astype(np.float64) - nr[:6]) / nr[6:12]
            yo, _ = M.forward(th.astype(np.float64), xs)
            yo = yo * nr[14:16] + nr[12:14]
            e16 = np.max(np.abs(y16 - yo) / nr[14:16])
            e32 = np.max(np.abs(y32 - yo) / nr[14:16])
            say("rows", n, "max |y16 - oracle| / std", e16, "fp32", e32)
            assert e16 <= 1e-2 and e32 <= 1e-4
        ctx.test_hook(cm.HOOK_MLP_MFMA, 0)


def test_gpu_mlp_harvested_rows_match_state(gpu):
    """Rows harvested by one orbit phase: one per sampled star, inputs computed from
    the star's state and targets equal to log of the apsides the phase wrote."""
    cm = gpu
    say("mlp harvest")
    cl = cm.make_plummer(20000, seed=3)
    with ctx_for(cm, cl) as ctx:
        ctx.upload(cl)
        ctx.init_cluster()
        d0 = ctx.download()
        ctx.mlp_init(seed=1, max_rows=1 << 16)
        ctx.mlp_harvest(True)
        ctx.orbits(step=0)
        d = ctx.download()
        rows = ctx.mlp_rows()
    N = cl.n
    sampled = (d["flags"][1:N + 1] & 256) != 0
    sampled &= (d["flags"][1:N + 1] & (32 | 16)) == 0        # not deferred, not Brent
    say("rows", rows.shape[0], "sampled stars", int(sampled.sum()))
    assert rows.shape[0] == int(sampled.sum())
    j = np.nonzero(sampled)[0] + 1
    rp = np.sort(np.log(d["r_peri"][j].astype(np.float32)))
    assert np.allclose(np.sort(rows[:, 6]), rp, rtol=1e-6, atol=1e-6)
    r, phi, m = d0["r"], d0["phi"], d0["m"]
    E = d0["E"][j] + m[j] / r[j] / cl.n
    want = np.stack([E, d0["J"][j], np.log(r[j]), phi[j],
                     np.log(np.maximum((phi[j + 1] - phi[j - 1]) / (r[j + 1] - r[j - 1]) * r[j] ** 2, 1e-30)),
                     m[j]], 1).astype(np.float32)
    for c in range(6):
        assert np.allclose(np.sort(rows[:, c]), np.sort(want[:, c]), rtol=1e-5, atol=1e-6), c


def _train(cm, ctx, steps, iters):
    ctx.mlp_init(seed=1, max_rows=1 << 21)
    ctx.mlp_harvest(True)
    for _ in range(steps):
        ctx.step()
    ctx.mlp_harvest(False)
    rows = ctx.mlp_rows()
    ctx.mlp_fit_norm()
    val = rows[::97][:8000]
    e0 = np.sqrt(np.mean((ctx.mlp_predict(val[:, :6]) - val[:, 6:]) ** 2))
    loss = ctx.mlp_train(iters, batch=4096, lr=2e-3, seed=3)
    e1 = np.sqrt(np.mean((ctx.mlp_predict(val[:, :6]) - val[:, 6:]) ** 2))
    return rows.shape[0], e0, e1, loss


def test_gpu_mlp_learns_apsides_and_keeps_orbits_exact(gpu):
    """Train on rows harvested from two steps of a Kroupa Plummer 1e5 run, then run
    the orbit phase of the next step twice from the same state: plain and with
    learned brackets.  Every orbit must be bit-identical (the bisection decides the
    index); the proposals must be accepted for most stars."""
    cm = gpu
    say("mlp training + learned brackets")
    cl = cm.make_plummer(100000, seed=2, kroupa=True)
    with ctx_for(cm, cl) as ctx:
        ctx.upload(cl)
        ctx.init_cluster()
        nrows, e0, e1, loss = _train(cm, ctx, 2, 600)
        say("rows", nrows, "rms log r error before", e0, "after", e1, "loss", loss)
        assert e1 < 0.25 * e0
        state = ctx.download()
        sc = ctx.scalars()
        ctx.set_learned_brackets(False)
        ctx.energy_conservation1()
        plain_stats = ctx.orbits(step=7)
        plain = ctx.download()
        ctx.upload(state)
        ctx.set_scalars(**{k: sc[k] for k in ("n_max", "n_max_new", "Mtotal", "Rtidal", "Dt", "Prev_Dt",
                                             "TotalTime", "TidalMassLoss", "TidalMassLoss_old",
                                             "OldTidalMassLoss")})
        ctx.set_learned_brackets(True, margin=max(3.0 * e1, 0.01))
        ctx.energy_conservation1()
        learned_stats = ctx.orbits(step=7)
        learned = ctx.download()
        # perfect proposals (the apsides just found) with a tiny margin: nearly every
        # search starts from a bracket of a few nodes
        ctx.test_hook(cm.HOOK_PERFECT_BRACKETS, 1)
        ctx.upload(state)
        ctx.set_scalars(**{k: sc[k] for k in ("n_max", "n_max_new", "Mtotal", "Rtidal", "Dt", "Prev_Dt",
                                             "TotalTime", "TidalMassLoss", "TidalMassLoss_old",
                                             "OldTidalMassLoss")})
        ctx.set_learned_brackets(True, margin=1e-6)
        ctx.energy_conservation1()
        perfect_stats = ctx.orbits(step=7)
        perfect = ctx.download()
        ctx.test_hook(cm.HOOK_PERFECT_BRACKETS, 0)
    N = plain.n
    for k in ("rnew", "vrnew", "vtnew", "r_peri", "r_apo", "X"):
        assert np.array_equal(plain[k][1:N + 1], learned[k][1:N + 1]), k
        assert np.array_equal(plain[k][1:N + 1], perfect[k][1:N + 1]), k
    ph, pm = perfect_stats["n_bracket_hit"], perfect_stats["n_bracket_fallback"]
    say("perfect proposals accepted", ph, "rejected", pm)
    assert ph > 0.95 * (ph + pm)
    hit, miss = learned_stats["n_bracket_hit"], learned_stats["n_bracket_fallback"]
    say("learned brackets accepted", hit, "rejected", miss)
    assert hit > 0.8 * (hit + miss)
    assert plain_stats["n_bracket_hit"] == 0
