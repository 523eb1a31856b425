"""Whole steps on small and ragged cluster sizes (partial waves, blocks and tiles in
every kernel) against the oracle with the device's sin/cos and FP64 Q: after 3 steps
every star is bit-identical and N_MAX, Dt agree.  N = 8 and 16 are below the
density and dispersion windows (2 AVEKERNEL + 2 = 42 stars), where the reference
indexes outside the star array; the device and the oracle both clamp those windows
to the array there."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n", [8, 16, 41, 63, 64, 65, 257, 1000, 2049, 12345])
def test_gpu_small_and_ragged_steps(gpu, O, n):
    cm = gpu
    say("three steps, N =", n)
    cl = cm.make_plummer(n, seed=n)
    o = O.Oracle(cl.a, cl.n, orbit_r=cl.rtid, rng_mode=O.RNG_PER_STAR, q_mode=O.Q_FP64, libm_mode=O.LIBM_DD)
    o.init()
    with cm.Context(cl.n, cm.default_params(cl.n, orbit_r=cl.rtid)) as ctx:
        ctx.upload(cl)
        ctx.init_cluster()
        for s in range(3):
            ctx.step()
            assert o.step() == 0, s
        d = ctx.download()
        sc = ctx.scalars()
    assert sc["n_max"] == o.N and sc["Dt"] == o.sc("Dt")
    gi = {int(i): k + 1 for k, i in enumerate(d["id"][1:o.N + 1])}
    idx = np.array([gi[int(i)] for i in o["id"][1:o.N + 1]])
    for k in ("r", "vr", "vt", "E", "J", "m", "phi"):
        assert np.array_equal(d[k][idx], o[k][1:o.N + 1]), k
