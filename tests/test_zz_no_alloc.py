"""cmcg_calc_pericenter_and_apocenter_noalloc (stages its inputs and outputs in a
fixed device array instead of allocating) against cmcg_calc_pericenter_and_apocenter:
bit-identical apsides on two calls in a row, the second with fewer stars, and the
size limit reported as an error."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _inputs(o, first, count):
    j = np.arange(first, first + count)
    E = o["E"][j] + o["m"][j] / o["r"][j] / o.s.N_STAR
    return E, o["J"][j].copy()


def test_gpu_apsides_noalloc_matches_allocating_entry(gpu, O):
    cm = gpu
    say("apsides: no-allocation entry point vs the allocating one")
    cl = cm.make_king(100000, 7.0, seed=12)
    o = O.Oracle(cl.a, cl.n, orbit_r=cl.rtid, q_mode=O.Q_FP64)
    o.init()
    with cm.Context(cl.n, cm.default_params(cl.n, orbit_r=cl.rtid, seed=1)) as ctx:
        h = cl.copy()
        for k in ("phi", "E", "J"):
            h[k] = o[k]
        ctx.upload(h)
        ctx.set_scalars(n_max=o.N, n_max_new=o.N, Mtotal=o.sc("Mtotal"), Rtidal=o.sc("Rtidal"))
        cap = cm.Context.APSIDES_NOALLOC_MAX
        for first, count in ((1, cap), (50001, 777)):
            E, J = _inputs(o, first, count)
            want = ctx.apsides(first, E, J)
            got = ctx.apsides(first, E, J, noalloc=True)
            say("first", first, "count", count, "circular", int(want["circular"].sum()))
            for k in want:
                assert np.array_equal(got[k], want[k]), (first, count, k)
        # a third call with the first size again still sees only its own inputs
        E, J = _inputs(o, 3, cap)
        assert np.array_equal(ctx.apsides(3, E, J, noalloc=True)["rp"], ctx.apsides(3, E, J)["rp"])
        E, J = _inputs(o, 1, cap + 1)
        with pytest.raises(cm.CmcgError) as ei:
            ctx.apsides(1, E, J, noalloc=True)
        assert ei.value.code == cm.E_CAPACITY
