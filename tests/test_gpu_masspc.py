"""comp_mass_percent on the device (cmcg_comp_mass_percent) against the oracle's
restatement of cmc_utils.c:872-982: the indices, radii, mean masses and the four
running sums bit for bit (their sequential roundings are reproduced), the densities
within 2 ulps (the device cubes r with one rounding where the reference calls
pow(r, 3)).  Also through the chains' fallback paths (test hook)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PC = np.array([1e-4, 3e-4, 1e-3, 3e-3, 0.01, 0.03, 0.05, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 0.95, 0.99,
               0.99, 0.99])


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _compare(d, o):
    assert d["reached"] == int(np.sum(o["no_star_r"] >= 0))
    for k in ("no_star_r", "mass_r", "ave_mass_r", "ke_rad_r", "ke_tan_r", "v2_rad_r", "v2_tan_r"):
        assert np.array_equal(d[k], o[k]), k
    a, b = d["densities_r"], o["densities_r"]
    ok = np.isnan(a) == np.isnan(b)
    fin = ~np.isnan(a)
    assert ok.all() and np.all(np.abs(a[fin] - b[fin]) <= 2 * np.spacing(np.abs(b[fin])))


@pytest.mark.parametrize("case", ["plummer_kroupa_1e5", "king_1e6"])
def test_gpu_comp_mass_percent(gpu, O, case):
    cm = gpu
    say("comp_mass_percent", case)
    cl = cm.make_plummer(100000, seed=4, kroupa=True) if case == "plummer_kroupa_1e5" else \
        cm.make_king(1000000, 7.0, seed=2)
    o = O.Oracle(cl.a, cl.n, orbit_r=cl.rtid)
    o.potential_calculate()
    want = o.comp_mass_percent(PC)
    with cm.Context(cl.n, cm.default_params(cl.n, orbit_r=cl.rtid, seed=1)) as ctx:
        ctx.upload(cl)
        ctx.potential()
        got = ctx.comp_mass_percent(PC)
        ctx.test_hook(cm.HOOK_CHAIN_K_DROP, 50)
        got_fb = ctx.comp_mass_percent(PC)
    say(case, "fractions reached", got["reached"], "last N", got["no_star_r"][got["reached"] - 1])
    _compare(got, want)
    _compare(got_fb, want)
