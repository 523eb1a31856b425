"""A central point mass (cenma.m > 0, loss cone off): the potential, two whole steps
and comp_mass_percent with MASS_PC_BH_INCLUDE on the device against the oracle with
the same central mass.  The central mass enters the potential as -cenma.m madhoc / r
(cmc_utils.c:675-736, 112-165), the orbits through it, and the enclosed-mass chain
of comp_mass_percent from cenma.m madhoc (cmc_utils.c:878).

Bars as for the plain cluster: with the device's sin/cos and FP64 Q in the oracle,
every star bit-identical after the steps; comp_mass_percent indices, radii and sums
bit for bit, densities within 2 ulps.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("frac", [0.001, 0.05])
def test_gpu_steps_with_central_mass(gpu, O, frac):
    cm = gpu
    cl = cm.make_plummer(100000, seed=5)
    cen = frac * cl.n                                   # stored units (m * N_STAR)
    say("two steps with a central mass of", frac, "of the cluster")
    o = O.Oracle(cl.a, cl.n, orbit_r=cl.rtid, rng_mode=O.RNG_PER_STAR, q_mode=O.Q_FP64, libm_mode=O.LIBM_DD)
    o.s.cenma_m = cen
    o.init()
    p = cm.default_params(cl.n, orbit_r=cl.rtid)
    p.cenma_m = cen
    with cm.Context(cl.n, p) as ctx:
        ctx.upload(cl)
        ctx.init_cluster()
        for s in range(2):
            ctx.step()
            assert o.step() == 0, s
        d = ctx.download()
        sc = ctx.scalars()
    say("n_max", sc["n_max"], o.N, "Etot", sc["Etot"], o.sc("Etot"))
    assert sc["n_max"] == o.N
    gi = {int(i): k + 1 for k, i in enumerate(d["id"][1:o.N + 1])}
    idx = np.array([gi[int(i)] for i in o["id"][1:o.N + 1]])
    for k in ("r", "vr", "vt", "E", "J", "phi"):
        same = int(np.sum(d[k][idx] == o[k][1:o.N + 1]))
        say(k, "bit-equal:", same, "of", o.N)
        assert same == o.N, k
    # the total energy is a block-reduced sum on the device and a sequential one in the
    # oracle; with the central mass its terms span more decades (the -cenma m / r of the
    # inner stars), so the two orders differ by a few 1e-12
    assert sc["Etot"] == pytest.approx(o.sc("Etot"), rel=1e-10)


def test_gpu_potential_and_mass_percent_with_central_mass(gpu, O):
    cm = gpu
    cl = cm.make_king(300000, 7.0, seed=3)
    cen = 0.02 * cl.n
    say("potential and comp_mass_percent (BH included) with a central mass")
    o = O.Oracle(cl.a, cl.n, orbit_r=cl.rtid)
    o.s.cenma_m = cen
    o.potential_calculate()
    pc = np.array([0.001, 0.01, 0.015, 0.02, 0.025, 0.05, 0.1, 0.5, 0.9])
    want = o.comp_mass_percent(pc, bh_include=1, minimum_r=1e-3)
    p = cm.default_params(cl.n, orbit_r=cl.rtid)
    p.cenma_m = cen
    with cm.Context(cl.n, p) as ctx:
        ctx.upload(cl)
        ctx.potential()
        d = ctx.download()
        got = ctx.comp_mass_percent(pc, bh_include=1, minimum_r=1e-3)
    n = cl.n
    assert np.array_equal(d["phi"][1:n + 1], o["phi"][1:n + 1])
    say("fractions inside the central mass:", int(np.sum(want["no_star_r"] == 0)), "reached", got["reached"])
    assert int(np.sum(want["no_star_r"] == 0)) >= 3
    assert got["reached"] == int(np.sum(want["no_star_r"] >= 0))
    for k in ("no_star_r", "mass_r", "ave_mass_r", "ke_rad_r", "ke_tan_r", "v2_rad_r", "v2_tan_r"):
        assert np.array_equal(got[k], want[k]), k
    a, b = got["densities_r"], want["densities_r"]
    assert np.all(np.abs(a - b) <= 2 * np.spacing(np.abs(b)))
