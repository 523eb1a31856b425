"""The oracle's comp_mass_percent (cmc_utils.c:872-982, one rank) pinned by what the
quantities mean: each Lagrange radius is the radius of the first star at which the
enclosed mass exceeds the fraction of Mtotal, the sums inside it agree with pairwise
summation to rounding, at most one fraction is assigned per star, and fractions the
central mass already exceeds get MINIMUM_R."""
import numpy as np


def _cluster(cm, O, n=20000, seed=3, cenma=0.0):
    cl = cm.make_plummer(n, seed=seed, kroupa=True)
    o = O.Oracle(cl.a, cl.n, orbit_r=cl.rtid)
    o.s.cenma_m = cenma
    o.potential_calculate()
    return cl, o


def test_mass_percent_meaning(cm, O):
    cl, o = _cluster(cm, O)
    N = o.sc("N_MAX")
    pc = np.array([1e-4, 1e-3, 0.01, 0.05, 0.1, 0.3, 0.5, 0.7, 0.9, 0.99])
    r = o.comp_mass_percent(pc)
    m, rr, vr, vt = (cl[k][1:N + 1] for k in ("m", "r", "vr", "vt"))
    mad = o.sc("madhoc")
    Mt = o.sc("Mtotal")
    cum = np.cumsum(m / cl.n)
    for i, f in enumerate(pc):
        k = int(r["no_star_r"][i])
        assert 1 <= k <= N
        # first star past the fraction (pairwise sums may move the crossing by one star)
        k_np = int(np.argmax(cum > f * Mt)) + 1
        assert abs(k - k_np) <= 1
        assert r["mass_r"][i] == rr[k - 1]
        mprev = r["ave_mass_r"][i] * Mt * k
        assert abs(mprev - cum[k - 1]) <= 1e-12 * cum[k - 1]
        assert abs(r["densities_r"][i] - mprev / (4.0 / 3.0 * np.pi * rr[k - 1] ** 3)) <= 1e-12 * r["densities_r"][i]
        for key, terms in (("ke_rad_r", 0.5 * m * mad * vr * vr), ("ke_tan_r", 0.5 * m * mad * vt * vt),
                           ("v2_rad_r", vr * vr), ("v2_tan_r", vt * vt)):
            want = np.sum(terms[:k])
            assert abs(r[key][i] - want) <= 1e-12 * want, key
    assert np.all(np.diff(r["no_star_r"]) > 0)


def test_one_fraction_per_star(cm, O):
    """Equal fractions are all exceeded at the same star, but the loop advances one
    fraction per star: the indices are consecutive."""
    cl, o = _cluster(cm, O)
    r = o.comp_mass_percent([0.3, 0.3, 0.3])
    k = r["no_star_r"]
    assert k[1] == k[0] + 1 and k[2] == k[0] + 2


def test_central_mass_included(cm, O):
    """MASS_PC_BH_INCLUDE: fractions below the central mass's share get MINIMUM_R and
    N = 0; the rest start the running sum from the central mass."""
    cl, o = _cluster(cm, O, cenma=1000.0)            # stored units: 1000 / N of the total
    share = 1000.0 * o.sc("madhoc") / o.sc("Mtotal")
    pc = np.array([0.2 * share, 0.5 * share, 2.0 * share, 0.5])
    r = o.comp_mass_percent(pc, bh_include=1, minimum_r=1e-6)
    assert r["mass_r"][0] == r["mass_r"][1] == 1e-6
    assert r["no_star_r"][0] == r["no_star_r"][1] == 0
    assert r["no_star_r"][2] >= 1 and r["no_star_r"][3] > r["no_star_r"][2]
    r0 = o.comp_mass_percent(pc, bh_include=0)
    assert r0["no_star_r"][0] >= 1                  # without the central mass every fraction needs stars
