"""The oracle's central_calculate + clusdyn_calculate (cmc_utils.c:1239-1345,
1614-1624) pinned by closed forms, not by restating its loops:

  - a uniform-density sphere of equal masses (r_k^3 = k/N): every local density
    estimate is 5 masses over the volume of 6 shells, rho_j = 5/(8 pi); the
    density-weighted density with the 4/5 bias and factor-2 corrections is then
    exactly 1/pi, the weighted averages are plain means, and with N a power of two
    the half-mass loop stops at star N/2 (nave = N/2 + 1);
  - the same sphere with one heavy star at index s: only the five estimates whose
    mass window (jmin+1 .. jmin+5, jmin = i-3) holds s rise, by (5 + d)/5, which
    fixes the window's width and alignment through rhojsum, rho and rc;
  - unequal masses: nave is the star after the first one whose enclosed mass
    reaches half of Mtotal, found here with exact rational sums;
  - the reference's exit condition (N_STAR <= 2J) is reported.
"""
import math
from fractions import Fraction

import numpy as np
import pytest


def sphere(n, heavy=None, d=0.0, v=(0.3, 0.4)):
    r = np.zeros(n + 2)
    r[1:n + 1] = np.cbrt(np.arange(1, n + 1) / n)
    r[0], r[n + 1] = 1e-20, 1e10
    m = np.zeros(n + 2)
    m[1:n + 1] = 1.0
    if heavy is not None:
        m[heavy] += d
    vr = np.zeros(n + 2)
    vt = np.zeros(n + 2)
    vr[1:n + 1], vt[1:n + 1] = v
    return {"r": r, "m": m, "vr": vr, "vt": vt, "id": np.arange(n + 2)}


def central(O, arrays, n):
    o = O.Oracle(arrays, n, orbit_r=1e6)
    o.init()
    return o, o.central_calculate()


def test_uniform_sphere_closed_form(O):
    n = 4096
    a = sphere(n)
    o, c = central(O, a, n)
    assert o.sc("Mtotal") == 1.0
    nave = n // 2 + 1
    assert c["nave"] == nave
    assert c["rh"] == a["r"][nave]
    b = 5.0 / (8.0 * math.pi)                       # 5 m madhoc / (4 pi/3 * 6/N)
    assert c["rhojsum"] == pytest.approx(nave * b, rel=1e-11)
    assert c["rhoj2sum"] == pytest.approx(nave * b * b, rel=1e-11)
    assert c["rho"] == pytest.approx(1.0 / math.pi, rel=1e-11)
    assert c["v_rms"] == pytest.approx(0.5, rel=1e-14)
    ri = a["r"][1:nave + 1]
    assert c["rc"] == pytest.approx(ri.mean(), rel=1e-11)
    assert c["rc_nb"] == pytest.approx(math.sqrt(np.mean(ri * ri)), rel=1e-11)
    assert c["m_ave"] == pytest.approx(1.0 / n, rel=1e-12)
    assert c["n"] == pytest.approx(n / math.pi, rel=1e-11)
    assert c["rc_spitzer"] == pytest.approx(math.sqrt(3 * 0.25 / 4.0), rel=1e-11)
    assert c["N_core"] == pytest.approx(4 * math.pi / 3 * c["rc"] ** 3 * c["n"] / 2, rel=1e-12)
    assert c["N_core_nb"] == pytest.approx(4 * math.pi / 3 * c["rc_nb"] ** 3 * c["n"] / 2, rel=1e-12)
    assert c["Trc"] == pytest.approx(0.065 * 0.125 / (1.0 / math.pi / n), rel=1e-11)


def test_heavy_star_window(O):
    n, s, d = 4096, 700, 995.0                      # star s weighs 996 mean masses
    a = sphere(n, heavy=s, d=d)
    o, c = central(O, a, n)
    # half-mass index with total (n + d)/n: first k with (k + d)/n >= (n + d)/(2n), s <= k
    k = math.ceil((n + d) / 2 - d)
    assert c["nave"] == k + 1
    nave = k + 1
    b = 5.0 / (8.0 * math.pi)
    boosted = range(s - 2, s + 3)                   # i with i-3+1 <= s <= i-3+5
    w = np.full(nave + 1, b)
    w[0] = 0.0
    for i in boosted:
        w[i] = b * (5.0 + d) / 5.0
    assert c["rhojsum"] == pytest.approx(w.sum(), rel=1e-11)
    assert c["rhoj2sum"] == pytest.approx(np.sum(w * w), rel=1e-11)
    assert c["rho"] == pytest.approx(1.6 * np.sum(w * w) / w.sum(), rel=1e-11)
    r = a["r"][:nave + 1]
    assert c["rc"] == pytest.approx(np.sum(w * r) / w.sum(), rel=1e-11)
    mm = a["m"][:nave + 1] / n
    assert c["m_ave"] == pytest.approx(np.sum(w * mm) / w.sum(), rel=1e-11)
    # a window shifted by one star would move rc by more than the tolerance
    w2 = np.full(nave + 1, b)
    w2[0] = 0.0
    for i in range(s - 3, s + 2):
        w2[i] = b * (5.0 + d) / 5.0
    assert abs(np.sum(w2 * r) / w2.sum() - c["rc"]) > 1e-8 * c["rc"]


def test_unequal_masses_half_mass_index(O):
    n = 5000
    rng = np.random.default_rng(3)
    a = sphere(n)
    a["m"][1:n + 1] = rng.uniform(0.5, 1.5, n)
    o, c = central(O, a, n)
    thr = Fraction(o.sc("Mtotal")) / 2
    acc = Fraction(0)
    k = None
    for j in range(1, n + 1):
        acc += Fraction(a["m"][j]) / n
        if acc >= thr:
            k = j
            break
    prev = acc - Fraction(a["m"][k]) / n
    assert float(acc - thr) > 1e-9 and float(thr - prev) > 1e-9   # no rounding tie near the threshold
    assert c["nave"] == k + 1
    assert c["rh"] == a["r"][k + 1]


def test_exit_condition(O):
    n = 12
    with pytest.raises(RuntimeError):
        central(O, sphere(n), n)
