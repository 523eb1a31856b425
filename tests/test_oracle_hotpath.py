"""Pins the CPU oracle's hot-path restatement against what the mathematics fixes:
closed forms, brute force on small inputs, invariants, and a textbook special case.
Each check is chosen so that a plausible transcription error (dropped term, wrong
sign or index) fails it."""
import ctypes as C

import mpmath
import numpy as np
import pytest
from scipy import stats


def small_cluster(cm, n=400, seed=3, kroupa=False):
    return cm.make_plummer(n, seed=seed, kroupa=kroupa)


def make_oracle(O, cl, **kw):
    kw.setdefault("orbit_r", cl.rtid)
    return O.Oracle(cl.a, cl.n, **kw)


# ------------------------------------------------------------------ potential
def test_potential_matches_exact_closed_form(cm, O):
    """phi_k = Mtotal/r_{N+1} - sum_{i>k} a_i/r_i - (Mtotal - sum_{i>k} a_i)/r_k with
    a_i = fl(m_i/N_STAR): the telescoped form of cmc_utils.c:717-727, evaluated with
    60-digit arithmetic on the same double inputs."""
    cl = small_cluster(cm, 1500, seed=11, kroupa=True)
    o = make_oracle(O, cl)
    o.potential_calculate()
    N = o.N
    mpmath.mp.dps = 60
    r = [mpmath.mpf(float(x)) for x in o["r"]]
    a = [mpmath.mpf(float(x) / float(N)) for x in o["m"]]     # fl(m/N) as the reference rounds
    Mt = mpmath.mpf(o.sc("Mtotal"))
    T = mpmath.mpf(0)
    S = mpmath.mpf(0)
    worst = 0.0
    for k in range(N, 0, -1):
        exact = Mt / mpmath.mpf(1e10) - S - (Mt - T) / r[k]
        rel = abs((mpmath.mpf(float(o["phi"][k])) - exact) / exact)
        worst = max(worst, float(rel))
        T += a[k]
        S += a[k] / r[k]
    assert worst < 1e-13, worst
    assert o["phi"][N + 1] == 0.0 and o["r"][N + 1] == 1e10
    assert o["phi"][0] == o["phi"][1]           # no central mass


def test_potential_mass_and_count(cm, O):
    cl = small_cluster(cm, 1000, seed=2)
    cl["r"][995:1001] = 1e10                     # stripped tail
    o = make_oracle(O, cl)
    n = o.potential_calculate()
    assert n == 994
    assert abs(o.sc("Mtotal") - 994 / 1000) < 1e-15
    assert abs(o.sc("Rtidal") - cl.rtid * (994 / 1000) ** (1 / 3)) < 1e-12 * cl.rtid


def test_potential_plummer_profile(cm, O):
    """Large-N potential approaches the analytic Plummer potential -1/sqrt(r^2+a^2),
    a = 3 pi/16 in Henon units (the generator scales exactly to PE = -1/2)."""
    cl = cm.make_plummer(100000, seed=1)
    o = make_oracle(O, cl)
    o.potential_calculate()
    r, phi = o["r"][1:o.N + 1], o["phi"][1:o.N + 1]
    a = 3 * np.pi / 16
    sel = (r > 0.1) & (r < 5)
    ana = -1 / np.sqrt(r[sel] ** 2 + a * a)
    assert np.median(np.abs(phi[sel] / ana - 1)) < 0.02


def test_findzero_r_bruteforce(O):
    rng = np.random.default_rng(0)
    n = 300
    r = np.sort(rng.choice(np.round(rng.random(150) * 10, 3), n))     # duplicates on purpose
    arr = {"r": np.concatenate([[1e-20], r, [1e10]]), "m": np.ones(n + 2)}
    o = O.Oracle(arr, n)
    for x in list(rng.random(200) * 11) + list(r[:20]):
        k = O.lib.orc_FindZero_r(C.byref(o.s), 1, n + 1, float(x))
        cand = [j for j in range(1, n + 2) if o["r"][j] < x]
        assert k == (max(cand) if cand else 1), x


def test_potential_interpolant_is_exact_at_nodes(cm, O):
    cl = small_cluster(cm, 500, seed=4)
    o = make_oracle(O, cl)
    o.potential_calculate()
    for k in [1, 2, 50, 250, 499, 500]:
        assert o.potential(o["r"][k]) == pytest.approx(o["phi"][k], rel=1e-14, abs=1e-15)
    # between nodes the interpolant is linear in 1/r
    k = 100
    r0, r1 = o["r"][k], o["r"][k + 1]
    x = 2.0 / (1 / r0 + 1 / r1)
    assert o.potential(x) == pytest.approx(0.5 * (o["phi"][k] + o["phi"][k + 1]), rel=1e-12)


# ------------------------------------------------------------------ apsides
def _q_all(O, o, j, E, J):
    return np.array([O.lib.orc_function_q(C.byref(o.s), j, float(o["r"][k]), float(o["phi"][k]), E, J)
                     for k in range(o.N + 2)])


@pytest.mark.parametrize("qmode", [0, 1])
def test_apsides_indices_bruteforce(cm, O, qmode):
    """kmin/kmax of calc_orbit_new equal the last negative-Q index before and after
    a positive-Q index, found by evaluating Q at every node."""
    cl = small_cluster(cm, 300, seed=5)
    o = make_oracle(O, cl, q_mode=qmode)
    o.init()
    checked = 0
    for j in range(1, o.N + 1, 7):
        E = o["E"][j] + o["m"][j] / o["r"][j] / o.s.N_STAR
        J = o["J"][j]
        if E >= 0:
            continue
        q = _q_all(O, o, j, E, J)
        res = o.apsides(j, E, J)
        if res["circular"]:
            continue
        pos = np.nonzero(q > 0)[0]
        kt = j if q[j] > 0 else pos[0]
        kmin = max(k for k in range(0, kt + 1) if q[k] < 0)
        kmax = max(k for k in range(kt, o.N + 2) if q[k] > 0)
        assert (res["kmin"], res["kmax"]) == (kmin, kmax), j
        assert o["r"][kmin] <= res["rp"] <= o["r"][kmin + 1]
        assert o["r"][kmax] <= res["ra"] <= o["r"][kmax + 1]
        checked += 1
    assert checked > 20


def test_apsides_kepler(O):
    """A dominant central mass: the apsides are the Kepler ones, a(1 -+ e)."""
    n = 200
    rng = np.random.default_rng(9)
    r = np.sort(rng.uniform(0.5, 3.0, n))
    mc = 1.0e9                                    # central mass in stored units (x madhoc)
    arr = {"r": np.concatenate([[1e-20], r, [1e10]]), "m": np.concatenate([[0], np.full(n, 1e-9), [0]])}
    o = O.Oracle(arr, n, n_star=n)
    o.s.cenma_m = mc
    o.potential_calculate()
    M = mc / n
    for j in range(5, n, 17):
        rj = o["r"][j]
        vc = np.sqrt(M / rj)
        vr, vt = 0.3 * vc, 0.9 * vc
        E = 0.5 * (vr * vr + vt * vt) - M / rj
        J = rj * vt
        a = -M / (2 * E)
        e = np.sqrt(1 + 2 * E * J * J / (M * M))
        Eo = o["phi"][j] + 0.5 * (vr * vr + vt * vt) + o["m"][j] / rj / n
        res = o.apsides(j, Eo, J)
        assert res["rp"] == pytest.approx(a * (1 - e), rel=1e-6)
        if a * (1 + e) < r[-1]:
            assert res["ra"] == pytest.approx(a * (1 + e), rel=1e-6)


def test_longdouble_vs_fp64_index_flips_are_rare(cm, O):
    """The reference evaluates Q in x87 long double; the device evaluates it in FP64.
    The kmin/kmax they select differ only where Q at a node is within rounding of 0,
    and the apsides then still agree closely."""
    cl = cm.make_king(20000, 7.0, seed=3)
    o = make_oracle(O, cl)
    o.init()
    flips = 0
    worst = 0.0
    tot = 0
    for j in range(1, o.N + 1, 3):
        E = o["E"][j] + o["m"][j] / o["r"][j] / o.s.N_STAR
        if E >= 0:
            continue
        J = o["J"][j]
        o.s.q_mode = 0
        a = o.apsides(j, E, J)
        o.s.q_mode = 1
        b = o.apsides(j, E, J)
        tot += 1
        if (a["kmin"], a["kmax"]) != (b["kmin"], b["kmax"]):
            flips += 1
            continue
        if not a["circular"]:
            worst = max(worst, abs(a["rp"] / b["rp"] - 1), abs(a["ra"] / b["ra"] - 1))
    assert flips <= max(2, tot // 1000), (flips, tot)
    assert worst < 1e-12


# ------------------------------------------------------------------ find_root_vr
def _exact_interval_root(o, j, k, E, J, dps=50, with_conditioning=False):
    """Root of v_r^2(r) = Q(r) on [r_k, r_{k+1}] in 50-digit arithmetic, from the
    double inputs: inside one interval the potential is the Henon interpolant
    phi_k + (phi_{k+1} - phi_k)(1/r_k - 1/r)/(1/r_k - 1/r_{k+1}) (phi_0 - cenma/r below
    r_1, cmc_nr.c:239-269) and the star's own term is m_j/r or m_j/r_j (r_j is a
    node), so Q is a quadratic in u = 1/r, Q = A + B u + C u^2, solved exactly."""
    mp = mpmath.mp
    mp.dps = dps
    f = lambda x: mpmath.mpf(float(x))
    N_STAR = o.s.N_STAR
    rk, rk1 = f(o["r"][k]), f(o["r"][k + 1])
    rj, mj = f(o["r"][j]), f(o["m"][j])
    if k == 0:
        # pot = phi_0 - cenma madhoc / r  -> constant part phi_0, 1/r part -cenma/N
        p0, p1 = f(o["phi"][0]), -f(o.s.cenma_m) / N_STAR
    else:
        pk, pk1 = f(o["phi"][k]), f(o["phi"][k + 1])
        slope = (pk1 - pk) / (1 / rk - 1 / rk1)         # pot = pk + slope (1/rk - u)
        p0, p1 = pk + slope / rk, -slope
    if k + 1 <= j:       # r <= r_{k+1} <= r_j: own term m_j / r_j
        s0, s1 = mj / rj / N_STAR, mpmath.mpf(0)
    else:                # r >= r_k >= r_j: own term m_j / r
        s0, s1 = mpmath.mpf(0), mj / N_STAR
    # Q(u) = 2(E - p0 - s0) - 2(p1 + s1) u - J^2 u^2
    A, B, Cq = 2 * (f(E) - p0 - s0), -2 * (p1 + s1), -f(J) ** 2
    disc = B * B - 4 * A * Cq
    assert disc >= 0, "no real root"
    roots = [(-B + sg * mpmath.sqrt(disc)) / (2 * Cq) for sg in (1, -1)] if Cq != 0 else [-A / B]
    rs = [1 / u for u in roots if u > 0 and rk <= 1 / u <= rk1]
    assert len(rs) == 1, (k, rs)
    if not with_conditioning:
        return rs[0]
    # how far rounding of the double evaluation of Q moves its sign change: each of
    # its terms is rounded a few times, |dQ| <~ 2^-48 (|2E| + |2 pot| + |2 self| +
    # (J/r)^2), and near the root dr = dQ / |dQ/dr|
    u = 1 / rs[0]
    dqdr = abs((B + 2 * Cq * u) * u * u)
    scale = 2 * abs(f(E)) + 2 * abs(p0 + p1 * u) + 2 * abs(s0 + s1 * u) + f(J) ** 2 * u * u
    return rs[0], float(scale * mpmath.mpf(2) ** -48 / dqdr) if dqdr > 0 else float("inf")


def _brent_tol(x):
    # gsl_root_test_interval(lo, hi, APSIDES_PRECISION, APSIDES_PRECISION) at
    # cmc_nr.c:540 with APSIDES_PRECISION = 1e-11: |hi - lo| < 1e-11 + 1e-11 min(|lo|, |hi|)
    return 1e-11 + 1e-11 * abs(x)


def test_find_root_vr_matches_exact_root(cm, O):
    """find_root_vr (cmc_nr.c:500-603) on the apsis intervals of ordinary stars: the
    Brent root equals the exact root of the interval's quadratic to the reference's
    convergence tolerance, and v_r^2 at the returned apsis is >= 0 (:587-600)."""
    cl = small_cluster(cm, 3000, seed=21, kroupa=True)
    o = make_oracle(O, cl, q_mode=1)
    o.init()
    checked = 0
    for j in range(3, o.N + 1, 97):
        E = o["E"][j] + o["m"][j] / o["r"][j] / o.s.N_STAR
        J = o["J"][j]
        res = o.apsides(j, E, J)
        if res["circular"]:
            continue
        for k in (res["kmin"], res["kmax"]):
            nb = o.s.n_brent
            got = O.lib.orc_find_root_vr(C.byref(o.s), j, k, E, J)
            assert o.s.n_brent == nb + 1 and o.s.error == 0
            want = _exact_interval_root(o, j, k, E, J)
            assert abs(got - float(want)) <= _brent_tol(float(want)), (j, k, got, want)
            q = O.lib.orc_function_q(C.byref(o.s), j, got, o.potential(got), E, J)
            assert q >= 0.0 or abs(got - o["r"][k]) < 2.3e-16 or abs(got - o["r"][k + 1]) < 2.3e-16
            checked += 1
        # the closed form of calc_orbit_new agrees with the exact root too
        # (a and b of set_a_b carry the cancellation of (U_k1 - U_k) / (1/r_k1 - 1/r_k))
        assert abs(res["rp"] - float(_exact_interval_root(o, j, res["kmin"], E, J))) <= 1e-10 * res["rp"]
    assert checked > 40


def test_find_root_vr_reached_by_near_tangent_orbits(cm, O):
    """calc_orbit_new reaches find_root_vr where rounding makes inside_sqrt < 0 and
    the double root -J^2/a falls outside the apsis interval (cmc_orbit.c:459-474,
    485-500): nearly circular orbits whose tangency sits on a node
    (cases.near_tangent_stars).  Each Brent apsis equals the exact root of the
    interval's quadratic to the reference's tolerance, or, where that quadratic has
    no real root, lies inside the interval with v_r^2 >= 0 there (:587-600)."""
    from cases import near_tangent_stars
    cl = small_cluster(cm, 5000, seed=22)
    o = make_oracle(O, cl, q_mode=1)
    o.init()
    js = near_tangent_stars(o, list(range(2, o.N - 1, 7)))
    n0 = o.s.n_brent
    brent_stars = exact_checked = tight = 0
    for j in js:
        E = o["E"][j] + o["m"][j] / o["r"][j] / o.s.N_STAR
        J = o["J"][j]
        nb = o.s.n_brent
        o.s.flags[j] = 0
        res = o.apsides(j, E, J)
        assert o.s.error == 0
        if o.s.n_brent == nb:
            continue
        brent_stars += 1
        assert o["flags"][j] & 16
        used = 0
        for key, k in (("rp", res["kmin"]), ("ra", res["kmax"])):
            if res["circular"]:
                used += 1
                break
            x = res[key]
            if x != O.lib.orc_find_root_vr(C.byref(o.s), j, k, E, J):
                continue                               # this apsis came from the closed form
            used += 1
            assert o["r"][k] - 2.3e-16 <= x <= o["r"][k + 1] + 2.3e-16
            try:
                want, dr = _exact_interval_root(o, j, k, E, J, with_conditioning=True)
            except AssertionError:
                continue                               # the quadratic has no real root here
            want = float(want)
            # a near-double root is ill-conditioned: the bound grows as 1/|dQ/dr|
            assert abs(x - want) <= _brent_tol(want) + dr, (j, key, x, want, dr)
            exact_checked += 1
            tight += dr < 0.05 * (o["r"][k + 1] - o["r"][k])   # the check still beats the interval
        assert used >= 1
    assert brent_stars > 100 and o.s.n_brent - n0 >= brent_stars
    assert exact_checked > 50 and tight > exact_checked // 2


def test_find_root_vr_catches_a_loose_solver(cm, O):
    """The exact-root check is tight enough to reject a solver that stops early:
    the interval midpoint, or the end of a bracket 1e-9 wide, fails it."""
    cl = small_cluster(cm, 3000, seed=21, kroupa=True)
    o = make_oracle(O, cl, q_mode=1)
    o.init()
    j = 1500
    E = o["E"][j] + o["m"][j] / o["r"][j] / o.s.N_STAR
    J = o["J"][j]
    res = o.apsides(j, E, J)
    k = res["kmax"]
    want = float(_exact_interval_root(o, j, k, E, J))
    mid = 0.5 * (o["r"][k] + o["r"][k + 1])
    assert abs(mid - want) > _brent_tol(want)
    assert abs((want + 1e-9 * want) - want) > _brent_tol(want)


# ------------------------------------------------------------------ relaxation
def test_relaxation_pair_invariants(cm, O):
    cl = cm.make_plummer(2000, seed=6, kroupa=True)
    o = make_oracle(O, cl)
    o.init()
    vr0, vt0, m = o["vr"].copy(), o["vt"].copy(), o["m"].copy()
    dt = O.lib.orc_simul_relax_new(C.byref(o.s))
    o.relax(dt)
    assert o.s.error == 0
    N = o.N
    for k in range(1, N - N % 2, 2):
        kp = k + 1
        ke0 = m[k] * (vr0[k] ** 2 + vt0[k] ** 2) + m[kp] * (vr0[kp] ** 2 + vt0[kp] ** 2)
        ke1 = m[k] * (o["vr"][k] ** 2 + o["vt"][k] ** 2) + m[kp] * (o["vr"][kp] ** 2 + o["vt"][kp] ** 2)
        assert ke1 == pytest.approx(ke0, rel=1e-12)
        assert o["Y"][k] == o["Y"][kp]
    E = o["phi"][1:N] + 0.5 * (o["vr"][1:N] ** 2 + o["vt"][1:N] ** 2)
    assert np.allclose(o["E"][1:N], E, rtol=1e-15, atol=0)
    assert np.allclose(o["J"][1:N], o["r"][1:N] * o["vt"][1:N], rtol=1e-15, atol=0)
    if N % 2:
        assert o["vr"][N] == vr0[N]                # unpaired last star untouched


def _kick_setup(n, seed, unequal):
    """A shell of n stars between r = 1 and 2 with random velocities (and masses)."""
    rng = np.random.default_rng(seed)
    r = np.sort(1.0 + rng.random(n))
    vr, vt = rng.normal(size=n) * 0.3, np.abs(rng.normal(size=n)) * 0.3 + 0.05
    m = rng.uniform(0.2, 3.0, n) if unequal else np.ones(n)
    m *= n / m.sum()                                 # stored masses are m * N_STAR
    pad = lambda x, a, b: np.concatenate([[a], x, [b]])
    return {"r": pad(r, 1e-20, 1e10), "m": pad(m, 0.0, 0.0), "vr": pad(vr, 0.0, 0.0), "vt": pad(vt, 0.0, 0.0)}


def _applied_angles(arr, o, k, kp):
    """The angle by which the oracle turned the pair's relative velocity w, recovered
    from its outputs alone: w_new's radial part is vr'_kp - vr'_k; its transverse part
    follows from the new transverse speeds and the conserved centre-of-mass velocity
    up to a reflection about that velocity, so both candidates are returned (with the
    two angles, |w_new| and the conditioning of the reflection)."""
    m_k, m_kp = arr["m"][k], arr["m"][kp]
    mu, mup = m_k / (m_k + m_kp), m_kp / (m_k + m_kp)
    ph = o["Y"][k] * 2.0 * np.pi
    vt_k, vr_k, vt_kp, vr_kp = arr["vt"][k], arr["vr"][k], arr["vt"][kp], arr["vr"][kp]
    v_t, vp_t = np.array([vt_k, 0.0]), np.array([vt_kp * np.cos(ph), vt_kp * np.sin(ph)])
    w = np.array([vp_t[0] - v_t[0], vp_t[1] - v_t[1], vr_kp - vr_k])
    W = np.linalg.norm(w)
    wn3 = o["vr"][kp] - o["vr"][k]
    c = mu * v_t + mup * vp_t                        # centre-of-mass transverse velocity
    wt2 = W * W - wn3 * wn3                          # |w_new_t|^2 if |w_new| = W
    cw = (o["vt"][kp] ** 2 - o["vt"][k] ** 2 - (mu - mup) * wt2) / 2.0
    c2 = c @ c
    a = cw / np.sqrt(c2)
    b2 = wt2 - a * a
    chat = c / np.sqrt(c2)
    cperp = np.array([-chat[1], chat[0]])
    out = []
    for sgn in (1.0, -1.0):
        wt = a * chat + sgn * np.sqrt(max(b2, 0.0)) * cperp
        wn = np.array([wt[0], wt[1], wn3])
        out.append(np.arctan2(np.linalg.norm(np.cross(w, wn)), w @ wn))
    # |w_new| from the first new speed (an independent check of |w_new| = |w|)
    wt_from_k = c - mup * (a * chat + np.sqrt(max(b2, 0.0)) * cperp)
    return out, W, b2 / max(wt2, 1e-300), (o["vt"][k], np.linalg.norm(wt_from_k))


def _beta_expected(arr, n, k, kp, W, dt, p=20):
    """beta = (pi/2) sqrt(dt/Trel12), Trel12 = (pi/32) W^3 / (N_STAR n_local
    ((m_k + m_kp) madhoc)^2), n_local = 2p (3/4pi) / (r_kmax^3 - r_kmin^3) over the
    window k-p .. k+p of calc_n_local (cmc_dynamics.c:429-430,444-446,
    cmc_dynamics_helper.c:768-786)."""
    kmin, kmax = k - p, k + p
    if kmin < 0:
        kmin, kmax = 0, 2 * p + 1
    elif kmax > n:
        kmin, kmax = n - 2 * p - 1, n
    r = arr["r"]
    nloc = 2.0 * p * 3.0 / (4.0 * np.pi * (r[kmax] ** 3 - r[kmin] ** 3))
    mm = (arr["m"][k] + arr["m"][kp]) / n
    trel = (np.pi / 32.0) * W ** 3 / (n * nloc * mm * mm)
    return min(np.pi / 2, (np.pi / 2) * np.sqrt(dt / trel))


@pytest.mark.parametrize("unequal", [False, True])
def test_relaxation_kick_angle(O, unequal):
    """The oracle's relaxation turns each pair's relative velocity w by the angle
    beta of cmc_dynamics.c:429-446 and keeps |w| and the pair momentum: the angle is
    measured between w and the w_new recovered from the oracle's output velocities,
    and compared with beta computed here from the pair's W, masses and the density
    window.  Then beta^2 is proportional to dt (same draws, dt x 4 doubles every
    unclamped angle) and a very long dt clamps every angle at pi/2 exactly."""
    n = 400
    arr = _kick_setup(n, 7 + unequal, unequal)
    runs = {}
    for dt in (2e-4, 8e-4, 1e6):
        o = O.Oracle(arr, n)
        o.potential_calculate()
        o.relax(dt)
        assert o.s.error == 0
        runs[dt] = o
    checked = 0
    ratios = []
    for k in range(1, n, 2):
        kp = k + 1
        m_k, m_kp = arr["m"][k], arr["m"][kp]
        # radial momentum of the pair (exactly conserved up to rounding)
        for o in runs.values():
            p0 = m_k * arr["vr"][k] + m_kp * arr["vr"][kp]
            p1 = m_k * o["vr"][k] + m_kp * o["vr"][kp]
            assert p1 == pytest.approx(p0, rel=1e-12, abs=1e-13 * (m_k + m_kp))
        got = {}
        for dt, o in runs.items():
            cands, W, cond, (vtk_out, vtk_rec) = _applied_angles(arr, o, k, kp)
            beta = _beta_expected(arr, n, k, kp, W, dt)
            if cond < 1e-4:
                continue                              # reflection ill-conditioned: skip the pair
            err = min(abs(c - beta) for c in cands)
            assert err <= 1e-9 * max(1.0, beta), (k, dt, cands, beta)
            got[dt] = beta
        if len(got) == 3:
            checked += 1
            assert got[1e6] == np.pi / 2
            if got[8e-4] < np.pi / 2:
                ratios.append(got[8e-4] / got[2e-4])
    assert checked > n // 4
    assert len(ratios) > 10 and np.allclose(ratios, 2.0, rtol=1e-12)
    # the clamp: w_new is perpendicular to w for every pair
    o = runs[1e6]
    for k in range(1, n, 2):
        cands, W, cond, _ = _applied_angles(arr, o, k, k + 1)
        if cond >= 1e-4:
            assert min(abs(c - np.pi / 2) for c in cands) <= 1e-9


def test_relaxation_kick_angle_catches_a_wrong_constant(O):
    """The check above fails for a beta that is off by a plausible mistake (Trel12
    without the pi/32, or sqrt(dt/Trel) replaced by dt/Trel): the recovered angles
    match only the correct formula."""
    n = 400
    arr = _kick_setup(n, 9, True)
    o = O.Oracle(arr, n)
    o.potential_calculate()
    dt = 2e-4
    o.relax(dt)
    miss_const = miss_power = 0
    for k in range(1, n, 2):
        cands, W, cond, _ = _applied_angles(arr, o, k, k + 1)
        if cond < 1e-4:
            continue
        beta = _beta_expected(arr, n, k, k + 1, W, dt)
        trel = (np.pi / 2) ** 2 * dt / beta ** 2 if beta < np.pi / 2 else None
        if trel is None:
            continue
        wrong_const = min(np.pi / 2, (np.pi / 2) * np.sqrt(dt / (trel * 32.0 / np.pi)))
        wrong_power = min(np.pi / 2, (np.pi / 2) * dt / trel)
        miss_const += min(abs(c - wrong_const) for c in cands) > 1e-6
        miss_power += min(abs(c - wrong_power) for c in cands) > 1e-6
    assert miss_const > 50 and miss_power > 50


# ------------------------------------------------------------------ sampling
def test_sampling_density_ks(cm, O):
    """Accepted radii of one star over many independent draws follow the orbital
    time density ~ 1/|v_r(r)| on [r_p, r_a] (cmc_evolution_thr.c:857-888)."""
    cl = small_cluster(cm, 200, seed=8)
    o = make_oracle(O, cl, q_mode=1)
    o.init()
    j = 120
    E = o["E"][j] + o["m"][j] / o["r"][j] / o.s.N_STAR
    J = o["J"][j]
    ap = o.apsides(j, E, J)
    assert not ap["circular"]
    samples = []
    for t in range(2500):
        o.s.step = t
        o.orbits()
        samples.append(o["rnew"][j])
    rmin, rmax = ap["rp"], ap["ra"]
    s = np.linspace(-1, 1, 20001)[1:-1]
    rr = 0.5 * (rmin + rmax) + 0.25 * (rmax - rmin) * (3 * s - s ** 3)
    drds = 0.25 * (rmax - rmin) * (3 - 3 * s * s)
    pot = np.array([o.potential(x) for x in rr]) + o["m"][j] / np.maximum(rr, o["r"][j]) / o.s.N_STAR
    Q = 2 * E - 2 * pot - J * J / rr / rr
    dens = drds / np.sqrt(np.maximum(Q, 1e-300))
    cdf = np.concatenate([[0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(s))])
    cdf /= cdf[-1]
    res = stats.kstest(samples, lambda x: np.interp(x, rr, cdf))
    assert res.pvalue > 1e-3, res


# ------------------------------------------------------------------ sort
def test_sort_matches_lexsort(O):
    rng = np.random.default_rng(12)
    n = 5000
    r = rng.choice(rng.random(3000) * 5, n)         # many ties
    r[rng.choice(n, 40, replace=False)] = 1e10       # stripped stars
    ids = rng.permutation(n) + 1
    arr = {"r": np.concatenate([[1e-20], r, [1e10]]), "m": np.concatenate([[0], rng.random(n), [0]]),
           "id": np.concatenate([[0], ids, [0]]), "vr": np.concatenate([[0], np.arange(n, dtype=float), [0]])}
    o = O.Oracle(arr, n)
    nn = o.sort()
    order = np.lexsort((ids, r))
    live = order[r[order] < 1e10]
    assert nn == live.size == n - 40
    assert np.array_equal(o["id"][1:nn + 1], ids[live])
    assert np.array_equal(o["vr"][1:nn + 1], np.arange(n, dtype=float)[live])


# ------------------------------------------------------------------ timestep and sigma
def test_timestep_and_sigma_direct(cm, O):
    cl = cm.make_plummer(3000, seed=13, kroupa=True)
    o = make_oracle(O, cl)
    o.init()
    N, p = o.N, 20
    r, m, vr, vt = o["r"], o["m"] / N, o["vr"], o["vt"]
    best = np.inf
    for si in range(1 + p, N - p, 2 * p):
        k = np.arange(si - p, si + p)
        Mv2 = np.sum(m[k] * (vr[k] ** 2 + vt[k] ** 2)) / (2 * p)
        M = np.sum(m[k]) / (2 * p)
        M2 = np.sum(m[k] ** 2) / (2 * p)
        W = 4 * np.sqrt(Mv2 / M) / np.sqrt(3 * np.pi)
        nl = (2.0 * p) * 3 / (4 * np.pi * (r[si + p] ** 3 - r[si - p] ** 3))
        dt = (2 * 1.412 / np.pi) ** 2 * (np.pi / 32) * W ** 3 / (N * nl * 4 * M2)
        best = min(best, dt)
    assert O.lib.orc_simul_relax_new(C.byref(o.s)) == pytest.approx(best, rel=1e-12)
    O.lib.orc_calc_sigma_r(C.byref(o.s))
    for si in [1, 5, 21, 500, N - 3, N]:
        lo, hi = si - p, si + p - 1
        if lo < 1:
            lo, hi = 1, 2 * p
        elif hi > N:
            lo, hi = N - 2 * p + 1, N
        k = np.arange(lo, hi + 1)
        sig = np.sqrt(np.sum(m[k] * (vr[k] ** 2 + vt[k] ** 2)) / np.sum(m[k]))
        assert o["sigma"][si] == pytest.approx(sig, rel=1e-11)


# ------------------------------------------------------------------ whole step
def test_step_energy_conservation(cm, O):
    """Stodolkiewicz's scheme keeps Etot (cmc_io.c:2347 definition) nearly constant."""
    cl = cm.make_plummer(20000, seed=1)
    o = make_oracle(O, cl)
    o.init()
    e0 = o.sc("Etot")
    for _ in range(3):
        assert o.step() == 0
    assert abs((o.sc("Etot") + o.sc("Eoops") - e0) / e0) < 2e-5


def test_per_star_streams_match_reference_sequence_statistically(cm, O):
    """The device's random-number layout (one substream per pair and per star) against
    the reference's single sequential stream (REF_SEQ): after three steps of the same
    Plummer 1e5 cluster the distributions of r, v_r and v_t agree (two-sample KS),
    and so do the energy changes."""
    cl = cm.make_plummer(100000, seed=1)
    runs = {}
    for mode in (O.RNG_REF_SEQ, O.RNG_PER_STAR):
        o = make_oracle(O, cl, rng_mode=mode, q_mode=O.Q_FP64)
        o.init()
        e0 = o.sc("Etot")
        for _ in range(3):
            assert o.step() == 0
        runs[mode] = (o, (o.sc("Etot") + o.sc("Eoops") - e0) / abs(e0))
    a, da = runs[O.RNG_REF_SEQ]
    b, db = runs[O.RNG_PER_STAR]
    for k in ("r", "vr", "vt"):
        x, y = a[k][1:a.N + 1], b[k][1:b.N + 1]
        res = stats.ks_2samp(x, y)
        assert res.pvalue > 1e-3, (k, res)
        # and the two runs are different realisations, not the same draws
        assert not np.array_equal(np.sort(x), np.sort(y))
    assert abs(da - db) < 1e-4, (da, db)


def test_set_velocities3_excess_chain(O):
    """Hand-built case for the excess-energy chain (cmc_utils.c:268-310): star 2 has
    vnew2 <= 0 and its excess is dumped into star 3 (capacity 0.8 * KE)."""
    n = 4
    arr = {"r": np.array([1e-20, 1.0, 2.0, 3.0, 4.0, 1e10]), "m": np.array([0, 1.0, 1.0, 1.0, 1.0, 0]),
           "vr": np.array([0, 0.1, 0.2, 0.3, 0.1, 0]), "vt": np.array([0, 0.1, 0.1, 0.1, 0.1, 0])}
    o = O.Oracle(arr, n)
    o.potential_calculate()
    for k in ("rOld", "vrold", "vtold", "Uoldrold", "Uoldrnew"):
        pass
    o["rOld"][1:5] = o["r"][1:5]
    o["vrold"][1:5] = o["vr"][1:5]
    o["vtold"][1:5] = o["vt"][1:5]
    o["Uoldrold"][1:5] = 0.0
    o["Uoldrnew"][1:5] = 0.0
    o["Uoldrnew"][2] = 1.0                      # makes vnew2 of star 2 negative
    vr0, vt0 = o["vr"].copy(), o["vt"].copy()
    O.lib.orc_set_velocities3(C.byref(o.s))
    # star 1: unchanged (alpha = 1, no excess yet)
    assert o["vr"][1] == pytest.approx(vr0[1], rel=1e-15)
    # star 2 keeps its velocity; excess = 0.5 (v2^2 - vnew2) m
    vnew2 = vr0[2] ** 2 + vt0[2] ** 2 + 2 * 0.5 * (0.0 - 1.0)
    exc = 0.5 * (vr0[2] ** 2 + vt0[2] ** 2 - vnew2)
    assert o["vr"][2] == vr0[2]
    v2 = vr0[3] ** 2 + vt0[3] ** 2
    dump = min(0.8 * 0.5 * v2, exc)
    ratio = np.sqrt((v2 - 2 * dump) / v2)
    assert o["vr"][3] == pytest.approx(vr0[3] * ratio, rel=1e-13)
    left = exc - dump
    v2b = vr0[4] ** 2 + vt0[4] ** 2
    dump2 = min(0.8 * 0.5 * v2b, left)
    assert o["vt"][4] == pytest.approx(vt0[4] * np.sqrt((v2b - 2 * dump2) / v2b), rel=1e-13)


# ------------------------------------------------------------------ generators
def test_generators_units(cm, O):
    cl = cm.make_plummer(20000, seed=1)
    o = make_oracle(O, cl)
    o.init()
    assert o.sc("Epot") == pytest.approx(-0.5, rel=1e-6)
    assert o.sc("Ekin") == pytest.approx(0.25, rel=1e-12)
    k = cm.make_king(20000, 7.0, seed=1, kroupa=False, segregate=False)
    o = make_oracle(O, k)
    o.init()
    assert o.sc("Epot") == pytest.approx(-0.5, rel=1e-6)
    E = o["E"][1:o.N + 1]
    assert (E < 0).all()
    ks = cm.make_king(20000, 7.0, seed=1)
    o = make_oracle(O, ks)
    o.init()
    assert (o["E"][1:o.N + 1] < 0).all()
    # heaviest 1% in the inner 10%
    m = ks["m"][1:ks.n + 1]
    heavy = np.argsort(-m)[: ks.n // 100]
    assert heavy.max() < ks.n // 10


def test_kroupa_imf(cm):
    m = cm.kroupa_masses(200000, seed=4)
    assert m.min() >= 0.08 and m.max() <= 150.0
    w1 = (0.08 ** -0.3 - 0.5 ** -0.3) / 0.3
    w2 = 0.5 * (0.5 ** -1.3 - 150 ** -1.3) / 1.3
    assert np.mean(m < 0.5) == pytest.approx(w1 / (w1 + w2), abs=0.005)
    hi = m[m > 1.0]
    # power-law slope of the upper segment by maximum likelihood: alpha = 1 + n / sum ln(m/mmin)
    alpha = 1 + hi.size / np.sum(np.log(hi / 1.0))
    assert alpha == pytest.approx(2.3, abs=0.05)
