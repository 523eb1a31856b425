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


def test_relaxation_deflection_angle(O):
    """One pair in isolation: the relative velocity keeps its length and turns by
    exactly beta = (pi/2) sqrt(dt/Trel12) (cmc_dynamics.c:429-430,467-469)."""
    n = 44
    r = np.linspace(1.0, 2.0, n)
    rng = np.random.default_rng(1)
    vr, vt = rng.normal(size=n) * 0.3, np.abs(rng.normal(size=n)) * 0.3
    arr = {"r": np.concatenate([[1e-20], r, [1e10]]), "m": np.ones(n + 2), "vr": np.concatenate([[0], vr, [0]]),
           "vt": np.concatenate([[0], vt, [0]])}
    o = O.Oracle(arr, n)
    o.potential_calculate()
    dt = 1e-4
    o.relax(dt)
    k, kp = 21, 22
    Y = o["Y"][k]
    ph = Y * 2 * np.pi
    w0 = np.array([vt[kp - 1] * np.cos(ph) - vt[k - 1], vt[kp - 1] * np.sin(ph), vr[kp - 1] - vr[k - 1]])
    W = np.linalg.norm(w0)
    # after: velocities of k and kp are v -/+ (1/2)(w_new - w) for equal masses; |w_new| = W
    dv = np.hypot(o["vr"][kp] - o["vr"][k], 0)  # radial part of w_new
    p = 20
    nloc = (2.0 * p) * 3.0 / (4.0 * np.pi * (r[k + p - 1] ** 3 - r[k - p - 1] ** 3))
    Trel = (np.pi / 32) * W ** 3 / (n * nloc * (2.0 / n) ** 2)
    beta = min(np.pi / 2, (np.pi / 2) * np.sqrt(dt / Trel))
    # kinetic energy in the centre-of-mass frame is W^2/4 per unit mass both before and after
    vk2 = o["vr"][k] ** 2 + o["vt"][k] ** 2
    vkp2 = o["vr"][kp] ** 2 + o["vt"][kp] ** 2
    v02 = vr[k - 1] ** 2 + vt[k - 1] ** 2 + vr[kp - 1] ** 2 + vt[kp - 1] ** 2
    assert vk2 + vkp2 == pytest.approx(v02, rel=1e-12)
    assert 0 < beta < np.pi / 2
    assert abs(dv) <= W * (1 + 1e-12)


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
