"""Constructed inputs shared by the oracle tests and the GPU tests."""
import numpy as np


def near_tangent_stars(o, js, rel_j=(0.0, 1e-16, 4e-16, 1e-15), vr_choices=(0.0, 1e-9, 1e-8)):
    """Turn stars js of an initialised oracle state into nearly circular orbits whose
    tangency sits on their own node: J = sqrt(-a r_j) makes -J^2/a (the double root
    that calc_orbit_new uses when inside_sqrt < 0, cmc_orbit.c:459-474) equal r_j for
    the interval just inside r_j, and a tiny v_r leaves Q at the node within rounding
    of zero.  Rounding then often gives inside_sqrt < 0 with -J^2/a just outside the
    interval, which is the reference's path into find_root_vr.  Writes vr, vt, E, J
    (E without the star's own term, as after set_star_EJ) and returns the stars."""
    N = o.N
    out = []
    for i, j in enumerate(js):
        rj = o["r"][j]
        k = j - 1
        rk, rk1 = o["r"][k], o["r"][k + 1]
        s = o["m"][j] / o.s.N_STAR
        Uk = o["phi"][k] + (s / rk if rk >= rj else s / rj)
        Uk1 = o["phi"][k + 1] + (s / rk1 if rk1 >= rj else s / rj)
        a = (Uk1 - Uk) / (1 / rk1 - 1 / rk)
        if not (a < 0) or j < 2 or j > N - 1:
            continue
        J = np.sqrt(-a * rj) * (1.0 + rel_j[i % len(rel_j)])
        vt = J / rj
        vr = vr_choices[(i // len(rel_j)) % len(vr_choices)] * vt
        o["vr"][j], o["vt"][j] = vr, vt
        o["E"][j] = o["phi"][j] + 0.5 * (vr * vr + vt * vt)
        o["J"][j] = rj * vt                       # set_star_EJ, cmc_dynamics_helper.c:878-882
        out.append(j)
    return out


def near_tangent_sampleable(O, cl, js, max_drop=64, **kw):
    """near_tangent_stars on a fresh oracle state of cl, leaving out the stars whose
    orbit the reference cannot sample (its get_positions stops with N_TRY exceeded on
    some of these degenerate orbits): returns the initialised oracle (orbit phase not
    run), the constructed stars and the ones left out."""
    drop = []
    for _ in range(max_drop):
        o = O.Oracle(cl.a, cl.n, orbit_r=cl.rtid, rng_mode=O.RNG_PER_STAR, q_mode=O.Q_FP64, **kw)
        o.init()
        used = near_tangent_stars(o, [j for j in js if j not in drop])
        probe = O.Oracle(cl.a, cl.n, orbit_r=cl.rtid, rng_mode=O.RNG_PER_STAR, q_mode=O.Q_FP64, **kw)
        probe.init()
        near_tangent_stars(probe, [j for j in js if j not in drop])
        probe.s.step = 0
        probe.orbits()
        if probe.s.error == 0:
            return o, used, drop
        drop.append(int(probe.s.error_index))
    raise RuntimeError("too many near-tangent stars the reference cannot sample")
