"""central_calculate + clusdyn_calculate on the device against the oracle
(cmc_utils.c:1239-1345, 1614-1624): nave, rh and every core quantity bit-equal
(the six density-weighted sums are the reference's sequential recurrences, and the
derived values use the same correctly rounded divisions and square roots).

Cases: Plummer 1e5 with a Kroupa spectrum and King W0=7 1e6 (segregated) straight
from their generators; King 1e6 again with the chains' classification bound widened
by 2^60 (test hook), which sends most elements through the fallback paths; the
reference's CI input after 3 steps on both sides.
"""
import os
import sys

import pytest

pytestmark = pytest.mark.gpu

H5 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ci_input.hdf5")


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _compare(d, o, case):
    say(case, "nave", d["nave"], o["nave"], "rh", d["rh"], "rc", d["rc"], o["rc"], "rho", d["rho"], o["rho"])
    assert d["nave"] == o["nave"]
    for k, v in o.items():
        assert d[k] == v, (case, k, d[k], v)


@pytest.mark.parametrize("case,drop", [("plummer_kroupa_1e5", 0), ("king_1e6", 0), ("king_1e6", 60)])
def test_gpu_central_matches_oracle(gpu, O, case, drop):
    cm = gpu
    say("central_calculate", case, "chain scale drop", drop)
    cl = cm.make_plummer(100000, seed=4, kroupa=True) if case.startswith("plummer") else \
        cm.make_king(1000000, seed=2)
    o = O.Oracle(cl.a, cl.n, orbit_r=cl.rtid)
    o.init()
    with cm.Context(cl.n, cm.default_params(cl.n, orbit_r=cl.rtid)) as ctx:
        ctx.upload(cl)
        ctx.init_cluster()
        if drop:
            ctx.test_hook(cm.HOOK_CHAIN_K_DROP, drop)
        d = ctx.central_calculate()
    _compare(d, o.central_calculate(), case)


def test_gpu_central_after_steps_ci_input(gpu, O):
    cm = gpu
    if not os.path.exists(cm.IO_LIB_PATH):
        pytest.fail("libcmcio.so not built")
    cl = cm.read_hdf5(H5)
    say("central_calculate, reference CI input after 3 steps")
    o = O.Oracle(cl.a, cl.n, orbit_r=cl.rtid, rng_mode=O.RNG_PER_STAR, q_mode=O.Q_FP64, libm_mode=O.LIBM_DD)
    o.init()
    with cm.Context(cl.n, cm.default_params(cl.n, orbit_r=cl.rtid)) as ctx:
        ctx.upload(cl)
        ctx.init_cluster()
        for s in range(3):
            ctx.step()
            assert o.step() == 0, s
        d = ctx.central_calculate()
    _compare(d, o.central_calculate(), "ci_input_3_steps")


def test_gpu_dyn_line_through_reference_names(gpu, O):
    """The reference-named host entry points (central_calculate, compute_energy_new)
    and the .dyn.dat row they feed: the columns the hot path owns equal the oracle's
    values printed with the reference's %.8g."""
    import ctypes as C
    cm = gpu
    say("dyn.dat row through the host shim")
    cl = cm.make_plummer(30000, seed=4, kroupa=True)
    h = cm.hostlib
    p = cm.default_params(cl.n, orbit_r=cl.rtid, seed=1)
    assert h.cmc_hotpath_init(C.byref(p), cl.n, C.byref(cl.soa()), 0) == 0
    try:
        for f in ("calc_potential_new", "compute_energy_new", "central_calculate", "clusdyn_calculate"):
            getattr(h, f).restype = None
            getattr(h, f)()
        h.cmc_hotpath_dyn_header.restype = C.c_char_p
        h.cmc_hotpath_dyn_line.argtypes = [C.c_char_p, C.c_size_t]
        buf = C.create_string_buffer(2048)
        nch = h.cmc_hotpath_dyn_line(buf, 2048)
        line = buf.value.decode()
        head = h.cmc_hotpath_dyn_header().decode().splitlines()
    finally:
        h.cmc_hotpath_finalize()
    say(line.strip())
    assert nch == len(line) and line.endswith("\n")
    assert head[0] == "# Dynamical information [code units]" and head[1].count("#") == 28
    col = line.split()
    assert len(col) == 28
    o = O.Oracle(cl.a, cl.n, orbit_r=cl.rtid)
    o.init()
    c = o.central_calculate()
    g = "%.8g"
    assert int(col[3]) == cl.n
    # bit-exact quantities: the same printed digits
    want = {5: o.sc("Mtotal"), 7: c["N_core"], 8: c["rc"], 21: c["rh"], 22: c["rho"], 23: c["rc_spitzer"],
            24: c["v_rms"], 25: c["rc_nb"], 28: c["N_core_nb"]}
    for k, v in want.items():
        assert col[k - 1] == g % v, (k, col[k - 1], g % v)
    # energies are block-reduced on the device (within 1e-12 of the oracle's sums)
    for k, v in {6: -2.0 * o.sc("Ekin") / o.sc("Epot"), 10: o.sc("Etot"), 11: o.sc("Ekin"), 12: o.sc("Epot"),
                 20: o.sc("Etot") + o.sc("Eoops")}.items():
        assert float(col[k - 1]) == pytest.approx(v, rel=1e-7), k
    for k in (13, 14, 15, 17, 18, 26, 27):
        assert float(col[k - 1]) == 0.0


def test_gpu_central_reports_the_reference_exit(gpu):
    """N_STAR <= 2J (J = 6): the reference exits in central_calculate; the device
    returns CMCG_E_STATE and leaves the context usable."""
    cm = gpu
    cl = cm.make_plummer(12, seed=3)
    say("central_calculate with N = 12")
    with cm.Context(cl.n, cm.default_params(cl.n, orbit_r=cl.rtid)) as ctx:
        ctx.upload(cl)
        ctx.init_cluster()
        with pytest.raises(cm.CmcgError) as e:
            ctx.central_calculate()
        assert e.value.code == cm.E_STATE
        ctx.step()
