"""Pins the oracle's taus113 restatement (src/libs/taus113-v2.c) against the
reference's own known answers (tests/golden/rng_golden.json, written from the
reference file compiled from source) and against brute-force stepping."""
import ctypes as C
import json
import os

import numpy as np
import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "rng_golden.json")))


def _draw_ints(O, st, n):
    return [int(O.lib.orc_t113_int(C.byref(st))) for _ in range(n)]


def test_seed_states_and_draws(O):
    for seed, rec in GOLD["seeds"].items():
        st = O.seed_state(int(seed))
        assert st.tolist() == rec["state"], seed
        s2 = O.T113(); s2.z[:] = st.z[:]
        assert _draw_ints(O, s2, 16) == rec["ints"], seed
        s3 = O.T113(); s3.z[:] = st.z[:]
        dbls = [O.lib.orc_t113_dbl(C.byref(s3)) for _ in range(8)]
        assert dbls == rec["dbls"], seed


def test_reference_testjumping_known_answer(O):
    # taus113-v2.c:272-292: seed 123456, draw after 2^20 steps = 3511752030
    st = O.seed_state(123456)
    assert st.tolist() == [3168676977, 3290129958, 1237851892, 1596210003]
    jp = (C.c_uint32 * 4)(*GOLD["jpoly_2_20"])
    j = O.lib.orc_t113_jump_poly(st, jp)
    assert _draw_ints(O, j, 1)[0] == 3511752030 == GOLD["testjumping_123456_after_2p20"]
    m = O.lib.orc_t113_jump(st, 1 << 20, 0)
    assert _draw_ints(O, m, 1)[0] == 3511752030


@pytest.mark.parametrize("which,shift", [("jump20", 20), ("jump80", 80)])
def test_matrix_jump_matches_reference_jump_polynomials(O, which, shift):
    """The GF(2) matrix powers reproduce the reference's JPoly_2_20 / JPoly_2_80
    jumps draw for draw (the jumped states may differ in bits the generator never
    reads, so the comparison is on draws)."""
    for seed, rec in GOLD["seeds"].items():
        st = O.seed_state(int(seed))
        lo, hi = (1 << shift, 0) if shift < 64 else (0, 1 << (shift - 64))
        j = O.lib.orc_t113_jump(st, lo, hi)
        assert _draw_ints(O, j, 4) == rec[which + "_ints"], seed


def test_jump_equals_stepping_bruteforce(O):
    rng = np.random.default_rng(5)
    for seed in [1, 77, 99991]:
        base = O.seed_state(seed)
        for n in list(rng.integers(0, 5000, 6)) + [0, 1, 2, 3, 1023, 1024, 4097]:
            s = O.T113(); s.z[:] = base.z[:]
            for _ in range(int(n)):
                O.lib.orc_t113_int(C.byref(s))
            j = O.lib.orc_t113_jump(base, int(n), 0)
            assert _draw_ints(O, j, 3) == _draw_ints(O, s, 3), (seed, n)


def test_stream_layout_step0_relaxation_is_reference_sequence(O):
    """Pair p of step 0 starts at offset 3p: exactly the reference's sequential
    stream consumed 3 draws per pair (cmc_dynamics_helper.c:839, cmc_dynamics.c:406,465)."""
    seeded = O.seed_state(1234)
    seq = O.T113(); seq.z[:] = seeded.z[:]
    for p in range(50):
        st = O.lib.orc_stream_relax(C.byref(seeded), 0, p)
        assert _draw_ints(O, st, 3) == _draw_ints(O, seq, 3), p


def test_orbit_streams_disjoint_windows(O):
    """Orbit substreams are 2^20 apart, more than the 2*N_TRY+1 = 100001 draws a star can take."""
    seeded = O.seed_state(7)
    a = O.lib.orc_stream_orbit(C.byref(seeded), 3, 10)
    b = O.lib.orc_stream_orbit(C.byref(seeded), 3, 11)
    a2 = O.lib.orc_t113_jump(a, 1 << 20, 0)
    assert _draw_ints(O, a2, 4) == _draw_ints(O, b, 4)
    assert 2 * 50000 + 1 < (1 << 20)


def test_phases_draw_from_their_layout_substreams(cm, O):
    """The oracle's relaxation and orbit phases walk the substream layout (3 draws per
    pair, 2^20 per orbit slot) instead of jumping from the seed for every star; the
    draws they use are the layout's: Y of pair p is the first draw of
    orc_stream_relax(step, p), and X of a star accepted on its first try is the first
    draw of orc_stream_orbit(step, slot)."""
    cl = cm.make_plummer(3001, seed=4)
    o = O.Oracle(cl.a, cl.n, orbit_r=cl.rtid, rng_mode=O.RNG_PER_STAR, q_mode=O.Q_FP64, seed=99)
    o.init()
    o.s.step = 5
    dt = O.lib.orc_simul_relax_new(C.byref(o.s))
    o.relax(dt)
    seeded = O.seed_state(99)
    for p in list(range(0, 40)) + list(range(1400, 1500)):
        st = O.lib.orc_stream_relax(C.byref(seeded), 5, p)
        assert o["Y"][2 * p + 1] == O.lib.orc_t113_dbl(C.byref(st)), p
    o.s.step = 6
    o.orbits()
    first = 0
    for j in range(1, o.N + 1):
        if o["ntry"][j] == 1:
            st = O.lib.orc_stream_orbit(C.byref(seeded), 6, j)
            assert o["X"][j] == O.lib.orc_t113_dbl(C.byref(st)), j
            first += 1
    assert first > 1000
