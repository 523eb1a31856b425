"""The reference's MPI layout on P threads (oracle/orc_mpi.c), the CPU baseline that is
timed beside the device: at one rank it is the serial step in the reference's draw
order bit for bit; at several ranks each rank draws from the seeded stream jumped
rank * 2^80 with the reference's own jump polynomial, the sample sort gives the
serial (r, id) order, and the evolution agrees statistically with the serial one."""
import ctypes as C
import json
import os

import numpy as np
import pytest
from scipy import stats

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "rng_golden.json")))
FIELDS = ("r", "m", "phi", "vr", "vt", "E", "J", "id")


def _oracle(O, cl, **kw):
    kw.setdefault("rng_mode", O.RNG_REF_SEQ)
    kw.setdefault("q_mode", O.Q_FP64)
    return O.Oracle(cl.a, cl.n, orbit_r=cl.rtid, **kw)


def test_find_limits_matches_reference_partition(O):
    """findLimits (cmc_utils.c:2299-2333): chunks in multiples of 20, the first
    blocks % P ranks one block larger, the remainder to the last rank."""
    for N, P in [(100000, 7), (20001, 4), (45, 3), (1000, 16)]:
        st, en = (C.c_int64 * P)(), (C.c_int64 * P)()
        O.lib.orc_find_limits.argtypes = [C.c_int64, C.c_int64, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        O.lib.orc_find_limits(N, 20, P, st, en)
        blocks, rem = divmod(N, 20)
        sizes = [en[i] - st[i] + 1 for i in range(P)]
        assert st[0] == 1 and en[P - 1] == N
        assert all(st[i + 1] == en[i] + 1 for i in range(P - 1))
        want = [(blocks // P + (1 if i < blocks % P else 0)) * 20 for i in range(P)]
        want[-1] += rem
        assert sizes == want, (N, P)


def test_one_rank_is_the_serial_step(cm, O):
    cl = cm.make_plummer(20000, seed=3, kroupa=True)
    a = _oracle(O, cl)
    a.init()
    b = _oracle(O, cl)
    b.init()
    R = b.ranks(1)
    for _ in range(3):
        assert a.step() == 0
        assert R.step() == 0
    assert a.N == b.N
    for k in FIELDS:
        assert np.array_equal(a[k][1:a.N + 1], b[k][1:b.N + 1]), k
    for k in ("Etot", "Eescaped", "TidalMassLoss", "max_r", "Dt"):
        assert a.sc(k) == b.sc(k), k


def test_rank_streams_are_the_reference_2p80_jumps(cm, O):
    """Rank i's first relaxation draw (Y of the first pair of its chunk, step 0) is
    the first draw of the seeded state advanced by i applications of the reference's
    JPoly_2_80 jump (taus113-v2.c:74, set_rng_states cmc_utils.c:2364-2374)."""
    P = 4
    cl = cm.make_plummer(40000, seed=5)
    o = _oracle(O, cl, seed=1234)
    o.init()
    R = o.ranks(P)
    st, en = (C.c_int64 * P)(), (C.c_int64 * P)()
    O.lib.orc_find_limits.argtypes = [C.c_int64, C.c_int64, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    O.lib.orc_find_limits(o.N, 20, P, st, en)
    assert R.step() == 0
    jp = (C.c_uint32 * 4)(*GOLD["jpoly_2_80"])
    cur = O.seed_state(1234)
    for i in range(P):
        probe = O.T113(); probe.z[:] = cur.z[:]
        y = O.lib.orc_t113_dbl(C.byref(probe))
        # the first pair of rank i's chunk: its Y is the rank's first draw.  The sort
        # moved the stars, so find that pair by the draw itself among all Y values.
        assert np.any(o["Y"][1:o.N + 1] == y), i
        cur = O.lib.orc_t113_jump_poly(cur, jp)


def test_sample_sort_matches_serial_sort(O):
    rng = np.random.default_rng(12)
    n = 60000
    r = rng.choice(rng.random(9000) * 5, n)           # many ties
    r[rng.choice(n, 300, replace=False)] = 1e10
    ids = rng.permutation(n) + 1
    arr = {"r": np.concatenate([[1e-20], r, [1e10]]), "m": np.concatenate([[0], rng.random(n), [0]]),
           "id": np.concatenate([[0], ids, [0]]), "vr": np.concatenate([[0], np.arange(n, dtype=float), [0]])}
    for P in (2, 3, 8):
        a, b = O.Oracle(arr, n), O.Oracle(arr, n)
        na = a.sort()
        nb = O.lib.orc_sort_by_r_ranks(C.byref(b.s), P)
        assert na == nb == n - 300
        for k in ("r", "id", "m", "vr"):
            assert np.array_equal(a[k][1:na + 1], b[k][1:nb + 1]), (P, k)


def test_several_ranks_agree_statistically(cm, O):
    """Eight ranks against one: different draws, the same physics (two-sample KS of
    r, v_r, v_t after three steps of Plummer 1e5; energy changes within 1e-4)."""
    cl = cm.make_plummer(100000, seed=1)
    runs = {}
    for P in (1, 8):
        o = _oracle(O, cl)
        o.init()
        e0 = o.sc("Etot")
        R = o.ranks(P)
        for _ in range(3):
            assert R.step() == 0
        runs[P] = (o, (o.sc("Etot") + o.sc("Eoops") - e0) / abs(e0))
    a, da = runs[1]
    b, db = runs[8]
    for k in ("r", "vr", "vt"):
        x, y = a[k][1:a.N + 1], b[k][1:b.N + 1]
        assert stats.ks_2samp(x, y).pvalue > 1e-3, k
        assert not np.array_equal(x, y)
    assert abs(da - db) < 1e-4
