"""Checkpoint and restart of the device state (the counterpart of the reference's
save_restart_file / load_restart_file, src/cmc/cmc_io.c:3829-4031): a run restored
from a checkpoint must continue bit for bit as the run that saved it."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("r", "m", "phi", "vr", "vt", "E", "J", "id", "rnew", "vrnew", "vtnew", "X", "r_peri", "r_apo")
SCALARS = ("n_max", "n_max_new", "Mtotal", "Rtidal", "max_r", "Dt", "TotalTime", "TidalMassLoss", "Eescaped",
           "Jescaped", "Etidal", "Etot", "Ekin", "Epot", "step")


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _same(a, b):
    n = a.n
    assert n == b.n
    for k in FIELDS:
        assert np.array_equal(a[k][: n + 2], b[k][: n + 2]), k


def test_gpu_checkpoint_restart_is_bit_exact(gpu):
    cm = gpu
    say("checkpoint: Kroupa Plummer 2e4, 3 steps, save, 4 more steps; restore in a new context, 4 steps")
    cl = cm.make_plummer(20000, seed=5, kroupa=True)
    params = cm.default_params(cl.n, orbit_r=cl.rtid, seed=9)
    with cm.Context(cl.n, params) as a:
        a.upload(cl)
        a.init_cluster()
        for _ in range(3):
            a.step()
        buf = a.checkpoint()
        p, n = cm.checkpoint_info(buf)
        assert n == a.scalars()["n_max_new"] and p.seed == 9 and p.n_star == cl.n
        for _ in range(4):
            a.step()
        sa, da = a.scalars(), a.download()
    with cm.Context(cl.n, params) as b:
        b.restore(buf)
        assert b.scalars()["step"] == 3
        for _ in range(4):
            b.step()
        sb, db = b.scalars(), b.download()
    say("N after 7 steps", sa["n_max"], "Etot", sa["Etot"], sb["Etot"])
    for k in SCALARS:
        assert sa[k] == sb[k], k
    _same(da, db)

    # a context with other parameters refuses it; so does a truncated buffer
    with cm.Context(cl.n, cm.default_params(cl.n, orbit_r=cl.rtid, seed=10)) as c:
        with pytest.raises(cm.CmcgError):
            c.restore(buf)
        with pytest.raises(cm.CmcgError):
            c.restore(buf[: buf.size // 2])


def test_gpu_host_shim_restart_file(gpu, tmp_path):
    """cmc_hotpath_save_restart / cmc_hotpath_load_restart through the C host side:
    two steps, save, two more; reload the file (context created from the file's
    parameters) and step twice again: identical state."""
    cm = gpu
    say("host shim restart file")
    cl = cm.make_plummer(10000, seed=6)
    h = cm.hostlib
    p = cm.default_params(cl.n, orbit_r=cl.rtid, seed=2)
    h.cmc_hotpath_step.argtypes = [C.c_void_p]
    h.cmc_hotpath_ctx.restype = C.c_void_p
    path = str(tmp_path / "hot.restart.bin").encode()

    class Borrowed(cm.Context):
        """The shim's context, seen through the Python wrapper without owning it."""

        def __init__(self):
            self.h = C.c_void_p(h.cmc_hotpath_ctx())

        def close(self):
            self.h = C.c_void_p()

    assert h.cmc_hotpath_init(C.byref(p), cl.n, C.byref(cl.soa()), 0) == 0
    try:
        ctx = Borrowed()
        ctx.init_cluster()
        for _ in range(2):
            assert h.cmc_hotpath_step(None) >= 0
        assert h.cmc_hotpath_save_restart(path) == 0
        for _ in range(2):
            assert h.cmc_hotpath_step(None) >= 0
        first = ctx.download()
        s1 = ctx.scalars()
        assert h.cmc_hotpath_load_restart(path, 0, 0) == 0
        ctx = Borrowed()
        assert ctx.scalars()["step"] == 2
        for _ in range(2):
            assert h.cmc_hotpath_step(None) >= 0
        second = ctx.download()
        s2 = ctx.scalars()
        ctx.close()
    finally:
        h.cmc_hotpath_finalize()
    for k in SCALARS:
        assert s1[k] == s2[k], k
    _same(first, second)
