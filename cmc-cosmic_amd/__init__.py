"""Python binding of the MI355X Henon-step hot path (libcmcg.so / libcmchost.so).

The library is a C ABI (include/cmcg.h); this module only wraps it with ctypes and
numpy so tests and the benchmark can drive it.  It never falls back to anything:
if the HIP library is missing, importing raises.

Directory name `cmc-cosmic_amd` is not a Python identifier; load it with
``load_package()`` from ``tests/conftest.py`` / ``bench.py`` (importlib by path).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcmcg.so")
HOST_LIB_PATH = os.path.join(HERE, "libcmchost.so")

if not os.path.exists(LIB_PATH) or not os.path.exists(HOST_LIB_PATH):
    raise ImportError(
        "cmc-cosmic_amd: libcmcg.so / libcmchost.so not built; run __graft_entry__.build() "
        "(make -C cmc-cosmic_amd)")

lib = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
hostlib = C.CDLL(HOST_LIB_PATH)

# ---------------------------------------------------------------- error codes
OK, E_NAN, E_NTRY, E_BRACKET, E_WZERO, E_WPZERO, E_BRENT, E_TIES = 0, -1, -2, -3, -4, -5, -6, -7
E_HIP, E_ARG, E_CAPACITY, E_NOMEM, E_STATE = -10, -11, -12, -13, -14

SF_CIRCULAR, SF_REMOVED_E, SF_REMOVED_RT, SF_DESTROYED = 1, 2, 4, 8
SF_BRENT, SF_LINSCAN, SF_STRIPPED, SF_BRACKET_HIT, SF_SAMPLED = 16, 32, 64, 128, 256
PF_LEARNED_BRACKETS = 1

lib.cmcg_error_string.restype = C.c_char_p
lib.cmcg_error_string.argtypes = [C.c_int]


class CmcgError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        super().__init__(f"{where}: {lib.cmcg_error_string(code).decode()} ({code})")


def _chk(rc, where):
    if rc != 0:
        raise CmcgError(rc, where)
    return rc


# ---------------------------------------------------------------- structs
class Params(C.Structure):
    _fields_ = [("n_star", C.c_int64), ("avekernel", C.c_int64), ("theta_se_max", C.c_double),
                ("gamma", C.c_double), ("orbit_r", C.c_double), ("cenma_m", C.c_double),
                ("seed", C.c_uint64), ("tidal_strip", C.c_int32), ("flags", C.c_int32)]


DFIELDS = ["r", "m", "phi", "vr", "vt", "E", "J"]
DFIELDS2 = ["rnew", "vrnew", "vtnew", "X", "Y", "r_peri", "r_apo",
            "rOld", "Uoldrold", "Uoldrnew", "vrold", "vtold", "EI", "sigma", "pot_rnew"]
FIELD_TYPES = {**{k: np.float64 for k in DFIELDS + DFIELDS2}, "id": np.int64,
               "ntry": np.int32, "flags": np.int32}
ALL_FIELDS = DFIELDS + ["id"] + DFIELDS2 + ["ntry", "flags"]

_soa_fields = [("n", C.c_int64)]
for _k in DFIELDS:
    _soa_fields.append((_k, C.POINTER(C.c_double)))
_soa_fields.append(("id", C.POINTER(C.c_int64)))
for _k in DFIELDS2:
    _soa_fields.append((_k, C.POINTER(C.c_double)))
_soa_fields += [("ntry", C.POINTER(C.c_int32)), ("flags", C.POINTER(C.c_int32))]


class HostSoa(C.Structure):
    _fields_ = _soa_fields


class Scalars(C.Structure):
    _fields_ = [("n_max", C.c_int64), ("n_max_new", C.c_int64)] + [
        (k, C.c_double) for k in ["Mtotal", "Rtidal", "max_r", "Dt", "Prev_Dt", "TotalTime", "TidalMassLoss",
                                  "TidalMassLoss_old", "OldTidalMassLoss", "Eescaped", "Jescaped", "Etidal",
                                  "Eoops", "Etot", "Ekin", "Epot"]] + [("step", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class OrbitStats(C.Structure):
    _fields_ = [(k, C.c_int64) for k in ["n_circular", "n_removed_E", "n_removed_rt", "n_destroyed", "n_brent",
                                         "n_linscan", "tries_total", "tries_max", "n_bracket_hit",
                                         "n_bracket_fallback"]] + [
        (k, C.c_double) for k in ["max_r", "tidal_mass_loss", "E_escaped", "J_escaped"]] + [
        ("error", C.c_int32), ("pad", C.c_int32), ("error_index", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class MassPc(C.Structure):
    _fields_ = [(k, C.POINTER(C.c_double)) for k in ("mass_r", "ave_mass_r", "densities_r", "ke_rad_r", "ke_tan_r",
                                                    "v2_rad_r", "v2_tan_r")] + [
        ("no_star_r", C.POINTER(C.c_int64)), ("reached", C.c_int32), ("pad", C.c_int32)]


MASS_PC_KEYS = ("mass_r", "ave_mass_r", "densities_r", "ke_rad_r", "ke_tan_r", "v2_rad_r", "v2_tan_r")

CENTRAL_KEYS = ("rh", "rhojsum", "rhoj2sum", "rho", "v_rms", "rc", "rc_nb", "m_ave", "n", "rc_spitzer",
                "N_core", "N_core_nb", "Trc")


class Central(C.Structure):
    _fields_ = [("nave", C.c_int64)] + [(k, C.c_double) for k in CENTRAL_KEYS]


_P = C.POINTER
_ctx = C.c_void_p
_sigs = {
    "cmcg_create": [_P(_ctx), C.c_int, C.c_int64, _P(Params)],
    "cmcg_destroy": [_ctx],
    "cmcg_upload": [_ctx, _P(HostSoa)],
    "cmcg_download": [_ctx, _P(HostSoa)],
    "cmcg_get_scalars": [_ctx, _P(Scalars)],
    "cmcg_set_scalars": [_ctx, _P(Scalars)],
    "cmcg_sync": [_ctx],
    "cmcg_checkpoint_size": [_ctx, _P(C.c_int64)],
    "cmcg_checkpoint_save": [_ctx, C.c_void_p, C.c_int64],
    "cmcg_checkpoint_info": [C.c_void_p, C.c_int64, _P(Params), _P(C.c_int64)],
    "cmcg_checkpoint_load": [_ctx, C.c_void_p, C.c_int64],
    "cmcg_sort_by_r": [_ctx, _P(C.c_int64)],
    "cmcg_potential": [_ctx, _P(C.c_int64), _P(C.c_double), _P(C.c_double)],
    "cmcg_potential_at": [_ctx, C.c_int64, _P(C.c_double), _P(C.c_double)],
    "cmcg_simul_relax": [_ctx, _P(C.c_double)],
    "cmcg_calc_timestep": [_ctx, _P(C.c_double)],
    "cmcg_sigma_r": [_ctx],
    "cmcg_relax": [_ctx, C.c_double, C.c_uint64],
    "cmcg_orbits": [_ctx, C.c_uint64, _P(OrbitStats)],
    "cmcg_calc_pericenter_and_apocenter": [_ctx, C.c_int64, C.c_int64, _P(C.c_double), _P(C.c_double),
                                           _P(C.c_double), _P(C.c_double), _P(C.c_double), _P(C.c_double),
                                           _P(C.c_int64), _P(C.c_int64), _P(C.c_int32)],
    "cmcg_calc_pericenter_and_apocenter_noalloc": [_ctx, C.c_int64, C.c_int64, _P(C.c_double), _P(C.c_double),
                                                   _P(C.c_double), _P(C.c_double), _P(C.c_double),
                                                   _P(C.c_double), _P(C.c_int64), _P(C.c_int64), _P(C.c_int32)],
    "cmcg_find_root_vr": [_ctx, C.c_int64, _P(C.c_int64), _P(C.c_int64), _P(C.c_double), _P(C.c_double),
                          _P(C.c_double)],
    "cmcg_find_zero_Q": [_ctx, C.c_int64, _P(C.c_int64), _P(C.c_int64), _P(C.c_int64), _P(C.c_double),
                         _P(C.c_double), _P(C.c_int64)],
    "cmcg_energy_conservation1": [_ctx],
    "cmcg_energy_conservation2": [_ctx],
    "cmcg_tidally_strip_stars": [_ctx],
    "cmcg_compute_intermediate_energy": [_ctx],
    "cmcg_set_velocities3": [_ctx],
    "cmcg_compute_energy": [_ctx],
    "cmcg_init_cluster": [_ctx],
    "cmcg_step": [_ctx, _P(OrbitStats)],
    "cmcg_step_graph": [_ctx, _P(OrbitStats)],
    "cmcg_autotune": [_ctx, C.c_int, _P(C.c_int32), _P(C.c_int32), _P(C.c_int32), _P(C.c_double), _P(C.c_double)],
    "cmcg_test_hook": [_ctx, C.c_int32, C.c_int64],
    "cmcg_comp_mass_percent": [_ctx, _P(C.c_double), C.c_int32, C.c_int32, C.c_double, C.c_double, _P(MassPc)],
    "cmcg_central_calculate": [_ctx, _P(Central)],
    "cmcg_orbit_variants": [_ctx, C.c_int32, _P(C.c_int32), _P(C.c_int32), _P(C.c_int32), _P(C.c_int32), _P(C.c_int32),
                            _P(C.c_double)],
    "cmcg_ensemble_step": [_P(_ctx), C.c_int, _P(OrbitStats), _P(C.c_int32)],
    "cmcg_ensemble_step_concurrent": [_P(_ctx), C.c_int, _P(OrbitStats), _P(C.c_int32)],
    "cmcg_ensemble_stats": [_P(C.c_int64), _P(C.c_int64)],
    "cmcg_autotune_relax": [_ctx, C.c_int, _P(C.c_int32), _P(C.c_double), _P(C.c_double)],
    "cmcg_rng_draws": [_ctx, C.c_int, C.c_uint64, C.c_int64, C.c_int64, C.c_int, _P(C.c_double)],
    "cmcg_timer_enable": [_ctx, C.c_int],
    "cmcg_timer_get": [_ctx, C.c_char_p, _P(C.c_double), _P(C.c_int64)],
    "cmcg_timer_reset": [_ctx],
    "cmcg_mlp_init": [_ctx, C.c_uint64, C.c_int64],
    "cmcg_mlp_get_params": [_ctx, _P(C.c_float), _P(C.c_float)],
    "cmcg_mlp_set_params": [_ctx, _P(C.c_float), _P(C.c_float)],
    "cmcg_mlp_harvest": [_ctx, C.c_int],
    "cmcg_mlp_rows": [_ctx, _P(C.c_int64), _P(C.c_float)],
    "cmcg_mlp_fit_norm": [_ctx],
    "cmcg_mlp_loss_grad": [_ctx, _P(C.c_float), C.c_int64, _P(C.c_double), _P(C.c_float)],
    "cmcg_mlp_train": [_ctx, C.c_int64, C.c_int64, C.c_float, C.c_uint64, _P(C.c_double)],
    "cmcg_mlp_predict": [_ctx, _P(C.c_float), C.c_int64, _P(C.c_float)],
    "cmcg_set_learned_brackets": [_ctx, C.c_int, C.c_double],
}
MLP_NPARAM, MLP_NNORM = 4738, 16
HOOK_CHAIN_K_DROP = 1
HOOK_ORBIT_VARIANT = 2
HOOK_MLP_MFMA = 3
HOOK_PERFECT_BRACKETS = 4
HOOK_COLD_PHASES = 5
HOOK_RELAX_SHAPE = 6
for _name, _args in _sigs.items():
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = C.c_int

EXPORTED = sorted(_sigs) + ["cmcg_error_string"]

hostlib.cmc_ic_plummer.argtypes = [C.c_int64, C.c_uint64, C.c_int, _P(HostSoa), _P(C.c_double)]
hostlib.cmc_ic_king.argtypes = [C.c_int64, C.c_double, C.c_uint64, C.c_int, C.c_int, _P(HostSoa), _P(C.c_double)]
hostlib.cmc_ic_kroupa.argtypes = [C.c_int64, C.c_uint64, _P(C.c_double)]
for _f in (hostlib.cmc_ic_plummer, hostlib.cmc_ic_king, hostlib.cmc_ic_kroupa):
    _f.restype = C.c_int


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


def checkpoint_info(buf):
    """(Params, N) stored in a checkpoint buffer; raises CmcgError if malformed."""
    buf = np.ascontiguousarray(buf, dtype=np.uint8)
    p, n = Params(), C.c_int64()
    _chk(lib.cmcg_checkpoint_info(buf.ctypes.data, buf.size, C.byref(p), C.byref(n)), "cmcg_checkpoint_info")
    return p, n.value


# ---------------------------------------------------------------- host cluster
class Cluster:
    """Host SoA copy of a cluster: arrays of n+2 entries (sentinels at 0, n+1)."""

    def __init__(self, n, fields=None):
        self.n = int(n)
        self.a = {}
        for k in (fields or ALL_FIELDS):
            self.a[k] = np.zeros(self.n + 2, dtype=FIELD_TYPES[k])
        self.rtid = None

    def __getitem__(self, k):
        return self.a[k]

    def __setitem__(self, k, v):
        self.a[k] = np.ascontiguousarray(v, dtype=FIELD_TYPES[k])

    def soa(self, fields=None):
        s = HostSoa()
        s.n = self.n
        for k in ALL_FIELDS:
            arr = self.a.get(k)
            if fields is not None and k not in fields:
                arr = None
            if arr is not None:
                ct = {np.float64: C.c_double, np.int64: C.c_int64, np.int32: C.c_int32}[FIELD_TYPES[k]]
                setattr(s, k, _ptr(arr, ct))
        return s

    def copy(self):
        c = Cluster(self.n, fields=list(self.a))
        for k, v in self.a.items():
            c.a[k] = v.copy()
        c.rtid = self.rtid
        return c


def make_plummer(n, seed=1, kroupa=False):
    c = Cluster(n)
    rt = C.c_double()
    _chk(hostlib.cmc_ic_plummer(n, seed, int(kroupa), C.byref(c.soa()), C.byref(rt)), "cmc_ic_plummer")
    c.rtid = rt.value
    return c


def make_king(n, w0=7.0, seed=1, kroupa=True, segregate=True):
    c = Cluster(n)
    rt = C.c_double()
    _chk(hostlib.cmc_ic_king(n, float(w0), seed, int(kroupa), int(segregate), C.byref(c.soa()), C.byref(rt)),
         "cmc_ic_king")
    c.rtid = rt.value
    return c


def kroupa_masses(n, seed=1):
    m = np.zeros(n)
    _chk(hostlib.cmc_ic_kroupa(n, seed, _ptr(m, C.c_double)), "cmc_ic_kroupa")
    return m


# ---------------------------------------------------------------- HDF5 input (libcmcio.so)
IO_LIB_PATH = os.path.join(HERE, "libcmcio.so")
_iolib = None


class Hdf5Header(C.Structure):
    _fields_ = [("nobj", C.c_int64), ("nbinary", C.c_int64), ("rows", C.c_int64), ("cols", C.c_int64),
                ("rvir", C.c_double), ("rtid", C.c_double), ("mclus", C.c_double), ("z", C.c_double),
                ("cenma_m", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CmcIoError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        super().__init__(f"{where}: {_io().cmc_io_error_string(code).decode()} ({code})")


def _io():
    """The HDF5 reader library, loaded on first use (it links the HDF5 C library,
    which nothing else here needs)."""
    global _iolib
    if _iolib is None:
        if not os.path.exists(IO_LIB_PATH):
            raise ImportError("cmc-cosmic_amd: libcmcio.so not built (needs the HDF5 C library; "
                              "make -C cmc-cosmic_amd HDF5_DIR=...)")
        io = C.CDLL(IO_LIB_PATH)
        io.cmc_io_error_string.restype = C.c_char_p
        io.cmc_io_error_string.argtypes = [C.c_int]
        io.cmc_hdf5_read_header.argtypes = [C.c_char_p, C.POINTER(Hdf5Header)]
        io.cmc_hdf5_read_objects.argtypes = [C.c_char_p, C.POINTER(Hdf5Header), C.POINTER(HostSoa),
                                             C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                             C.POINTER(C.c_int64)]
        _iolib = io
    return _iolib


def read_hdf5_header(path):
    h = Hdf5Header()
    rc = _io().cmc_hdf5_read_header(os.fsencode(path), C.byref(h))
    if rc:
        raise CmcIoError(rc, "cmc_hdf5_read_header")
    return h


def read_hdf5(path, columns=False):
    """Cluster from the reference's HDF5 input file: id, m (scaled by NOBJ), r, vr, vt
    in slots 0..n+1, `rtid` = RTID (the initial orbit_r), `cenma_m` = the central
    mass and `header` = the file's attributes.  With columns=True also returns a
    dict of the stellar type `k`, radius `Reff` and binary index `binind` columns."""
    h = read_hdf5_header(path)
    cl = Cluster(h.nobj)
    extra = {"k": np.zeros(h.nobj + 2, np.int32), "Reff": np.zeros(h.nobj + 2),
             "binind": np.zeros(h.nobj + 2, np.int64)}
    rc = _io().cmc_hdf5_read_objects(os.fsencode(path), C.byref(h), C.byref(cl.soa()),
                                     _ptr(extra["k"], C.c_int32), _ptr(extra["Reff"], C.c_double),
                                     _ptr(extra["binind"], C.c_int64))
    if rc:
        raise CmcIoError(rc, "cmc_hdf5_read_objects")
    cl.rtid = h.rtid
    cl.cenma_m = h.cenma_m
    cl.header = h.as_dict()
    return (cl, extra) if columns else cl


def default_params(n_star, orbit_r=1e3, seed=1, theta_se_max=1.412, avekernel=20, gamma=0.1, tidal_strip=1,
                   flags=0):
    p = Params()
    p.n_star = int(n_star)
    p.avekernel = int(avekernel)
    p.theta_se_max = float(theta_se_max)
    p.gamma = float(gamma)
    p.orbit_r = float(orbit_r)
    p.cenma_m = 0.0
    p.seed = int(seed)
    p.tidal_strip = int(tidal_strip)
    p.flags = int(flags)
    return p


# ---------------------------------------------------------------- device context
class Context:
    def __init__(self, n_cap, params, device=0):
        self.h = C.c_void_p()
        self.params = params
        _chk(lib.cmcg_create(C.byref(self.h), device, int(n_cap), C.byref(params)), "cmcg_create")

    def close(self):
        if self.h:
            lib.cmcg_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, cl, fields=None):
        _chk(lib.cmcg_upload(self.h, C.byref(cl.soa(fields))), "cmcg_upload")

    def download(self, n_hint=None, fields=None):
        sc = self.scalars()
        n = sc["n_max_new"] if n_hint is None else n_hint
        cl = Cluster(n, fields=fields)
        s = cl.soa()
        _chk(lib.cmcg_download(self.h, C.byref(s)), "cmcg_download")
        cl.n = s.n
        return cl

    def scalars(self):
        s = Scalars()
        rc = lib.cmcg_get_scalars(self.h, C.byref(s))
        if rc == E_HIP:
            _chk(rc, "cmcg_get_scalars")
        d = s.as_dict()
        d["error"] = rc
        return d

    def set_scalars(self, **kw):
        s = Scalars()
        lib.cmcg_get_scalars(self.h, C.byref(s))
        for k, v in kw.items():
            setattr(s, k, v)
        _chk(lib.cmcg_set_scalars(self.h, C.byref(s)), "cmcg_set_scalars")

    def sync(self):
        return lib.cmcg_sync(self.h)

    def checkpoint(self):
        """The device state as a byte array (cmcg_checkpoint_save)."""
        n = C.c_int64()
        _chk(lib.cmcg_checkpoint_size(self.h, C.byref(n)), "cmcg_checkpoint_size")
        buf = np.empty(n.value, dtype=np.uint8)
        _chk(lib.cmcg_checkpoint_save(self.h, buf.ctypes.data, n.value), "cmcg_checkpoint_save")
        return buf

    def restore(self, buf):
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        _chk(lib.cmcg_checkpoint_load(self.h, buf.ctypes.data, buf.size), "cmcg_checkpoint_load")

    def sort_by_r(self):
        n = C.c_int64()
        _chk(lib.cmcg_sort_by_r(self.h, C.byref(n)), "cmcg_sort_by_r")
        return n.value

    def potential(self):
        n, mt, rt = C.c_int64(), C.c_double(), C.c_double()
        _chk(lib.cmcg_potential(self.h, C.byref(n), C.byref(mt), C.byref(rt)), "cmcg_potential")
        return n.value, mt.value, rt.value

    def potential_at(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        out = np.zeros_like(r)
        _chk(lib.cmcg_potential_at(self.h, r.size, _ptr(r, C.c_double), _ptr(out, C.c_double)), "cmcg_potential_at")
        return out

    def simul_relax(self):
        d = C.c_double()
        _chk(lib.cmcg_simul_relax(self.h, C.byref(d)), "cmcg_simul_relax")
        return d.value

    def calc_timestep(self):
        d = C.c_double()
        _chk(lib.cmcg_calc_timestep(self.h, C.byref(d)), "cmcg_calc_timestep")
        return d.value

    def sigma_r(self):
        _chk(lib.cmcg_sigma_r(self.h), "cmcg_sigma_r")

    def relax(self, dt, step=0):
        _chk(lib.cmcg_relax(self.h, float(dt), int(step)), "cmcg_relax")

    def orbits(self, step=0, check=True):
        st = OrbitStats()
        rc = lib.cmcg_orbits(self.h, int(step), C.byref(st))
        if check:
            _chk(rc, "cmcg_orbits")
        return st.as_dict()

    APSIDES_NOALLOC_MAX = 32768

    def apsides(self, first, E, J, noalloc=False):
        """calc_orbit_new for slots first.. (cmcg_calc_pericenter_and_apocenter, or its
        variant that allocates nothing when noalloc is set)."""
        E = np.ascontiguousarray(E, dtype=np.float64)
        J = np.ascontiguousarray(J, dtype=np.float64)
        n = E.size
        rp, ra, dp, da = (np.zeros(n) for _ in range(4))
        kmin, kmax = np.zeros(n, np.int64), np.zeros(n, np.int64)
        circ = np.zeros(n, np.int32)
        name = "cmcg_calc_pericenter_and_apocenter" + ("_noalloc" if noalloc else "")
        _chk(getattr(lib, name)(
            self.h, int(first), n, _ptr(E, C.c_double), _ptr(J, C.c_double), _ptr(rp, C.c_double),
            _ptr(ra, C.c_double), _ptr(dp, C.c_double), _ptr(da, C.c_double), _ptr(kmin, C.c_int64),
            _ptr(kmax, C.c_int64), _ptr(circ, C.c_int32)), name)
        return dict(rp=rp, ra=ra, dQdrp=dp, dQdra=da, kmin=kmin, kmax=kmax, circular=circ)

    def find_zero_Q(self, j, kmin, kmax, E, J):
        j, kmin, kmax = (np.ascontiguousarray(x, dtype=np.int64) for x in (j, kmin, kmax))
        E, J = (np.ascontiguousarray(x, dtype=np.float64) for x in (E, J))
        out = np.zeros(j.size, np.int64)
        _chk(lib.cmcg_find_zero_Q(self.h, j.size, _ptr(j, C.c_int64), _ptr(kmin, C.c_int64), _ptr(kmax, C.c_int64),
                                  _ptr(E, C.c_double), _ptr(J, C.c_double), _ptr(out, C.c_int64)),
             "cmcg_find_zero_Q")
        return out

    def find_root_vr(self, j, k, E, J):
        """find_root_vr (cmc_nr.c:500-603) for arrays of (star j, interval k, E, J)."""
        j, k = (np.ascontiguousarray(x, dtype=np.int64) for x in (j, k))
        E, J = (np.ascontiguousarray(x, dtype=np.float64) for x in (E, J))
        out = np.zeros(j.size, np.float64)
        _chk(lib.cmcg_find_root_vr(self.h, j.size, _ptr(j, C.c_int64), _ptr(k, C.c_int64), _ptr(E, C.c_double),
                                   _ptr(J, C.c_double), _ptr(out, C.c_double)), "cmcg_find_root_vr")
        return out

    def energy_conservation1(self):
        _chk(lib.cmcg_energy_conservation1(self.h), "cmcg_energy_conservation1")

    def energy_conservation2(self):
        _chk(lib.cmcg_energy_conservation2(self.h), "cmcg_energy_conservation2")

    def tidally_strip_stars(self):
        _chk(lib.cmcg_tidally_strip_stars(self.h), "cmcg_tidally_strip_stars")

    def compute_intermediate_energy(self):
        _chk(lib.cmcg_compute_intermediate_energy(self.h), "cmcg_compute_intermediate_energy")

    def set_velocities3(self):
        _chk(lib.cmcg_set_velocities3(self.h), "cmcg_set_velocities3")

    def compute_energy(self):
        _chk(lib.cmcg_compute_energy(self.h), "cmcg_compute_energy")

    def init_cluster(self):
        _chk(lib.cmcg_init_cluster(self.h), "cmcg_init_cluster")

    def step(self, check=True):
        st = OrbitStats()
        rc = lib.cmcg_step(self.h, C.byref(st))
        if check:
            _chk(rc, "cmcg_step")
        return st.as_dict()

    def autotune(self, reps=3):
        v, th, cs = C.c_int32(), C.c_int32(), C.c_int32()
        best, default = C.c_double(), C.c_double()
        _chk(lib.cmcg_autotune(self.h, int(reps), C.byref(v), C.byref(th), C.byref(cs), C.byref(best),
                               C.byref(default)), "cmcg_autotune")
        out = {"variant": v.value, "threads": th.value, "coarse": cs.value, "ms_best": best.value,
               "ms_default": default.value}
        out["variants"] = self.orbit_variants()
        return out

    def autotune_relax(self, reps=3):
        """cmcg_autotune_relax: the relaxation pass's block size, timed on this state."""
        th, best, default = C.c_int32(), C.c_double(), C.c_double()
        _chk(lib.cmcg_autotune_relax(self.h, int(reps), C.byref(th), C.byref(best), C.byref(default)),
             "cmcg_autotune_relax")
        return {"threads": th.value, "ms_best": best.value, "ms_default": default.value}

    def comp_mass_percent(self, pc, bh_include=0, initial_total_mass=1.0, minimum_r=0.0, fill=np.nan):
        """cmcg_comp_mass_percent: dict of arrays (entries not reached = fill) and the count reached."""
        pc = np.ascontiguousarray(pc, dtype=np.float64)
        out = {k: np.full(pc.size, fill) for k in MASS_PC_KEYS}
        out["no_star_r"] = np.full(pc.size, -1, dtype=np.int64)
        o = MassPc(*[_ptr(out[k], C.c_double) for k in MASS_PC_KEYS], _ptr(out["no_star_r"], C.c_int64), 0, 0)
        _chk(lib.cmcg_comp_mass_percent(self.h, _ptr(pc, C.c_double), int(pc.size), int(bh_include),
                                        float(initial_total_mass), float(minimum_r), C.byref(o)),
             "cmcg_comp_mass_percent")
        out["reached"] = o.reached
        return out

    def central_calculate(self):
        """cmcg_central_calculate: dict of the core quantities, the half-mass radius rh
        and nave (central_calculate + clusdyn_calculate)."""
        o = Central()
        _chk(lib.cmcg_central_calculate(self.h, C.byref(o)), "cmcg_central_calculate")
        d = {k: getattr(o, k) for k in CENTRAL_KEYS}
        d["nave"] = o.nave
        return d

    def test_hook(self, hook, value):
        _chk(lib.cmcg_test_hook(self.h, int(hook), int(value)), "cmcg_test_hook")

    def orbit_variants(self):
        """Launch variants of the orbit sampler with their last autotune times."""
        cnt = C.c_int32()
        _chk(lib.cmcg_orbit_variants(self.h, 0, C.byref(cnt), None, None, None, None, None), "cmcg_orbit_variants")
        n = cnt.value
        th, cs, cp, mb = (np.zeros(n, np.int32) for _ in range(4))
        ms = np.zeros(n, np.float64)
        _chk(lib.cmcg_orbit_variants(self.h, n, C.byref(cnt), _ptr(th, C.c_int32), _ptr(cs, C.c_int32),
                                     _ptr(cp, C.c_int32), _ptr(mb, C.c_int32), _ptr(ms, C.c_double)),
             "cmcg_orbit_variants")
        return [{"threads": int(th[i]), "coarse": int(cs[i]), "cap": int(cp[i]), "min_blocks": int(mb[i]),
                 "ms": float(ms[i])} for i in range(n)]

    def step_graph(self, check=True):
        st = OrbitStats()
        rc = lib.cmcg_step_graph(self.h, C.byref(st))
        if check:
            _chk(rc, "cmcg_step_graph")
        return st.as_dict()

    def rng_draws(self, phase, step, first, count, ndraw):
        out = np.zeros(count * ndraw)
        _chk(lib.cmcg_rng_draws(self.h, phase, step, first, count, ndraw, _ptr(out, C.c_double)), "cmcg_rng_draws")
        return out.reshape(count, ndraw)

    def timer_enable(self, on=True):
        lib.cmcg_timer_enable(self.h, int(on))

    def timer_get(self, name):
        ms, calls = C.c_double(), C.c_int64()
        lib.cmcg_timer_get(self.h, name.encode(), C.byref(ms), C.byref(calls))
        return ms.value, calls.value

    def timer_reset(self):
        lib.cmcg_timer_reset(self.h)

    # ---------------------------------------------------------- learned brackets
    def mlp_init(self, seed=1, max_rows=1 << 22):
        _chk(lib.cmcg_mlp_init(self.h, int(seed), int(max_rows)), "cmcg_mlp_init")

    def mlp_params(self):
        th = np.zeros(MLP_NPARAM, np.float32)
        nr = np.zeros(MLP_NNORM, np.float32)
        _chk(lib.cmcg_mlp_get_params(self.h, _ptr(th, C.c_float), _ptr(nr, C.c_float)), "cmcg_mlp_get_params")
        return th, nr

    def mlp_set_params(self, theta, norm):
        th = np.ascontiguousarray(theta, np.float32)
        nr = np.ascontiguousarray(norm, np.float32)
        assert th.size == MLP_NPARAM and nr.size == MLP_NNORM
        _chk(lib.cmcg_mlp_set_params(self.h, _ptr(th, C.c_float), _ptr(nr, C.c_float)), "cmcg_mlp_set_params")

    def mlp_harvest(self, on=True):
        _chk(lib.cmcg_mlp_harvest(self.h, int(on)), "cmcg_mlp_harvest")

    def mlp_rows(self, copy=True):
        n = C.c_int64()
        _chk(lib.cmcg_mlp_rows(self.h, C.byref(n), None), "cmcg_mlp_rows")
        if not copy:
            return n.value
        rows = np.zeros((n.value, 8), np.float32)
        if n.value:
            _chk(lib.cmcg_mlp_rows(self.h, C.byref(n), _ptr(rows, C.c_float)), "cmcg_mlp_rows")
        return rows

    def mlp_fit_norm(self):
        _chk(lib.cmcg_mlp_fit_norm(self.h), "cmcg_mlp_fit_norm")

    def mlp_loss_grad(self, rows):
        rows = np.ascontiguousarray(rows, np.float32)
        assert rows.ndim == 2 and rows.shape[1] == 8
        loss = C.c_double()
        g = np.zeros(MLP_NPARAM, np.float32)
        _chk(lib.cmcg_mlp_loss_grad(self.h, _ptr(rows, C.c_float), rows.shape[0], C.byref(loss), _ptr(g, C.c_float)),
             "cmcg_mlp_loss_grad")
        return loss.value, g

    def mlp_train(self, iters, batch=4096, lr=1e-3, seed=1):
        loss = C.c_double()
        _chk(lib.cmcg_mlp_train(self.h, int(iters), int(batch), float(lr), int(seed), C.byref(loss)), "cmcg_mlp_train")
        return loss.value

    def mlp_predict(self, x):
        x = np.ascontiguousarray(x, np.float32)
        assert x.ndim == 2 and x.shape[1] == 6
        y = np.zeros((x.shape[0], 2), np.float32)
        _chk(lib.cmcg_mlp_predict(self.h, _ptr(x, C.c_float), x.shape[0], _ptr(y, C.c_float)), "cmcg_mlp_predict")
        return y

    def set_learned_brackets(self, on=True, margin=0.05):
        _chk(lib.cmcg_set_learned_brackets(self.h, int(on), float(margin)), "cmcg_set_learned_brackets")


def ensemble_step(ctxs, check=True, concurrent=False):
    """One step of every context's cluster: merged into one segmented launch per
    kernel (default), or each cluster's step graph on its own stream (concurrent)."""
    n = len(ctxs)
    arr = (_ctx * n)(*[c.h for c in ctxs])
    stats = (OrbitStats * n)()
    errs = (C.c_int32 * n)()
    name = "cmcg_ensemble_step_concurrent" if concurrent else "cmcg_ensemble_step"
    rc = getattr(lib, name)(arr, n, stats, errs)
    if check:
        _chk(rc, name)
    return [stats[i].as_dict() for i in range(n)], [int(errs[i]) for i in range(n)]


def ensemble_stats():
    """(merged launches, recorded launches) of the last merged ensemble step."""
    a, b = C.c_int64(), C.c_int64()
    _chk(lib.cmcg_ensemble_stats(C.byref(a), C.byref(b)), "cmcg_ensemble_stats")
    return int(a.value), int(b.value)
