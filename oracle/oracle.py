"""ctypes binding of the CPU oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this, as
the checker or the timed CPU baseline.  The state is a numpy SoA mirror of
cmc_oracle.h's orc_state.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_TAUS = os.path.join(HERE, "_ref", "libtaus113_ref.so")

lib = C.CDLL(LIB_PATH)

DPTR = C.POINTER(C.c_double)
I64P = C.POINTER(C.c_int64)
I32P = C.POINTER(C.c_int32)


class T113(C.Structure):
    _fields_ = [("z", C.c_uint32 * 4)]

    def tolist(self):
        return [int(x) for x in self.z]


D_FIELDS_A = ["r", "m", "phi", "vr", "vt", "E", "J"]
D_FIELDS_B = ["rnew", "vrnew", "vtnew", "X", "Y", "r_peri", "r_apo", "rOld", "Uoldrold", "Uoldrnew",
              "vrold", "vtold", "EI", "pot_rnew", "sigma"]
SCALARS = ["madhoc", "cenma_m", "Mtotal", "Rtidal", "orbit_r", "max_r", "Dt", "Prev_Dt", "TotalTime",
           "TidalMassLoss", "TidalMassLoss_old", "OldTidalMassLoss", "Eescaped", "Jescaped", "Etidal", "Eoops",
           "Etot", "Ekin", "Epot", "theta_se_max", "gamma"]


class State(C.Structure):
    _fields_ = ([("N_STAR", C.c_int64), ("N_MAX", C.c_int64), ("N_MAX_NEW", C.c_int64), ("cap", C.c_int64)]
                + [(k, DPTR) for k in D_FIELDS_A] + [("id", I64P)] + [(k, DPTR) for k in D_FIELDS_B]
                + [("ntry", I32P), ("flags", I32P)] + [(k, C.c_double) for k in SCALARS]
                + [("avekernel", C.c_int64), ("seeded", T113), ("seq", T113), ("rng_mode", C.c_int32),
                   ("q_mode", C.c_int32), ("step", C.c_uint64)]
                + [(k, C.c_int64) for k in ["n_brent", "n_circular", "n_removed_E", "n_removed_tidal",
                                            "n_removed_mass", "n_linscan"]]
                + [("error", C.c_int32), ("error_index", C.c_int64), ("n_excess_events", C.c_int64),
                   ("n_excess_dumps", C.c_int64), ("libm_mode", C.c_int32), ("range_lo", C.c_int64),
                   ("range_hi", C.c_int64), ("skip_sigma", C.c_int32)])


class OrbitRS(C.Structure):
    _fields_ = [("rp", C.c_double), ("ra", C.c_double), ("dQdrp", C.c_double), ("dQdra", C.c_double),
                ("kmax", C.c_int64), ("kmin", C.c_int64), ("circular_flag", C.c_int32)]


_SP = C.POINTER(State)
_sigs = {
    "orc_t113_seed": (None, [C.c_uint64, C.POINTER(T113)]),
    "orc_t113_int": (C.c_uint32, [C.POINTER(T113)]),
    "orc_t113_dbl": (C.c_double, [C.POINTER(T113)]),
    "orc_t113_jump_poly": (T113, [T113, C.POINTER(C.c_uint32)]),
    "orc_t113_jump": (T113, [T113, C.c_uint64, C.c_uint64]),
    "orc_stream_relax": (T113, [C.POINTER(T113), C.c_uint64, C.c_uint64]),
    "orc_stream_orbit": (T113, [C.POINTER(T113), C.c_uint64, C.c_uint64]),
    "orc_FindZero_r": (C.c_int64, [_SP, C.c_int64, C.c_int64, C.c_double]),
    "orc_potential": (C.c_double, [_SP, C.c_double]),
    "orc_potential_calculate": (C.c_int64, [_SP]),
    "orc_function_q": (C.c_double, [_SP, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double]),
    "orc_find_zero_Q": (C.c_int64, [_SP, C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_double]),
    "orc_get_positive_Q_index": (C.c_int64, [_SP, C.c_int64, C.c_double, C.c_double]),
    "orc_calc_orbit_new": (OrbitRS, [_SP, C.c_int64, C.c_double, C.c_double]),
    "orc_find_root_vr": (C.c_double, [_SP, C.c_int64, C.c_int64, C.c_double, C.c_double]),
    "orc_get_positions": (C.c_double, [_SP]),
    "orc_calc_n_local": (C.c_double, [_SP, C.c_int64, C.c_int64, C.c_int64]),
    "orc_calc_sigma_r": (None, [_SP]),
    "orc_simul_relax_new": (C.c_double, [_SP]),
    "orc_dynamics_apply": (None, [_SP, C.c_double]),
    "orc_sort_by_r": (C.c_int64, [_SP]),
    "orc_energy_conservation1": (None, [_SP]),
    "orc_energy_conservation2": (None, [_SP]),
    "orc_tidally_strip_stars": (None, [_SP]),
    "orc_compute_intermediate_energy": (None, [_SP]),
    "orc_set_velocities3": (None, [_SP]),
    "orc_compute_energy": (None, [_SP]),
    "orc_calc_timestep": (C.c_double, [_SP]),
    "orc_init": (None, [_SP]),
    "orc_dd_sincos": (None, [C.c_double, DPTR, DPTR]),
    "orc_dd_trig_table": (None, [C.c_int, DPTR]),
    "orc_comp_mass_percent": (None, [C.c_void_p, DPTR, C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p]),
    "orc_central_calculate": (C.c_int, [C.c_void_p, C.c_void_p]),
    "orc_step": (None, [_SP]),
    "orc_ranks_new": (C.c_void_p, [_SP, C.c_int]),
    "orc_ranks_free": (None, [C.c_void_p]),
    "orc_step_ranks": (None, [C.c_void_p]),
    "orc_sort_by_r_ranks": (C.c_int64, [_SP, C.c_int]),
}
for _n, (_r, _a) in _sigs.items():
    _f = getattr(lib, _n)
    _f.restype = _r
    _f.argtypes = _a

RNG_REF_SEQ, RNG_PER_STAR = 0, 1
Q_LONG_DOUBLE, Q_FP64 = 0, 1
LIBM_SYSTEM, LIBM_LONG_DOUBLE, LIBM_DD, LIBM_SINCOS = 0, 1, 2, 3


def seed_state(seed):
    st = T113()
    lib.orc_t113_seed(seed, C.byref(st))
    return st


class MassPcOut(C.Structure):
    _fields_ = [(k, DPTR) for k in ("mass_r", "ave_mass_r", "densities_r", "ke_rad_r", "ke_tan_r", "v2_rad_r",
                                    "v2_tan_r")] + [("no_star_r", C.POINTER(C.c_int64))]


CENTRAL_KEYS = ("rh", "rhojsum", "rhoj2sum", "rho", "v_rms", "rc", "rc_nb", "m_ave", "n", "rc_spitzer",
                "N_core", "N_core_nb", "Trc")


class CentralOut(C.Structure):
    _fields_ = [("nave", C.c_int64)] + [(k, C.c_double) for k in CENTRAL_KEYS]


MASS_PC_KEYS = ("mass_r", "ave_mass_r", "densities_r", "ke_rad_r", "ke_tan_r", "v2_rad_r", "v2_tan_r")


class Oracle:
    """Oracle cluster state with numpy-owned arrays."""

    def __init__(self, cluster_arrays, n, n_star=None, orbit_r=1e3, seed=1, rng_mode=RNG_PER_STAR,
                 q_mode=Q_LONG_DOUBLE, theta_se_max=1.412, avekernel=20, gamma=0.1, libm_mode=LIBM_SYSTEM):
        self.n = int(n)
        cap = self.n + 2
        self.arr = {}
        for k in D_FIELDS_A + D_FIELDS_B:
            src = cluster_arrays.get(k)
            a = np.zeros(cap, np.float64)
            if src is not None:
                a[:] = np.asarray(src, np.float64)[:cap]
            self.arr[k] = a
        ids = cluster_arrays.get("id")
        self.arr["id"] = np.zeros(cap, np.int64) if ids is None else np.array(ids[:cap], np.int64)
        self.arr["ntry"] = np.zeros(cap, np.int32)
        self.arr["flags"] = np.zeros(cap, np.int32)
        s = State()
        s.N_STAR = int(n_star or n)
        s.N_MAX = self.n
        s.N_MAX_NEW = self.n
        s.cap = cap
        for k in D_FIELDS_A + D_FIELDS_B:
            setattr(s, k, self.arr[k].ctypes.data_as(DPTR))
        s.id = self.arr["id"].ctypes.data_as(I64P)
        s.ntry = self.arr["ntry"].ctypes.data_as(I32P)
        s.flags = self.arr["flags"].ctypes.data_as(I32P)
        s.madhoc = 1.0 / float(s.N_STAR)
        s.orbit_r = float(orbit_r)
        s.theta_se_max = float(theta_se_max)
        s.gamma = float(gamma)
        s.avekernel = int(avekernel)
        s.seeded = seed_state(seed)
        s.seq = seed_state(seed)
        s.rng_mode = int(rng_mode)
        s.q_mode = int(q_mode)
        s.libm_mode = int(libm_mode)
        self.s = s

    def __getitem__(self, k):
        return self.arr[k]

    @property
    def N(self):
        return int(self.s.N_MAX)

    def sc(self, k):
        return getattr(self.s, k)

    def call(self, name, *args):
        return getattr(lib, name)(C.byref(self.s), *args)

    def init(self):
        lib.orc_init(C.byref(self.s))

    def step(self):
        lib.orc_step(C.byref(self.s))
        return self.s.error

    def ranks(self, P):
        """The reference's MPI layout on P threads (orc_step_ranks): returns a stepper."""
        return Ranks(self, P)

    def potential_calculate(self):
        return lib.orc_potential_calculate(C.byref(self.s))

    def comp_mass_percent(self, pc, bh_include=0, initial_total_mass=1.0, minimum_r=0.0, fill=np.nan):
        """comp_mass_percent (cmc_utils.c:872-982): dict of arrays, entries not reached = fill."""
        pc = np.ascontiguousarray(pc, dtype=np.float64)
        out = {k: np.full(pc.size, fill) for k in MASS_PC_KEYS}
        out["no_star_r"] = np.full(pc.size, -1, dtype=np.int64)
        o = MassPcOut(*[out[k].ctypes.data_as(DPTR) for k in MASS_PC_KEYS],
                      out["no_star_r"].ctypes.data_as(C.POINTER(C.c_int64)))
        lib.orc_comp_mass_percent(C.byref(self.s), pc.ctypes.data_as(DPTR), int(pc.size), int(bh_include),
                                  float(initial_total_mass), float(minimum_r), C.byref(o))
        return out

    def central_calculate(self):
        """central_calculate + clusdyn_calculate (cmc_utils.c:1239-1345, 1614-1624):
        dict of the core quantities and rh; raises where the reference exits."""
        o = CentralOut()
        if lib.orc_central_calculate(C.byref(self.s), C.byref(o)) != 0:
            raise RuntimeError("central_calculate: N_STAR <= 12 or nave >= N_STAR - 6")
        d = {k: getattr(o, k) for k in CENTRAL_KEYS}
        d["nave"] = o.nave
        return d

    def potential(self, r):
        return lib.orc_potential(C.byref(self.s), float(r))

    def sort(self):
        return lib.orc_sort_by_r(C.byref(self.s))

    def relax(self, dt):
        lib.orc_dynamics_apply(C.byref(self.s), float(dt))

    def orbits(self):
        return lib.orc_get_positions(C.byref(self.s))

    def apsides(self, j, E, J):
        o = lib.orc_calc_orbit_new(C.byref(self.s), int(j), float(E), float(J))
        return dict(rp=o.rp, ra=o.ra, dQdrp=o.dQdrp, dQdra=o.dQdra, kmin=o.kmin, kmax=o.kmax,
                    circular=o.circular_flag)


class Ranks:
    """P-rank step of an Oracle (its REF_SEQ stream seeds rank 0; rank i is jumped i*2^80)."""

    def __init__(self, o, P):
        self.o = o
        self.h = lib.orc_ranks_new(C.byref(o.s), int(P))
        if not self.h:
            raise ValueError("bad rank count")

    def step(self):
        lib.orc_step_ranks(self.h)
        return self.o.s.error

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_ranks_free(self.h)
            self.h = None


def dd_sincos(x):
    s, c = C.c_double(), C.c_double()
    lib.orc_dd_sincos(float(x), C.byref(s), C.byref(c))
    return s.value, c.value


def dd_trig_table(j):
    """(sin hi, sin lo, cos hi, cos lo) of j pi/128 from the double-double sin/cos table."""
    out = (C.c_double * 4)()
    lib.orc_dd_trig_table(int(j), out)
    return tuple(out)


def ref_taus_lib():
    """The reference's own taus113-v2.c compiled from source into oracle/_ref (or None)."""
    if not os.path.exists(REF_TAUS):
        return None
    L = C.CDLL(REF_TAUS)

    class RefState(C.Structure):
        _fields_ = [("z", C.c_ulong * 4)]

    L.reset_rng_t113_new.argtypes = [C.c_ulong, C.POINTER(RefState)]
    L.rng_t113_int_new.argtypes = [C.POINTER(RefState)]
    L.rng_t113_int_new.restype = C.c_ulong
    L.rng_t113_dbl_new.argtypes = [C.POINTER(RefState)]
    L.rng_t113_dbl_new.restype = C.c_double
    L.rng_t113_jump.argtypes = [RefState, C.POINTER(C.c_uint)]
    L.rng_t113_jump.restype = RefState
    L.RefState = RefState
    return L
