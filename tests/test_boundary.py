"""The C-ABI libraries load and export every function their headers declare
(no GPU needed: only symbol lookup, no compute calls)."""
import ctypes as C
import os
import re

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    """Names of the functions a header declares (prototype lines, not typedefs/macros)."""
    txt = open(os.path.join(REPO, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set()
    for line in txt.splitlines():
        line = line.strip()
        if not line or line.startswith(("#", "typedef", "}", "extern \"C\"")):
            continue
        m = re.match(r"^(?:const\s+)?[A-Za-z_][\w]*(?:\s+[\w]+)*\s*\**\s*\b([A-Za-z_]\w*)\s*\(", line)
        if m:
            names.add(m.group(1))
    return names


def test_cmcg_exports_all_declared_symbols(cm):
    names = {n for n in declared("cmcg.h") if n.startswith("cmcg_")}
    assert len(names) >= 30
    for n in sorted(names):
        assert hasattr(cm.lib, n), n
    assert set(cm.EXPORTED) <= names


def test_hotpath_exports_reference_names(cm):
    names = declared("cmc_hotpath.h")
    ref_names = {"qsorts_new", "post_sort_comm", "calc_potential_new", "potential_calculate", "potential",
                 "dynamics_apply", "new_orbits_calculate", "get_positions", "calc_orbit_new",
                 "calc_pericenter_and_apocenter", "find_zero_Q", "simul_relax_new", "calc_timestep",
                 "energy_conservation1", "energy_conservation2", "energy_conservation3",
                 "tidally_strip_stars", "ComputeIntermediateEnergy", "compute_energy_new"}
    assert ref_names <= names
    for n in sorted(n for n in names if n not in ("cmc_hotpath_error_handler",)):
        if n in ("cmc_hotpath_globals",):
            continue
        assert hasattr(cm.hostlib, n), n


def test_error_strings(cm):
    for code in [0, -1, -2, -3, -4, -5, -6, -7, -10, -11, -12, -13, -14, -99]:
        assert cm.lib.cmcg_error_string(code)


def test_struct_layouts(cm):
    # field counts / sizes mirrored in the binding
    assert C.sizeof(cm.Params) == 8 * 7 + 8
    assert C.sizeof(cm.HostSoa) == 8 + 8 * 25
    assert C.sizeof(cm.Scalars) == 8 * 2 + 8 * 16 + 8
    assert C.sizeof(cm.OrbitStats) == 8 * 10 + 8 * 4 + 8 + 8


def test_io_exports_declared_symbols(cm):
    import pytest
    if not os.path.exists(cm.IO_LIB_PATH):
        pytest.skip("libcmcio.so not built (no HDF5 C library)")
    names = declared("cmc_io.h")
    assert {"cmc_hdf5_read_header", "cmc_hdf5_read_objects", "cmc_io_error_string"} <= names
    for n in sorted(names):
        assert hasattr(cm._io(), n), n
    assert C.sizeof(cm.Hdf5Header) == 8 * 4 + 8 * 5
