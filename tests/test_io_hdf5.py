"""Reader of the reference's HDF5 cluster file (libcmcio.so, include/cmc_io.h) on the
reference's own CI input (tests/golden/ci_input.hdf5 = ci/input.hdf5: Plummer,
N = 10000, 1967 binaries, written by COSMIC's InitialCMCTable.write).

The checks do not go through the HDF5 library a second time:
  - the object table is contiguous in the file (h5dump's storage layout, kept in
    tests/golden/ci_input_layout.json by make_ci_input_layout.py), so numpy reads
    the same doubles straight from the file's bytes; every column the reader returns
    must equal them bit for bit, with the mass scaled by NOBJ and the sentinels of
    load_fits_file_data (src/cmc/cmc_fits.c:121-136);
  - the raw view itself is pinned by what the file must hold: ids 1..N, masses that
    sum to 1, r ascending, the outer sentinel row, and the virial kinetic energy
    1/4 of a Plummer model in N-body units (sum m (vr^2 + vt^2) / 2), which only the
    right m, vr and vt columns give;
  - the attributes equal h5dump's full-precision print.
Error paths (missing file, not HDF5, wrong n, a table that breaks the layout) return
the documented codes.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
H5 = os.path.join(HERE, "golden", "ci_input.hdf5")
LAYOUT = json.load(open(os.path.join(HERE, "golden", "ci_input_layout.json")))


@pytest.fixture(scope="module")
def io(cm):
    if not os.path.exists(cm.IO_LIB_PATH):
        pytest.skip("libcmcio.so not built (no HDF5 C library)")
    return cm


def raw_table(path=H5):
    with open(path, "rb") as f:
        f.seek(LAYOUT["offset"])
        b = f.read(LAYOUT["size"])
    return np.frombuffer(b, "<f8").reshape(LAYOUT["shape"])


def test_raw_view_is_the_object_table():
    assert LAYOUT["layout"] == "CONTIGUOUS" and LAYOUT["dtype"] == "H5T_IEEE_F64LE"
    assert LAYOUT["columns"] == ["id", "k", "m", "Reff", "r", "vr", "vt", "binind"]
    t = raw_table()
    n = LAYOUT["attributes"]["NOBJ"]["value"]
    assert t.shape == (n + 2, 8)
    c = {k: t[:, i] for i, k in enumerate(LAYOUT["columns"])}
    assert np.array_equal(c["id"][1:n + 1], np.arange(1, n + 1))
    assert abs(np.sum(c["m"][1:n + 1]) - 1.0) < 1e-12
    assert np.all(np.diff(c["r"][1:n + 1]) > 0)
    assert c["r"][n + 1] >= 1e30 and c["m"][n + 1] == 0.0
    ke = 0.5 * np.sum(c["m"][1:n + 1] * (c["vr"][1:n + 1] ** 2 + c["vt"][1:n + 1] ** 2))
    assert abs(ke - 0.25) < 1e-12
    assert int(np.sum(c["binind"][1:n + 1] > 0)) == LAYOUT["attributes"]["NBINARY"]["value"]


def test_header_matches_h5dump(io):
    h = io.read_hdf5_header(H5)
    a = LAYOUT["attributes"]
    assert h.nobj == a["NOBJ"]["value"] and h.nbinary == a["NBINARY"]["value"]
    assert (h.rows, h.cols) == tuple(LAYOUT["shape"])
    assert h.rvir == float(a["RVIR"]["value"])          # int64 in the file, read as double
    assert h.rtid == a["RTID"]["value"] and h.mclus == a["MCLUS"]["value"] and h.z == a["Z"]["value"]
    assert h.cenma_m == raw_table()[0, 2] * h.nobj


def test_objects_match_file_bytes(io):
    cl, ex = io.read_hdf5(H5, columns=True)
    t = raw_table()
    n = cl.n
    assert n == LAYOUT["attributes"]["NOBJ"]["value"]
    sl = slice(1, n + 1)
    assert np.array_equal(cl["id"][sl], t[sl, 0].astype(np.int64))
    assert np.array_equal(cl["m"][sl], t[sl, 2] * float(n))
    for k, col in (("r", 4), ("vr", 5), ("vt", 6)):
        assert np.array_equal(cl[k][sl], t[sl, col]), k
    assert np.array_equal(ex["k"][sl], t[sl, 1].astype(np.int32))
    assert np.array_equal(ex["Reff"][sl], t[sl, 3])
    assert np.array_equal(ex["binind"][sl], t[sl, 7].astype(np.int64))
    # sentinels and the central mass (cmc_fits.c:127-136)
    assert cl["r"][0] == 1e-20 and cl["r"][n + 1] == 1e10
    assert cl["m"][0] == 0.0 and cl["m"][n + 1] == 0.0
    assert cl.rtid == LAYOUT["attributes"]["RTID"]["value"] and cl.cenma_m == t[0, 2] * n
    # the scaled masses sum to N (Mtotal = 1 in N-body units after the 1/N of the step)
    assert abs(np.sum(cl["m"][sl]) - n) < 1e-8 * n


def test_error_codes(io, tmp_path):
    with pytest.raises(io.CmcIoError) as e:
        io.read_hdf5(str(tmp_path / "missing.hdf5"))
    assert e.value.code == -1
    with pytest.raises(io.CmcIoError) as e:
        io.read_hdf5(os.path.join(HERE, "golden", "ci_input_layout.json"))
    assert e.value.code == -1
    # caller's n different from NOBJ
    import ctypes as C
    h = io.read_hdf5_header(H5)
    small = io.Cluster(h.nobj - 1)
    rc = io._io().cmc_hdf5_read_objects(os.fsencode(H5), C.byref(h), C.byref(small.soa()), None, None, None)
    assert rc == -3


@pytest.mark.parametrize("fault", ["r_not_ascending", "negative_mass"])
def test_table_breaking_layout_is_rejected(io, tmp_path, fault):
    """A copy of the CI file with two radii swapped, or one mass made negative, in
    the table's bytes: the reader refuses it (the reference would run on)."""
    b = bytearray(open(H5, "rb").read())
    t = np.frombuffer(b, "<f8", count=LAYOUT["size"] // 8, offset=LAYOUT["offset"]).reshape(LAYOUT["shape"]).copy()
    if fault == "r_not_ascending":
        t[100, 4], t[101, 4] = t[101, 4], t[100, 4]
    else:
        t[50, 2] = -t[50, 2]
    b[LAYOUT["offset"]:LAYOUT["offset"] + LAYOUT["size"]] = t.astype("<f8").tobytes()
    p = tmp_path / "bad.hdf5"
    p.write_bytes(bytes(b))
    with pytest.raises(io.CmcIoError) as e:
        io.read_hdf5(str(p))
    assert e.value.code == -6


def test_missing_object_table_is_rejected(io, tmp_path):
    """A file holding only the binary table (made with the HDF5 library's h5copy)."""
    h5copy = shutil.which("h5copy") or "/opt/conda/bin/h5copy"
    if not os.path.exists(h5copy):
        pytest.skip("h5copy not available")
    p = tmp_path / "binaries_only.hdf5"
    subprocess.run([h5copy, "-i", H5, "-o", str(p), "-s", "CLUS_BINARY_DATA", "-d", "CLUS_BINARY_DATA"],
                   check=True, capture_output=True)
    with pytest.raises(io.CmcIoError) as e:
        io.read_hdf5(str(p))
    assert e.value.code == -2
