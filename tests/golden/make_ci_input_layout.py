"""Writes ci_input_layout.json: what the HDF5 library's own dump tool (h5dump)
prints about tests/golden/ci_input.hdf5 (a copy of the reference's ci/input.hdf5,
the input of its CI run): the byte offset and shape of the contiguous object table
CLUS_OBJ_DATA/block0_values, its column names, and its scalar attributes at full
precision.  The reader tests use it to check the C reader against a plain numpy
view of the file's bytes.

    python tests/golden/make_ci_input_layout.py [path/to/h5dump]
"""
import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
H5 = os.path.join(HERE, "ci_input.hdf5")
DSET = "CLUS_OBJ_DATA/block0_values"


def dump(h5dump, *args):
    return subprocess.run([h5dump, *args, H5], check=True, capture_output=True, text=True).stdout


def main():
    h5dump = sys.argv[1] if len(sys.argv) > 1 else "/opt/conda/bin/h5dump"
    out = {"file": "ci_input.hdf5", "dataset": DSET}
    hdr = dump(h5dump, "-p", "-H", "-d", DSET)
    out["layout"] = re.search(r"STORAGE_LAYOUT \{\s*(\w+)", hdr).group(1)
    out["offset"] = int(re.search(r"OFFSET (\d+)", hdr).group(1))
    out["size"] = int(re.search(r"SIZE (\d+)", hdr).group(1))
    out["dtype"] = re.search(r"DATATYPE\s+(H5T_\w+)", hdr).group(1)
    out["shape"] = [int(x) for x in re.search(r"SIMPLE \{ \( (\d+), (\d+) \)", hdr).groups()]
    items = dump(h5dump, "-d", "CLUS_OBJ_DATA/block0_items")
    out["columns"] = re.findall(r'"(\w+)"', re.search(r"DATA \{\s*\(0\): (.*)", items).group(1))
    attrs = {}
    for a in ("NOBJ", "NBINARY", "RVIR", "RTID", "MCLUS", "Z"):
        t = dump(h5dump, "-m", "%.17g", "-a", f"{DSET}/{a}")
        v = re.search(r"\(0\): (\S+)", t).group(1)
        typ = re.search(r"DATATYPE\s+(H5T_\w+)", t).group(1)
        attrs[a] = {"type": typ, "value": int(v) if "STD_I" in typ else float(v)}
    out["attributes"] = attrs
    with open(os.path.join(HERE, "ci_input_layout.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
