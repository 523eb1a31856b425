"""Writes tests/golden/rng_golden.json from the reference's own RNG.

The reference file src/libs/taus113-v2.c is compiled from source by
`make -C oracle ref` into oracle/_ref/libtaus113_ref.so (outputs stay out of git);
this script calls it through ctypes and records known-answer vectors:
seeded states, the first draws (integer and double), and the states / draws after
the reference's own jump polynomials JPoly_2_20 and JPoly_2_80.  The testjumping()
known answer of src/libs/taus113-v2.c:272-292 (seed 123456) is included.
"""
import ctypes as C
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import oracle  # noqa: E402

L = oracle.ref_taus_lib()
if L is None:
    sys.exit("oracle/_ref/libtaus113_ref.so missing: run `make -C oracle ref` first")
S = L.RefState
jp20 = (C.c_uint * 4).in_dll(L, "JPoly_2_20")
jp80 = (C.c_uint * 4).in_dll(L, "JPoly_2_80")

out = {"source": "src/libs/taus113-v2.c compiled from the reference tree", "seeds": {}}
for seed in [0, 1, 2, 1234, 123456, 4294967295]:
    st = S()
    L.reset_rng_t113_new(seed, C.byref(st))
    rec = {"state": [int(x) for x in st.z]}
    s2 = S(); s2.z[:] = st.z[:]
    rec["ints"] = [int(L.rng_t113_int_new(C.byref(s2))) for _ in range(16)]
    s3 = S(); s3.z[:] = st.z[:]
    rec["dbls"] = [float(L.rng_t113_dbl_new(C.byref(s3))) for _ in range(8)]
    for name, jp in (("jump20", jp20), ("jump80", jp80)):
        j = L.rng_t113_jump(st, jp)
        rec[name + "_state"] = [int(x) for x in j.z]
        rec[name + "_ints"] = [int(L.rng_t113_int_new(C.byref(j))) for _ in range(4)]
    out["seeds"][str(seed)] = rec
out["jpoly_2_20"] = [int(x) for x in jp20]
out["jpoly_2_80"] = [int(x) for x in jp80]
# testjumping(): the draw after 2^20 steps from seed 123456, by stepping
st = S(); L.reset_rng_t113_new(123456, C.byref(st))
for _ in range(1 << 20):
    L.rng_t113_int_new(C.byref(st))
out["testjumping_123456_after_2p20"] = int(L.rng_t113_int_new(C.byref(st)))
with open(os.path.join(HERE, "rng_golden.json"), "w") as f:
    json.dump(out, f, indent=1)
print("wrote", os.path.join(HERE, "rng_golden.json"), out["testjumping_123456_after_2p20"])
