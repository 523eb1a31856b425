"""Energy drift of the oracle over 100 Henon steps (Plummer N=2e4, seed 1, per-star
RNG streams, long-double Q) under the reference's own math-library variants: the C
library's sin() and cos() (the documented, unoptimised build), sincos() for
same-argument pairs (what an optimising gcc makes of the same source) and long
double sin/cos rounded to double; and under the device's double-double sin/cos.
The largest difference among the reference's variants is the drift's sensitivity
to last-bit libm differences, the floor against which the device's drift is
compared with the reference's.

    python tests/golden/make_drift_golden.py   ->  tests/golden/drift_golden.json
"""
import importlib.util
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

spec = importlib.util.spec_from_file_location("cmc_amd", os.path.join(ROOT, "cmc-cosmic_amd", "__init__.py"))
cm = importlib.util.module_from_spec(spec)
spec.loader.exec_module(cm)

N, STEPS, SEED = 20000, 100, 1


def drift(libm):
    cl = cm.make_plummer(N, seed=SEED)
    o = O.Oracle(cl.a, cl.n, orbit_r=cl.rtid, rng_mode=O.RNG_PER_STAR, q_mode=O.Q_LONG_DOUBLE, libm_mode=libm)
    o.init()
    e0 = o.sc("Etot")
    for _ in range(STEPS):
        assert o.step() == 0
    return (o.sc("Etot") + o.sc("Eoops") - e0) / abs(e0), o.N


if __name__ == "__main__":
    out = {"n": N, "steps": STEPS, "seed": SEED}
    for name, lm in (("system", O.LIBM_SYSTEM), ("sincos", O.LIBM_SINCOS), ("long_double", O.LIBM_LONG_DOUBLE),
                     ("dd", O.LIBM_DD)):
        d, n = drift(lm)
        out[name] = {"drift": d, "n_live": n}
        print(name, d, n, flush=True)
    ref = [out[k]["drift"] for k in ("system", "sincos", "long_double")]
    out["libm_spread"] = max(ref) - min(ref)
    with open(os.path.join(HERE, "drift_golden.json"), "w") as f:
        json.dump(out, f, indent=1)
