"""Benchmark: stars/s per Henon step on one MI355X.

Workload (the configuration the metric is quoted on): King W0 = 7 cluster,
N = 4e6, Kroupa IMF (alpha 1.3 on 0.08-0.5, 2.3 on 0.5-150 Msun), heaviest 1% of
masses placed in the inner 10% of radii, FP64, plain GPU step (sort, potential,
relaxation, orbit sampling and the step glue between them), seeded synthetic data.

Prints ONE JSON line.  value = live stars processed per second of wall time over
the timed steps (inputs resident in HBM).  roofline: the phase with the largest
device time, its algorithmic bytes per launch over its HIP-event time.  cpu_baseline:
the CPU oracle (a restatement of the reference's serial path, reference RNG order,
x87 long-double Q) timed on a bounded sample on this host.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0

# algorithmic bytes per live star for each phase (8-byte doubles and ids)
ALG_BYTES = {
    "sort": 160.0,        # key + 72 B payload, read once and written once
    "potential": 24.0,    # read r, m; write phi
    "relax": 72.0,        # read r, m, phi, vr, vt; write vr, vt, E, J
    "orbits": 96.0,       # read E, J, r, m, vr, vt; write rnew, vrnew, vtnew, r_peri, r_apo, X
}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def load_package():
    spec = importlib.util.spec_from_file_location("cmc_cosmic_amd", os.path.join(REPO, "cmc-cosmic_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def cpu_baseline(cl):
    """One Henon step of the oracle (serial C restatement of the reference path,
    reference RNG order, x87 long-double Q) on the benchmark cluster itself."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    o = O.Oracle(cl.a, cl.n, orbit_r=cl.rtid, rng_mode=O.RNG_REF_SEQ, q_mode=O.Q_LONG_DOUBLE)
    o.init()
    n0 = o.N
    t = time.perf_counter()
    o.step()
    dt = time.perf_counter() - t
    return {"value": n0 / dt, "unit": "stars/s", "cores": 1, "kind": "port",
            "sample": f"one Henon step of the same {cl.n:.0e}-star King W0=7 cluster on one host core "
                      f"(CPU oracle: serial, reference RNG order, long-double Q), {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=4_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs the hot path on one GPU; --gpus above 1 is not supported "
                                   "(multi-GPU sharding is a design in DESIGN.md only)"}))
        sys.exit(2)

    cm = load_package()
    t0 = time.perf_counter()
    cl = cm.make_king(args.n, 7.0, seed=1, kroupa=True, segregate=True)
    log(f"generated King W0=7 N={args.n} in {time.perf_counter() - t0:.1f} s, rtid={cl.rtid:.3f}")
    params = cm.default_params(cl.n, orbit_r=cl.rtid, seed=1)
    ctx = cm.Context(cl.n, params)
    ctx.upload(cl)
    ctx.init_cluster()
    for i in range(args.warmup):
        ctx.step()
    log(f"warmup {args.warmup} steps done, N={ctx.scalars()['n_max']}")
    ctx.timer_reset()
    ctx.timer_enable(True)
    live = []
    times = []
    ctx.sync()
    t_start = time.perf_counter()
    for i in range(args.steps):
        n_live = ctx.scalars()["n_max"] if i == 0 else live_next
        ts = time.perf_counter()
        st = ctx.step()
        te = time.perf_counter()
        times.append(te - ts)
        live.append(n_live)
        live_next = ctx.scalars()["n_max"]
    ctx.sync()
    total = time.perf_counter() - t_start
    ctx.timer_enable(False)
    log(f"timed {args.steps} steps in {total:.3f} s")
    stars = float(np.sum(live))
    phases = {}
    for name in ["timestep", "relax", "ec1", "orbits", "glue_pre_sort", "ie", "sort", "potential", "glue_post_sort"]:
        ms, calls = ctx.timer_get(name)
        if calls:
            per = ms / calls
            ent = {"ms": per}
            if name in ALG_BYTES:
                gbs = ALG_BYTES[name] * np.mean(live) / (per * 1e-3) / 1e9
                ent.update({"alg_bytes_per_star": ALG_BYTES[name], "alg_GBs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS})
            phases[name] = ent
    dom = max((k for k in phases if k in ALG_BYTES), key=lambda k: phases[k]["ms"])
    d = phases[dom]
    roofline = {"kernel": dom, "bound": "hbm", "achieved": d["alg_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": d["alg_GBs"] / HBM_PEAK_GBS, "traffic": None,
                "alg_bytes_per_launch": ALG_BYTES[dom] * float(np.mean(live))}
    cpu = None
    if not args.no_cpu_baseline:
        log(f"cpu baseline: oracle step on the same {cl.n} stars")
        cpu = cpu_baseline(cl)
        log(f"cpu baseline {cpu['value']:.3e} stars/s")
    ms_per_step = total / args.steps * 1e3
    out = {
        "metric": "stars/s per Henon step at N=4e6 (plain vs learned-bracket), % HBM peak",
        "value": stars / total,
        "unit": "stars/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "median_ms_per_step": float(np.median(times) * 1e3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded King W0=7 generator)",
        "config": {"workload": "King W0=7, N=4e6, Kroupa IMF, heaviest 1% in inner 10%, plain GPU Henon step",
                   "n_stars": args.n, "n_live_mean": float(np.mean(live)), "seed": 1, "step": "plain"},
        "roofline": roofline,
        "phases": phases,
        "cpu_baseline": cpu,
        "last_step_stats": st,
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
