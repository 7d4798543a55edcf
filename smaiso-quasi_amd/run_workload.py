#!/usr/bin/env python3
"""Run one of the named workloads on the GPU and report throughput and accuracy.

  python smaiso-quasi_amd/run_workload.py interface256 --prec fp64 --steps 1000
  python smaiso-quasi_amd/run_workload.py tilted512 --compare --steps 1000
  python smaiso-quasi_amd/run_workload.py sphere1024 --steps 500 --compare-at 50

--compare runs the FP64 path beside the mixed one and reports the relative L2
difference of psi, rho and v (as a vector) and the relative energy and mass
differences every --every steps; --compare-at K compares at step K only (the
FP64 handle is then freed and the mixed run continues alone).  --divv-switch
S overrides the workload's divv_switch (the step after which D is computed,
quasi.cpp:1014-1965), so a short run reaches the later regime.  --autotune picks the plain
passes' tile and radix variants first, as bench.py does.  Output: one JSON
object per report line on stdout.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import ics  # noqa: E402
import qsm  # noqa: E402


def rel(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / max(np.sum(b * b), 1e-300)))


def compare(a, b):
    out = {f: rel(a.field(f), b.field(f)) for f in ("PSI", "RHO")}
    num = den = 0.0
    for f in ("VX", "VY", "VZ"):
        x, y = a.field(f), b.field(f)
        num += float(np.sum((x - y) ** 2))
        den += float(np.sum(y * y))
    out["V"] = float(np.sqrt(num / max(den, 1e-300)))
    da, db = a.diag(), b.diag()
    out["energy"] = abs(da["energy"] - db["energy"]) / abs(db["energy"])
    out["mass"] = abs(da["mass"] - db["mass"]) / abs(db["mass"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=sorted(ics.WORKLOADS))
    ap.add_argument("--prec", choices=("fp64", "mixed"), default="mixed")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--compare", action="store_true")
    ap.add_argument("--compare-at", type=int, default=0)
    ap.add_argument("--divv-switch", type=int, default=None)
    ap.add_argument("--autotune", action="store_true",
                    help="pick the plain passes' tile and radix variants first (as bench.py does)")
    args = ap.parse_args()
    w = ics.WORKLOADS[args.workload]
    steps = args.steps or w["steps"]
    over = {} if args.divv_switch is None else {"divv_switch": args.divv_switch}
    d = ics.params_dict(args.workload, **over)
    p = qsm.Params(**{k: d[k] for k, _ in qsm.Params._fields_})
    t0 = time.time()
    psi = ics.workload_psi(args.workload)
    prec = qsm.MIXED if args.prec == "mixed" else qsm.FP64
    sim = qsm.Sim(p, prec)
    if args.autotune:
        sim.autotune()
    sim.set_psi(psi)
    ref = None
    if args.compare or args.compare_at:
        ref = qsm.Sim(p, qsm.FP64)
        ref.set_psi(psi)
    del psi
    print(json.dumps({"workload": args.workload, "grid": w["shape"], "prec": args.prec,
                      "setup_s": round(time.time() - t0, 1)}), flush=True)
    pts = float(np.prod(w["shape"]))
    done = 0
    chunk = args.compare_at if args.compare_at else args.every
    while done < steps:
        k = min(chunk, steps - done)
        t = time.time()
        sim.step(k)
        dt = time.time() - t
        done += k
        line = {"step": done, "updates_per_s": pts * k / dt, "ms_per_step": dt / k * 1e3}
        if ref is not None:
            ref.step(k)
            line["vs_fp64"] = compare(sim, ref)
            if args.compare_at:
                ref.close()
                ref = None
                chunk = args.every
        line["diag"] = sim.diag()
        print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
