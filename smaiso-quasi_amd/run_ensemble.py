#!/usr/bin/env python3
"""Run the stability-ensemble workload on the GPU: 96 quasi-stb instances with
grid edges drawn from {32, 48, 64, 96, 128} per axis, each a flat interface
perturbed by one mode of wavenumber Qi ~ U(0.05, 1) (quasi-stb.cpp:529-552),
advanced together as batched kernels.

  python smaiso-quasi_amd/run_ensemble.py --prec mixed --steps 500
  python smaiso-quasi_amd/run_ensemble.py --prec fp64 --kernels 4

Reports ms/step and grid-point updates/s over the steps after the first
3 + --graph-steps (the single-step signatures and the capture of the
multi-step graph).  --autotune picks tile and radix variants per pencil
length first.  --kernels K times K more steps eagerly with HIP events around
every batched launch and prints the kernel table (device ms per step, mean
us per launch, algorithmic bytes per launch and their rate).  Output: JSON
lines on stdout.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import ics  # noqa: E402
import qsm  # noqa: E402



def symbol(k, fp32):
    t = "float" if fp32 else "double"
    if k["ld"] == 63:
        return "qsm::k_finish_batch (FP64 mean-mode sums)"
    # forward x passes that also run the x inverse of what they store
    kern = "k_pass_batch_sx" if qsm.STORE_OPS[k["st"]].endswith("_x") else "k_pass_batch"
    return (f"void qsm::{kern}<{t}, {k['n']}, {k['axis']}, {'true' if k['inv'] else 'false'}, "
            f"{k['ld']}, {k['st']}, {k['variant']}>")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--prec", choices=("fp64", "mixed"), default="mixed")
    ap.add_argument("--count", type=int, default=96)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--graph-steps", type=int, default=16)
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--kernels", type=int, default=0)
    args = ap.parse_args()
    prec = qsm.MIXED if args.prec == "mixed" else qsm.FP64
    members = ics.ensemble_members(args.count, seed=args.seed)
    params = [qsm.Params(**{k: d[k] for k, _ in qsm.Params._fields_}) for d, _ in members]
    pts = float(sum(d["nx"] * d["ny"] * d["nz"] for d, _ in members))
    e = qsm.Ensemble(params, prec)
    if args.autotune:
        e.autotune()
    e.graph_steps(args.graph_steps)
    for i, (_, psi) in enumerate(members):
        e.set_psi(i, psi)
    warm = min(args.steps, 3 + args.graph_steps)
    e.step(warm)
    t0 = time.perf_counter()
    e.step(args.steps - warm)
    dt = time.perf_counter() - t0
    timed = args.steps - warm
    print(json.dumps({"workload": f"{args.count} stb instances", "prec": args.prec,
                      "grid_points": pts, "steps": args.steps, "steps_timed": timed,
                      "graph_steps": args.graph_steps, "autotune": args.autotune,
                      "ms_per_step": dt / max(1, timed) * 1e3,
                      "updates_per_s": pts * timed / dt if timed else None}), flush=True)
    if args.kernels:
        e.timing(True)
        e.step(args.kernels)
        rank = 0
        while True:
            try:
                k = e.timing_kernel(rank)
            except qsm.QsmError:
                break
            rank += 1
            per = k["ms"] / k["launches"] * 1e3
            print(json.dumps({"kernel": symbol(k, prec == qsm.MIXED),
                              "load_op": ("finish" if k["ld"] == 63 else qsm.LOAD_OPS[k["ld"]]),
                              "store_op": qsm.STORE_OPS[k["st"]],
                              "ms_per_step": k["ms"] / args.kernels,
                              "launches_per_step": k["launches"] / args.kernels,
                              "us_per_launch": per,
                              "alg_bytes_per_launch": k["bytes"] / k["launches"],
                              "alg_GBs": k["bytes"] / (k["ms"] * 1e-3) / 1e9 if k["ms"] else 0.0}),
                  flush=True)
        e.timing(False)
    e.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
