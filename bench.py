#!/usr/bin/env python3
"""Benchmark of the smectic/isotropic quasi-incompressible time step on one MI355X.

Workload (metric: grid-point updates/s at 512^3, mixed vs FP64 on one GPU):
  tilted512 -- 512^3 grid, smectic slab tilted 15 degrees about x with an
  undulated interface, seeded noise 1e-4; quasi step (24 r2r transforms per
  step), 1000 steps by default: the divergence and pressure terms switch on
  after step 200 (divv_switch), so the timed steps cover both regimes in the
  workload's proportions.  The mixed-precision path is the headline `value`;
  the FP64 path runs beside it on the same input, and the relative L2
  differences of psi, rho and u (as a vector) and the relative free-energy
  and mass differences between the two at the end are reported
  (`accuracy_mixed_vs_fp64`).

A step is one pass of the hot path (quasi.cpp:1017-1965) over the whole grid.
Inputs are resident in HBM before timing starts; timing is wall clock around
exactly `--steps` steps, with a device synchronisation on both sides (the
steps before and after the switch are timed as two back-to-back segments,
whose sum is the reported time).

Also reported:
  roofline      the pass kernel with the largest share of device time (chosen
                over the warmup steps, which time every launch): algorithmic
                HBM bytes of its launches in the timed region / their HIP-event
                time (events recorded on the stream the passes run on, around
                that kernel's launches only)
  passes_GBs    algorithmic bytes / device time of every pass launch of the
                warmup steps (of the timed steps when --warmup is 0)
  cpu_baseline  the CPU oracle (oracle/, FP64 C restatement of the same step)
                on this host's cores, on a bounded 64^3 sample
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "smaiso-quasi_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per grid-point update of the quasi step (24 transforms,
# every pointwise stage fused into a transform pass): 194 field sweeps
SWEEPS_PER_STEP = 194

LOAD_OPS = ["plain", "zero", "coef_f", "coef_g", "grad_x", "grad_y", "grad_z", "lapgrad_x",
            "lapgrad_y", "lapgrad_z", "sh", "rhostar", "divv",
            "stress", "pressure", "vel0", "vel1", "vel2", "scale", "nonlin", "nonlin_p",
            "stress_stb", "nonlin_stb", "nonlin_stb0", "stress3", "stress_stb3"]
# timer categories 40 + g of the x-pass group kernels (one kernel, several outputs)
GROUPS = ["psi", "rho", "proj", "stb_psi", "vel"]


def load_op_name(code):
    if code < len(LOAD_OPS):
        return LOAD_OPS[code]
    g = code - 40
    return f"group {GROUPS[g]}" if 0 <= g < len(GROUPS) else f"op {code}"
STORE_OPS = ["plain", "filter", "update", "update_stb"]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(symbol):
    """HBM bytes per launch of `symbol` measured with rocprofv3 counters
    (FETCH_SIZE + WRITE_SIZE, each calibrated on a copy kernel of known size)
    and recorded in profiles/traffic.json; None when that kernel was not
    measured."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(symbol)
    return None if k is None else float(k["bytes_per_launch"])


def make_sim(qsm, ics, name, prec, shape):
    d = ics.params_dict(name)
    d["nx"], d["ny"], d["nz"] = shape
    p = qsm.Params(**{k: d[k] for k, _ in qsm.Params._fields_})
    return qsm.Sim(p, prec)


def time_steps(sim, steps, first, switch):
    """Wall time of `steps` steps starting at step `first`, in two segments
    split at the divergence switch: returns (total, pre, n_pre, post, n_post)."""
    n_pre = max(0, min(steps, switch - first + 1))
    sim.step(0)  # synchronises the stream
    t0 = time.perf_counter()
    if n_pre:
        sim.step(n_pre)  # returns after a stream synchronisation
    t1 = time.perf_counter()
    if steps - n_pre:
        sim.step(steps - n_pre)
    t2 = time.perf_counter()
    return t2 - t0, t1 - t0, n_pre, t2 - t1, steps - n_pre


def ensemble_run(qsm, ics, steps):
    """Stability ensemble: 96 quasi-stb instances, grid edges from
    {32, 48, 64, 96, 128} per axis, perturbation wavenumbers U(0.05, 1),
    `steps` steps in FP64 and in mixed precision, each as batched kernels
    replayed from one HIP graph per step signature."""
    members = ics.ensemble_members(96, seed=4)
    params = [qsm.Params(**{k: d[k] for k, _ in qsm.Params._fields_}) for d, _ in members]
    pts = float(sum(d["nx"] * d["ny"] * d["nz"] for d, _ in members))
    res = {"workload": "96 stb instances, edges {32,48,64,96,128}, Qi ~ U(0.05,1), amplitude 1e-3",
           "steps": steps, "grid_points": pts}
    growth = {}
    for prec, tag in ((qsm.FP64, "fp64"), (qsm.MIXED, "mixed")):
        e = qsm.Ensemble(params, prec)
        e.autotune()
        for i, (_, psi) in enumerate(members):
            e.set_psi(i, psi)
        e.step(2)  # first two steps capture the step signatures
        t0 = time.perf_counter()
        e.step(steps - 2)
        dt = time.perf_counter() - t0
        growth[tag] = e.growth()
        e.close()
        res[tag] = {"updates_per_s": pts * (steps - 2) / dt, "ms_per_step": dt / (steps - 2) * 1e3}
    gf, gm = growth["fp64"], growth["mixed"]
    res["growth_fp64_first8"] = [float(v) for v in gf[:8]]
    res["growth_mixed_vs_fp64_max_rel"] = float(np.max(np.abs(gm - gf)) / np.max(np.abs(gf)))
    return res


def cpu_baseline(target_s=15.0):
    """FP64 CPU oracle on a bounded sample: 64^3 interface workload, P = 4 halo map."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import ics
    oracle.build()
    d = ics.params_dict("interface64")
    st = oracle.State(oracle.make_params(**d))
    st.set_psi(ics.workload_psi("interface64"))
    st.step(1)
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < target_s:
        st.step(5)
        steps += 5
    dt = time.perf_counter() - t0
    pts = d["nx"] * d["ny"] * d["nz"]
    return {"value": pts * steps / dt, "unit": "grid-pt updates/s", "cores": oracle.lib().orc_num_threads(),
            "kind": "port",
            "sample": f"oracle FP64 quasi step, 64^3 interface workload, {steps} steps in {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=512, help="cubic grid edge (default 512)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ens-steps", type=int, default=500)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path is single-GPU by design)", file=sys.stderr)
        return 2

    import qsm
    import ics
    shape = (args.size,) * 3
    name = "tilted512"
    log(f"building initial condition {shape}")
    psi0 = ics.workload_psi(name, shape=shape)

    res = {}
    fields = {}
    top = None
    for prec, tag in ((qsm.MIXED, "mixed"), (qsm.FP64, "fp64")):
        if prec == qsm.FP64 and args.no_fp64:
            continue
        log(f"{tag}: create + setup")
        sim = make_sim(qsm, ics, name, prec, shape)
        sim.autotune()
        sim.set_psi(psi0)
        # the warmup steps are timed per launch (every kernel): they give the
        # aggregate pass bandwidth and the kernel with the largest device
        # time, whose launches alone are then timed with HIP events through
        # the timed steps (an event pair per launch of every kernel would
        # add ~0.4 ms per step to the timed region)
        sim.timing(True)
        sim.step(args.warmup)
        kt = sim.timing_get(0)
        if args.warmup > 0:
            w = sim.timing_kernel(0)
            sim.timing_select(w["axis"], w["inv"], w["ld"], w["st"], w["variant"])
        else:
            sim.timing(True)
        log(f"{tag}: timing {args.steps} steps")
        dt, t_pre, n_pre, t_post, n_post = time_steps(sim, args.steps, args.warmup + 1,
                                                      sim.params.divv_switch)
        if args.warmup == 0:
            kt = sim.timing_get(0)
        if prec == qsm.MIXED:
            top = sim.timing_kernel(0)
        sim.timing(False)
        # the fields compared between the precisions (FP64 host copies; only
        # when both runs happen, since each 512^3 copy is 1 GiB)
        fields[tag] = {f: sim.field(f) for f in (("PSI", "RHO", "VX", "VY", "VZ") if not args.no_fp64
                                                 else ("PSI",))}
        diag = sim.diag()
        sim.close()
        res[tag] = {"s": dt, "kt": kt, "diag": diag,
                    "regimes": {"steps_before_switch": n_pre,
                                "ms_per_step_before_switch": t_pre / n_pre * 1e3 if n_pre else None,
                                "steps_after_switch": n_post,
                                "ms_per_step_after_switch": t_post / n_post * 1e3 if n_post else None}}
        log(f"{tag}: {dt / args.steps * 1e3:.2f} ms/step")

    pts = float(np.prod(shape))
    m = res["mixed"]
    value = pts * args.steps / m["s"]
    k = top
    achieved = k["bytes"] / (k["ms"] * 1e-3) / 1e9 if k["ms"] > 0 else 0.0
    n_ax = shape[k["axis"]]
    if k["ld"] >= 40:  # x-pass group kernel
        symbol = f"void qsm::k_group<float, {n_ax}, {k['ld'] - 40}>(qsm::GroupArgs)"
    elif LOAD_OPS[k["ld"]] in ("stress3", "stress_stb3"):  # multi-output z pass
        symbol = f"void qsm::k_multi<float, {n_ax}, {k['ld']}>(qsm::MultiArgs)"
    else:
        symbol = (f"void qsm::k_pass<float, {n_ax}, {k['axis']}, {'true' if k['inv'] else 'false'}, "
                  f"{k['ld']}, {k['st']}, {k['variant']}>(qsm::PassArgs)")
    direction = "inverse" if k["inv"] else "forward"
    out = {
        "metric": "grid-pt updates/s at 512^3, mixed vs FP64 on 1 GPU; % HBM roofline; psi rel-L2",
        "value": value,
        "unit": "grid-pt updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["s"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (tilted-slab initial condition, seeded noise)",
        "config": {"workload": f"{name}: quasi step, {shape[0]}x{shape[1]}x{shape[2]}, mixed precision "
                               "(FP32 fields/transforms, FP64 psihat, coefficients, mean modes)",
                   "grid": list(shape), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(symbol),
                     "kernel": symbol,
                     "pass": f"{direction} {'xyz'[k['axis']]} pass, load op {load_op_name(k['ld'])}, "
                             f"store op {STORE_OPS[k['st']]}",
                     "avg_launch_ms": k["ms"] / max(1, k["launches"]),
                     "launches": k["launches"],
                     "bytes_per_launch": k["bytes"] / max(1, k["launches"])},
        "regimes": m["regimes"],
        "passes_GBs": m["kt"]["bytes"] / (m["kt"]["ms"] * 1e-3) / 1e9 if m["kt"]["ms"] > 0 else 0.0,
        "step_roofline": {"bytes_per_update": SWEEPS_PER_STEP * 4,
                          "achieved_GBs": value * SWEEPS_PER_STEP * 4 / 1e9,
                          "frac": value * SWEEPS_PER_STEP * 4 / 1e9 / HBM_PEAK_GBS},
    }
    if "fp64" in res:
        f = res["fp64"]
        v64 = pts * args.steps / f["s"]
        a, b = fields["mixed"], fields["fp64"]
        out["fp64"] = {"value": v64, "ms_per_step": f["s"] / args.steps * 1e3, "regimes": f["regimes"],
                       "step_roofline_frac": v64 * SWEEPS_PER_STEP * 8 / 1e9 / HBM_PEAK_GBS}
        out["mixed_over_fp64"] = value / v64
        def rel(fs):
            num = sum(float(np.sum((a[f] - b[f]) ** 2)) for f in fs)
            return float(np.sqrt(num / sum(float(np.sum(b[f] ** 2)) for f in fs)))
        out["psi_rel_l2_mixed_vs_fp64"] = rel(("PSI",))
        dm, df = res["mixed"]["diag"], f["diag"]
        out["accuracy_mixed_vs_fp64"] = {
            "psi_rel_l2": rel(("PSI",)), "rho_rel_l2": rel(("RHO",)),
            "u_rel_l2": rel(("VX", "VY", "VZ")),  # velocity as a vector
            "energy_rel": abs(dm["energy"] - df["energy"]) / abs(df["energy"]),
            "mass_rel": abs(dm["mass"] - df["mass"]) / abs(df["mass"])}
        out["steps_compared"] = args.warmup + args.steps
    if not args.no_ensemble:
        log("ensemble")
        out["ensemble"] = ensemble_run(qsm, ics, args.ens_steps)
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
