#!/usr/bin/env python3
"""Benchmark of the smectic/isotropic quasi-incompressible time step on one MI355X.

Workload (metric: grid-point updates/s at 512^3, mixed vs FP64 on one GPU):
  tilted512 -- 512^3 grid, smectic slab tilted 15 degrees about x with an
  undulated interface, seeded noise 1e-4; quasi step, 1000 steps in the
  workload.  The divergence D switches on after step 200 (divv_switch) and
  the pressure term after step 201, so 20% of the workload's steps run the
  early regime (18 transforms) and 80% the late one (23 transforms).  The
  mixed-precision path is the headline `value`; the FP64 path runs beside it
  on the same input, and the relative L2 differences of psi, rho and u (as a
  vector) and the relative free-energy and mass differences between the two
  at the end are reported (`accuracy_mixed_vs_fp64`).

A step is one pass of the hot path (quasi.cpp:1017-1965) over the whole grid.
Inputs are resident in HBM before timing starts.  Exactly `--steps` steps
are timed with wall clock and a device synchronisation on both sides, in two
windows that keep the workload's proportions for any step count: round(0.2
K) steps of the early regime right after the warmup, then (after an untimed
advance to step divv_switch + 1) the remaining steps in the late regime.
`value` = grid points x K / (time of both windows); `regimes` gives each
window's ms/step and `value_workload_weighted` their 200/800 mix.

Also reported:
  roofline      the pass kernel with the largest share of device time (chosen
                over the warmup steps, which time every launch): algorithmic
                HBM bytes of its launches in the timed windows / their
                HIP-event time (events recorded on the stream the passes run
                on, around that kernel's launches only)
  step_traffic  per regime, the bytes one step moves: algorithmic (sum over
                the step's launches of the fields each reads and writes) and
                measured (rocprofv3 FETCH_SIZE + WRITE_SIZE per launch of each
                kernel, profiles/traffic.json, times its launches per step),
                each over the window's ms/step and as a fraction of 8 TB/s
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

from qsm import LOAD_OPS, STORE_OPS  # noqa: E402  (operator code names; no library load)

HBM_PEAK_GBS = 8000.0
# tilted512 runs 1000 steps with divv_switch = 200: the share of early steps
EARLY_SHARE = 0.2

# timer categories 40 + g of the x-pass group kernels (one kernel, several outputs)
GROUPS = ["psi", "rho", "proj", "stb_psi", "vel"]


def load_op_name(code):
    if code < len(LOAD_OPS):
        return LOAD_OPS[code]
    g = code - 40
    return f"group {GROUPS[g]}" if 0 <= g < len(GROUPS) else f"op {code}"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_table():
    """{kernel symbol: HBM bytes per launch} measured with rocprofv3 counters
    (FETCH_SIZE + WRITE_SIZE, each calibrated on a copy kernel of known size)
    and recorded in profiles/traffic.json (empty when absent)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return {}
    return {k: float(v["bytes_per_launch"]) for k, v in t.get("kernels", {}).items()}


def measured_traffic(symbol):
    """Measured HBM bytes per launch of `symbol`, or None when that kernel was
    not measured."""
    return traffic_table().get(symbol)


def kernel_symbol(k, shape, fp32):
    """The kernel's name as rocprofv3 reports it, from the timer's fields."""
    t = "float" if fp32 else "double"
    n_ax = shape[k["axis"]]
    if k["ld"] >= 40:  # x-pass group kernel
        return f"void qsm::k_group<{t}, {n_ax}, {k['ld'] - 40}>(qsm::GroupArgs)"
    if k["ld"] < len(LOAD_OPS) and LOAD_OPS[k["ld"]] in ("stress3", "stress_stb3", "stress3_divv"):
        # the quasi stress passes complete lap grad psi's inverse z passes
        # (LZ 1), with the divergence grad rho's too (LZ 2)
        lz = {"stress3": 1, "stress_stb3": 0, "stress3_divv": 2}[LOAD_OPS[k["ld"]]]
        return f"void qsm::k_multi<{t}, {n_ax}, {k['ld']}, {lz}>(qsm::MultiArgs)"
    if k["variant"] >= 4:  # software-pipelined persistent plain pass
        return (f"void qsm::k_pass_pipe<{t}, {n_ax}, {k['axis']}, {'true' if k['inv'] else 'false'}, "
                f"{k['variant']}>(qsm::PassArgs, int)")
    # forward x passes that also run the x inverse of what they store
    kern = "k_pass_sx" if STORE_OPS[k["st"]] in ("filter_x", "update_x", "update_stb_x") else "k_pass"
    return (f"void qsm::{kern}<{t}, {n_ax}, {k['axis']}, {'true' if k['inv'] else 'false'}, "
            f"{k['ld']}, {k['st']}, {k['variant']}>(qsm::PassArgs)")


def kernel_table(sim, steps):
    """Every kernel timed since the last reset: symbol, launches, device ms
    and algorithmic bytes, per step."""
    out = []
    rank = 0
    while True:
        try:
            k = sim.timing_kernel(rank)
        except Exception:  # past the last timed kernel
            break
        rank += 1
        out.append({"symbol": kernel_symbol(k, sim.shape, sim.prec == 1),
                    "symbol_v0": kernel_symbol(dict(k, variant=0), sim.shape, sim.prec == 1),
                    "variant": k["variant"],
                    "launches": k["launches"] / steps, "ms": k["ms"] / steps,
                    "alg_bytes": k["bytes"] / steps})
    return out


def step_traffic(table, ms_per_step, measured):
    """Bytes one step moves, algorithmic and measured (falling back to the
    algorithmic bytes of kernels without a counter measurement), over the
    step's wall time."""
    alg = sum(k["alg_bytes"] for k in table)
    meas = 0.0
    missing, as_default = [], []
    for k in table:
        b = measured.get(k["symbol"])
        if b is None and k.get("variant", 0) >= 2:
            # radix variants 2, 3 change only the in-LDS radix plan, not the
            # access pattern: the default tile's measurement stands for them
            b = measured.get(k["symbol_v0"])
            if b is not None:
                as_default.append(k["symbol"])
        if b is None:
            missing.append(k["symbol"])
            meas += k["alg_bytes"]
        else:
            meas += b * k["launches"]
    gbs = lambda b: b / (ms_per_step * 1e-3) / 1e9
    return {"alg_bytes_per_step": alg, "alg_GBs": gbs(alg), "alg_frac": gbs(alg) / HBM_PEAK_GBS,
            "measured_bytes_per_step": meas, "measured_GBs": gbs(meas),
            "measured_frac": gbs(meas) / HBM_PEAK_GBS,
            "kernels_without_counters": missing,
            "kernels_counted_as_default_plan": as_default,
            "kernel_ms_per_step": sum(k["ms"] for k in table)}


def make_sim(qsm, ics, name, prec, shape):
    d = ics.params_dict(name)
    d["nx"], d["ny"], d["nz"] = shape
    p = qsm.Params(**{k: d[k] for k, _ in qsm.Params._fields_})
    return qsm.Sim(p, prec)


def timed(sim, steps):
    """Wall time of `steps` steps, with a device synchronisation on both sides."""
    sim.step(0)  # synchronises the stream
    t0 = time.perf_counter()
    if steps:
        sim.step(steps)  # returns after a stream synchronisation
    return time.perf_counter() - t0


def run_precision(qsm, ics, name, prec, shape, psi0, args):
    """One precision's run: warmup (every launch timed), an early-regime
    window, an untimed advance past the switches, a late-regime window, one
    more step with every launch timed.  Returns times, kernel tables, the
    dominant kernel's event timing and the final fields."""
    tag = "mixed" if prec == qsm.MIXED else "fp64"
    log(f"{tag}: create + setup")
    sim = make_sim(qsm, ics, name, prec, shape)
    sim.autotune()
    sim.set_psi(psi0)
    switch = sim.params.divv_switch
    k_pre = min(int(round(EARLY_SHARE * args.steps)), max(0, switch - args.warmup))
    k_post = args.steps - k_pre
    # warmup: every launch timed (eager); gives the early-regime kernel table
    # and the kernel with the largest device time, whose launches alone are
    # then timed with HIP events through the timed windows (an event pair
    # per launch of every kernel would add ~0.4 ms per step)
    # (with --warmup 0 nothing runs before the first window: the windows then
    # time no kernel, and the dominant kernel and its event timing come from
    # the one fully timed late step at the end)
    early, sel = [], None
    if args.warmup:
        sim.timing(True)
        sim.step(args.warmup)
        early = kernel_table(sim, args.warmup)
        top = sim.timing_kernel(0)
        sel = (top["axis"], top["inv"], top["ld"], top["st"], top["variant"])
    log(f"{tag}: timing {k_pre} early + {k_post} late steps")
    if sel:
        sim.timing_select(*sel)
    else:
        sim.timing(False)
    t_pre = timed(sim, k_pre)
    k_top = [sim.timing_kernel(0)] if sel and k_pre else []
    # untimed advance (graph replay) to step divv_switch + 1: every later
    # step computes D and the pressure term
    sim.timing(False)
    done = args.warmup + k_pre
    if done < switch + 1:
        sim.step(switch + 1 - done)
        done = switch + 1
    if sel:
        sim.timing_select(*sel)
    t_post = timed(sim, k_post)
    if sel and k_post:
        k_top.append(sim.timing_kernel(0))
    done += k_post
    # one late step with every launch timed: the late-regime kernel table
    sim.timing(True)
    sim.step(1)
    late = kernel_table(sim, 1)
    if not k_top:
        k_top = [sim.timing_kernel(0)]
    done += 1
    sim.timing(False)
    if not early:
        early = late
    fields = {f: sim.field(f) for f in (("PSI", "RHO", "VX", "VY", "VZ") if not args.no_fp64
                                        else ("PSI",))}
    diag = sim.diag()
    sim.close()
    total = t_pre + t_post
    log(f"{tag}: {total / args.steps * 1e3:.2f} ms/step")
    ms_pre = t_pre / k_pre * 1e3 if k_pre else None
    ms_post = t_post / k_post * 1e3 if k_post else None
    top_acc = None
    if k_top:
        top_acc = dict(k_top[0])
        for extra in k_top[1:]:
            for key in ("ms", "launches", "bytes"):
                top_acc[key] += extra[key]
    return {"s": total, "diag": diag, "fields": fields, "top": top_acc, "steps_run": done,
            "regimes": {"steps_before_switch": k_pre, "ms_per_step_before_switch": ms_pre,
                        "steps_after_switch": k_post, "ms_per_step_after_switch": ms_post},
            "tables": {"early": early, "late": late}}


ENS_GRAPH_STEPS = 16


def ensemble_run(qsm, ics, steps):
    """Stability ensemble: 96 quasi-stb instances, grid edges from
    {32, 48, 64, 96, 128} per axis, perturbation wavenumbers U(0.05, 1),
    `steps` steps in FP64 and in mixed precision, each as batched kernels
    (tile and radix variants chosen by the autotuner) replayed from HIP
    graphs of ENS_GRAPH_STEPS steps per launch.  The first 3 + ENS_GRAPH_STEPS
    steps (the single-step signatures and the capture of the block graph)
    are untimed; the rest are timed."""
    members = ics.ensemble_members(96, seed=4)
    params = [qsm.Params(**{k: d[k] for k, _ in qsm.Params._fields_}) for d, _ in members]
    pts = float(sum(d["nx"] * d["ny"] * d["nz"] for d, _ in members))
    warm = 3 + ENS_GRAPH_STEPS
    res = {"workload": "96 stb instances, edges {32,48,64,96,128}, Qi ~ U(0.05,1), amplitude 1e-3",
           "steps": steps, "steps_timed": steps - warm, "graph_steps": ENS_GRAPH_STEPS,
           "grid_points": pts}
    growth = {}
    for prec, tag in ((qsm.FP64, "fp64"), (qsm.MIXED, "mixed")):
        e = qsm.Ensemble(params, prec)
        e.autotune()
        e.graph_steps(ENS_GRAPH_STEPS)
        for i, (_, psi) in enumerate(members):
            e.set_psi(i, psi)
        e.step(warm)  # single-step signatures, then the block graph's capture
        t0 = time.perf_counter()
        e.step(steps - warm)
        dt = time.perf_counter() - t0
        growth[tag] = e.growth()
        e.close()
        res[tag] = {"updates_per_s": pts * (steps - warm) / dt,
                    "ms_per_step": dt / (steps - warm) * 1e3}
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
    measured = traffic_table()

    res = {}
    for prec in (qsm.MIXED, qsm.FP64):
        if prec == qsm.FP64 and args.no_fp64:
            continue
        res["mixed" if prec == qsm.MIXED else "fp64"] = run_precision(qsm, ics, name, prec, shape,
                                                                       psi0, args)

    pts = float(np.prod(shape))

    def summary(r):
        g = r["regimes"]
        out = {"value": pts * args.steps / r["s"], "ms_per_step": r["s"] / args.steps * 1e3,
               "regimes": g}
        if g["ms_per_step_before_switch"] and g["ms_per_step_after_switch"]:
            w = EARLY_SHARE * g["ms_per_step_before_switch"] + (1 - EARLY_SHARE) * g["ms_per_step_after_switch"]
            out["ms_per_step_workload_weighted"] = w
            out["value_workload_weighted"] = pts / (w * 1e-3)
        out["step_traffic"] = {}
        for reg, key in (("early", "ms_per_step_before_switch"), ("late", "ms_per_step_after_switch")):
            if g[key]:
                out["step_traffic"][reg] = step_traffic(r["tables"][reg], g[key], measured)
        return out

    m = res["mixed"]
    ms = summary(m)
    k = m["top"]
    achieved = k["bytes"] / (k["ms"] * 1e-3) / 1e9 if k and k["ms"] > 0 else 0.0
    symbol = kernel_symbol(k, shape, True)
    direction = "inverse" if k["inv"] else "forward"
    out = {
        "metric": "grid-pt updates/s at 512^3, mixed vs FP64 on 1 GPU; % HBM roofline; psi rel-L2",
        "value": ms["value"],
        "unit": "grid-pt updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (tilted-slab initial condition, seeded noise)",
        "config": {"workload": f"{name}: quasi step, {shape[0]}x{shape[1]}x{shape[2]}, mixed precision "
                               "(FP32 fields/transforms, FP64 psihat, coefficients, mean modes); "
                               "timed steps split 20/80 between the regimes before and after "
                               "the divergence switch, as in the 1000-step workload",
                   "grid": list(shape), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured.get(symbol),
                     "kernel": symbol,
                     "pass": f"{direction} {'xyz'[k['axis']]} pass, load op {load_op_name(k['ld'])}, "
                             f"store op {STORE_OPS[k['st']]}",
                     "avg_launch_ms": k["ms"] / max(1, k["launches"]),
                     "launches": k["launches"],
                     "bytes_per_launch": k["bytes"] / max(1, k["launches"])},
        "regimes": ms["regimes"],
        "value_workload_weighted": ms.get("value_workload_weighted"),
        "ms_per_step_workload_weighted": ms.get("ms_per_step_workload_weighted"),
        "step_traffic": ms["step_traffic"],
        "steps_run": m["steps_run"],
    }
    if "fp64" in res:
        f = res["fp64"]
        fs = summary(f)
        out["fp64"] = {"value": fs["value"], "ms_per_step": fs["ms_per_step"], "regimes": fs["regimes"],
                       "value_workload_weighted": fs.get("value_workload_weighted"),
                       "step_traffic": fs["step_traffic"]}
        out["mixed_over_fp64"] = ms["value"] / fs["value"]
        a, b = m["fields"], f["fields"]

        def rel(names):
            num = sum(float(np.sum((a[n] - b[n]) ** 2)) for n in names)
            return float(np.sqrt(num / sum(float(np.sum(b[n] ** 2)) for n in names)))
        out["psi_rel_l2_mixed_vs_fp64"] = rel(("PSI",))
        dm, df = m["diag"], f["diag"]
        out["accuracy_mixed_vs_fp64"] = {
            "psi_rel_l2": rel(("PSI",)), "rho_rel_l2": rel(("RHO",)),
            "u_rel_l2": rel(("VX", "VY", "VZ")),  # velocity as a vector
            "energy_rel": abs(dm["energy"] - df["energy"]) / abs(df["energy"]),
            "mass_rel": abs(dm["mass"] - df["mass"]) / abs(df["mass"])}
        out["steps_compared"] = m["steps_run"]
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
