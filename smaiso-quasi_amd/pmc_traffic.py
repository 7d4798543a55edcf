#!/usr/bin/env python3
"""HBM bytes per launch of each pass kernel from two rocprofv3 counter runs.

  rocprofv3 --pmc FETCH_SIZE --kernel-trace --output-format csv -d F -o bench -- python3 bench.py ...
  rocprofv3 --pmc WRITE_SIZE --kernel-trace --output-format csv -d W -o bench -- python3 bench.py ...
  python smaiso-quasi_amd/pmc_traffic.py F/bench_counter_collection.csv \\
      W/bench_counter_collection.csv --copy-bytes 1073741824 -o profiles/traffic.json

FETCH_SIZE and WRITE_SIZE count KiB per dispatch.  Each is scaled by what
the same counter reads for a plain copy kernel of known size: the largest
`__amd_rocclr_copyBuffer` dispatch of the run (the psihat -> psihat_old
device copy at set-up, `--copy-bytes` read and the same written), so the
factor maps counter units to bytes moved.  Output: per kernel symbol, the
mean calibrated read + write bytes per dispatch, the dispatch count and the
factors.
"""
import argparse
import collections
import csv
import json

COPY = "__amd_rocclr_copyBuffer"


def per_dispatch(path, counter):
    """{dispatch id: (kernel name, summed counter value)} of one counter run."""
    out = {}
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            if r.get("Counter_Name") != counter:
                continue
            d = r["Dispatch_Id"]
            name, v = out.get(d, (r["Kernel_Name"], 0.0))
            out[d] = (name, v + float(r["Counter_Value"]))
    if not out:
        raise SystemExit(f"{path}: no {counter} rows")
    return out


def by_kernel(disp):
    acc = collections.defaultdict(list)
    for name, v in disp.values():
        acc[name].append(v)
    return acc


def calibration(acc, counter, copy_bytes):
    vals = acc.get(COPY)
    if not vals:
        raise SystemExit(f"no {COPY} dispatch in the {counter} run to calibrate against")
    kib = max(vals)
    return copy_bytes / (kib * 1024.0), kib


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("fetch_csv")
    ap.add_argument("write_csv")
    ap.add_argument("--copy-bytes", type=float, required=True,
                    help="bytes read (and written) by the largest copyBuffer dispatch")
    ap.add_argument("-o", "--out", required=True)
    ap.add_argument("--command", default="", help="the profiled command, recorded in the output")
    a = ap.parse_args()
    fa = by_kernel(per_dispatch(a.fetch_csv, "FETCH_SIZE"))
    wa = by_kernel(per_dispatch(a.write_csv, "WRITE_SIZE"))
    ff, fk = calibration(fa, "FETCH_SIZE", a.copy_bytes)
    wf, wk = calibration(wa, "WRITE_SIZE", a.copy_bytes)
    kernels = {}
    for name in sorted(set(fa) & set(wa)):
        if not name.startswith("void qsm::"):
            continue
        rd = sum(fa[name]) / len(fa[name]) * 1024.0 * ff
        wr = sum(wa[name]) / len(wa[name]) * 1024.0 * wf
        kernels[name] = {"bytes_per_launch": rd + wr, "read_bytes": rd, "write_bytes": wr,
                         "dispatches": min(len(fa[name]), len(wa[name]))}
    out = {"method": "rocprofv3 FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each scaled by "
                     f"bytes/KiB of the largest {COPY} dispatch ({a.copy_bytes:.0f} B read and written)",
           "command": a.command,
           "calibration": {"fetch_kib_copy": fk, "fetch_factor": ff, "write_kib_copy": wk,
                           "write_factor": wf},
           "kernels": kernels}
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)
    for name, k in sorted(kernels.items(), key=lambda kv: -kv[1]["bytes_per_launch"])[:12]:
        print(f"{k['bytes_per_launch'] / 2**20:10.1f} MiB/launch (read {k['read_bytes'] / 2**20:8.1f}, "
              f"write {k['write_bytes'] / 2**20:8.1f})  {name}")


if __name__ == "__main__":
    main()
