"""Writes tests/golden/interface256_s{N}.npz: CPU-oracle reference of the FP64
256^3 interface workload (the drop-in configuration) after N steps.

The oracle (oracle/qsm_oracle.c, FP64 restatement of quasi.cpp's step) runs
the full 256^3 grid; the fixture keeps psi, rho, vx, vy, vz on every 8th grid
point in each direction (32^3 samples per field: a uniform subsample whose
relative L2 estimates the full-field one), the full-field L2 norms, and the
diagnostics L1, mass, energy.  Snapshots at 10, 100, 200, 500, 1000 steps.

Run:  OMP_NUM_THREADS=6 python tests/golden/make_interface256_golden.py [max_steps]
(about 6-8 s per step on 6-8 Xeon cores)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "smaiso-quasi_amd")]

import ics  # noqa: E402
import oracle  # noqa: E402

STRIDE = 8
SNAPS = (10, 100, 200, 500, 1000)


def main():
    max_steps = int(sys.argv[1]) if len(sys.argv) > 1 else 1000
    oracle.build()
    d = ics.params_dict("interface256")
    st = oracle.State(oracle.make_params(**d))
    st.set_psi(ics.workload_psi("interface256"))
    t0 = time.time()
    done = 0
    for snap in SNAPS:
        if snap > max_steps:
            break
        st.step(snap - done)
        done = snap
        out = {"steps": np.array(done), "stride": np.array(STRIDE)}
        for k, v in d.items():
            out["param_" + k] = np.array(v)
        for f in ("PSI", "RHO", "VX", "VY", "VZ"):
            a = st.field(f)
            out[f] = np.ascontiguousarray(a[::STRIDE, ::STRIDE, ::STRIDE])
            out[f + "_norm"] = np.array(np.sqrt(np.sum(a * a)))
        dg = st.diag()
        for k, v in dg.items():
            out["diag_" + k] = np.array(v)
        path = os.path.join(HERE, f"interface256_s{done}.npz")
        np.savez_compressed(path, **out)
        print(f"step {done}: {time.time() - t0:.0f} s, wrote {path}", flush=True)


if __name__ == "__main__":
    main()
