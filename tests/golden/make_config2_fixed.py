"""Writes config2_fixed_64.json: the oracle's two-pass trajectories of the first
64 binaries of the 1024-binary workload (q ~ U[1,10], chi ~ U[-0.9,0.9],
log10 M ~ U[1,2], r0 ~ U[15,40] M, seed 2), both passes at fixed step (pass 1 at
h = 10 M, halved for good on a reject; pass 2 the reference's walk-back and
0.05 M steps), rtol = atol = 1e-10, per binary.  A fixed-step sequence is
deterministic, so the device must give the same step counts and stop codes.
Recorded per binary: stored points, pass-1 steps are not recorded by the
oracle, stop code, and (t, r, phi, pr, pphi) at rows 0, L/4, L/2, 3L/4, L-1.
Oracle only (CPU, about a minute); run from the repository root:
    python tests/golden/make_config2_fixed.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from eobgpu_amd import configs  # noqa: E402

N_BIN, H1, MAX_STEP = 64, 10.0, 65536


def main():
    p = configs.uneven_batch(n=1024, seed=2)
    idx = np.arange(N_BIN)
    add = O.make_additional_args(p["m1"][idx], p["m2"][idx], p["chi1"][idx], p["chi2"][idx])
    co = O.fac_coeffs(add)
    pphi0, pr0, om0, add = O.initial_conditions(p["r0"][idx], add, co)
    cond = np.array([p["r0"][idx], np.zeros(N_BIN), pr0, pphi0])
    t, traj, ns, si = O.run_trajectory(cond, add, co, max_step=MAX_STEP, rtol=1e-10, atol=1e-10, hInit=H1, fix_step1=True)
    rows = []
    for b in range(N_BIN):
        L = int(ns[b])
        sel = sorted({0, L // 4, L // 2, (3 * L) // 4, L - 1})
        rows.append({"rows": sel, "t": [float(t[b, r]) for r in sel],
                     "state": [[float(traj[b, c, r]) for c in range(4)] for r in sel]})
    out = {"workload": "first 64 binaries of uneven_batch(n=1024, seed=2)", "hInit": H1, "max_step": MAX_STEP,
           "rtol": 1e-10, "cond": cond.tolist(), "num_steps": ns.tolist(), "stop_info": si.tolist(), "samples": rows}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "config2_fixed_64.json"), "w") as fh:
        json.dump(out, fh)
    print("counts", ns.min(), ns.max(), "codes", dict(zip(*np.unique(si, return_counts=True))))


if __name__ == "__main__":
    main()
