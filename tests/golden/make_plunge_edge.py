"""Writes plunge_edge.json: the three binaries of the 1024-binary workload
(q ~ U[1,10], chi ~ U[-0.9,0.9], log10 M ~ U[1,2], r0 ~ U[15,40] M, seed 2)
that the oracle's adaptive two-pass integration at rtol = atol = 1e-10 ends with
stop code -1 (the next stage leaves the Hamiltonian's domain at r ~ 2.3-2.7
before any stop condition holds): their indices, initial states, the oracle's
adaptive stored points and codes, and their two-pass trajectories with pass 1
at fixed step (h = 10 M, halved for good on a reject; pass 2 the reference's
walk-back and 0.05 M steps) -- stored points, stop codes and (t, r, phi, pr,
pphi) at rows 0, L/4, L/2, 3L/4, L-1.  Oracle only (CPU, seconds); run from the
repository root:
    python tests/golden/make_plunge_edge.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from eobgpu_amd import configs  # noqa: E402

H1, MAX_STEP = 10.0, 65536


def main():
    p = configs.uneven_batch(n=1024, seed=2)
    add_all = O.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"])
    co_all = O.fac_coeffs(add_all)
    pphi0, pr0, om0, add_all = O.initial_conditions(p["r0"], add_all, co_all)
    cond_all = np.array([p["r0"], np.zeros(1024), pr0, pphi0])
    # the adaptive codes of all 1024 (each binary as a batch of one)
    si_all = np.zeros(1024, dtype=np.int64)
    ns_all = np.zeros(1024, dtype=np.int64)
    for s in range(0, 1024, 64):
        sl = slice(s, s + 64)
        _, _, ns, si = O.run_trajectory(cond_all[:, sl], add_all[:, sl], co_all[sl], max_step=4096, rtol=1e-10, atol=1e-10)
        si_all[sl], ns_all[sl] = si, ns
    idx = np.nonzero(si_all == -1)[0]
    cond = cond_all[:, idx]
    add = add_all[:, idx]
    co = co_all[idx]
    t, traj, ns, si = O.run_trajectory(cond, add, co, max_step=MAX_STEP, rtol=1e-10, atol=1e-10, hInit=H1, fix_step1=True)
    rows = []
    for b in range(idx.size):
        L = int(ns[b])
        sel = sorted({0, L // 4, L // 2, (3 * L) // 4, L - 1})
        rows.append({"rows": sel, "t": [float(t[b, r]) for r in sel],
                     "state": [[float(traj[b, c, r]) for c in range(4)] for r in sel]})
    out = {"workload": "uneven_batch(n=1024, seed=2)", "index": idx.tolist(), "hInit": H1, "max_step": MAX_STEP, "rtol": 1e-10,
           "cond": cond.tolist(), "adaptive_num_steps": ns_all[idx].tolist(), "adaptive_stop_info": si_all[idx].tolist(),
           "adaptive_histogram": {str(int(k)): int(v) for k, v in zip(*np.unique(si_all, return_counts=True))},
           "num_steps": ns.tolist(), "stop_info": si.tolist(), "samples": rows}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "plunge_edge.json"), "w") as fh:
        json.dump(out, fh)
    print("binaries", idx.tolist(), "fixed counts", ns.tolist(), "codes", si.tolist())


if __name__ == "__main__":
    main()
