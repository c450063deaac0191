"""Headline benchmark: likelihood evaluations per second for a batch of 1024
uneven aligned-spin binaries on one MI355X.

One step = one full pass of the batched hot path over the batch:
  per-binary tables -> circular initial state at r0 -> two-pass DOP853
  trajectory (persistent work-stealing kernel) -> (2,2) mode on the knots ->
  not-a-knot splines -> uniform resample of h+ (N = 2^20 samples, dt = 1/4096 s)
  -> row FFT (hipFFT) -> frequency-domain likelihood against synthetic data
  (injection of binary 0 + Gaussian noise, analytic PSD S(f) = (f/f0)^-4 + 1 + (f/f0)^2).
value = n_binaries / step time.  The same pass with the reference's batching model
(one launch per attempted step, lane per binary, host-synchronised compaction)
is timed beside it as plain_batch, and the oracle (CPU restatement) is timed on a
bounded sample of the same binaries as cpu_baseline.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# FP64 operations per attempted DOP853 step (12 right-hand sides + stages +
# error norm + controller), counted by eobgpu_amd/csrc/flop_count.cpp on the
# library's own Hamiltonian and flux code.
STEP_FLOPS = 55679
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--log2N", type=int, default=20)
    ap.add_argument("--group", type=int, default=64)
    ap.add_argument("--plain-group", type=int, default=1)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-binaries", type=int, default=48)
    ap.add_argument("--max-step", type=int, default=4096)
    ap.add_argument("--rtol", type=float, default=1e-10)
    return ap.parse_args()


class Pipeline:
    def __init__(self, n, log2N, max_step, rtol, seed=2):
        import torch
        from eobgpu_amd import configs, pipeline as P
        self.torch, self.P = torch, P
        self.params = configs.uneven_batch(n=n, seed=seed)
        p = self.params
        self.add = P.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"])
        self.n = n
        self.N = 1 << log2N
        self.dt = 1.0 / 4096
        self.F = self.N // 2 + 1
        self.df = 1.0 / (self.N * self.dt)
        self.M = p["m1"] + p["m2"]
        self.rtol = rtol
        self.b = P.EOBBatch(self.add, max_step=max_step)
        dev = self.b.device
        self.hp = torch.empty((n, self.N), dtype=torch.float64, device=dev)
        self.ht = torch.empty((n, self.F), dtype=torch.complex128, device=dev)
        self.logl = torch.empty(n, dtype=torch.float64, device=dev)
        self.partial = torch.empty(int(P.lib().eobhip_fd_likelihood_partial_doubles(n, self.F)), dtype=torch.float64, device=dev)
        self.amp = np.full(n, 1.0)
        self.data = None

    def run(self, mode="persistent", group=64, events=None):
        P, b = self.P, self.b
        rec = (lambda k: events[k].record()) if events is not None else (lambda k: None)
        rec("t0")
        P.call("eobhip_bin_setup", b.add.data_ptr(), b.n, 1, b.bindata.data_ptr(), b.modes.data_ptr(), P.current_stream())
        b.initial_conditions(self.params["r0"])
        rec("t1")
        b.integrate(rtol=self.rtol, atol=self.rtol, two_pass=True, mode=mode, group=group)
        rec("t2")
        b.hlm22_knots(out=getattr(b, "knots", None))
        b.build_splines()
        b.resample_hplus(self.M, self.dt, self.N, amp=self.amp, out=self.hp)
        rec("t3")
        P.rfft_batched(self.hp, out=self.ht)
        rec("t4")
        if self.data is not None:
            P.fd_likelihood(self.ht, self.data, self.df, partial=self.partial, out=self.logl)
        rec("t5")

    def make_data(self, snr=20.0, seed=7):
        """d = injection (binary 0, scaled to the target SNR) + noise with the analytic PSD."""
        torch = self.torch
        self.run()
        torch.cuda.synchronize()
        f = np.arange(self.F) * self.df
        S = self.P.analytic_psd(f)
        w = np.where(f > 0, 1.0 / S, 0.0)
        h0 = self.ht[0].cpu().numpy()
        rho = np.sqrt(4 * self.df * np.sum(np.abs(h0) ** 2 * w))
        scale = snr / rho if rho > 0 else 1.0
        self.amp = np.full(self.n, scale)
        rng = np.random.default_rng(seed)
        sig = np.sqrt(np.where(f > 0, S, 0.0) / (4 * self.df))
        noise = sig * (rng.normal(size=self.F) + 1j * rng.normal(size=self.F)) / np.sqrt(2)
        self.data = torch.tensor(h0 * scale + noise, device=self.b.device)


def time_pipeline(pipe, mode, group, steps, warmup):
    torch = pipe.torch
    for _ in range(warmup):
        pipe.run(mode=mode, group=group)
    torch.cuda.synchronize()
    keys = ["t0", "t1", "t2", "t3", "t4", "t5"]
    stage = {k: 0.0 for k in ("setup_ic", "trajectory", "waveform", "fft", "likelihood")}
    t0 = time.perf_counter()
    for _ in range(steps):
        ev = {k: torch.cuda.Event(enable_timing=True) for k in keys}
        pipe.run(mode=mode, group=group, events=ev)
        torch.cuda.synchronize()
        for name, (a, c) in zip(stage, zip(keys[:-1], keys[1:])):
            stage[name] += ev[a].elapsed_time(ev[c])
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    return wall, {k: v / steps for k, v in stage.items()}


def cpu_baseline(pipe, nb):
    """Oracle (CPU restatement, 1 thread) on the first nb binaries: trajectory + (2,2) knots."""
    from oracle import oracle as O
    p = pipe.params
    idx = np.arange(nb)
    add = O.make_additional_args(p["m1"][idx], p["m2"][idx], p["chi1"][idx], p["chi2"][idx])
    co = O.fac_coeffs(add)
    t0 = time.perf_counter()
    pphi0, pr0, om0, add = O.initial_conditions(p["r0"][idx], add, co)
    cond = np.array([p["r0"][idx], np.zeros(nb), pr0, pphi0])
    t, traj, ns, si = O.run_trajectory(cond, add, co, max_step=pipe.b.max_step, rtol=pipe.rtol, atol=pipe.rtol)
    O.get_hlm22(traj, ns, add, co)
    el = time.perf_counter() - t0
    return nb / el, el


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the batched path shards by binary and "
              "this build measures the one-GPU share.", file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.init()
    pipe = Pipeline(args.n, args.log2N, args.max_step, args.rtol)
    pipe.make_data()
    print(f"[bench] data ready, n={args.n} N=2^{args.log2N}", file=sys.stderr, flush=True)

    wall, stages = time_pipeline(pipe, "persistent", args.group, args.steps, args.warmup)
    b = pipe.b
    ns = b.num_steps.cpu().numpy()
    att = b.attempts.cpu().numpy()
    p1 = b.pass1_steps.cpu().numpy()
    si = b.stop_info.cpu().numpy()
    logl = pipe.logl.cpu().numpy()
    print(f"[bench] persistent: {wall*1e3:.2f} ms/step stages {json.dumps({k: round(v, 3) for k, v in stages.items()})}",
          file=sys.stderr, flush=True)

    plain = None
    if not args.no_plain:
        wall_p, stages_p = time_pipeline(pipe, "plain", args.plain_group, args.plain_steps, 0)
        plain = {"value": args.n / wall_p, "ms_per_step": wall_p * 1e3, "group": args.plain_group,
                 "trajectory_ms": stages_p["trajectory"], "speedup_of_persistent": wall_p / wall}
        print(f"[bench] plain: {wall_p*1e3:.2f} ms/step", file=sys.stderr, flush=True)

    # roofline of the dominant kernel
    dom = max(stages, key=stages.get)
    total_attempts = float(att.sum())
    if dom == "trajectory":
        achieved = total_attempts * STEP_FLOPS / (stages["trajectory"] * 1e-3) / 1e12
        roof = {"bound": "fp64-vector", "kernel": "k_integrate_persistent", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                "algorithmic": f"{STEP_FLOPS} FP64 ops per attempted step x {int(total_attempts)} attempted steps"}
    else:
        nbytes = {"fft": 2 * args.n * pipe.N * 8 + args.n * pipe.F * 16,
                  "waveform": args.n * pipe.N * 8,
                  "likelihood": args.n * pipe.F * 16 + pipe.F * 16}.get(dom, 0)
        achieved = nbytes / (stages[dom] * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS, "traffic": None}

    cpu = None
    if not args.no_cpu:
        rate, el = cpu_baseline(pipe, args.cpu_binaries)
        cpu = {"value": rate, "unit": "binaries/s (trajectory + (2,2) knots, no resample/FFT)", "cores": 1, "kind": "port",
               "sample": f"first {args.cpu_binaries} binaries of the 1024-binary batch, oracle restatement, {el:.1f} s"}

    out = {
        "metric": "likelihood evals/s @1024 uneven binaries", "value": args.n / wall, "unit": "likelihood evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "1024 aligned-spin binaries, q~U[1,10], chi~U[-0.9,0.9], log10 M~U[1,2], r0~U[15,40] M; "
                               "two-pass DOP853 rtol=atol=1e-10; (2,2) resampled to 2^20 samples at 1/4096 s; FD likelihood",
                   "global_batch": args.n, "seq_len": pipe.N, "parallelism": "none (1 GPU)", "group_lanes": args.group},
        "stages_ms": stages, "steps_per_binary": {"min": int(ns.min()), "median": float(np.median(ns)), "max": int(ns.max())},
        "attempted_steps_total": int(total_attempts), "stop_codes": {str(k): int(v) for k, v in zip(*np.unique(si, return_counts=True))},
        "logl_finite": bool(np.all(np.isfinite(logl))),
        "plain_batch": plain, "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
