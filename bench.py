"""Headline benchmark: likelihood evaluations per second for a batch of 1024
uneven aligned-spin binaries on one MI355X.

One step = one full pass of the batched hot path over the batch:
  per-binary tables -> circular initial state at r0 -> two-pass DOP853
  trajectory (persistent work-stealing kernel) -> (2,2) mode on the knots ->
  not-a-knot splines -> h+ on the uniform grid (N = 2^20 samples, dt = 1/4096 s)
  -> frequency-domain likelihood against synthetic data (injection of binary 0
  + Gaussian noise, analytic PSD S(f) = (f/f0)^-4 + 1 + (f/f0)^2), evaluated
  from each template's support (one M-point FFT, M >= 2 x support) instead of
  an N-point FFT per template; the full-length path is timed beside it and the
  two log-likelihoods are compared.
value = n_binaries / step time.  The reference's batching model (one launch per
attempted step, lane per binary, host-synchronised compaction) is timed beside
it as plain_batch, the continuous-batching server on Poisson arrivals as
served, and the oracle (CPU restatement) on a bounded sample of the same
binaries as cpu_baseline.

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
STEP_FLOPS = 50663
# HBM traffic of one integrator launch on the headline batch (rocprofv3 FETCH_SIZE
# and WRITE_SIZE, each scaled by the counter's reading for a 1 GiB device copy;
# profiles/pipeline_uneven1024.md): 28.70 MB read + 27.81 MB written.
INTEGRATOR_TRAFFIC_BYTES = (28.70 + 27.81) * 1e6
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
    ap.add_argument("--cpu-binaries", type=int, default=256)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=1.0)
    ap.add_argument("--serve-streams", type=int, default=4)
    ap.add_argument("--max-step", type=int, default=4096)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--pass", dest="pass_kind", choices=("streamed", "chunked", "one-batch"), default="streamed",
                    help="the headline pass: streamed (eobgpu_amd.pipeline.StreamedPass: one integrator launch, downstream "
                         "work per group of finished binaries), chunked (ChunkedPass) or the one-batch pass")
    ap.add_argument("--chunk-fractions", default="0.1,0.2,0.7",
                    help="ChunkedPass shares (longest binaries first), timed beside the headline pass; '1' skips it")
    ap.add_argument("--sustained", type=int, default=4,
                    help="batches of n fresh binaries integrated back to back in one streamed pass (a SIMD freed by a "
                         "short binary starts the next batch's binaries), reported beside the headline; 0 skips it")
    ap.add_argument("--no-imr", action="store_true",
                    help="skip the IMR generator figure (five modes with NQC and ringdown) reported beside the headline")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip timing the pass captured as one HIP graph (pipeline.GraphPass, full-length likelihood)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device to device (a kernel of known traffic, for scaling memory counters)")
    ap.add_argument("--workload", choices=("uneven", "spline", "capacity", "imr", "detector", "ic", "modes"), default="uneven",
                    help="uneven: the headline 1024-binary pipeline; spline: 512 synthetic (2,2) modes with 5k-60k knots "
                         "through splines + likelihood; capacity: 16384 binaries with r0 in [30, 60] M through the integrator; "
                         "imr: the headline binaries' IMR templates (five modes, NQC and ringdown) at 16384 Hz; detector: the "
                         "headline binaries' detector-frame signals in three interferometers; ic: the reference's "
                         "initial-condition solvers as drop-ins on the headline binaries; modes: the headline binaries' "
                         "five modes through the mode-sum path and the TDInterp_wrap2 drop-in")
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
        self.lik = None

    def run(self, mode="persistent", group=64, events=None, direct=False, block=256, grid_groups=0):
        P, b = self.P, self.b
        rec = (lambda k: events[k].record()) if events is not None else (lambda k: None)
        rec("t0")
        P.call("eobhip_bin_setup", b.add.data_ptr(), b.n, 1, b.bindata.data_ptr(), b.modes.data_ptr(), P.current_stream())
        b.initial_conditions(self.params["r0"])
        rec("t1")
        b.integrate(rtol=self.rtol, atol=self.rtol, two_pass=True, mode=mode, group=group, block=block, grid_groups=grid_groups)
        rec("t2")
        b.hlm22_knots(out=getattr(b, "knots", None))
        b.build_splines()
        rec("t3")
        if direct or self.lik is None:
            # full-length path: h+ on N samples, N-point FFT, sum over N/2+1 bins
            b.resample_hplus(self.M, self.dt, self.N, amp=self.amp, out=self.hp)
            P.rfft_batched(self.hp, out=self.ht)
            if self.data is not None:
                P.fd_likelihood(self.ht, self.data, self.df, partial=self.partial, out=self.logl)
        else:
            self.lik(b, self.M, self.dt, amp=self.amp, out=self.logl)
        rec("t4")

    def make_data(self, snr=20.0, seed=7):
        """d = injection (binary 0, scaled to the target SNR) + noise with the analytic PSD."""
        torch = self.torch
        self.run(direct=True)
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
        self.lik = self.P.SupportLikelihood(self.data, self.N, self.df)


def time_pipeline(pipe, mode, group, steps, warmup, direct=False, block=256, grid_groups=0):
    torch = pipe.torch
    for _ in range(warmup):
        pipe.run(mode=mode, group=group, direct=direct, block=block, grid_groups=grid_groups)
    torch.cuda.synchronize()
    keys = ["t0", "t1", "t2", "t3", "t4"]
    stage = {k: 0.0 for k in ("setup_ic", "trajectory", "knots_splines", "likelihood")}
    t0 = time.perf_counter()
    for _ in range(steps):
        ev = {k: torch.cuda.Event(enable_timing=True) for k in keys}
        pipe.run(mode=mode, group=group, events=ev, direct=direct, block=block, grid_groups=grid_groups)
        torch.cuda.synchronize()
        for name, (a, c) in zip(stage, zip(keys[:-1], keys[1:])):
            stage[name] += ev[a].elapsed_time(ev[c])
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    return wall, {k: v / steps for k, v in stage.items()}


def time_chunked(pipe, fractions, steps, warmup, group=64, block=64):
    """The same pass as Pipeline.run (support-length likelihood) with the batch in
    stream-overlapped chunks of the given shares; returns (wall per pass, log L,
    trajectory time of the longest chunk, chunk sizes)."""
    torch, P = pipe.torch, pipe.P
    p = pipe.params
    cp = P.ChunkedPass(p["m1"], p["m2"], p["chi1"], p["chi2"], p["r0"], pipe.lik, pipe.M, pipe.dt, amp=pipe.amp,
                       max_step=pipe.b.max_step, fractions=fractions, rtol=pipe.rtol)
    for _ in range(warmup):
        cp.run(group=group, block=block)
    torch.cuda.synchronize()
    t_traj = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        ev = {"start": torch.cuda.Event(enable_timing=True), "traj_end_longest": torch.cuda.Event(enable_timing=True)}
        ev["start"].record()
        cp.run(group=group, block=block, events=ev)
        torch.cuda.synchronize()
        t_traj += ev["start"].elapsed_time(ev["traj_end_longest"])
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    return wall, cp.out.cpu().numpy(), t_traj / steps, [len(g) for g in cp.groups]


def time_streamed(pipe, steps, warmup, group=64, block=64, grid_groups=0, resident=True):
    """The same pass as Pipeline.run (support-length likelihood) as a StreamedPass:
    one integrator launch that flags each finished binary, knots, splines and
    likelihood queued per group of finished binaries.  Every timed step
    evaluates a new batch of binaries drawn from the same distribution (seeds
    1000 + k), and the pass builds their additional args, tables and initial
    state on the device.  resident: every batch's parameters are in device
    memory before the timed steps and a step takes them with
    ``set_params_device`` (one device copy); otherwise ``set_params`` orders
    each batch on the host and copies it to the device inside the step.
    Returns (wall per pass, log L of the batch the one-batch pass ran,
    trajectory time, downstream groups of the last pass, host ingestion time
    per pass in ms, attempted steps per pass, per-step wall and GPU times)."""
    torch, P = pipe.torch, pipe.P
    from eobgpu_amd import configs
    p = pipe.params
    sp = P.StreamedPass(p["m1"], p["m2"], p["chi1"], p["chi2"], p["r0"], pipe.lik, pipe.M, pipe.dt, amp=pipe.amp,
                        max_step=pipe.b.max_step, rtol=pipe.rtol)
    sp.run(group=group, block=block, grid_groups=grid_groups)
    torch.cuda.synchronize()
    logl_same = sp.out.cpu().numpy().copy()
    props = [configs.uneven_batch(n=pipe.n, seed=1000 + k) for k in range(warmup + steps)]
    if resident:
        amp = np.broadcast_to(np.asarray(pipe.amp, dtype=np.float64), (pipe.n,))
        dev_props = [torch.tensor(np.stack([q["m1"], q["m2"], q["chi1"], q["chi2"], q["r0"], q["m1"] + q["m2"], amp]),
                                  dtype=torch.float64, device=sp.out.device) for q in props]
        torch.cuda.synchronize()

        def ingest(k):
            sp.set_params_device(dev_props[k])
    else:
        def ingest(k):
            q = props[k]
            sp.set_params(q["m1"], q["m2"], q["chi1"], q["chi2"], q["r0"], amp=pipe.amp)

    for k in range(warmup):
        ingest(k)
        sp.run(group=group, block=block, grid_groups=grid_groups)
    torch.cuda.synchronize()
    t_traj = t_ing = 0.0
    attempts = []
    per_step = []
    evs = [{"start": torch.cuda.Event(enable_timing=True), "traj_end_longest": torch.cuda.Event(enable_timing=True),
            "end": torch.cuda.Event(enable_timing=True)} for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        ev = evs[k]
        tk = time.perf_counter()
        ev["start"].record()
        ingest(warmup + k)
        t_ing += time.perf_counter() - tk
        sp.run(group=group, block=block, grid_groups=grid_groups, events=ev)
        ev["end"].record()
        torch.cuda.synchronize()
        per_step.append((time.perf_counter() - tk) * 1e3)
    wall = (time.perf_counter() - t0) / steps
    for k in range(steps):
        t_traj += evs[k]["start"].elapsed_time(evs[k]["traj_end_longest"])
    gpu_ms = [evs[k]["start"].elapsed_time(evs[k]["end"]) for k in range(steps)]
    attempts.append(int(sp.batch.attempts.sum().item()))
    return wall, logl_same, t_traj / steps, len(sp.groups), t_ing / steps * 1e3, attempts, per_step, gpu_ms


def time_sustained(pipe, batches, steps, warmup, group=64, block=256):
    """Back-to-back batches: `batches` fresh sets of n binaries (seeds 2000 +
    batches k + j) in one StreamedPass, so that the SIMDs freed by short
    binaries start other binaries instead of idling.  Two queue orders: batch
    by batch (each batch longest expected first; the next batch's binaries
    fill the SIMDs the previous one frees) and longest expected first over all
    of them (the batches given together).  Each binary's log L is its
    single-batch value (binaries are independent).  Returns {order: (wall per
    pass, evals/s)} and the ingestion time per pass."""
    torch, P = pipe.torch, pipe.P
    from eobgpu_amd import configs
    n = pipe.n
    sets = []
    for k in range(warmup + steps):
        ps = [configs.uneven_batch(n=n, seed=2000 + batches * k + j) for j in range(batches)]
        cat = {key: np.concatenate([q[key] for q in ps]) for key in ("m1", "m2", "chi1", "chi2", "r0")}
        major = np.concatenate([j * n + np.argsort(-P.expected_length(q["r0"], q["m1"], q["m2"], q["chi1"], q["chi2"]), kind="stable")
                                for j, q in enumerate(ps)])
        glob = np.argsort(-P.expected_length(cat["r0"], cat["m1"], cat["m2"], cat["chi1"], cat["chi2"]), kind="stable")
        sets.append((cat, {"batch_major": major, "longest_first": glob}))
    cat, orders = sets[0]
    sp = P.StreamedPass(cat["m1"], cat["m2"], cat["chi1"], cat["chi2"], cat["r0"], pipe.lik, cat["m1"] + cat["m2"], pipe.dt,
                        max_step=pipe.b.max_step, rtol=pipe.rtol, order=orders["batch_major"])
    out = {}
    t_ing = 0.0
    for name in ("batch_major", "longest_first"):
        for k in range(warmup):
            c, o = sets[k]
            sp.set_params(c["m1"], c["m2"], c["chi1"], c["chi2"], c["r0"], order=o[name])
            sp.run(group=group, block=block)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            c, o = sets[warmup + k]
            ti = time.perf_counter()
            sp.set_params(c["m1"], c["m2"], c["chi1"], c["chi2"], c["r0"], order=o[name])
            t_ing += time.perf_counter() - ti
            sp.run(group=group, block=block)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / steps
        out[name] = (wall, batches * n / wall)
    return out, t_ing / (2 * steps) * 1e3


def time_graph(pipe, steps, warmup):
    """The one-batch pass with the full-length likelihood captured as one HIP graph
    (pipeline.GraphPass): every timed pass takes a fresh batch (seeds 3000 + k)
    through set_params and one graph launch; the same launches issued one by one
    (run_eager) are timed beside it.  Returns ms per pass for both."""
    torch, P = pipe.torch, pipe.P
    from eobgpu_amd import configs
    gp = P.GraphPass(pipe.n, pipe.data, pipe.N, pipe.dt, df=pipe.df, max_step=pipe.b.max_step, rtol=pipe.rtol)
    props = [configs.uneven_batch(n=pipe.n, seed=3000 + k) for k in range(warmup + steps)]

    def put(q):
        gp.set_params(q["m1"], q["m2"], q["chi1"], q["chi2"], q["r0"], amp=pipe.amp)

    put(props[0])
    gp.capture()
    out = {}
    for label, fn in (("graph", gp.run), ("eager", gp.run_eager)):
        for k in range(warmup):
            put(props[k])
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            put(props[warmup + k])
            fn()
        torch.cuda.synchronize()
        out[label] = (time.perf_counter() - t0) / steps * 1e3
    return out


def cpu_info():
    """(threads this process may use, capped at 16 -- the host cores that go with
    one GPU on an 8-GPU MI355X node; CPU model)."""
    import os
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return min(n, 16), n, model


def cpu_baseline(pipe, nb, threads=1):
    """The whole pass on the CPU for the first nb binaries: the oracle's initial
    conditions, two-pass trajectory and (2,2) knots, the oracle's not-a-knot
    splines of the three knot series evaluated on the uniform grid (up to each
    template's last knot), h+, numpy's rfft of N samples and the likelihood sum
    against the same data.  threads > 1 splits the binaries over a thread pool
    (the oracle's C code and numpy's FFT run without the interpreter lock).
    Returns (likelihood evals/s, seconds, log L)."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    p = pipe.params
    data = pipe.data.cpu().numpy()
    parts = [c for c in np.array_split(np.arange(nb), threads) if c.size]

    def work(idx):
        logl, _, _ = O.pipeline_logl(p["m1"][idx], p["m2"][idx], p["chi1"][idx], p["chi2"][idx], p["r0"][idx], data, pipe.dt,
                                     pipe.N, amp=pipe.amp[idx], max_step=pipe.b.max_step, rtol=pipe.rtol, atol=pipe.rtol)
        return logl

    O.lib()
    t0 = time.perf_counter()
    if threads == 1:
        logl = work(np.arange(nb))
    else:
        with ThreadPoolExecutor(max_workers=threads) as ex:
            logl = np.concatenate(list(ex.map(work, parts)))
    el = time.perf_counter() - t0
    return nb / el, el, logl


def serve_section(args):
    """Poisson arrivals (200/s) of proposal batches of U{1..256} binaries drawn as the
    16-binary small workload (q ~ U[1,5], M = 30, chi ~ U[-0.5,0.5], r0 = 20), each a
    full waveform + likelihood at N = 2^12, dt = 1/2048 s; continuous batching vs one
    request at a time with the one-launch-per-step integrator."""
    import torch
    from eobgpu_amd import configs, pipeline as P
    from eobgpu_amd.serving import Request, Server, summarize
    N, dt = 1 << 12, 1.0 / 2048
    F = N // 2 + 1
    df = 1.0 / (N * dt)
    inj = configs.small_batch(n=1, seed=11)
    b = P.EOBBatch(P.make_additional_args(inj["m1"], inj["m2"], inj["chi1"], inj["chi2"]), max_step=2048)
    b.initial_conditions(inj["r0"])
    b.integrate(rtol=1e-10, atol=1e-10, two_pass=True)
    b.hlm22_knots()
    b.build_splines()
    h0 = P.rfft_batched(b.resample_hplus(inj["m1"] + inj["m2"], dt, N))[0].cpu().numpy()
    f = np.arange(F) * df
    S = P.analytic_psd(f)
    rng = np.random.default_rng(17)
    sig = np.sqrt(np.where(f > 0, S, 0.0) / (4 * df))
    data = h0 + sig * (rng.normal(size=F) + 1j * rng.normal(size=F)) / np.sqrt(2)
    reqs = [Request(i, t, p) for i, (t, p) in enumerate(configs.served_requests(duration_s=args.serve_seconds))]
    srv = Server(data, df=df, dt=dt, N=N, capacity=8192, max_step=2048, streams=args.serve_streams)
    srv.warm()
    srv.run([Request(-1, 0.0, configs.small_batch(n=64, seed=900))])
    out = {"requests": len(reqs), "N": N, "dt": dt, "streams": args.serve_streams}
    for policy in ("continuous", "plain"):
        res = srv.run(reqs, policy=policy)
        out[policy] = summarize(res)
        print(f"[bench] served ({policy}): {json.dumps(out[policy])}", file=sys.stderr, flush=True)
    return out


def workload_spline(args):
    """512 binaries' (2,2) modes on 5k-60k non-uniform knots (chirp spacing
    t_k = T (1 - (1 - u_k)^(3/8)), u_k sorted uniform), splined and resampled to
    2^20 samples at 1/4096 s, likelihood against injection + noise (analytic PSD)."""
    import torch
    from eobgpu_amd import pipeline as P
    n, N, dt = 512, 1 << args.log2N, 1.0 / 4096
    F, df = N // 2 + 1, 1.0 / (N * dt)
    Msun = 30.0
    from eobgpu_amd import configs
    ts, phis, res, ims, K = configs.chirp_modes(n=n, seed=3, N=N, dt=dt, Msun=Msun)
    rng = np.random.default_rng(33)
    kmax = int(K.max())
    b = P.EOBBatch(P.make_additional_args(np.full(n, 0.6), np.full(n, 0.4), np.zeros(n), np.zeros(n)), max_step=kmax)
    dense = np.zeros((n, kmax, 5))
    knots = np.zeros((n, kmax, 3))
    for i in range(n):
        dense[i, :K[i], 0] = ts[i]
        knots[i, :K[i], 0] = phis[i]
        knots[i, :K[i], 1] = res[i]
        knots[i, :K[i], 2] = ims[i]
    del ts, phis, res, ims
    b.dense.copy_(torch.from_numpy(dense))
    b.knots = torch.from_numpy(knots).to(b.device)
    b.num_steps.copy_(torch.from_numpy(K.astype(np.int32)))
    del dense, knots
    b.build_splines()
    hp = b.resample_hplus(np.full(n, Msun), dt, N)
    h0 = P.rfft_batched(hp[:1].contiguous())[0].cpu().numpy()
    del hp
    f = np.arange(F) * df
    S = P.analytic_psd(f)
    w = np.where(f > 0, 1.0 / S, 0.0)
    rho = np.sqrt(4 * df * np.sum(np.abs(h0) ** 2 * w))
    sig = np.sqrt(np.where(f > 0, S, 0.0) / (4 * df))
    data = h0 * (20.0 / rho) + sig * (rng.normal(size=F) + 1j * rng.normal(size=F)) / np.sqrt(2)
    lik = P.SupportLikelihood(data, N, df)
    logl = torch.empty(n, dtype=torch.float64, device=b.device)

    def step():
        b.build_splines()
        lik(b, np.full(n, Msun), dt, out=logl)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    t0 = time.perf_counter()
    spl = lk = 0.0
    for _ in range(args.steps):
        ev[0].record()
        b.build_splines()
        ev[1].record()
        lik(b, np.full(n, Msun), dt, out=logl)
        ev[2].record()
        torch.cuda.synchronize()
        spl += ev[0].elapsed_time(ev[1])
        lk += ev[1].elapsed_time(ev[2])
    wall = (time.perf_counter() - t0) / args.steps
    spl /= args.steps
    lk /= args.steps
    knots_total = int(K.sum())
    # spline build: 8 (t) + 24 (three y) read + 96 (coefficients) written per knot; support likelihood:
    # h on M_b samples written and read by the FFT, M_b/2+1 bins read back
    mb = (1 << lik.mlog2.cpu().numpy().astype(np.int64))
    lik_bytes = float(np.sum(8 * mb * 2 + 16 * (mb // 2 + 1)))
    out = {"metric": "likelihood evals/s (512 synthetic (2,2) modes, 5k-60k knots, N=2^20)", "value": n / wall,
           "unit": "likelihood evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "512 synthetic chirp (2,2) modes, K ~ U{5000..60000} knots, M = 30 Msun, N = 2^20, dt = 1/4096 s",
                      "knots_total": knots_total},
           "stages_ms": {"spline_build": spl, "likelihood": lk},
           "roofline": {"spline_build": {"bound": "hbm", "achieved": knots_total * 128 / (spl * 1e-3) / 1e9, "peak": PEAK_HBM_GBS,
                                         "unit": "GB/s", "frac": knots_total * 128 / (spl * 1e-3) / 1e9 / PEAK_HBM_GBS},
                        "likelihood": {"bound": "hbm", "achieved": lik_bytes / (lk * 1e-3) / 1e9, "peak": PEAK_HBM_GBS,
                                       "unit": "GB/s", "frac": lik_bytes / (lk * 1e-3) / 1e9 / PEAK_HBM_GBS}},
           "logl_finite": bool(torch.isfinite(logl).all().item())}
    print(json.dumps(out))


def workload_capacity(args):
    """16384 binaries, r0 ~ U[30, 60] M: the two-pass trajectory of the whole
    batch with dense per-step output kept on the device."""
    import torch
    from eobgpu_amd import autotune, configs, pipeline as P
    p = configs.capacity_batch()
    n = p["m1"].size
    add = P.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"])
    max_step = max(args.max_step, 8192)
    tuned = autotune.tune(add, p["r0"], max_step=max_step, sample=2048, rtol=args.rtol, atol=args.rtol)
    b = P.EOBBatch(add, max_step=max_step)
    b.initial_conditions(p["r0"])
    cond = b.cond.clone()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    times = []
    for it in range(args.warmup + args.steps):
        b.cond = cond
        ev[0].record()
        b.integrate(rtol=args.rtol, atol=args.rtol, two_pass=True, group=tuned["group"], block=tuned["block"],
                    grid_groups=tuned["grid_groups"])
        ev[1].record()
        torch.cuda.synchronize()
        if it >= args.warmup:
            times.append(ev[0].elapsed_time(ev[1]))
    ms = float(np.mean(times))
    att = float(b.attempts.sum().item())
    ns = b.num_steps.cpu().numpy()
    si = b.stop_info.cpu().numpy()
    ach = att * STEP_FLOPS / (ms * 1e-3) / 1e12
    out = {"metric": "binaries/s (16384-binary capacity batch, trajectory)", "value": n / (ms * 1e-3), "unit": "binaries/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "16384 binaries, q~U[1,10], chi~U[-0.9,0.9], log10 M~U[1,2], r0~U[30,60] M, two-pass DOP853 1e-10",
                      "max_step": max_step, "dense_bytes": int(b.dense.numel() * 8), "group": tuned["group"],
                      "block": tuned["block"], "grid_groups": tuned["grid_groups"]},
           "steps_per_binary": {"min": int(ns.min()), "median": float(np.median(ns)), "max": int(ns.max())},
           "attempted_steps_total": int(att), "stop_codes": {str(k): int(v) for k, v in zip(*np.unique(si, return_counts=True))},
           "roofline": {"bound": "fp64-vector", "kernel": "k_integrate_persistent", "achieved": ach, "peak": PEAK_FP64_TFLOPS,
                        "unit": "TFLOP/s", "frac": ach / PEAK_FP64_TFLOPS}}
    del b
    torch.cuda.empty_cache()
    out["served_by_footprint"] = capacity_served(p, args)
    print(json.dumps(out))


def run_imr(args, steps, warmup):
    """The waveform generator of the headline batch: for each of the 1024 uneven
    binaries, fresh parameters ingested on the device, initial conditions, the
    two-pass trajectory, the five modes on the knots, their splines, the NQC and
    ringdown parameters and the IMR template summed over the modes on 2^20
    samples at 1/16384 s (SEOBNRv4PHM.__call__ with compute_full_waveform and
    BBHWaveformTD's mode sum, eob/seobnrv4phm.py:2628-2726, 2275-2556, 720)."""
    import torch
    from eobgpu_amd import configs, pipeline as P
    n = args.n
    N = 1 << args.log2N
    dt = 1.0 / 16384
    ps = [configs.uneven_batch(n=n, seed=2 + k) for k in range(warmup + steps)]
    b = P.EOBBatch.from_params(ps[0]["m1"], ps[0]["m2"], ps[0]["chi1"], ps[0]["chi2"], max_step=args.max_step)
    names = ("setup_ic", "trajectory", "knots_splines", "imr_params", "imr_template")
    times, stages = [], []
    for it, p in enumerate(ps):
        M = p["m1"] + p["m2"]
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(names) + 1)]
        ev[0].record()
        b.set_params(p["m1"], p["m2"], p["chi1"], p["chi2"])
        b.initial_conditions(p["r0"])
        ev[1].record()
        b.integrate(rtol=args.rtol, atol=args.rtol, two_pass=True)
        ev[2].record()
        b.hlm_modes_knots(P.REFERENCE_MODES)
        b.build_mode_splines()
        ev[3].record()
        prm = b.imr_params(M, dt)
        ev[4].record()
        hp, hc = b.imr_template(M, dt, N, params=prm)
        ev[5].record()
        torch.cuda.synchronize()
        if it >= warmup:
            times.append(ev[0].elapsed_time(ev[-1]))
            stages.append([ev[k].elapsed_time(ev[k + 1]) for k in range(len(names))])
        del hp, hc
    ms = float(np.mean(times))
    st = np.mean(np.array(stages), axis=0)
    bini = b.imr["bini"].cpu().numpy()
    nqc = b.imr["nqc"].cpu().numpy()
    ok = np.isfinite(nqc).all(axis=(1, 2)) & (bini[:, 1] >= 0)
    out = {"metric": "IMR waveforms/s (1024 uneven binaries, five modes with NQC and ringdown)", "value": n / (ms * 1e-3),
           "unit": "waveforms/s", "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "1024 aligned-spin binaries, q~U[1,10], chi~U[-0.9,0.9], log10 M~U[1,2], r0~U[15,40] M, fresh "
                                  "parameters each step; modes (2,2),(2,1),(3,3),(4,4),(5,5); NQC + ringdown; 2^%d samples at "
                                  "1/16384 s" % args.log2N, "global_batch": n, "seq_len": N},
           "stages_ms": dict(zip(names, [float(x) for x in st])),
           "attached_binaries": int(ok.sum()), "ringdown_samples": {"min": int(bini[ok, 2].min()), "max": int(bini[ok, 2].max())}}
    del b
    torch.cuda.empty_cache()
    return out


def workload_imr(args):
    """``--workload imr``: the IMR generator's line (run_imr)."""
    print(json.dumps(run_imr(args, args.steps, args.warmup)))


def workload_detector(args, nifo=3):
    """``--workload detector``: the detector-frame signals of the headline batch
    (BBHWaveformTD's "detector_fd" branch, eob/seobnrv4phm.py:597-634, 742-767):
    h+ and hx of the 1024 uneven binaries on 2^20 samples at 1/4096 s
    (resample_hpc), their real FFTs, the antenna patterns and delays of three
    interferometers for fresh sky positions, and signal [n][3][N/2+1] =
    (F+ h+~ + Fx hx~) e^{-2 pi i f dt} (k_detector_signal).  The trajectory,
    knots and splines are built once before the timed steps; the detectors are
    synthetic (orthonormal arm pairs, vertices on a 6 400 km sphere)."""
    import torch
    from eobgpu_amd import configs, detector, pipeline as P
    n, N, dt = args.n, 1 << args.log2N, 1.0 / 4096
    F = N // 2 + 1
    p = configs.uneven_batch(n=n, seed=2)
    b = P.EOBBatch(P.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"]), max_step=args.max_step)
    b.initial_conditions(p["r0"])
    b.integrate(rtol=args.rtol, atol=args.rtol, two_pass=True)
    b.hlm22_knots()
    b.build_splines()
    M = p["m1"] + p["m2"]
    rng = np.random.default_rng(23)
    D, V = [], []
    for _ in range(nifo):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        D.append(0.5 * (np.outer(q[:, 0], q[:, 0]) - np.outer(q[:, 1], q[:, 1])))
        V.append(q[:, 2] * 6.4e6)
    D, V = np.array(D), np.array(V)
    sig = torch.empty((n, nifo, F), dtype=torch.complex128, device=b.device)
    names = ["hplus_hcross", "rfft", "antenna", "signal"]
    times, stages = [], []
    for it in range(args.warmup + args.steps):
        ra, dec = rng.uniform(0, 2 * np.pi, n), np.arcsin(rng.uniform(-1, 1, n))
        psi, gmst = rng.uniform(0, np.pi, n), rng.uniform(0, 2 * np.pi, n)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(names) + 1)]
        torch.cuda.synchronize()
        ev[0].record()
        hp, hc = detector.resample_hpc(b, M, dt, N)
        ev[1].record()
        hpf, hcf = P.rfft_batched(hp), P.rfft_batched(hc)
        ev[2].record()
        Fp, Fc, tdel = detector.get_detector_information(ra, dec, gmst, psi, D, V, device=b.device)
        ev[3].record()
        detector.detector_signal(hpf, hcf, Fp, Fc, tdel, 1.0 / (N * dt), out=sig)
        ev[4].record()
        torch.cuda.synchronize()
        if it >= args.warmup:
            times.append(ev[0].elapsed_time(ev[-1]))
            stages.append([ev[k].elapsed_time(ev[k + 1]) for k in range(len(names))])
        del hp, hc, hpf, hcf
    ms = float(np.mean(times))
    st = dict(zip(names, [float(x) for x in np.mean(np.array(stages), axis=0)]))
    sig_bytes = n * F * (2 + nifo) * 16          # h+~ and hx~ read, nifo signals written
    ach = sig_bytes / (st["signal"] * 1e-3) / 1e9
    out = {"metric": "detector-frame signals/s (1024 uneven binaries x 3 interferometers)", "value": n / (ms * 1e-3),
           "unit": "binaries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "1024 aligned-spin binaries (trajectory, knots and splines built once), h+ and hx on 2^%d samples "
                                  "at 1/4096 s, real FFTs, antenna patterns and delays of 3 synthetic interferometers for fresh "
                                  "sky positions, detector-frame signals" % args.log2N, "global_batch": n, "seq_len": N},
           "stages_ms": st,
           "roofline": {"bound": "hbm", "kernel": "k_detector_signal", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": ach / PEAK_HBM_GBS, "algorithmic": "%d B per (binary, bin): h+~ and hx~ read, 3 signals written" %
                        ((2 + nifo) * 16)}}
    print(json.dumps(out))


def workload_ic(args):
    """``--workload ic``: the reference's initial-condition solvers as drop-ins on
    the headline batch (SEOBNRv4PHM's start at f = 20 Hz, eob/seobnrv4phm.py:1966-2070):
    root_find_all (discrete Newton on IC_cons for r and pphi, src/EOB.cu:2971-3031)
    then root_find_scalar_all (false position on IC_diss for pr in [-3e-2, 0],
    :2778-2834), each a synchronous call as the reference's are; beside them the
    batched circular start at r0 that the pipeline uses (k_initial_conditions)."""
    import torch
    from eobgpu_amd import configs, pyEOB, pipeline as P
    n, dev = args.n, "cuda"
    p = configs.uneven_batch(n=n, seed=2)
    omega0 = 20.0 * ((p["m1"] + p["m2"]) * P.MTSUN_SI * np.pi)
    add = P.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"], omega0=omega0)

    def T(x):
        return torch.tensor(np.ascontiguousarray(x, dtype=np.float64).reshape(-1), device=dev)
    m1, m2, c1, c2 = add[0], add[1], add[2], add[3]
    eta = m1 * m2 / (m1 + m2) ** 2
    chiS, chiA = 0.5 * (c1 + c2), 0.5 * (c1 - c2)
    ode = pyEOB.SEOBNRv5Class()
    ode.update_information(T(m1), T(m2), T(eta), T((1 - 2 * eta) * chiS + (m1 - m2) / (m1 + m2) * chiA), T(chiS), T(chiA), 1, n,
                           T(add[10:80].T))
    rg = omega0 ** (-2.0 / 3)
    x0, add0 = T(np.array([rg, np.sqrt(rg)])), T(add)
    tadd = add0.clone()
    xo, argbuf = torch.zeros(2 * n, dtype=torch.float64, device=dev), torch.zeros(4 * n, dtype=torch.float64, device=dev)
    bounds = T(np.tile(np.array([-3e-2, 0.0]), (n, 1)).T)
    pr = torch.zeros(n, dtype=torch.float64, device=dev)
    zero = torch.zeros(n, dtype=torch.float64, device=dev)
    b = P.EOBBatch(add, max_step=64)
    t_rf, t_sc, t_circ = [], [], []
    for it in range(args.warmup + args.steps):
        tadd.copy_(add0)   # IC_diss writes dH/dpphi into the omega row
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pyEOB.root_find_all(xo, x0, argbuf, tadd, 1000, 1e-12, n, 2, 4, 80)
        t1 = time.perf_counter()
        x = xo.reshape(2, n)
        a4 = torch.stack([x[0], zero, zero, x[1]]).reshape(-1).contiguous()
        ode.root_find_scalar_all(pr, bounds, a4, tadd, 1000, 1e-12, n, 4, 80)
        t2 = time.perf_counter()
        b.initial_conditions(p["r0"])
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        if it >= args.warmup:
            t_rf.append((t1 - t0) * 1e3)
            t_sc.append((t2 - t1) * 1e3)
            t_circ.append((t3 - t2) * 1e3)
    x = xo.reshape(2, n).cpu().numpy()
    prv = pr.cpu().numpy()
    ode.deallocate_information()
    ms = float(np.mean(t_rf) + np.mean(t_sc))
    out = {"metric": "initial conditions/s (1024 uneven binaries, the reference's solvers as drop-ins)", "value": n / (ms * 1e-3),
           "unit": "binaries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "1024 aligned-spin binaries, q~U[1,10], chi~U[-0.9,0.9], log10 M~U[1,2]; start at f = 20 Hz: "
                                  "root_find_all (r, pphi) then root_find_scalar_all (pr)", "global_batch": n},
           "stages_ms": {"root_find_all": float(np.mean(t_rf)), "root_find_scalar_all": float(np.mean(t_sc))},
           "batched_circular_start_ms": float(np.mean(t_circ)),
           "finite": bool(np.isfinite(x).all() and np.isfinite(prv).all()),
           "pr_range": [float(prv.min()), float(prv.max())]}
    print(json.dumps(out))


def workload_modes(args):
    """``--workload modes``: the five modes of the headline batch through the
    reference's mode-sum path (get_hlms and the mode splines, then the sum
    over modes of (re + i im) e^{-i m phi} on the data's time grid):
    k_hlm_modes_knots, the mode splines, k_resample_modes (each binary on its
    own 2^20-sample grid at 1/4096 s) and the TDInterp_wrap2 drop-in
    (k_td_interp2, src/interpolate.cu:441-549) on one common grid of 2^20
    samples in seconds with its interval indices from torch.searchsorted, as
    the reference's caller computes them.  The trajectory is integrated once
    before the timed steps."""
    import torch
    from eobgpu_amd import configs, pytdinterp, pipeline as P
    n, N, dt = args.n, 1 << args.log2N, 1.0 / 4096
    p = configs.uneven_batch(n=n, seed=2)
    M = p["m1"] + p["m2"]
    b = P.EOBBatch(P.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"]), max_step=args.max_step)
    b.initial_conditions(p["r0"])
    b.integrate(rtol=args.rtol, atol=args.rtol, two_pass=True)
    modes = P.REFERENCE_MODES
    nm = len(modes)
    ls = torch.tensor([l for l, _ in modes], dtype=torch.int32, device=b.device)
    ms = torch.tensor([m for _, m in modes], dtype=torch.int32, device=b.device)
    names = ["knots", "splines", "resample_modes", "td2_layout", "td2"]
    times, stages = [], []
    out = torch.empty((n, N), dtype=torch.complex128, device=b.device)
    dataTime = torch.arange(N, dtype=torch.float64, device=b.device) * dt
    Mt = torch.tensor(M, dtype=torch.float64, device=b.device)
    for it in range(args.warmup + args.steps):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(names) + 1)]
        torch.cuda.synchronize()
        ev[0].record()
        b.hlm_modes_knots(modes)
        ev[1].record()
        coef = b.build_mode_splines()
        ev[2].record()
        hp, hc = b.resample_modes(M, dt, N)
        ev[3].record()
        # the reference's layout [bin][2M + 1][old_length] and its caller's inputs:
        # knot times in seconds, lengths and interval indices on the common grid
        ns = b.num_steps.to(torch.int64)
        L = b.max_step
        t = b.dense[:, :, 0] * (Mt * P.MTSUN_SI)[:, None]
        valid = torch.arange(L, device=b.device)[None, :] < ns[:, None]
        ts = torch.where(valid, t, torch.full_like(t, float("inf"))).contiguous()
        c = coef.view(n, L, 2 * nm + 1, 4).permute(3, 0, 2, 1).contiguous()   # [4][n][2M + 1][L]
        sc = (1.0 / (Mt * P.MTSUN_SI))[:, None, None]                            # the splines' x in seconds
        c[1] *= sc
        c[2] *= sc * sc
        c[3] *= sc * sc * sc
        tlast = ts.gather(1, (ns - 1).clamp(min=0)[:, None])[:, 0]
        lengths = torch.clamp(torch.searchsorted(dataTime, tlast, right=True), max=N).to(torch.int32)
        inds = (torch.searchsorted(ts, dataTime.expand(n, N).contiguous(), right=True) - 1)
        inds = torch.minimum(inds.clamp(min=0), (ns - 2).clamp(min=0)[:, None]).to(torch.int32).contiguous()
        ev[4].record()
        pytdinterp.TDInterp_wrap2(out, dataTime, ts, c[0], c[1], c[2], c[3], L, ns.to(torch.int32), N, n, nm, ls, ms, inds, lengths, N,
                                  legacy_stride=False)
        ev[5].record()
        torch.cuda.synchronize()
        if it >= args.warmup:
            times.append(ev[0].elapsed_time(ev[3]))
            stages.append([ev[k].elapsed_time(ev[k + 1]) for k in range(len(names))])
        del hp, hc, c, inds, ts, t
    ms_ = float(np.mean(times))
    st = dict(zip(names, [float(x) for x in np.mean(np.array(stages), axis=0)]))
    res = {"metric": "five-mode templates/s (1024 uneven binaries, the reference's mode-sum path)", "value": n / (ms_ * 1e-3),
           "unit": "templates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "1024 aligned-spin binaries (trajectory integrated once), modes (2,2),(2,1),(3,3),(4,4),(5,5) on "
                                  "the knots, their splines and the mode sum on 2^%d samples at 1/4096 s (value: knots, splines, "
                                  "k_resample_modes); TDInterp_wrap2 on the same grid in seconds timed beside" % args.log2N,
                      "global_batch": n, "seq_len": N},
           "stages_ms": st, "td2_finite": bool(torch.isfinite(torch.view_as_real(out)).all().item())}
    print(json.dumps(res))


def capacity_served(p, args, budget_gb=6.0, N=1 << 12, dt=1.0 / 2048):
    """The same 16384 binaries as 64 requests of 256 arriving at once, served
    under one memory budget in two ways: admission by projected footprint (size
    classes of 1024-8192 trajectory rows, each binary in the smallest class its
    projected length fits, classes sized by the projected shares) and uniform
    slots of the longest class (8192 rows, what admission by slot count needs
    to be safe).  The full chain per binary, likelihood at N = 2^12; refills as
    slots free up.  Reports for both: binaries resident at once, throughput, and
    the binaries run again after filling their rows."""
    import torch
    from eobgpu_amd.serving import Request, Server, classes_for_budget, summarize, projected_rows, assign_classes, SlotPool
    rng = np.random.default_rng(5)
    F = N // 2 + 1
    data = (rng.normal(size=F) + 1j * rng.normal(size=F)) * 1e-3
    rows = (1024, 2048, 4096, 8192)
    proj = projected_rows(p)
    share = np.bincount(assign_classes(proj, rows), minlength=len(rows)).astype(np.float64)
    budget = budget_gb * (1 << 30)
    n = p["m1"].size
    reqs = [Request(i, 0.0, {k: v[i * 256:(i + 1) * 256] for k, v in p.items()}) for i in range(n // 256)]
    out = {"budget_bytes": int(budget), "binaries": n, "class_share_projected": share.astype(int).tolist()}
    uniform_cap = int(budget // SlotPool.bytes_per_binary_for(rows[-1], N))
    for name, classes in (("footprint_classes", classes_for_budget(budget, N, class_rows=rows,
                                                                   weights=np.maximum(share, 1.0) * np.array(rows))),
                          ("uniform_8192_rows", [(rows[-1], uniform_cap)])):
        srv = Server(data, df=1.0 / (N * dt), dt=dt, N=N, classes=classes, streams=args.serve_streams)
        srv.warm()
        srv.run(reqs[:1])   # first use of the kernels and buffers
        t0 = time.perf_counter()
        res = srv.run(reqs)
        wall = time.perf_counter() - t0
        s = summarize(res)
        out[name] = {"classes": [{"max_step": ms, "capacity": c} for ms, c in classes],
                     "resident_binaries": int(sum(c for _, c in classes)), "pool_bytes": int(srv.memory_bytes()),
                     "seconds": wall, "binaries_per_s": n / wall, "p50_ms": s["p50_ms"], "p99_ms": s["p99_ms"],
                     "reruns": int(srv.reruns), "logl_finite": bool(all(np.isfinite(r.logl).all() for r in res))}
        print(f"[bench] capacity served ({name}): {json.dumps(out[name])}", file=sys.stderr, flush=True)
        del srv, res
        torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the batched path shards by binary and "
              "this benchmark measures one GPU.", file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.init()
    if args.copy_probe:
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 8, "copies": 3}))
        return
    if args.workload == "spline":
        return workload_spline(args)
    if args.workload == "capacity":
        return workload_capacity(args)
    if args.workload == "imr":
        return workload_imr(args)
    if args.workload == "detector":
        return workload_detector(args)
    if args.workload == "ic":
        return workload_ic(args)
    if args.workload == "modes":
        return workload_modes(args)
    pipe = Pipeline(args.n, args.log2N, args.max_step, args.rtol)
    pipe.make_data()
    print(f"[bench] data ready, n={args.n} N=2^{args.log2N}", file=sys.stderr, flush=True)

    tuned = None
    group, block, grid = args.group, 256, 0
    if not args.no_tune:
        from eobgpu_amd import autotune
        tuned = autotune.tune(pipe.add, pipe.params["r0"], max_step=args.max_step, rtol=args.rtol, atol=args.rtol)
        group, block, grid = tuned["group"], tuned["block"], tuned["grid_groups"]
        print(f"[bench] autotune: group {group} block {block} grid_groups {grid} ({len(tuned['table'])} shapes)",
              file=sys.stderr, flush=True)
    wall, stages = time_pipeline(pipe, "persistent", group, args.steps, args.warmup, block=block, grid_groups=grid)
    b = pipe.b
    ns = b.num_steps.cpu().numpy()
    att = b.attempts.cpu().numpy()
    p1 = b.pass1_steps.cpu().numpy()
    si = b.stop_info.cpu().numpy()
    logl = pipe.logl.cpu().numpy()
    print(f"[bench] persistent: {wall*1e3:.2f} ms/step stages {json.dumps({k: round(v, 3) for k, v in stages.items()})}",
          file=sys.stderr, flush=True)

    # the full-length likelihood path (N-sample h+, N-point FFT, N/2+1-bin sum) on
    # the same batch, for its time and for the agreement of the two results
    logl_support = logl.copy()
    wall_d, stages_d = time_pipeline(pipe, "persistent", group, 1, 0, direct=True, block=block, grid_groups=grid)
    logl_direct = pipe.logl.cpu().numpy()
    full_length = {"ms_per_step": wall_d * 1e3, "likelihood_ms": stages_d["likelihood"],
                   "max_rel_diff_logl": float(np.max(np.abs(logl_support - logl_direct) / np.abs(logl_direct)))}
    print(f"[bench] full-length likelihood path: {json.dumps(full_length)}", file=sys.stderr, flush=True)

    chunked = streamed = None
    headline_wall = wall
    # the streamed passes before the chunked one: a ChunkedPass timed first left the
    # streamed pass slower in the same process
    if args.pass_kind == "streamed":
        wall_s, logl_s, traj_s, ngroups, ingest_ms, att_s, per_step, gpu_ms = time_streamed(pipe, args.steps, args.warmup, group=group,
                                                                                           block=block, grid_groups=grid)
        # the same steps with each batch's parameters handed over from host memory
        # inside the step (ordered on the host, one copy over PCIe): not the headline
        wall_h, _, _, _, ingest_h, _, per_step_h, _ = time_streamed(pipe, args.steps, args.warmup, group=group, block=block,
                                                                    grid_groups=grid, resident=False)
        streamed = {"ms_per_step": wall_s * 1e3, "trajectory_ms": traj_s, "downstream_groups": ngroups,
                    "params": "in device memory before the timed steps (set_params_device)",
                    "ingest_ms": ingest_ms, "fresh_batches": True,
                    "per_step_ms": [round(x, 3) for x in per_step], "per_step_gpu_ms": [round(x, 3) for x in gpu_ms],
                    "attempted_steps_last_pass": att_s[-1],
                    "one_batch_ms_per_step": wall * 1e3,
                    "max_rel_diff_logl_vs_one_batch": float(np.max(np.abs(logl_s - logl_support) / np.abs(logl_support))),
                    "host_params": {"ms_per_step": wall_h * 1e3, "evals_per_s": args.n / wall_h, "host_ingest_ms": ingest_h,
                                    "per_step_ms": [round(x, 3) for x in per_step_h]}}
        headline_wall = wall_s
        print(f"[bench] streamed pass: {json.dumps(streamed)}", file=sys.stderr, flush=True)

    sustained = None
    if args.sustained > 0 and args.pass_kind == "streamed":
        res, ing = time_sustained(pipe, args.sustained, max(1, min(args.steps, 3)), 1, group=group, block=block)
        sustained = {"batches": args.sustained, "binaries_per_pass": args.sustained * args.n, "fresh_batches": True,
                     "by_order": {o: {"ms_per_pass": w * 1e3, "evals_per_s": v} for o, (w, v) in res.items()},
                     "host_ingest_ms": ing, "single_batch_evals_per_s": args.n / headline_wall}
        print(f"[bench] sustained ({args.sustained} batches back to back): {json.dumps(sustained)}", file=sys.stderr, flush=True)

    fractions = [float(x) for x in args.chunk_fractions.split(",") if x.strip()]
    if len(fractions) > 1 or args.pass_kind == "chunked":
        wall_c, logl_c, traj_c, sizes = time_chunked(pipe, fractions, args.steps, args.warmup, group=group)
        chunked = {"fractions": fractions, "chunk_sizes": sizes, "ms_per_step": wall_c * 1e3,
                   "trajectory_longest_chunk_ms": traj_c, "one_batch_ms_per_step": wall * 1e3,
                   "max_rel_diff_logl_vs_one_batch": float(np.max(np.abs(logl_c - logl_support) / np.abs(logl_support)))}
        if args.pass_kind == "chunked":
            headline_wall = wall_c
        print(f"[bench] chunked pass: {json.dumps(chunked)}", file=sys.stderr, flush=True)

    graph = None
    if not args.no_graph:
        g = time_graph(pipe, max(1, min(args.steps, 3)), 1)
        graph = {"ms_per_pass_graph": g["graph"], "ms_per_pass_eager": g["eager"], "fresh_batches": True,
                 "likelihood": "full length (N-point transforms)", "one_batch_full_length_ms": full_length["ms_per_step"]}
        print(f"[bench] graph-captured pass: {json.dumps(graph)}", file=sys.stderr, flush=True)

    plain = None
    if not args.no_plain:
        # the reference's batching model: one launch per attempted step, one lane per
        # binary; and the same per-step launches with this library's 64-lane groups
        plain = {}
        for name, grp in (("lane_per_binary", args.plain_group), ("group64", 64)):
            wall_p, stages_p = time_pipeline(pipe, "plain", grp, args.plain_steps, 0)
            plain[name] = {"value": args.n / wall_p, "ms_per_step": wall_p * 1e3, "group": grp,
                           "trajectory_ms": stages_p["trajectory"], "speedup_of_persistent": wall_p / headline_wall}
            print(f"[bench] plain ({name}): {wall_p*1e3:.2f} ms/step", file=sys.stderr, flush=True)

    served = None
    if not args.no_serve:
        served = serve_section(args)

    # the IMR generator on the same kind of batch (five modes, NQC and ringdown)
    imr = None
    if not args.no_imr:
        r = run_imr(args, 2, 1)
        imr = {"waveforms_per_s": r["value"], "ms_per_step": r["ms_per_step"], "stages_ms": r["stages_ms"],
               "attached_binaries": r["attached_binaries"], "config": r["config"]["workload"]}
        print(f"[bench] IMR generator: {json.dumps(imr)}", file=sys.stderr, flush=True)

    # roofline of the dominant kernel
    dom = max(stages, key=stages.get)
    total_attempts = float(att.sum())
    if dom == "trajectory":
        achieved = total_attempts * STEP_FLOPS / (stages["trajectory"] * 1e-3) / 1e12
        roof = {"bound": "fp64-vector", "kernel": "k_integrate_persistent", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": INTEGRATOR_TRAFFIC_BYTES,
                "algorithmic": f"{STEP_FLOPS} FP64 ops per attempted step x {int(total_attempts)} attempted steps"}
    else:
        # support path: h on M_b samples written and read by the FFT, the M_b/2+1 bins read back
        mb = (1 << pipe.lik.mlog2.cpu().numpy().astype(np.int64)) if pipe.lik is not None else np.zeros(1)
        nbytes = {"likelihood": float(np.sum(8 * mb * 2 + 16 * (mb // 2 + 1)))}.get(dom, 0)
        achieved = nbytes / (stages[dom] * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS, "traffic": None}

    cpu = None
    if not args.no_cpu:
        threads, nproc, model = cpu_info()
        k1 = max(1, args.cpu_binaries // 4)
        rate1, el1, logl1 = cpu_baseline(pipe, k1, threads=1)
        k = args.cpu_binaries
        rate, el, logl_cpu = cpu_baseline(pipe, k, threads=threads)
        print(f"[bench] cpu baseline: {rate:.1f} evals/s on {threads} threads, {rate1:.1f} on 1 ({model})", file=sys.stderr,
              flush=True)
        cpu = {"value": rate, "unit": "likelihood evals/s", "cores": threads, "kind": "port",
               "sample": f"whole pass (tables, initial conditions, trajectory, knots, splines, h+ on 2^{args.log2N} samples, "
                         f"rfft, likelihood) for the first {k} binaries of the batch on {threads} threads (oracle restatement + "
                         f"numpy, binaries split over a thread pool), {el:.1f} s",
               "cpu_model": model, "nproc": nproc,
               "single_thread": {"value": rate1, "binaries": k1, "seconds": el1},
               "max_rel_diff_logl_vs_gpu": float(np.max(np.abs(logl_cpu - logl_support[:k]) / np.abs(logl_support[:k])))}

    out = {
        "metric": "likelihood evals/s @1024 uneven binaries", "value": args.n / headline_wall, "unit": "likelihood evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": headline_wall * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "1024 aligned-spin binaries, q~U[1,10], chi~U[-0.9,0.9], log10 M~U[1,2], r0~U[15,40] M; "
                               "two-pass DOP853 rtol=atol=1e-10; (2,2) on 2^20 samples at 1/4096 s; FD likelihood (support-length transforms)",
                   "global_batch": args.n, "seq_len": pipe.N, "parallelism": "none (1 GPU)", "group_lanes": group,
                   "block": block, "grid_groups": grid, "pass": args.pass_kind},
        "autotune": None if tuned is None else {k: tuned[k] for k in ("group", "block", "grid_groups", "time_s", "table")},
        "stages_ms": stages, "one_batch_ms_per_step": wall * 1e3, "streamed": streamed, "chunked": chunked, "steps_per_binary": {"min": int(ns.min()), "median": float(np.median(ns)), "max": int(ns.max())},
        "attempted_steps_total": int(total_attempts), "stop_codes": {str(k): int(v) for k, v in zip(*np.unique(si, return_counts=True))},
        "logl_finite": bool(np.all(np.isfinite(logl))),
        "full_length_likelihood": full_length, "sustained": sustained, "graph_pass": graph, "plain_batch": plain, "served": served,
        "imr": imr,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
