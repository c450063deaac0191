"""Continuous-batching server for batched SEOBNR likelihoods on one GPU.

Requests are proposal batches (a few to a few hundred binaries) that arrive
over time.  Each binary's full chain runs on the device: table setup,
circular initial conditions, the two-pass DOP853 trajectory (persistent
kernel), the (2,2) mode on the knots, splines, the uniform h+, the FFT and
the likelihood against the resident data.

* Memory: one ``SlotPool`` holds every per-binary buffer for ``capacity``
  binaries (tables, trajectory, knots, spline coefficients, h+, h~+, log L).
  A request is admitted when a contiguous range of slots for all its
  binaries is free (admission by memory footprint); its slots are released
  when its result has been read.
* Overlap: admitted requests run on a small pool of streams (default
  priority), so a request that arrives while others are on the GPU starts
  at once on a free stream and its kernels overlap theirs.
* Results do not depend on the mix: every kernel treats each binary alone
  (the integrator's per-binary arithmetic does not depend on which lane group
  or which launch takes the binary; the FFT runs per request in fixed-size
  row chunks; the likelihood reduces each binary in a fixed order).

The ``plain`` policy is the comparison point: one request at a time, in
arrival order, with the one-launch-per-step integrator.
"""
from __future__ import annotations

import collections
import dataclasses
import time

import numpy as np
import torch

from . import pipeline as P
from ._lib import call, lib

MTSUN_SI = P.MTSUN_SI


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float                     # seconds after the start of the run
    params: dict                       # m1, m2 (Msun), chi1, chi2, r0 arrays of one length

    @property
    def nb(self):
        return int(np.asarray(self.params["m1"]).size)


@dataclasses.dataclass
class Result:
    rid: int
    nb: int
    logl: np.ndarray
    stop_info: np.ndarray
    arrival: float
    admitted: float
    done: float

    @property
    def latency(self):
        return self.done - self.arrival


class SlotPool:
    """Per-binary device buffers for ``capacity`` binaries, handed out as contiguous ranges."""

    def __init__(self, capacity, max_step, N, device="cuda"):
        L = lib()
        self.capacity, self.max_step, self.N = int(capacity), int(max_step), int(N)
        self.F = self.N // 2 + 1
        c, ms = self.capacity, self.max_step
        self.bd_bytes, self.mode_bytes = int(L.eobhip_bindata_bytes()), int(L.eobhip_modes_bytes())
        self.bindata = torch.empty(c * self.bd_bytes, dtype=torch.uint8, device=device)
        self.modes = torch.empty(c * self.mode_bytes, dtype=torch.uint8, device=device)
        self.dense = torch.zeros((c, ms, 5), dtype=torch.float64, device=device)
        self.knots = torch.zeros((c, ms, 3), dtype=torch.float64, device=device)
        self.coef = torch.zeros((c, ms, 12), dtype=torch.float64, device=device)
        self.spl_per_bin = int(L.eobhip_spline_scratch_doubles(1, ms))
        self.spl = torch.zeros(c * self.spl_per_bin, dtype=torch.float64, device=device)
        self.hp = torch.zeros((c, self.N), dtype=torch.float64, device=device)
        self.ht = torch.zeros((c, self.F), dtype=torch.complex128, device=device)
        self.num_steps = torch.zeros(c, dtype=torch.int32, device=device)
        self.stop_info = torch.zeros(c, dtype=torch.int32, device=device)
        self.pass1 = torch.zeros(c, dtype=torch.int32, device=device)
        self.attempts = torch.zeros(c, dtype=torch.int64, device=device)
        self.nsamp = torch.zeros(c, dtype=torch.int32, device=device)
        self.logl = torch.zeros(c, dtype=torch.float64, device=device)
        self.free = [(0, c)]                 # sorted (start, length)

    def bytes_per_binary(self):
        ms = self.max_step
        return self.bd_bytes + self.mode_bytes + 8 * ms * (5 + 3 + 12) + 8 * self.spl_per_bin + 8 * self.N + 16 * self.F + 4 * 5 + 8 * 2

    def alloc(self, nb):
        for i, (s, ln) in enumerate(self.free):
            if ln >= nb:
                if ln == nb:
                    self.free.pop(i)
                else:
                    self.free[i] = (s + nb, ln - nb)
                return s
        return None

    def release(self, start, nb):
        self.free.append((start, nb))
        self.free.sort()
        merged = []
        for s, ln in self.free:
            if merged and merged[-1][0] + merged[-1][1] == s:
                merged[-1] = (merged[-1][0], merged[-1][1] + ln)
            else:
                merged.append((s, ln))
        self.free = merged


class Server:
    """Runs a list of requests with continuous batching (``policy="continuous"``) or
    one at a time with the one-launch-per-step integrator (``policy="plain"``)."""

    def __init__(self, data_fd, df, dt, N, capacity=2048, max_step=2048, streams=4, rtol=1e-10, atol=1e-10, group=64,
                 f0=200.0, device="cuda"):
        self.pool = SlotPool(capacity, max_step, N, device)
        self.device = device
        self.data = torch.as_tensor(data_fd, dtype=torch.complex128, device=device).contiguous()
        self.df, self.dt, self.N, self.f0 = float(df), float(dt), int(N), float(f0)
        self.rtol, self.atol, self.group = rtol, atol, group
        self.streams = [torch.cuda.Stream(device=device) for _ in range(max(1, int(streams)))]
        self._stream_load = [0] * len(self.streams)

    # -------------------------------------------------------------- one request
    def _launch(self, req, start, stream, mode):
        pool, nb = self.pool, req.nb
        p = req.params
        sl = slice(start, start + nb)
        add = P.make_additional_args(np.asarray(p["m1"], dtype=np.float64), np.asarray(p["m2"], dtype=np.float64),
                                     np.asarray(p["chi1"], dtype=np.float64), np.asarray(p["chi2"], dtype=np.float64))
        M = np.asarray(p["m1"], dtype=np.float64) + np.asarray(p["m2"], dtype=np.float64)
        host = np.concatenate([add.ravel(), np.asarray(p["r0"], dtype=np.float64), self.dt / (M * MTSUN_SI)])
        with torch.cuda.stream(stream):
            s = stream.cuda_stream
            h_pin = torch.from_numpy(host).pin_memory()
            dev = h_pin.to(self.device, non_blocking=True)
            add_d = dev[:80 * nb]
            r0 = dev[80 * nb:81 * nb]
            dT_M = dev[81 * nb:82 * nb]
            bd = pool.bindata[start * pool.bd_bytes:(start + nb) * pool.bd_bytes]
            md = pool.modes[start * pool.mode_bytes:(start + nb) * pool.mode_bytes]
            call("eobhip_bin_setup", add_d.data_ptr(), nb, 1, bd.data_ptr(), md.data_ptr(), s)
            ic = torch.empty((3, nb), dtype=torch.float64, device=self.device)
            call("eobhip_initial_conditions", r0.data_ptr(), bd.data_ptr(), md.data_ptr(), nb, ic[0].data_ptr(), ic[1].data_ptr(),
                 ic[2].data_ptr(), 1000, 1e-12, s)
            cond = torch.stack([r0, torch.zeros_like(r0), ic[1], ic[0]]).contiguous()
            scratch = torch.empty(int(lib().eobhip_integrate_scratch_bytes(nb)), dtype=torch.uint8, device=self.device)
            it = P.ctypes.c_longlong(0)
            dense = pool.dense[sl]
            call("eobhip_integrate", cond.data_ptr(), bd.data_ptr(), md.data_ptr(), nb, pool.max_step, self.rtol, self.atol, 0.01,
                 10.0, 0.05, 1e13, 1, 0, None, None, 1, 0 if mode == "persistent" else 1,
                 self.group if mode == "persistent" else 1, 256, 0, dense.data_ptr(), pool.num_steps[sl].data_ptr(),
                 pool.stop_info[sl].data_ptr(), pool.pass1[sl].data_ptr(), pool.attempts[sl].data_ptr(), scratch.data_ptr(),
                 P.ctypes.byref(it), s)
            call("eobhip_hlm22_knots", dense.data_ptr(), pool.num_steps[sl].data_ptr(), bd.data_ptr(), md.data_ptr(), nb,
                 pool.max_step, pool.knots[sl].data_ptr(), s)
            call("eobhip_spline_build_batched", dense.data_ptr(), pool.knots[sl].data_ptr(), pool.num_steps[sl].data_ptr(), nb,
                 pool.max_step, pool.coef[sl].data_ptr(), pool.spl[start * pool.spl_per_bin:].data_ptr(), s)
            amp = torch.ones(nb, dtype=torch.float64, device=self.device)
            call("eobhip_resample_hplus", dense.data_ptr(), pool.coef[sl].data_ptr(), pool.num_steps[sl].data_ptr(), amp.data_ptr(),
                 dT_M.data_ptr(), nb, pool.max_step, self.N, pool.hp[sl].data_ptr(), pool.nsamp[sl].data_ptr(), s)
            call("eobhip_rfft_batched", pool.hp[sl].data_ptr(), pool.ht[sl].data_ptr(), nb, self.N, s)
            part = torch.empty(int(lib().eobhip_fd_likelihood_partial_doubles(nb, pool.F)), dtype=torch.float64, device=self.device)
            call("eobhip_fd_likelihood", pool.ht[sl].data_ptr(), self.data.data_ptr(), part.data_ptr(), nb, pool.F, self.df, self.f0,
                 pool.logl[sl].data_ptr(), s)
            out = torch.empty((2, nb), dtype=torch.float64, pin_memory=True)
            out[0].copy_(pool.logl[sl], non_blocking=True)
            out[1].copy_(pool.stop_info[sl].to(torch.float64), non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
            # keep temporaries alive until the event has passed
            keep = (h_pin, dev, ic, cond, scratch, amp, part)
        return ev, out, keep

    # ---------------------------------------------------------------- the loop
    def run(self, requests, policy="continuous", poll_s=2e-4):
        """Serve ``requests`` (their ``arrival`` times are honoured against the
        wall clock) and return the list of Results in completion order."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        for r in pending:
            if r.nb > self.pool.capacity:
                raise ValueError(f"request {r.rid} needs {r.nb} slots, the pool has {self.pool.capacity}")
        waiting = collections.deque()
        inflight = []
        results = []
        max_inflight = len(self.streams) if policy == "continuous" else 1
        mode = "persistent" if policy == "continuous" else "plain"
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        while pending or waiting or inflight:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            progressed = False
            for item in list(inflight):
                req, start, si, ev, out, keep, t_adm = item
                if ev.query():
                    t_done = time.perf_counter() - t0
                    o = out.numpy()
                    results.append(Result(req.rid, req.nb, o[0].copy(), o[1].astype(np.int32), req.arrival, t_adm, t_done))
                    self.pool.release(start, req.nb)
                    self._stream_load[si] -= 1
                    inflight.remove(item)
                    progressed = True
            while waiting and len(inflight) < max_inflight:
                req = waiting[0]
                start = self.pool.alloc(req.nb)
                if start is None:
                    break
                waiting.popleft()
                si = int(np.argmin(self._stream_load))
                self._stream_load[si] += 1
                t_adm = time.perf_counter() - t0
                ev, out, keep = self._launch(req, start, self.streams[si], mode)
                inflight.append((req, start, si, ev, out, keep, t_adm))
                progressed = True
            if not progressed:
                if not inflight and not waiting and pending:
                    time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - t0), 0.01)))
                else:
                    time.sleep(poll_s)
        torch.cuda.synchronize()
        return results


def summarize(results):
    lat = np.array([r.latency for r in results])
    nbin = sum(r.nb for r in results)
    span = max(r.done for r in results) - min(r.arrival for r in results)
    return {"requests": len(results), "binaries": int(nbin), "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3),
            "throughput_binaries_per_s": nbin / span if span > 0 else float("nan"),
            "throughput_requests_per_s": len(results) / span if span > 0 else float("nan")}
