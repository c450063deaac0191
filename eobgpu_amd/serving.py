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
* Overlap and batching: when a stream is free, every request waiting at that
  moment (up to a group size) is admitted together and runs as one launch
  group on that stream, so requests that arrive while others are on the GPU
  are batched into the next group and their kernels overlap the running
  groups' (a small pool of streams, default priority).
* Results do not depend on the mix: every kernel treats each binary alone
  (the integrator's per-binary arithmetic does not depend on which lane group
  or which launch takes the binary; the FFT runs per request with the same
  row chunks as when the request runs alone; the likelihood reduces each
  binary in a fixed order).

The ``plain`` policy is the comparison point: one request at a time, in
arrival order, with the one-launch-per-step integrator (64-lane groups by
default, so its results equal the continuous policy's bit for bit).
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
                 f0=200.0, device="cuda", plain_group=None):
        self.pool = SlotPool(capacity, max_step, N, device)
        self.device = device
        self.data = torch.as_tensor(data_fd, dtype=torch.complex128, device=device).contiguous()
        self.df, self.dt, self.N, self.f0 = float(df), float(dt), int(N), float(f0)
        self.rtol, self.atol, self.group = rtol, atol, group
        # lanes per binary of the plain policy: the same as the persistent one by
        # default, so that both policies give bit-identical results
        self.plain_group = group if plain_group is None else plain_group
        # streams whose kernels run concurrently (a stream that shares a hardware
        # queue with another would hold its requests behind the other's)
        self.streams = P.concurrent_streams(max(1, int(streams)), device=device)
        if not self.streams:
            raise RuntimeError("no stream available")
        self._stream_load = [0] * len(self.streams)

    def warm(self):
        """Create the FFT plans of every row-chunk size on every stream (plan
        creation allocates device memory, which is kept out of the served path)."""
        rows = 2 * 64 - 1      # 64 + 32 + 16 + 8 + 4 + 2 + 1
        h = torch.zeros((rows, self.N), dtype=torch.float64, device=self.device)
        ht = torch.empty((rows, self.N // 2 + 1), dtype=torch.complex128, device=self.device)
        for st in self.streams:
            with torch.cuda.stream(st):
                call("eobhip_rfft_batched", h.data_ptr(), ht.data_ptr(), rows, self.N, st.cuda_stream)
        torch.cuda.synchronize()

    # ---------------------------------------------------------- one launch group
    def _launch(self, reqs, start, stream, mode):
        """All device work of the requests ``reqs`` (admitted together, slots
        [start, start + sum nb)) on ``stream``: one set of launches for the
        tables, initial state, trajectories, knots, splines, h+ and the
        likelihood sums, and the FFT per request (so that every request's rows
        go through the same FFT calls as when it runs alone)."""
        pool = self.pool
        nbs = [r.nb for r in reqs]
        nb = sum(nbs)
        sl = slice(start, start + nb)
        cat = lambda key: np.concatenate([np.asarray(r.params[key], dtype=np.float64).reshape(-1) for r in reqs])
        m1, m2, c1, c2, r0h = cat("m1"), cat("m2"), cat("chi1"), cat("chi2"), cat("r0")
        add = P.make_additional_args(m1, m2, c1, c2)
        host = np.concatenate([add.ravel(), r0h, self.dt / ((m1 + m2) * MTSUN_SI)])
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
                 self.group if mode == "persistent" else self.plain_group, 256, 0, dense.data_ptr(), pool.num_steps[sl].data_ptr(),
                 pool.stop_info[sl].data_ptr(), pool.pass1[sl].data_ptr(), pool.attempts[sl].data_ptr(), scratch.data_ptr(),
                 P.ctypes.byref(it), s)
            call("eobhip_hlm22_knots", dense.data_ptr(), pool.num_steps[sl].data_ptr(), bd.data_ptr(), md.data_ptr(), nb,
                 pool.max_step, pool.knots[sl].data_ptr(), s)
            call("eobhip_spline_build_batched", dense.data_ptr(), pool.knots[sl].data_ptr(), pool.num_steps[sl].data_ptr(), nb,
                 pool.max_step, pool.coef[sl].data_ptr(), pool.spl[start * pool.spl_per_bin:].data_ptr(), s)
            amp = torch.ones(nb, dtype=torch.float64, device=self.device)
            call("eobhip_resample_hplus", dense.data_ptr(), pool.coef[sl].data_ptr(), pool.num_steps[sl].data_ptr(), amp.data_ptr(),
                 dT_M.data_ptr(), nb, pool.max_step, self.N, pool.hp[sl].data_ptr(), pool.nsamp[sl].data_ptr(), s)
            row = start
            for k in nbs:
                call("eobhip_rfft_batched", pool.hp[row].data_ptr(), pool.ht[row].data_ptr(), k, self.N, s)
                row += k
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
    def run(self, requests, policy="continuous", poll_s=2e-4, max_group=None):
        """Serve ``requests`` (their ``arrival`` times are honoured against the
        wall clock) and return the list of Results in completion order.

        continuous: whenever a stream is free, every waiting request that fits
        (in arrival order, up to ``max_group`` binaries, default the pool's
        capacity / number of streams) is admitted as one launch group on it.
        plain: one request at a time with the one-launch-per-step integrator."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        for r in pending:
            if r.nb > self.pool.capacity:
                raise ValueError(f"request {r.rid} needs {r.nb} slots, the pool has {self.pool.capacity}")
        if max_group is None:
            max_group = max(1, self.pool.capacity // len(self.streams))
        waiting = collections.deque()
        inflight = []
        results = []
        continuous = policy == "continuous"
        max_inflight = len(self.streams) if continuous else 1
        mode = "persistent" if continuous else "plain"
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        while pending or waiting or inflight:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            progressed = False
            for item in list(inflight):
                reqs, start, si, ev, out, keep, t_adm = item
                if ev.query():
                    t_done = time.perf_counter() - t0
                    o = out.numpy()
                    off = 0
                    for r in reqs:
                        results.append(Result(r.rid, r.nb, o[0, off:off + r.nb].copy(), o[1, off:off + r.nb].astype(np.int32),
                                              r.arrival, t_adm, t_done))
                        off += r.nb
                    self.pool.release(start, off)
                    self._stream_load[si] -= 1
                    inflight.remove(item)
                    progressed = True
            while waiting and len(inflight) < max_inflight:
                group, total = [], 0
                for r in waiting:
                    if group and (not continuous or total + r.nb > max_group):
                        break
                    group.append(r)
                    total += r.nb
                start = self.pool.alloc(total)
                while start is None and len(group) > 1:      # the whole group does not fit: admit fewer
                    total -= group.pop().nb
                    start = self.pool.alloc(total)
                if start is None:
                    break
                for _ in group:
                    waiting.popleft()
                si = int(np.argmin(self._stream_load))
                self._stream_load[si] += 1
                t_adm = time.perf_counter() - t0
                ev, out, keep = self._launch(group, start, self.streams[si], mode)
                inflight.append((group, start, si, ev, out, keep, t_adm))
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
