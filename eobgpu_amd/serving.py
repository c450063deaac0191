"""Continuous-batching server for batched SEOBNR likelihoods on one GPU.

Requests are proposal batches (a few to a few hundred binaries) that arrive
over time.  Each binary's full chain runs on the device: additional args and
tables from its parameters, circular initial conditions, the two-pass DOP853
trajectory (persistent kernel), the (2,2) mode on the knots, splines, the
uniform h+, the FFT and the likelihood against the resident data.

* Memory and admission by projected footprint.  Device memory is organised in
  size classes: a class is a ``SlotPool`` of ``capacity`` binaries whose dense
  output, knots and spline coefficients hold ``max_step`` rows each (the
  trajectory is the part of a binary's footprint that grows with its length).
  A binary goes to the smallest class whose rows cover its projected length
  (``rows_per_length`` x ``pipeline.expected_length``, the fitted attempted-
  step count, plus a margin), so a long binary does not make every slot
  large.  A part of a request (its binaries of one class) is admitted when a
  contiguous range of that class's slots is free; the slots return when its
  results have been read.  A binary whose trajectory fills its rows (the
  reference's driver stops a pass at max_step - 1) is run again in the next
  larger class before its request completes, so a wrong projection costs
  time, never a wrong result.
* Overlap and batching: when a stream is free, the waiting parts of the class
  of the oldest waiting part (in arrival order, up to a group size) are
  admitted together and run as one launch group on that stream, so requests
  that arrive while others are on the GPU are batched into the next group and
  their kernels overlap the running groups' (a small pool of streams, default
  priority).
* No allocation on the admission path: every per-binary buffer lives in the
  class pools and every per-group buffer (parameter staging in pinned memory,
  integrator scratch, likelihood partial sums, the pinned result rows) is
  allocated per stream when the server is built.
* Results do not depend on the mix: every kernel treats each binary alone
  (the integrator's per-binary arithmetic does not depend on which lane group
  or which launch takes the binary, nor on the rows it is given as long as it
  does not fill them; the FFT runs per request part with the same row chunks
  as when the request runs alone; the likelihood reduces each binary in a
  fixed order).

The ``plain`` policy is the comparison point: one request at a time, in
arrival order, with the one-launch-per-step integrator (64-lane groups by
default, so its results equal the continuous policy's bit for bit).
"""
from __future__ import annotations

import collections
import ctypes
import dataclasses
import time

import numpy as np
import torch

from . import pipeline as P
from ._lib import call, lib

MTSUN_SI = P.MTSUN_SI

# stored trajectory points per expected attempted step (pipeline.expected_length):
# 6.0-9.2 over the uneven, small and capacity workloads; 11 leaves a margin
ROWS_PER_LENGTH = 11.0
ROW_MARGIN = 64


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
    """Per-binary device buffers for ``capacity`` binaries of up to ``max_step``
    trajectory rows, handed out as contiguous ranges."""

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

    @staticmethod
    def bytes_per_binary_for(max_step, N):
        L = lib()
        F = N // 2 + 1
        return (int(L.eobhip_bindata_bytes()) + int(L.eobhip_modes_bytes())
                + 8 * max_step * (5 + 3 + 12) + 8 * int(L.eobhip_spline_scratch_doubles(1, max_step)) + 8 * N + 16 * F
                + 4 * 4 + 8 * 2)

    def bytes_per_binary(self):
        return self.bytes_per_binary_for(self.max_step, self.N)

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


def projected_rows(params, rows_per_length=ROWS_PER_LENGTH, margin=ROW_MARGIN):
    """Projected stored trajectory rows per binary from its parameters."""
    e = P.expected_length(params["r0"], params["m1"], params["m2"], params.get("chi1", 0.0), params.get("chi2", 0.0))
    return np.ceil(rows_per_length * np.atleast_1d(e)).astype(np.int64) + margin


def assign_classes(rows, class_rows):
    """Index of the smallest class whose rows cover each projection (the largest
    class for projections beyond every class)."""
    cr = np.asarray(class_rows)
    k = np.searchsorted(cr, np.asarray(rows), side="left")
    return np.minimum(k, cr.size - 1)


def classes_for_budget(budget_bytes, N, class_rows=(512, 1024, 2048, 4096, 8192), weights=None):
    """(max_step, capacity) per class from a device-memory budget: the budget is
    split over the classes by ``weights`` (equal by default) and each share
    buys as many binaries as fit."""
    w = np.ones(len(class_rows)) if weights is None else np.asarray(weights, dtype=np.float64)
    w = w / w.sum()
    out = []
    for ms, share in zip(class_rows, w):
        cap = int(budget_bytes * share // SlotPool.bytes_per_binary_for(ms, N))
        if cap > 0:
            out.append((int(ms), cap))
    if not out:
        raise ValueError("the budget holds no binary of any class")
    return out


class _Part:
    """The binaries of one request in one size class (idx into the request)."""

    def __init__(self, req, idx, cls):
        self.req, self.idx, self.cls = req, np.asarray(idx, dtype=np.int64), int(cls)

    @property
    def nb(self):
        return int(self.idx.size)


class _StreamBuffers:
    """Per-stream buffers of a launch group of up to ``cap`` binaries, allocated
    once; a group of nb binaries uses the first rows x nb elements of each as
    [rows][nb] (contiguous, as the kernels take them).  A stream runs one group
    at a time, so its next group reuses them."""

    def __init__(self, cap, N, device):
        L = lib()
        F = N // 2 + 1
        self.stage = torch.empty(7 * cap, dtype=torch.float64, pin_memory=True)      # m1 m2 chi1 chi2 r0 dT_M amp
        self.par = torch.empty(7 * cap, dtype=torch.float64, device=device)
        self.add = torch.empty(P.NADD * cap, dtype=torch.float64, device=device)
        self.cond = torch.empty(4 * cap, dtype=torch.float64, device=device)
        self.scratch = torch.empty(int(L.eobhip_integrate_scratch_bytes(cap)), dtype=torch.uint8, device=device)
        self.part = torch.empty(int(L.eobhip_fd_likelihood_partial_doubles(cap, F)), dtype=torch.float64, device=device)
        self.out = torch.empty(4 * cap, dtype=torch.float64, pin_memory=True)         # log L, stop code, points, pass-1 steps
        self.out_dev = torch.empty(4 * cap, dtype=torch.float64, device=device)

    def rows(self, t, k, nb):
        return t[:k * nb].view(k, nb)


class Server:
    """Runs a list of requests with continuous batching (``policy="continuous"``) or
    one at a time with the one-launch-per-step integrator (``policy="plain"``).

    classes: list of (max_step, capacity) size classes (default one class of
    ``capacity`` binaries of ``max_step`` rows); ``classes_for_budget`` sizes
    them from a memory budget."""

    def __init__(self, data_fd, df, dt, N, capacity=2048, max_step=2048, streams=4, rtol=1e-10, atol=1e-10, group=64,
                 f0=200.0, device="cuda", plain_group=None, classes=None, rows_per_length=ROWS_PER_LENGTH):
        cls = sorted((int(ms), int(c)) for ms, c in (classes if classes is not None else [(max_step, capacity)]))
        self.pools = [SlotPool(c, ms, N, device) for ms, c in cls]
        self.class_rows = [p.max_step for p in self.pools]
        self.pool = self.pools[0]
        self.device = device
        self.rows_per_length = float(rows_per_length)
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
        cap_max = max(p.capacity for p in self.pools)
        self._buf = [_StreamBuffers(cap_max, self.N, device) for _ in self.streams]
        self._busy = [False] * len(self.streams)
        self.reruns = 0    # binaries run again in a larger class (their rows filled up)

    @property
    def capacity(self):
        return max(p.capacity for p in self.pools)

    def memory_bytes(self):
        """Device bytes the class pools hold."""
        return sum(p.capacity * p.bytes_per_binary() for p in self.pools)

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

    def split(self, req):
        """The request's parts, one per size class its binaries project into."""
        rows = projected_rows(req.params, self.rows_per_length)
        k = assign_classes(rows, self.class_rows)
        return [_Part(req, np.flatnonzero(k == c), c) for c in np.unique(k)]

    # ---------------------------------------------------------- one launch group
    def _launch(self, parts, start, si, mode):
        """All device work of ``parts`` (one class, admitted together, slots
        [start, start + sum nb)) on stream ``si``: the parameters in one copy,
        one set of launches for the tables, initial state, trajectories, knots,
        splines, h+ and the likelihood sums, and the FFT per part (so that a
        request's rows go through the same FFT calls as when it runs alone)."""
        pool = self.pools[parts[0].cls]
        stream, buf = self.streams[si], self._buf[si]
        nb = sum(p.nb for p in parts)
        sl = slice(start, start + nb)
        stage = buf.rows(buf.stage, 7, nb)
        st = stage.numpy()
        off = 0
        for p in parts:
            q = p.req.params
            k = p.nb
            for j, key in enumerate(("m1", "m2", "chi1", "chi2", "r0")):
                st[j, off:off + k] = np.broadcast_to(np.asarray(q[key], dtype=np.float64), (p.req.nb,))[p.idx]
            off += k
        st[5] = self.dt / ((st[0] + st[1]) * MTSUN_SI)
        st[6] = 1.0
        par = buf.rows(buf.par, 7, nb)
        add = buf.rows(buf.add, P.NADD, nb)
        cond = buf.rows(buf.cond, 4, nb)
        with torch.cuda.stream(stream):
            s = stream.cuda_stream
            par.copy_(stage, non_blocking=True)
            bd = pool.bindata[start * pool.bd_bytes:(start + nb) * pool.bd_bytes]
            md = pool.modes[start * pool.mode_bytes:(start + nb) * pool.mode_bytes]
            call("eobhip_params_setup", par[0].data_ptr(), par[1].data_ptr(), par[2].data_ptr(), par[3].data_ptr(), nb, 1, 0.0, 0.0,
                 0.0, 0.0, 0.0, add.data_ptr(), bd.data_ptr(), md.data_ptr(), s)
            call("eobhip_initial_state", par[4].data_ptr(), bd.data_ptr(), md.data_ptr(), nb, cond.data_ptr(), add[4].data_ptr(),
                 1000, 1e-12, s)
            it = ctypes.c_longlong(0)
            dense = pool.dense[sl]
            call("eobhip_integrate", cond.data_ptr(), bd.data_ptr(), md.data_ptr(), nb, pool.max_step, self.rtol, self.atol, 0.01,
                 10.0, 0.05, 1e13, 1, 0, None, None, 1, 0 if mode == "persistent" else 1,
                 self.group if mode == "persistent" else self.plain_group, 256, 0, dense.data_ptr(), pool.num_steps[sl].data_ptr(),
                 pool.stop_info[sl].data_ptr(), pool.pass1[sl].data_ptr(), pool.attempts[sl].data_ptr(), buf.scratch.data_ptr(),
                 ctypes.byref(it), s)
            call("eobhip_hlm22_knots", dense.data_ptr(), pool.num_steps[sl].data_ptr(), bd.data_ptr(), md.data_ptr(), nb,
                 pool.max_step, pool.knots[sl].data_ptr(), s)
            call("eobhip_spline_build_batched", dense.data_ptr(), pool.knots[sl].data_ptr(), pool.num_steps[sl].data_ptr(), nb,
                 pool.max_step, pool.coef[sl].data_ptr(), pool.spl[start * pool.spl_per_bin:].data_ptr(), s)
            call("eobhip_resample_hplus", dense.data_ptr(), pool.coef[sl].data_ptr(), pool.num_steps[sl].data_ptr(),
                 par[6].data_ptr(), par[5].data_ptr(), nb, pool.max_step, self.N, pool.hp[sl].data_ptr(), pool.nsamp[sl].data_ptr(), s)
            row = start
            for p in parts:
                call("eobhip_rfft_batched", pool.hp[row].data_ptr(), pool.ht[row].data_ptr(), p.nb, self.N, s)
                row += p.nb
            call("eobhip_fd_likelihood", pool.ht[sl].data_ptr(), self.data.data_ptr(), buf.part.data_ptr(), nb, pool.F, self.df,
                 self.f0, pool.logl[sl].data_ptr(), s)
            od = buf.rows(buf.out_dev, 4, nb)
            od[0].copy_(pool.logl[sl])
            od[1].copy_(pool.stop_info[sl])
            od[2].copy_(pool.num_steps[sl])
            od[3].copy_(pool.pass1[sl])
            buf.rows(buf.out, 4, nb).copy_(od, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
        return ev

    # ---------------------------------------------------------------- the loop
    def run(self, requests, policy="continuous", poll_s=2e-4, max_group=None):
        """Serve ``requests`` (their ``arrival`` times are honoured against the
        wall clock) and return the list of Results in completion order.

        continuous: whenever a stream is free, the waiting parts of the oldest
        waiting part's class that fit (in arrival order, up to ``max_group``
        binaries, default the class capacity / number of streams) are admitted
        as one launch group on it.
        plain: one request part at a time with the one-launch-per-step integrator."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        for r in pending:
            for p in self.split(r):
                if p.nb > self.pools[p.cls].capacity:
                    raise ValueError(f"request {r.rid} needs {p.nb} slots of class {self.class_rows[p.cls]} rows, "
                                     f"the class has {self.pools[p.cls].capacity}")
        waiting = collections.deque()        # parts, in arrival order
        inflight = []
        results = []
        state = {}                           # rid -> [remaining binaries, logl, stop_info, t_admitted]
        continuous = policy == "continuous"
        max_inflight = len(self.streams) if continuous else 1
        mode = "persistent" if continuous else "plain"
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        while pending or waiting or inflight:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                state[r.rid] = [r.nb, np.zeros(r.nb), np.zeros(r.nb, dtype=np.int32), None]
                waiting.extend(self.split(r))
            progressed = False
            for item in list(inflight):
                parts, start, si, ev, t_adm = item
                if not ev.query():
                    continue
                t_done = time.perf_counter() - t0
                buf = self._buf[si]
                o = buf.rows(buf.out, 4, sum(p.nb for p in parts)).numpy()
                pool = self.pools[parts[0].cls]
                off = 0
                for p in parts:
                    k = p.nb
                    full = (o[2, off:off + k] >= pool.max_step) | (o[3, off:off + k] >= pool.max_step - 1)
                    st = state[p.req.rid]
                    st[3] = t_adm if st[3] is None else min(st[3], t_adm)
                    if full.any() and p.cls + 1 < len(self.pools):
                        # rows filled before a stop: run those binaries again in the next class
                        self.reruns += int(full.sum())
                        waiting.appendleft(_Part(p.req, p.idx[full], p.cls + 1))
                    keep_idx = ~full if p.cls + 1 < len(self.pools) else np.ones(k, dtype=bool)
                    st[1][p.idx[keep_idx]] = o[0, off:off + k][keep_idx]
                    st[2][p.idx[keep_idx]] = o[1, off:off + k][keep_idx].astype(np.int32)
                    st[0] -= int(keep_idx.sum())
                    if st[0] == 0:
                        r = p.req
                        results.append(Result(r.rid, r.nb, st[1], st[2], r.arrival, st[3], t_done))
                    off += k
                pool.release(start, off)
                self._busy[si] = False
                inflight.remove(item)
                progressed = True
            while waiting and len(inflight) < max_inflight:
                cls = waiting[0].cls
                pool = self.pools[cls]
                limit = max(1, pool.capacity // len(self.streams)) if max_group is None else max_group
                group, total = [], 0
                for p in waiting:
                    if p.cls != cls:
                        continue
                    if group and (not continuous or total + p.nb > limit):
                        break
                    group.append(p)
                    total += p.nb
                start = pool.alloc(total)
                while start is None and len(group) > 1:      # the whole group does not fit: admit fewer
                    total -= group.pop().nb
                    start = pool.alloc(total)
                if start is None:
                    break
                for p in group:
                    waiting.remove(p)
                si = self._busy.index(False)
                self._busy[si] = True
                t_adm = time.perf_counter() - t0
                ev = self._launch(group, start, si, mode)
                inflight.append((group, start, si, ev, t_adm))
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
