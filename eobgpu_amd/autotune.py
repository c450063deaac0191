"""Launch-shape autotuner for the persistent integrator kernel.

The persistent kernel has three launch knobs:

* ``group``: lanes per binary (64 = one binary per wavefront, each RHS spread
  over the wavefront; 16 = four binaries per wavefront, ...);
* ``block``: threads per workgroup (64, 128 or 256);
* ``grid_groups``: how many lane groups are launched (the refill
  granularity); every group takes binaries from the device queue until it is
  empty, so fewer groups than binaries means each group integrates several
  binaries in turn, and more resident groups than the hardware holds leaves
  the rest waiting for a slot.

``tune`` times every candidate on the caller's batch (or on a subsample of
it), keeps the fastest and caches the answer per (batch-size bucket, device
name).  Block size and grid size do not change any result: each binary's
arithmetic is the same whichever group, block or launch takes it (the GPU
tests check persistent == plain bit for bit).  The group width does change
the order in which the 35 flux modes are summed (a 64-lane tree versus
partial sums per lane), i.e. results at the rounding level, so ``tune``
keeps the caller's group (64 by default) unless asked to search groups too.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np
import torch

from . import pipeline as P

DEFAULT_GROUPS = (64,)
DEFAULT_BLOCKS = (256, 128, 64)

_cache: dict = {}


def _bucket(n):
    return 1 << max(0, int(np.ceil(np.log2(max(1, n)))))


def candidates(n, groups=DEFAULT_GROUPS, blocks=DEFAULT_BLOCKS, cu=256):
    """(group, block, grid_groups) triples worth timing for n binaries."""
    out = []
    for g in groups:
        for bl in blocks:
            if bl < g:
                continue
            per_block = bl // g
            # one wavefront per SIMD fits (4 per CU); launch all binaries as groups,
            # just enough groups to fill the resident slots and refill from the
            # queue, three quarters of them (three wavefronts per CU run each a
            # little faster than four), or half
            resident = cu * 4 * (64 // g)
            for gg in sorted({n, min(n, resident), min(n, max(per_block, resident * 3 // 4)),
                              min(n, max(per_block, min(n, resident) // 2))}):
                if gg % per_block and gg != n:
                    continue
                out.append((g, bl, gg))
    return out


def time_config(batch, cond, group, block, grid_groups, reps=1, **kw):
    """Median wall time (s) of ``reps`` integrations with one launch shape."""
    times = []
    for _ in range(reps):
        batch.cond = cond
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batch.integrate(group=group, block=block, grid_groups=grid_groups, **kw)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    return float(np.median(times))


def tune(add, r0, max_step=4096, sample=None, reps=1, rtol=1e-10, atol=1e-10, groups=DEFAULT_GROUPS, blocks=DEFAULT_BLOCKS,
         cache_file=None, seed=0):
    """Return {"group", "block", "grid_groups", "time_s", "table"} for this batch.

    add: additionalArgs [80][n]; r0: initial separations [n].  sample: time on
    this many binaries drawn from the batch (None = the whole batch)."""
    add = np.asarray(add, dtype=np.float64)
    r0 = np.broadcast_to(np.asarray(r0, dtype=np.float64), (add.shape[1],))
    n_full = add.shape[1]
    dev_name = torch.cuda.get_device_name(0)
    key = (_bucket(n_full), dev_name, tuple(groups), tuple(blocks))
    if key in _cache:
        return _cache[key]
    if cache_file and os.path.exists(cache_file):
        with open(cache_file) as f:
            disk = json.load(f)
        hit = disk.get(json.dumps([key[0], key[1], list(groups), list(blocks)]))
        if hit:
            _cache[key] = hit
            return hit
    if sample is not None and sample < n_full:
        idx = np.sort(np.random.default_rng(seed).choice(n_full, size=sample, replace=False))
        add, r0 = add[:, idx], r0[idx]
    n = add.shape[1]
    b = P.EOBBatch(add, max_step=max_step)
    b.initial_conditions(r0)
    cond = b.cond.clone()
    table = []
    for g, bl, gg in candidates(n, groups, blocks):
        t = time_config(b, cond, g, bl, gg, reps=reps, rtol=rtol, atol=atol, two_pass=True)
        table.append({"group": g, "block": bl, "grid_groups": gg, "time_s": t})
    best = min(table, key=lambda e: e["time_s"])
    # grid_groups scales with the batch when the tuned batch was a subsample
    gg = best["grid_groups"] if best["grid_groups"] < n else n_full
    res = {"group": best["group"], "block": best["block"], "grid_groups": gg, "time_s": best["time_s"], "tuned_on": n,
           "table": table}
    _cache[key] = res
    if cache_file:
        disk = {}
        if os.path.exists(cache_file):
            with open(cache_file) as f:
                disk = json.load(f)
        disk[json.dumps([key[0], key[1], list(groups), list(blocks)])] = res
        with open(cache_file, "w") as f:
            json.dump(disk, f, indent=1)
    return res
