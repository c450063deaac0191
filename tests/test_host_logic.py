"""CPU tests of the host-side logic: slot allocation of the server, the
autotuner's candidate launch shapes, the synthetic workloads, and bench.py's
refusal of more than one GPU."""
import os
import subprocess
import sys

import numpy as np
import pytest

from eobgpu_amd import configs
from eobgpu_amd.autotune import candidates
from eobgpu_amd.serving import Request, SlotPool, summarize, Result

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _pool(capacity):
    pool = SlotPool.__new__(SlotPool)     # the range bookkeeping alone, no device buffers
    pool.capacity = capacity
    pool.free = [(0, capacity)]
    return pool


def test_slot_pool_ranges():
    pool = _pool(10)
    a = SlotPool.alloc(pool, 4)
    b = SlotPool.alloc(pool, 4)
    assert (a, b) == (0, 4)
    assert SlotPool.alloc(pool, 3) is None          # 2 left: a 3-binary request waits
    SlotPool.release(pool, a, 4)
    assert SlotPool.alloc(pool, 3) == 0             # first fit
    SlotPool.release(pool, 0, 3)
    SlotPool.release(pool, 4, 4)
    assert pool.free == [(0, 10)]                   # neighbours merge back


def test_slot_pool_random_churn():
    rng = np.random.default_rng(0)
    pool = _pool(64)
    held = []
    for _ in range(2000):
        if held and rng.uniform() < 0.5:
            s, k = held.pop(int(rng.integers(len(held))))
            SlotPool.release(pool, s, k)
        else:
            k = int(rng.integers(1, 9))
            s = SlotPool.alloc(pool, k)
            if s is not None:
                held.append((s, k))
        used = np.zeros(64, dtype=int)
        for s, k in held:
            used[s:s + k] += 1
        assert used.max() <= 1                      # no slot handed out twice
        free = sum(k for _, k in pool.free)
        assert free + used.sum() == 64
    for s, k in held:
        SlotPool.release(pool, s, k)
    assert pool.free == [(0, 64)]


def test_request_and_summary():
    r = Request(3, 0.25, configs.small_batch(n=5, seed=1))
    assert r.nb == 5
    res = [Result(i, 2, np.zeros(2), np.zeros(2, dtype=np.int32), 0.0, 0.0, 0.01 * (i + 1)) for i in range(10)]
    s = summarize(res)
    assert s["requests"] == 10 and s["binaries"] == 20
    assert s["p50_ms"] <= s["p99_ms"] <= 100.0


@pytest.mark.parametrize("n", [1, 100, 1024, 16384])
def test_autotune_candidates(n):
    c = candidates(n)
    assert c, n
    for g, bl, gg in c:
        assert g in (64, 32, 16, 8) and bl in (64, 128, 256) and bl >= g
        assert 1 <= gg <= n


def test_workload_shapes():
    p = configs.uneven_batch()
    assert p["m1"].size == 1024 and np.all(p["m1"] >= p["m2"])
    assert 15.0 <= p["r0"].min() and p["r0"].max() <= 40.0
    q = p["m1"] / p["m2"]
    assert q.min() >= 1.0 and q.max() <= 10.0
    c = configs.capacity_batch()
    assert c["m1"].size == 16384 and 30.0 <= c["r0"].min() and c["r0"].max() <= 60.0
    reqs = configs.served_requests(duration_s=1.0)
    assert 150 < len(reqs) < 260                     # Poisson, mean 200/s
    assert all(1 <= r[1]["m1"].size <= 256 for r in reqs)


def test_bench_refuses_several_gpus():
    res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 2
    assert "one GPU" in res.stderr


def test_projection_classes_and_budget():
    from eobgpu_amd.serving import assign_classes, classes_for_budget, projected_rows
    p = configs.uneven_batch(n=64, seed=2)
    rows = projected_rows(p)
    assert rows.shape == (64,) and np.all(rows > 64)
    k = assign_classes(rows, [512, 1024, 2048, 4096])
    cr = np.array([512, 1024, 2048, 4096])
    assert np.all((cr[k] >= rows) | (k == 3))             # the class covers the projection, or it is the largest
    assert np.all((k == 0) | (cr[np.maximum(k - 1, 0)] < rows))   # and no smaller class would
    assert list(assign_classes([1, 512, 513, 10**6], [512, 1024])) == [0, 0, 1, 1]
    import eobgpu_amd._lib as L
    try:
        L.lib()
    except Exception:
        pytest.skip("library not built")
    cls = classes_for_budget(8 << 30, 1 << 12, class_rows=(512, 2048))
    assert [ms for ms, _ in cls] == [512, 2048] and cls[0][1] > cls[1][1] > 0
