"""CPU: the streamed pass's host schedule (pipeline.drain_finished) against a
simulated integrator that finishes binaries while the host runs, and a
simulated device on which a group's support sizes arrive some polls after
its launch."""
import numpy as np
import pytest

from eobgpu_amd.pipeline import drain_finished


class Sim:
    def __init__(self, n, nslot, seed, lag=3, stall=None):
        self.rng = np.random.default_rng(seed)
        self.done = np.zeros(n, dtype=np.int32)
        self.order = self.rng.permutation(n)   # finishing order
        self.next = 0
        self.nslot, self.lag, self.stall = nslot, lag, stall
        self.clock = 0
        self.in_flight = {}      # slot -> handle
        self.launched, self.completed = [], []

    def tick(self):
        # the integrator finishes 0..3 binaries per host step, up to `stall` of them
        self.clock += 1
        limit = self.done.size if self.stall is None else self.stall
        k = int(self.rng.integers(0, 4))
        for _ in range(k):
            if self.next < limit:
                self.done[self.order[self.next]] = 1
                self.next += 1

    def launch(self, new, slot):
        assert slot not in self.in_flight, "slot reused while its group is in flight"
        assert 0 <= slot < self.nslot
        h = {"id": len(self.launched), "slot": slot, "ready_at": self.clock + self.lag, "idx": new.copy()}
        self.in_flight[slot] = h
        self.launched.append(h)
        self.tick()
        return h

    def is_ready(self, h):
        self.tick()
        return self.clock >= h["ready_at"]

    def complete(self, h):
        assert self.in_flight.pop(h["slot"]) is h
        self.completed.append(h["id"])
        self.tick()

    def ended(self):
        self.tick()
        return self.next >= self.done.size if self.stall is None else self.clock > 10_000


@pytest.mark.parametrize("n,nslot,seed", [(1, 1, 0), (7, 2, 1), (200, 3, 2), (513, 1, 3), (64, 3, 4)])
def test_schedule_covers_every_binary_once_in_order(n, nslot, seed):
    sim = Sim(n, nslot, seed)
    groups = drain_finished(sim.done, nslot, sim.launch, sim.complete, sim.is_ready, sim.ended, poll_s=0.0)
    allidx = np.concatenate(groups)
    assert sorted(allidx.tolist()) == list(range(n))            # every binary downstream exactly once
    assert sim.completed == list(range(len(sim.launched)))      # second halves in launch order
    assert not sim.in_flight                                    # every group's second half queued
    for g in groups:
        assert np.all(np.diff(g) > 0)                          # a group lists its binaries once, ascending
    assert max(len(g) for g in groups) <= n


def test_schedule_raises_when_flags_are_missing():
    sim = Sim(50, 2, 5, stall=20)   # the integrator stops flagging after 20 binaries
    with pytest.raises(RuntimeError, match="without flagging"):
        drain_finished(sim.done, 2, sim.launch, sim.complete, sim.is_ready, sim.ended, poll_s=0.0)
    assert len(np.concatenate([h["idx"] for h in sim.launched])) == 20


@pytest.mark.parametrize("max_group", [1, 5, 64])
def test_schedule_group_limit(max_group):
    """Groups never exceed max_group binaries and still cover every binary once."""
    sim = Sim(300, 3, 7)
    sim.tick = _burst_tick(sim)
    groups = drain_finished(sim.done, 3, sim.launch, sim.complete, sim.is_ready, sim.ended, poll_s=0.0, max_group=max_group)
    assert sorted(np.concatenate(groups).tolist()) == list(range(300))
    assert max(len(g) for g in groups) <= max_group
    assert sim.completed == list(range(len(sim.launched)))


def _burst_tick(sim):
    """An integrator that finishes binaries in bursts of up to 40 (more than a group holds)."""
    def tick():
        sim.clock += 1
        for _ in range(int(sim.rng.integers(0, 41))):
            if sim.next < sim.done.size:
                sim.done[sim.order[sim.next]] = 1
                sim.next += 1
    return tick
