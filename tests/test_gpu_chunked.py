"""GPU: the stream-overlapped pass (ChunkedPass) gives every binary the numbers
of the one-batch pass: the same trajectory, bit for bit (the integrator treats
binaries independently), and the same log L up to the last bits (its
transforms run at other batch sizes)."""
import numpy as np
import pytest
import torch

from eobgpu_amd import configs
from eobgpu_amd import pipeline as P

pytestmark = pytest.mark.gpu

N, DT = 1 << 13, 1.0 / 2048


def test_chunked_pass_matches_one_batch(dev):
    n = 20
    p = configs.uneven_batch(n=n, seed=9)
    p["r0"] = np.linspace(11.0, 16.0, n)
    M = p["m1"] + p["m2"]
    amp = np.linspace(0.5, 1.5, n)
    rng = np.random.default_rng(4)
    F = N // 2 + 1
    data = (rng.normal(size=F) + 1j * rng.normal(size=F)) * 1e-2
    lik = P.SupportLikelihood(data, N, 1.0 / (N * DT), m_min_log2=6)

    b = P.EOBBatch(P.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"]), max_step=2048, device=dev)
    b.initial_conditions(p["r0"])
    b.integrate(rtol=1e-10, atol=1e-10, two_pass=True, block=64)
    b.hlm22_knots()
    b.build_splines()
    ref = lik(b, M, DT, amp=amp).cpu().numpy()

    cp = P.ChunkedPass(p["m1"], p["m2"], p["chi1"], p["chi2"], p["r0"], lik, M, DT, amp=amp, max_step=2048, fractions=(0.2, 0.3, 0.5))
    for _ in range(2):   # a second pass reuses the chunks' buffers
        got = cp.run().cpu().numpy()
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)
    # trajectories: each chunk's binaries against the same binaries of the one batch
    order = np.concatenate(cp.groups)
    assert sorted(order.tolist()) == list(range(n))
    for g, cb in zip(cp.groups, cp.batches):
        ns = cb.num_steps.cpu().numpy()
        np.testing.assert_array_equal(ns, b.num_steps.cpu().numpy()[g])
        L = int(ns.max())
        np.testing.assert_array_equal(cb.dense[:, :L].cpu().numpy(), b.dense[torch.as_tensor(g, device=dev)][:, :L].cpu().numpy())
    # the longest expected binaries went to the first chunk
    key = P.expected_length(p["r0"], p["m1"], p["m2"], p["chi1"], p["chi2"])
    assert key[cp.groups[0]].min() >= key[cp.groups[-1]].max()
