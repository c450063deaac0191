"""GPU: the whole pass captured as a HIP graph (pipeline.GraphPass) gives, on
every replay, the numbers of the same launches made one by one, for new
parameters copied in between replays; replays alternate with eager work and
synchronisations of the program's own."""
import numpy as np
import pytest
import torch

from eobgpu_amd import configs
from eobgpu_amd import pipeline as P

pytestmark = pytest.mark.gpu

N, DT = 1 << 13, 1.0 / 2048


def _batches(n):
    out = []
    for seed in (21, 22):
        p = configs.uneven_batch(n=n, seed=seed)
        p["r0"] = np.linspace(11.0, 15.0, n)
        out.append(p)
    return out


def _eager_chain(p, data, dev):
    b = P.EOBBatch(P.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"]), max_step=2048, device=dev)
    b.initial_conditions(p["r0"])
    b.integrate(rtol=1e-10, atol=1e-10, two_pass=True)
    b.hlm22_knots()
    b.build_splines()
    hp = b.resample_hplus(p["m1"] + p["m2"], DT, N, amp=0.8)
    ht = P.rfft_batched(hp)
    return P.fd_likelihood(ht, data, 1.0 / (N * DT)).cpu().numpy(), b.num_steps.cpu().numpy()


def test_graph_pass_replays_match_eager(dev):
    n = 12
    rng = np.random.default_rng(12)
    F = N // 2 + 1
    data = torch.tensor((rng.normal(size=F) + 1j * rng.normal(size=F)) * 1e-2, device=dev)
    batches = _batches(n)
    gp = P.GraphPass(n, data, N, DT, max_step=2048, device=dev)
    refs = []
    for p in batches:
        gp.set_params(p["m1"], p["m2"], p["chi1"], p["chi2"], p["r0"], amp=0.8)
        logl = gp.run_eager().cpu().numpy().copy()
        chain, ns = _eager_chain(p, data, dev)
        np.testing.assert_array_equal(logl, chain)          # the same kernels as the step-by-step pipeline
        np.testing.assert_array_equal(gp.batch.num_steps.cpu().numpy(), ns)
        refs.append((logl, ns))
    gp.capture()
    for k in range(5):
        p = batches[k % 2]
        gp.set_params(p["m1"], p["m2"], p["chi1"], p["chi2"], p["r0"], amp=0.8)
        got = gp.run()
        # eager work of the program's own between the replays
        junk = torch.randn(4096, device=dev).sum()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got.cpu().numpy(), refs[k % 2][0])
        np.testing.assert_array_equal(gp.batch.num_steps.cpu().numpy(), refs[k % 2][1])
        assert torch.isfinite(junk)
