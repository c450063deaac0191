"""GPU: the stream-overlapped passes (ChunkedPass, StreamedPass) give every binary the numbers
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


def test_streamed_pass_matches_one_batch(dev):
    """StreamedPass: one integrator launch that flags each finished binary, the
    downstream work queued per group of finished binaries; the same numbers as
    the one-batch pass."""
    n = 24
    p = configs.uneven_batch(n=n, seed=11)
    p["r0"] = np.linspace(11.0, 17.0, n)
    M = p["m1"] + p["m2"]
    amp = np.linspace(0.5, 1.5, n)
    rng = np.random.default_rng(5)
    F = N // 2 + 1
    data = (rng.normal(size=F) + 1j * rng.normal(size=F)) * 1e-2
    lik = P.SupportLikelihood(data, N, 1.0 / (N * DT), m_min_log2=6)

    b = P.EOBBatch(P.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"]), max_step=2048, device=dev)
    b.initial_conditions(p["r0"])
    b.integrate(rtol=1e-10, atol=1e-10, two_pass=True, block=64)
    b.hlm22_knots()
    b.build_splines()
    ref = lik(b, M, DT, amp=amp).cpu().numpy()

    sp = P.StreamedPass(p["m1"], p["m2"], p["chi1"], p["chi2"], p["r0"], lik, M, DT, amp=amp, max_step=2048)
    for _ in range(2):   # a second pass reuses the flags and the downstream buffers
        got = sp.run().cpu().numpy()
        torch.cuda.synchronize()
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)
        assert (sp.done.numpy() == 1).all()
        taken = np.concatenate(sp.groups)
        assert sorted(taken.tolist()) == list(range(n))   # every binary downstream exactly once
    # the batch is the caller's binaries longest-expected first; trajectories bit for bit
    key = P.expected_length(p["r0"], p["m1"], p["m2"], p["chi1"], p["chi2"])
    assert np.all(np.diff(key[sp.order]) <= 0)
    ns = sp.batch.num_steps.cpu().numpy()
    np.testing.assert_array_equal(ns, b.num_steps.cpu().numpy()[sp.order])
    L = int(ns.max())
    np.testing.assert_array_equal(sp.batch.dense[:, :L].cpu().numpy(),
                                  b.dense[torch.as_tensor(sp.order, device=dev)][:, :L].cpu().numpy())
    np.testing.assert_array_equal(sp.batch.attempts.cpu().numpy(), b.attempts.cpu().numpy()[sp.order])


def test_integrate_flags_need_pinned_memory(dev):
    b = P.EOBBatch(P.make_additional_args([20.0], [10.0], [0.1], [0.0]), max_step=256, device=dev)
    b.initial_conditions([12.0])
    with pytest.raises(ValueError):
        b.integrate(done=torch.zeros(1, dtype=torch.int32))            # pageable host memory
    with pytest.raises(ValueError):
        b.integrate(done=torch.zeros(1, dtype=torch.int32, device=dev))   # device memory
    with pytest.raises(ValueError):
        b.integrate(mode="plain", done=torch.zeros(1, dtype=torch.int32, pin_memory=True))


def test_gather_binaries(dev):
    """eobhip_gather_binaries against torch's index_select on a small batch (rows
    below num_steps, tables, amplitudes, positions) and dT_M against the host's
    expression, bit for bit."""
    n, ms = 6, 128
    p = configs.uneven_batch(n=n, seed=3)
    b = P.EOBBatch(P.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"]), max_step=ms, device=dev)
    gen = torch.Generator(device="cpu").manual_seed(0)
    b.dense.copy_(torch.randn(b.dense.shape, generator=gen, dtype=torch.float64))
    b.num_steps.copy_(torch.tensor([5, 128, 1, 77, 30, 2], dtype=torch.int32))
    M = torch.as_tensor(p["m1"] + p["m2"], device=dev)
    amp = torch.linspace(0.5, 2.0, n, dtype=torch.float64, device=dev)
    pos = torch.arange(100, 100 + n, dtype=torch.int64, device=dev)
    rows = P._Rows(n, ms, dev)
    idx = torch.tensor([4, 1, 3], dtype=torch.int64, device=dev)
    k, dt = idx.numel(), 1.0 / 4096
    rows.dense.zero_()
    P.call("eobhip_gather_binaries", idx.data_ptr(), k, n, ms, b.dense.data_ptr(), b.num_steps.data_ptr(), b.bindata.data_ptr(),
           rows.bd_bytes, b.modes.data_ptr(), rows.md_bytes, M.data_ptr(), amp.data_ptr(), pos.data_ptr(), dt, P.MTSUN_SI,
           rows.dense.data_ptr(), rows.num_steps.data_ptr(), rows.bindata.data_ptr(), rows.modes.data_ptr(), rows.dT_M.data_ptr(),
           rows.amp.data_ptr(), rows.pos.data_ptr(), P.current_stream())
    torch.cuda.synchronize()
    ns = b.num_steps[idx].cpu().numpy()
    np.testing.assert_array_equal(rows.num_steps[:k].cpu().numpy(), ns)
    for j in range(k):
        np.testing.assert_array_equal(rows.dense[j, :ns[j]].cpu().numpy(), b.dense[idx[j], :ns[j]].cpu().numpy())
        assert not rows.dense[j, ns[j]:].any()   # rows at and past num_steps are not copied
    np.testing.assert_array_equal(rows.bindata[:k].cpu().numpy(), b.bindata.view(n, -1)[idx].cpu().numpy())
    np.testing.assert_array_equal(rows.modes[:k].cpu().numpy(), b.modes.view(n, -1)[idx].cpu().numpy())
    np.testing.assert_array_equal(rows.amp[:k].cpu().numpy(), amp[idx].cpu().numpy())
    np.testing.assert_array_equal(rows.pos[:k].cpu().numpy(), pos[idx].cpu().numpy())
    np.testing.assert_array_equal(rows.dT_M[:k].cpu().numpy(), (dt / (M * P.MTSUN_SI))[idx].cpu().numpy())


def test_params_setup_matches_host(dev):
    """eobhip_params_setup builds the additional args and tables on the device bit
    for bit as the host's make_additional_args + eobhip_bin_setup (equal masses
    included, where the odd-m prefixes take their special values)."""
    p = configs.uneven_batch(n=40, seed=12)
    p["m2"][:3] = p["m1"][:3]
    add = P.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"])
    b_host = P.EOBBatch(add, max_step=16, device=dev)
    b_dev = P.EOBBatch.from_params(p["m1"], p["m2"], p["chi1"], p["chi2"], max_step=16, device=dev)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(b_dev.add.cpu().numpy(), add)
    np.testing.assert_array_equal(b_dev.bindata.cpu().numpy(), b_host.bindata.cpu().numpy())
    np.testing.assert_array_equal(b_dev.modes.cpu().numpy(), b_host.modes.cpu().numpy())
    # new parameters into the same buffers, from device tensors
    q = configs.uneven_batch(n=40, seed=13)
    b_dev.set_params(*(torch.tensor(q[k], device=dev) for k in ("m1", "m2", "chi1", "chi2")))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(b_dev.add.cpu().numpy(), P.make_additional_args(q["m1"], q["m2"], q["chi1"], q["chi2"]))


def test_streamed_pass_new_parameters(dev):
    """StreamedPass.set_params and set_params_device: a pass over a new batch in
    the same buffers gives that batch's one-batch numbers."""
    n = 16
    p = configs.uneven_batch(n=n, seed=14)
    q = configs.uneven_batch(n=n, seed=15)
    for d in (p, q):
        d["r0"] = np.linspace(11.0, 15.0, n)[::-1].copy()
    rng = np.random.default_rng(6)
    F = N // 2 + 1
    data = (rng.normal(size=F) + 1j * rng.normal(size=F)) * 1e-2
    lik = P.SupportLikelihood(data, N, 1.0 / (N * DT), m_min_log2=6)
    b = P.EOBBatch(P.make_additional_args(q["m1"], q["m2"], q["chi1"], q["chi2"]), max_step=2048, device=dev)
    b.initial_conditions(q["r0"])
    b.integrate(rtol=1e-10, atol=1e-10, two_pass=True, block=64)
    b.hlm22_knots()
    b.build_splines()
    Mq = q["m1"] + q["m2"]
    ref = lik(b, Mq, DT, amp=0.7).cpu().numpy()
    sp = P.StreamedPass(p["m1"], p["m2"], p["chi1"], p["chi2"], p["r0"], lik, p["m1"] + p["m2"], DT, amp=0.7, max_step=2048)
    sp.set_params(q["m1"], q["m2"], q["chi1"], q["chi2"], q["r0"], amp=0.7)
    got = sp.run().cpu().numpy()
    torch.cuda.synchronize()
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)
    ns = sp.batch.num_steps.cpu().numpy()
    np.testing.assert_array_equal(ns, b.num_steps.cpu().numpy()[sp.order])
    # the same batch from device memory (set_params_device): the caller's order
    par = torch.tensor(np.stack([q["m1"], q["m2"], q["chi1"], q["chi2"], q["r0"], Mq, np.full(n, 0.7)]), device=dev)
    sp.set_params_device(par)
    got = sp.run().cpu().numpy()
    torch.cuda.synchronize()
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)
    np.testing.assert_array_equal(sp.order, np.arange(n))
    np.testing.assert_array_equal(sp.batch.num_steps.cpu().numpy(), b.num_steps.cpu().numpy())
    with pytest.raises(ValueError):
        sp.set_params_device(par[:6])
