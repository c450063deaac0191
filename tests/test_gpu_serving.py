"""Continuous-batching server: every request's result is the same alone and in
the mix (bit for bit: each kernel treats each binary independently), the same
under the plain one-request-at-a-time policy, and the same as the batch
pipeline's; memory admission makes requests wait for slots rather than fail."""
import numpy as np
import pytest
import torch

from eobgpu_amd import configs
from eobgpu_amd import pipeline as P
from eobgpu_amd.serving import Request, Server, summarize

pytestmark = pytest.mark.gpu

N = 1 << 12
DT = 1.0 / 2048


def _req(rid, nb, seed, arrival):
    p = configs.small_batch(n=nb, seed=seed)
    p["r0"] = np.full(nb, 12.0)
    return Request(rid, arrival, p)


def _data(seed=3):
    rng = np.random.default_rng(seed)
    F = N // 2 + 1
    return (rng.normal(size=F) + 1j * rng.normal(size=F)) * 1e-3


def _server(capacity=64, streams=4):
    return Server(_data(), df=1.0 / (N * DT), dt=DT, N=N, capacity=capacity, max_step=2048, streams=streams)


def test_alone_and_in_the_mix(dev):
    reqs = [_req(0, 5, 101, 0.0), _req(1, 3, 102, 0.0005), _req(2, 7, 103, 0.001), _req(3, 2, 104, 0.001), _req(4, 9, 105, 0.002)]
    srv = _server()
    srv.warm()
    alone = {}
    for r in reqs:
        res = srv.run([Request(r.rid, 0.0, r.params)])
        assert len(res) == 1
        alone[r.rid] = res[0].logl
        assert np.all(np.isfinite(res[0].logl)), res[0]
    mix = srv.run(reqs)
    assert sorted(x.rid for x in mix) == [r.rid for r in reqs]
    for x in mix:
        np.testing.assert_array_equal(x.logl, alone[x.rid])
    # all at once, with a group limit that holds them: admitted as one launch group
    burst = srv.run([Request(r.rid, 0.0, r.params) for r in reqs], max_group=64)
    assert sorted(x.rid for x in burst) == [r.rid for r in reqs]
    assert len({x.admitted for x in burst}) == 1
    # the default limit (capacity / streams = 16 binaries) splits the same burst
    split = srv.run([Request(r.rid, 0.0, r.params) for r in reqs])
    assert len({x.admitted for x in split}) > 1
    for x in split:
        np.testing.assert_array_equal(x.logl, alone[x.rid])
    for x in burst:
        np.testing.assert_array_equal(x.logl, alone[x.rid])
    plain = srv.run(reqs, policy="plain")
    for x in plain:
        np.testing.assert_array_equal(x.logl, alone[x.rid])
    s = summarize(mix)
    assert s["requests"] == len(reqs) and s["p99_ms"] >= s["p50_ms"] > 0


def test_server_matches_batch_pipeline(dev):
    r = _req(0, 6, 201, 0.0)
    srv = _server()
    got = srv.run([r])[0].logl
    p = r.params
    b = P.EOBBatch(P.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"]), max_step=2048)
    b.initial_conditions(p["r0"])
    b.integrate(rtol=1e-10, atol=1e-10, two_pass=True)
    b.hlm22_knots()
    b.build_splines()
    hp = b.resample_hplus(p["m1"] + p["m2"], DT, N)
    ht = P.rfft_batched(hp)
    data = torch.as_tensor(_data(), dtype=torch.complex128, device=dev)
    ref = P.fd_likelihood(ht, data, 1.0 / (N * DT)).cpu().numpy()
    np.testing.assert_array_equal(got, ref)


def test_admission_waits_for_slots(dev):
    # capacity 8: the 6-binary requests cannot all be resident; they queue and all finish
    reqs = [_req(i, 6, 300 + i, 0.0) for i in range(4)]
    srv = _server(capacity=8)
    res = srv.run(reqs)
    assert sorted(x.rid for x in res) == [0, 1, 2, 3]
    assert srv.pool.free == [(0, 8)]
    with pytest.raises(ValueError):
        srv.run([_req(9, 9, 400, 0.0)])


def test_footprint_classes_and_reruns(dev):
    """Size classes: each binary runs in the smallest class its projected length
    fits; with a projection far too short the binaries fill their rows, run
    again in larger classes, and every result still equals the batch
    pipeline's (bit for bit)."""
    reqs = []
    for i, (nb, seed) in enumerate(((5, 501), (7, 502), (3, 503))):
        p = configs.small_batch(n=nb, seed=seed)
        p["r0"] = np.linspace(10.0, 16.0, nb)      # 231-332 stored points each
        reqs.append(Request(i, 0.0005 * i, p))
    data = _data()
    refs = {}
    for r in reqs:
        p = r.params
        b = P.EOBBatch(P.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"]), max_step=4096)
        b.initial_conditions(p["r0"])
        b.integrate(rtol=1e-10, atol=1e-10, two_pass=True)
        b.hlm22_knots()
        b.build_splines()
        ht = P.rfft_batched(b.resample_hplus(p["m1"] + p["m2"], DT, N))
        refs[r.rid] = P.fd_likelihood(ht, torch.as_tensor(data, dtype=torch.complex128, device=dev), 1.0 / (N * DT)).cpu().numpy()
    classes = [(256, 16), (1024, 16), (4096, 16)]
    for rpl, expect_reruns in ((11.0, False), (0.3, True)):
        srv = Server(data, df=1.0 / (N * DT), dt=DT, N=N, classes=classes, streams=4, rows_per_length=rpl)
        parts = [srv.split(r) for r in reqs]
        res = srv.run(reqs)
        assert sorted(x.rid for x in res) == [0, 1, 2]
        for x in res:
            np.testing.assert_array_equal(x.logl, refs[x.rid])
        assert (srv.reruns > 0) == expect_reruns, srv.reruns
        for pool in srv.pools:
            assert pool.free == [(0, pool.capacity)]
        if not expect_reruns:
            assert len({pt.cls for ps in parts for pt in ps}) >= 1
