"""GPU parity of the batch-coupled driver (the reference's batch semantics of
StoppingCriterion, eob/seobnrv4phm.py:82-150, inside DOPR853.integrate,
eob/new_Dormand_horiz.py:585-775) against the oracle's batch_coupled mode, and
of the reference's driver quirks (first-step reject, non-finite guard) on both
drivers."""
import numpy as np
import pytest
import torch

from eobgpu_amd import configs
from eobgpu_amd import pipeline as P

pytestmark = pytest.mark.gpu

_CACHE = {}


def _config1(oracle, n=16, seed=1):
    key = ("c1", n, seed)
    if key not in _CACHE:
        p = configs.small_batch(n=n, seed=seed)
        add = oracle.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"])
        co = oracle.fac_coeffs(add)
        pphi0, pr0, om0, add = oracle.initial_conditions(p["r0"], add, co)
        cond = np.array([p["r0"], np.zeros(n), pr0, pphi0])
        _CACHE[key] = (add, co, cond)
    return _CACHE[key]


def _run(add, cond, dev, max_step, **kw):
    b = P.EOBBatch(add, max_step=max_step, device=dev)
    b.set_initial_state(cond)
    it = b.integrate(**kw)
    torch.cuda.synchronize()
    t, traj, ns = b.trajectory_host()
    return t, traj, ns, b.stop_info.cpu().numpy(), b.pass1_steps.cpu().numpy(), it, b


# run_trajectory with both passes at fixed step (pass 1 at h = 5 M, then the
# walk-back and the 0.05 M pass): a deterministic step sequence, so counts and
# codes must be identical to the oracle's batch-coupled restatement.
@pytest.mark.parametrize("group", [64, 16, 1])
def test_coupled_two_pass_fixed_step_parity(oracle, dev, group):
    add, co, cond = _config1(oracle)
    max_step = 8192
    key = ("ref_fixed_coupled", max_step)
    if key not in _CACHE:
        _CACHE[key] = oracle.run_trajectory(cond, add, co, max_step=max_step, rtol=1e-10, atol=1e-10, hInit=5.0, fix_step1=True,
                                            batch_coupled=True)
    t_r, traj_r, ns_r, si_r = _CACHE[key]
    print(f"coupled fixed-step two-pass, group {group}", flush=True)
    t, traj, ns, si, _, it, _ = _run(add, cond, dev, max_step, rtol=1e-10, atol=1e-10, hInit=5.0, fix_step=True, mode="coupled",
                                     group=group)
    assert np.array_equal(ns, ns_r), (ns, ns_r)
    assert np.array_equal(si, si_r), (si, si_r)
    L = ns.max()
    np.testing.assert_allclose(t, t_r[:, :L], rtol=1e-11, atol=1e-9)
    for c in range(4):
        np.testing.assert_allclose(traj[:, c], traj_r[:, c, :L], rtol=1e-9, atol=1e-13)
    # the batch coupling matters on this workload: the per-binary driver differs
    per = _run(add, cond, dev, max_step, rtol=1e-10, atol=1e-10, hInit=5.0, fix_step=True, mode="persistent", group=group)
    assert (per[2] != ns).sum() >= 4 and (per[3] != si).sum() >= 4


def test_coupled_single_pass_and_restart(oracle, dev):
    """One DOPR853.integrate pass at fixed step (reference batch semantics), then a
    second call that continues from caller-chosen rows (the restart form)."""
    add, co, cond = _config1(oracle)
    max_step = 2048
    dL, dO, sn, si, it_r = oracle.integrate(cond, add, co, hInit=2.0, max_step=max_step, fix_step=True, rtol=1e-10, atol=1e-10,
                                            batch_coupled=True)
    t, traj, ns, sig, _, it, b = _run(add, cond, dev, max_step, rtol=1e-10, atol=1e-10, hInit=2.0, fix_step=True, two_pass=False,
                                      mode="coupled", group=64)
    assert np.array_equal(ns - 1, sn), (ns - 1, sn)
    assert np.array_equal(sig, si)
    assert it == it_r
    # restart: from row step_num - 40 of each binary, 0.05 M fixed steps
    sn0 = np.maximum(sn - 40, 0).astype(np.int32)
    dO2, dL2 = dO.copy(), dL.copy()
    dL3, dO3, sn3, si3, _ = oracle.integrate(cond, add, co, hInit=0.05, fix_step=True, rtol=1e-10, atol=1e-10, step_num=sn0,
                                             denseOutput=dO2, denseOutputLoc=dL2, batch_coupled=True)
    ar = np.arange(cond.shape[1])
    x0 = torch.tensor(dL[sn0, ar], device=dev)
    b.cond = torch.tensor(np.ascontiguousarray(dO[sn0, :, ar].T), device=dev)
    b.integrate(rtol=1e-10, atol=1e-10, hInit=0.05, fix_step=True, two_pass=False, mode="coupled", group=64, x0=x0,
                step0=torch.tensor(sn0, device=dev))
    torch.cuda.synchronize()
    ns2 = b.num_steps.cpu().numpy() - 1
    assert np.array_equal(ns2, sn3), (ns2, sn3)
    assert np.array_equal(b.stop_info.cpu().numpy(), si3)
    d = b.dense.cpu().numpy()
    for i in range(cond.shape[1]):
        np.testing.assert_allclose(d[i, :sn3[i] + 1, 1:], dO3[:sn3[i] + 1, :, i], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("mode", ["persistent", "plain", "coupled"])
def test_first_reject_and_nonfinite_guard(oracle, dev, mode):
    """A first attempt that the controller rejects ends the pass with no step
    (the stop flag starts True, eob/new_Dormand_horiz.py:583, 761), and a state
    that is not finite ends the binary with code -1 instead of the reference's
    endless rejections."""
    add, co, cond = _config1(oracle, n=4)
    cond = cond.copy()
    cond[0, 2] = np.nan
    # binaries 0 and 1: a first step of 800 M is rejected (the error stays finite;
    # from about 3000 M on the stages leave the Hamiltonian's domain instead)
    dL, dO, sn, si, _ = oracle.integrate(cond[:, :2], add[:, :2], co[:2], hInit=800.0, max_step=256, rtol=1e-10, atol=1e-10,
                                         batch_coupled=True)
    # (under the batch coupling binary 3 reads the NaN binary's empty omega column
    # in the third check, so its end differs from its end alone)
    dL2, dO2, sn2, si2, _ = oracle.integrate(cond[:, 2:], add[:, 2:], co[2:], hInit=5.0, fix_step=True, max_step=4096, rtol=1e-10,
                                             atol=1e-10, batch_coupled=(mode == "coupled"))
    assert list(sn) == [0, 0] and list(si) == [0, 0]   # oracle: the quirk
    assert sn2[0] == 0 and si2[0] == -1                # oracle: the guard
    print(f"first-reject / guard, mode {mode}", flush=True)
    a = _run(add[:, :2], cond[:, :2], dev, 256, rtol=1e-10, atol=1e-10, hInit=800.0, two_pass=False, mode=mode, group=64)
    c = _run(add[:, 2:], cond[:, 2:], dev, 4096, rtol=1e-10, atol=1e-10, hInit=5.0, fix_step=True, two_pass=False, mode=mode,
             group=64)
    assert list(a[2]) == [1, 1] and list(a[3]) == [0, 0]
    assert c[2][0] == 1 and c[3][0] == -1
    # the finite binary of the second call is unaffected by its NaN neighbour
    assert c[2][1] - 1 == sn2[1] and c[3][1] == si2[1]
