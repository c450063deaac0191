"""GPU parity of the DOPR853 drop-in (eob/new_Dormand_horiz.py:219-778 call
convention) against the oracle's restatement of the same driver.

Both runs use fixed steps (``fix_step=True``), which makes the accepted-step
sequence deterministic (the adaptive controller's accept/reject sequence is
sensitive to rounding; its parity is covered on matched times elsewhere)."""
import numpy as np
import pytest
import torch

from eobgpu_amd import configs
from eobgpu_amd.integrator import DOPR853, StoppingCriterion
from eobgpu_amd.pyEOB import SEOBNRv5Class

pytestmark = pytest.mark.gpu

MAX_STEP = 8192


def _inputs(oracle, n=12, seed=1):
    p = configs.small_batch(n=n, seed=seed)
    add = oracle.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"])
    co = oracle.fac_coeffs(add)
    pphi0, pr0, om0, add = oracle.initial_conditions(p["r0"], add, co)
    cond = np.array([p["r0"], np.zeros(n), pr0, pphi0])
    return add, co, cond


def _integrator(coupled):
    ode = SEOBNRv5Class()
    return DOPR853(ode, stopping_criterion=StoppingCriterion(ode, MAX_STEP, use_gpu=True), tmax=1e13, max_step=MAX_STEP,
                   batch_coupled=coupled)


# batch_coupled=True (the default) is the reference's batch semantics of the
# stopping criterion; False integrates each binary as a batch of one.  The
# oracle restates both.
@pytest.mark.parametrize("coupled", [True, False])
def test_dopr853_dropin_single_pass(oracle, dev, coupled):
    add, co, cond = _inputs(oracle)
    loc_r, dense_r, sn_r, si_r, _ = oracle.integrate(cond, add, co, hInit=5.0, max_step=MAX_STEP, fix_step=True,
                                                     batch_coupled=coupled)
    integ = _integrator(coupled)
    loc, dense, sn = integ.integrate(cond.copy(), add, hInit=5.0, fix_step=True)
    assert loc.shape == (MAX_STEP, cond.shape[1]) and dense.shape == (MAX_STEP, 4, cond.shape[1])
    sn = sn.cpu().numpy()
    assert np.array_equal(sn, sn_r), (sn, sn_r)
    assert np.array_equal(integ.stop_info.cpu().numpy(), si_r)
    L = int(sn.max()) + 1
    np.testing.assert_allclose(loc[:L].cpu().numpy(), loc_r[:L], rtol=1e-11, atol=1e-9)
    np.testing.assert_allclose(dense[:L].cpu().numpy(), dense_r[:L], rtol=1e-9, atol=1e-13)
    # rows past each binary's last step are untouched (zero)
    d = dense.cpu().numpy()
    for b in range(d.shape[2]):
        assert not d[sn[b] + 1:, :, b].any()


@pytest.mark.parametrize("coupled", [True, False])
def test_dopr853_dropin_restart_matches_two_pass(oracle, dev, coupled):
    """The reference's run_trajectory glue (eob/seobnrv4phm.py:2091-2127) around two
    drop-in calls: a coarse pass, walk back 10 M from t[step_num - 1], then a
    fixed 0.05 M pass restarted from the row before the walk-back point."""
    add, co, cond = _inputs(oracle)
    t_r, traj_r, ns_r, si_r = oracle.run_trajectory(cond, add, co, max_step=MAX_STEP, hInit=5.0, fix_step1=True,
                                                    batch_coupled=coupled)
    integ = _integrator(coupled)
    t, traj, step_num = integ.integrate(cond.copy(), add, hInit=5.0, fix_step=True)
    n = cond.shape[1]
    cols = torch.arange(n, device=t.device)
    t_stop = t[(step_num - 1).clamp(min=0), cols]
    idx = torch.argmin(torch.abs(t - (t_stop - 10.0)[None, :]), dim=0)
    rows = torch.arange(MAX_STEP, device=t.device)[:, None]
    t[rows >= idx[None, :]] = 0.0
    traj[(rows >= idx[None, :])[:, None, :].expand(-1, 4, -1)] = 0.0
    t, traj, step_num = integ.integrate(cond.copy(), add, step_num=idx - 1, denseOutput=traj, denseOutputLoc=t, hInit=0.05,
                                        fix_step=True)
    ns = (step_num + 1).cpu().numpy()
    assert np.array_equal(ns, ns_r), (ns, ns_r)
    assert np.array_equal(integ.stop_info.cpu().numpy(), si_r)
    L = int(ns.max())
    np.testing.assert_allclose(t[:L].cpu().numpy().T, t_r[:, :L], rtol=1e-11, atol=1e-9)
    np.testing.assert_allclose(traj[:L].cpu().numpy().transpose(2, 1, 0), traj_r[:, :, :L], rtol=1e-9, atol=1e-13)


def test_dopr853_rejects_generic_ode():
    with pytest.raises(TypeError):
        DOPR853(lambda *a: None)
    with pytest.raises(ValueError):
        DOPR853(None, use_gpu=False)
