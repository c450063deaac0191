"""GPU parity of the reference-layout drop-ins that a caller keeping the
reference's own Python driver would run: the per-stage DOP853 steps
(``dormandPrinceSteps``), ``error``, ``controllerSuccess``, the radiation-
reaction force, the initial-condition residuals and the (2,2) mode, each
against the oracle; and the reference's DOPR853.integrate loop
(eob/new_Dormand_horiz.py:585-775) restated here on top of those drop-ins and
the StoppingCriterion, against the oracle's batch-coupled driver."""
import numpy as np
import pytest
import torch

from eobgpu_amd import configs
from eobgpu_amd import pipeline as P
from eobgpu_amd import pyEOB, pydopr853
from eobgpu_amd.integrator import StoppingCriterion

pytestmark = pytest.mark.gpu


def _T(x, dev):
    return torch.tensor(np.ascontiguousarray(x), device=dev)


def _ode(add, dev):
    """A SEOBNRv5Class with the tables SEOBNRv4PHM.setup_class gives it (eob/seobnrv4phm.py:1922-1964)."""
    m1, m2, c1, c2 = add[0], add[1], add[2], add[3]
    eta = m1 * m2 / (m1 + m2) ** 2
    chiS, chiA = 0.5 * (c1 + c2), 0.5 * (c1 - c2)
    a = (1 - 2 * eta) * chiS + (m1 - m2) / (m1 + m2) * chiA
    n = add.shape[1]
    ode = pyEOB.SEOBNRv5Class()
    pref = torch.tensor(np.ascontiguousarray(add[10:80].T.reshape(-1)), device=dev)
    ode.update_information(_T(m1, dev), _T(m2, dev), _T(eta, dev), _T(a, dev), _T(chiS, dev), _T(chiA, dev), 1, n, pref)
    return ode


def _start(oracle, n, r0, seed=1):
    p = configs.small_batch(n=n, seed=seed)
    add = oracle.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"])
    co = oracle.fac_coeffs(add)
    pphi0, pr0, om0, add = oracle.initial_conditions(np.full(n, r0), add, co)
    cond = np.array([np.full(n, r0), np.zeros(n), pr0, pphi0])
    return add, co, cond


def test_dormand_prince_steps_error_controller(oracle, dev):
    """Stages 1..10 land in k1..k10, stage 11 in k2 and stage 12 in k3
    (src/DOPR853.cu:498, 516); error and controller follow on those buffers."""
    n = 24
    add, co, cond = _start(oracle, n, 14.0)
    rng = np.random.default_rng(7)
    # states along the inspiral and into the plunge, steps from 0.05 to 5 M
    # (every stage stays inside the Hamiltonian's domain)
    cond[0] = rng.uniform(4.0, 14.0, n)
    cond[2] = rng.uniform(-0.1, 0.0, n)
    cond[3] = np.sqrt(cond[0]) * rng.uniform(0.97, 1.1, n)
    h = 10.0 ** rng.uniform(np.log10(0.05), np.log10(5.0), n)
    k_r, yn_r, err_r = oracle.dopr_attempt(cond, h, add, co)
    ode = _ode(add, dev)
    ks = [torch.zeros(4 * n, dtype=torch.float64, device=dev) for _ in range(10)]
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    xc, ak, arg = torch.zeros_like(x), torch.zeros(4 * n, dtype=torch.float64, device=dev), torch.zeros(4 * n, dtype=torch.float64,
                                                                                                         device=dev)
    sol = _T(cond.reshape(-1), dev)
    ht = _T(h, dev)
    idx = torch.arange(n, dtype=torch.int32, device=dev)
    pydopr853.dormandPrinceSteps(x, xc, sol, ht, arg, _T(add.reshape(-1), dev), *ks, ak, 4, n, 80, ode, idx)
    got = [k.cpu().numpy().reshape(4, n) for k in ks]
    for s, buf in ((0, 0), (3, 3), (4, 4), (5, 5), (6, 6), (7, 7), (8, 8), (9, 9), (10, 1), (11, 2)):
        np.testing.assert_allclose(got[buf], k_r[s], rtol=5e-9, atol=1e-16, err_msg=f"stage {s + 1}")
    err = torch.zeros(n, dtype=torch.float64, device=dev)
    sol_new = torch.zeros(4 * n, dtype=torch.float64, device=dev)
    pydopr853.error(err, sol, sol_new, ht, ks[0], ks[1], ks[2], ks[5], ks[6], ks[7], ks[8], ks[9], n, 4)
    np.testing.assert_allclose(sol_new.cpu().numpy().reshape(4, n), yn_r, rtol=1e-9, atol=1e-15)
    # the error estimate is a difference of stages that cancels by up to ~1e11 at
    # these tolerances, so it is checked on the device's own stages: the norm of
    # src/DOPR853.cu:586-616 evaluated by the oracle on the same buffers
    kd = {k: got[i] for k, i in (("k1", 0), ("k2", 1), ("k3", 2), ("k6", 5), ("k7", 6), ("k8", 7), ("k9", 8), ("k10", 9))}
    err_o, sol_o = oracle.error(cond, h, kd)
    np.testing.assert_allclose(err.cpu().numpy(), err_o, rtol=1e-12)
    np.testing.assert_allclose(sol_new.cpu().numpy().reshape(4, n), sol_o, rtol=1e-15, atol=1e-15)
    errOld = np.full(n, 1e-4)
    prev = np.zeros(n, dtype=bool)
    f_r, eo_r, pr_r, h_r = oracle.controller(err.cpu().numpy(), errOld, prev, h)
    flag = torch.zeros(n, dtype=torch.uint8, device=dev)
    eo, prt = _T(errOld, dev), torch.zeros(n, dtype=torch.uint8, device=dev)
    pydopr853.controllerSuccess(flag, err, eo, prt, ht, x, n, 4)
    assert np.array_equal(flag.cpu().numpy().astype(bool), f_r)
    np.testing.assert_allclose(ht.cpu().numpy(), h_r, rtol=1e-14)
    ode.deallocate_information()


def test_rr_force_and_ic_cons(oracle, dev):
    rng = np.random.default_rng(8)
    n = 65
    add, co, _ = _start(oracle, n, 20.0)
    r = rng.uniform(2.5, 40.0, n)
    L = np.sqrt(r) * rng.uniform(0.95, 1.2, n)
    arg = np.array([r, rng.uniform(0, 6, n), rng.uniform(-0.5, 0.0, n), L])
    g_full = oracle.grad_H(arg, add)
    add_w = add.copy()
    add_w[4] = g_full[3]   # the ODE passes dH/dpphi as omega (src/EOB.cu:3112)
    F_r = oracle.rr_force(arg, add_w, co)
    ode = _ode(add, dev)
    force = torch.zeros((2, n), dtype=torch.float64, device=dev)
    grad = torch.zeros((4, n), dtype=torch.float64, device=dev)
    idx = torch.arange(n, dtype=torch.int32, device=dev)
    ode.RR_force(force, grad, _T(arg, dev), _T(add_w, dev), n, idx)
    np.testing.assert_allclose(force.cpu().numpy(), F_r, rtol=5e-10, atol=1e-20)
    arg0 = arg.copy()
    arg0[2] = 0.0
    g0 = oracle.grad_H(arg0, add)
    gg = grad.cpu().numpy()
    np.testing.assert_allclose(gg[[0, 3]], g0[[0, 3]], rtol=2e-10, atol=1e-16)
    assert not gg[1].any()
    ode.deallocate_information()

    # IC_cons: res = (dH/dpphi - omega0, dH/dr) at (r, 0, 0, pphi); args rewritten
    xin = np.array([r, L])
    args = torch.zeros((4, n), dtype=torch.float64, device=dev)
    res = torch.zeros((2, n), dtype=torch.float64, device=dev)
    grad = torch.zeros((4, n), dtype=torch.float64, device=dev)
    pyEOB.IC_cons(res, _T(xin, dev), args, _T(add, dev), grad, n)
    a_exp = np.array([r, np.zeros(n), np.zeros(n), L])
    np.testing.assert_array_equal(args.cpu().numpy(), a_exp)
    gi = oracle.grad_H(a_exp, add)
    rr = res.cpu().numpy()
    np.testing.assert_allclose(rr[0] + add[4], gi[3], rtol=2e-10)
    np.testing.assert_allclose(rr[1], gi[0], rtol=2e-10, atol=1e-16)
    np.testing.assert_allclose(grad.cpu().numpy()[[0, 3]], gi[[0, 3]], rtol=2e-10, atol=1e-16)


def test_spin_factorized_waveform_22(oracle, dev):
    """EOBGetSpinFactorizedWaveform(l=2, m=2) on knots (src/EOB.cu:1616-1677)."""
    rng = np.random.default_rng(9)
    nb, Lk = 6, 300
    add, co, _ = _start(oracle, nb, 20.0)
    m1, m2 = add[0], add[1]
    eta = m1 * m2 / (m1 + m2) ** 2
    r = np.sort(rng.uniform(2.5, 25.0, (nb, Lk)), axis=1)[:, ::-1].copy()
    L = np.sqrt(r) * 1.02
    pr = -0.01 * rng.uniform(0, 1, (nb, Lk))
    phi = rng.uniform(0, 60, (nb, Lk))
    v = np.empty_like(r); Hr = np.empty_like(r); vPhi = np.empty_like(r)
    for b in range(nb):
        a1 = add[:, b:b + 1].repeat(Lk, axis=1)
        st = np.array([r[b], phi[b], pr[b], L[b]])
        g = oracle.grad_H(st, a1)
        st0 = st.copy(); st0[2] = 0.0
        gc = oracle.grad_H(st0, a1)
        H = oracle.evaluate_H(r[b], phi[b], pr[b], L[b], a1[0], a1[1], a1[2], a1[3])
        v[b] = np.cbrt(g[3])
        Hr[b] = eta[b] * H
        vPhi[b] = 1.0 / (gc[3] ** 2 * r[b] ** 3)
    pf = add[12] + 1j * add[13]
    h22c = 0.0
    h_r = oracle.hlm22(r, phi, v, Hr, eta, vPhi, pf, co, h22c)
    ode = _ode(add, dev)
    out = torch.zeros(2 * nb * Lk, dtype=torch.float64, device=dev)
    pfi = _T(np.stack([pf.real, pf.imag], axis=1).reshape(-1), dev)
    idx = torch.arange(nb, dtype=torch.int32, device=dev)
    ode.EOBGetSpinFactorizedWaveform(out, _T(r, dev), _T(phi, dev), _T(pr, dev), _T(L, dev), _T(v, dev), _T(Hr, dev), _T(eta, dev),
                                     _T(vPhi, dev), 2, 2, pfi, h22c, nb, Lk, idx)
    o = out.cpu().numpy().reshape(nb, Lk, 2)
    h = o[..., 0] + 1j * o[..., 1]
    scale = np.abs(h_r).max(axis=1, keepdims=True)
    assert np.max(np.abs(h - h_r) / scale) < 1e-11
    ode.deallocate_information()


def _python_dopr853(ode, sc, cond, add, hInit, max_step, dev, tmax=1e13):
    """DOPR853.integrate with fix_step=True, restated from
    eob/new_Dormand_horiz.py:504-775 on the drop-ins (one attempted step per
    iteration for the active binaries, compacted; the criterion on the accepted
    ones in ascending order)."""
    n = cond.shape[1]
    f64 = dict(dtype=torch.float64, device=dev)
    sol = _T(cond, dev).clone()
    add_t = _T(add, dev)
    dense = torch.zeros((max_step, 4, n), **f64)
    loc = torch.zeros((max_step, n), **f64)
    dense[0] = sol
    step_num = torch.zeros(n, dtype=torch.int64, device=dev)
    x = torch.zeros(n, **f64)
    sc.setup(n)
    h_orig = torch.full((n,), float(hInit), **f64)
    h = h_orig.clone()
    err_old = torch.full((n,), 1e-4, **f64)
    prev = torch.zeros(n, dtype=torch.bool, device=dev)
    live = torch.ones(n, dtype=torch.bool, device=dev)
    stop = torch.ones(n, dtype=torch.bool, device=dev)
    stop_info = torch.zeros(n, dtype=torch.int64, device=dev)
    iters = 0
    while True:
        here = torch.nonzero(live).reshape(-1)
        m = int(here.numel())
        xT, hT = x[here].contiguous(), h[here].contiguous()
        solT = sol[:, here].contiguous()
        addT = add_t[:, here].contiguous()
        ks = [torch.zeros(4 * m, **f64) for _ in range(10)]
        pydopr853.dormandPrinceSteps(xT, torch.zeros(m, **f64), solT.reshape(-1), hT, torch.zeros(4 * m, **f64), addT.reshape(-1), *ks,
                                     torch.zeros(4 * m, **f64), 4, m, 80, ode, here.to(torch.int32))
        err = torch.zeros(m, **f64)
        sol_new = torch.zeros(4 * m, **f64)
        pydopr853.error(err, solT.reshape(-1), sol_new, hT, ks[0], ks[1], ks[2], ks[5], ks[6], ks[7], ks[8], ks[9], m, 4)
        h_old = hT.clone()
        flag = torch.zeros(m, dtype=torch.uint8, device=dev)
        eo = err_old[here].contiguous()
        pv = prev[here].to(torch.uint8).contiguous()
        pydopr853.controllerSuccess(flag, err, eo, pv, hT, xT, m, 4)
        ok = flag.bool()
        if bool((~ok).any()):
            h_orig[here[~ok]] /= 2
        hT = h_orig[here].clone()
        solT[:, ok] = sol_new.reshape(4, m)[:, ok]
        xT[ok] = xT[ok] + h_old[ok]
        x[here], h[here], sol[:, here], err_old[here], prev[here] = xT, hT, solT, eo, pv.bool()
        iu = here[ok]
        step_num[iu] += 1
        dense[step_num[iu], :, iu] = sol[:, iu].T
        loc[step_num[iu], iu] = x[iu]
        if iu.numel():
            st = sc(step_num, sol[:, iu], add_t[:, iu], iu)
            stop_info[iu] = st
            stop[iu] = st.bool()
        live[(x >= tmax) | (step_num >= max_step - 1) | stop] = False
        iters += 1
        if not bool(live.any()):
            break
    sc.reset()
    return loc, dense, step_num, stop_info, iters


def test_reference_driver_through_dropins(oracle, dev):
    """r0 = 12, fixed 5 M steps, the drop-in error's hard-coded tolerances
    (1e-12, src/DOPR853.cu:7-8): four of the twelve binaries end differently
    under the reference's batch coupling than alone (oracle)."""
    n, max_step = 12, 1024
    add, co, cond = _start(oracle, n, 12.0)
    dL, dO, sn_r, si_r, it_r = oracle.integrate(cond, add, co, hInit=5.0, fix_step=True, max_step=max_step, batch_coupled=True)
    per = oracle.integrate(cond, add, co, hInit=5.0, fix_step=True, max_step=max_step, batch_coupled=False)
    assert (per[2] != sn_r).sum() >= 2
    ode = _ode(add, dev)
    sc = StoppingCriterion(ode, max_step, use_gpu=True)
    loc, dense, sn, si, it = _python_dopr853(ode, sc, cond, add, 5.0, max_step, dev)
    assert np.array_equal(sn.cpu().numpy(), sn_r), (sn.cpu().numpy(), sn_r)
    assert np.array_equal(si.cpu().numpy(), si_r)
    assert it == it_r
    L = int(sn_r.max()) + 1
    np.testing.assert_allclose(dense[:L].cpu().numpy(), dO[:L], rtol=1e-9, atol=1e-13)
    # the coupled kernel driver on the same batch: same counts, codes, iterations
    b = P.EOBBatch(add, max_step=max_step, device=dev)
    b.set_initial_state(cond)
    itk = b.integrate(rtol=1e-12, atol=1e-12, hInit=5.0, fix_step=True, two_pass=False, mode="coupled", group=64)
    assert np.array_equal(b.num_steps.cpu().numpy() - 1, sn_r) and np.array_equal(b.stop_info.cpu().numpy(), si_r)
    assert itk == it_r
    ode.deallocate_information()


def test_initial_condition_solvers(oracle, dev):
    """root_find_all (discrete Newton on IC_cons from the reference's start
    r = omega0^(-2/3), pphi = sqrt(r) at f = 20 Hz, eob/seobnrv4phm.py:1966-2009),
    root_find_scalar_all (false position on IC_diss in [-3e-2, 0]) and the IC_diss
    kernel itself, against the oracle's restatements."""
    n = 32
    p = configs.uneven_batch(n=n, seed=17)
    mt = p["m1"] + p["m2"]
    omega0 = 20.0 * (mt * P.MTSUN_SI * np.pi)
    add = oracle.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"], omega0=omega0)
    co = oracle.fac_coeffs(add)
    rg = omega0 ** (-2.0 / 3)
    x0 = np.array([rg, np.sqrt(rg)])
    x_r = oracle.root_find_all(x0, add, 1000, 1e-12)
    xo = torch.zeros(2 * n, dtype=torch.float64, device=dev)
    pyEOB.root_find_all(xo, _T(x0.reshape(-1), dev), torch.zeros(4 * n, dtype=torch.float64, device=dev), _T(add.reshape(-1), dev),
                        1000, 1e-12, n, 2, 4, 80)
    x = xo.cpu().numpy().reshape(2, n)
    np.testing.assert_allclose(x, x_r, rtol=1e-10)
    # the root: dH/dr = 0 and dH/dpphi = omega0 at pr = 0
    g = oracle.grad_H(np.array([x[0], np.zeros(n), np.zeros(n), x[1]]), add)
    assert np.all(np.abs(g[0]) < 1e-12 * np.abs(g[3])) and np.allclose(g[3], omega0, rtol=1e-10)

    ode = _ode(add, dev)
    args = np.array([x_r[0], np.zeros(n), np.zeros(n), x_r[1]])
    bounds = np.tile(np.array([-3e-2, 0.0]), (n, 1)).T.copy()
    pr_r = oracle.root_find_scalar_all(bounds, args, add, co, 1000, 1e-12)
    pr = torch.zeros(n, dtype=torch.float64, device=dev)
    ode.root_find_scalar_all(pr, _T(bounds.reshape(-1), dev), _T(args.reshape(-1), dev), _T(add.reshape(-1), dev), 1000, 1e-12, n, 4, 80)
    # the false position stops at |IC_diss| < 1e-12 (absolute), which fixes pr to about 1e-8 relative
    np.testing.assert_allclose(pr.cpu().numpy(), pr_r, rtol=1e-8, atol=1e-13)

    rng = np.random.default_rng(18)
    prs = rng.uniform(-3e-2, -1e-3, n)
    out_r, args_r, add_r, g_r, gt_r, hs_r, F_r = oracle.ic_diss(prs, args, add, co)
    ta, tadd = _T(args, dev), _T(add, dev)
    out = torch.zeros(n, dtype=torch.float64, device=dev)
    gd = torch.zeros((4, n), dtype=torch.float64, device=dev)
    gt = torch.zeros((4, n), dtype=torch.float64, device=dev)
    hs = torch.zeros((16, n), dtype=torch.float64, device=dev)
    F = torch.zeros((2, n), dtype=torch.float64, device=dev)
    ode.IC_diss(out, _T(prs, dev), ta, tadd, gd, gt, hs, F, n)
    np.testing.assert_allclose(out.cpu().numpy(), out_r, rtol=1e-8)
    np.testing.assert_array_equal(ta.cpu().numpy(), args_r)                         # pr written into args row 2
    np.testing.assert_allclose(tadd.cpu().numpy()[4], add_r[4], rtol=2e-10)         # omega row := dH/dpphi
    np.testing.assert_allclose(gd.cpu().numpy(), g_r, rtol=2e-10, atol=1e-16)
    np.testing.assert_allclose(hs.cpu().numpy(), hs_r, rtol=3e-9, atol=1e-14)
    np.testing.assert_allclose(F.cpu().numpy(), F_r, rtol=5e-10, atol=1e-20)
    np.testing.assert_allclose(gt.cpu().numpy()[[0, 3]], gt_r[[0, 3]], rtol=2e-10, atol=1e-16)
    ode.deallocate_information()


def test_stopping_criterion_codes_2_to_4(oracle, dev):
    """Stop codes 2-4 of StoppingCriterion.__call__ (eob/seobnrv4phm.py:109-117)
    on chosen states, with the expected code read off the oracle's right-hand
    side: r < 1 gives 2 whatever the derivatives (no trajectory reaches it: the
    Hamiltonian leaves its domain above r = 1 for every spin, and the run ends
    with -1 first); r < 6 with dr/dt > 0 gives 3; r < 6 with dpr/dt > 0 and
    dr/dt <= 0 gives 4; r >= 6 gives 0 (step numbers <= 5: no omega check)."""
    n = 8
    add, co, cond = _start(oracle, n, 20.0)
    r = np.array([0.5, 0.8, 5.0, 4.0, 5.0, 4.0, 8.0, 12.0])
    L = np.sqrt(r) * np.array([1.0, 1.0, 1.0, 1.0, 3.0, 3.0, 1.0, 1.0])
    pr = np.array([-0.1, 0.0, 0.2, 0.1, -0.01, -0.01, 0.1, -0.1])
    st = np.array([r, np.zeros(n), pr, L])
    k = oracle.rhs(st, add, co)
    exp = np.zeros(n, dtype=np.int64)
    for i in range(n):
        if r[i] < 1.0:
            exp[i] = 2
        elif r[i] < 6.0 and k[0, i] > 0:
            exp[i] = 3
        elif r[i] < 6.0 and k[2, i] > 0:
            exp[i] = 4
    assert set(exp.tolist()) >= {0, 2, 3, 4}, (exp, k)
    ode = _ode(add, dev)
    sc = StoppingCriterion(ode, 64)
    sc.setup(n)
    stop = sc(torch.zeros(n, dtype=torch.int64, device=dev), _T(st, dev), _T(add, dev), torch.arange(n, device=dev))
    np.testing.assert_array_equal(stop.cpu().numpy(), exp)
    ode.deallocate_information()
