"""ORACLE — test infrastructure only.

numpy/ctypes front end for the CPU restatement in ``oracle_eob.cpp`` (built into
``oracle/build/liboracle.so`` by ``oracle/Makefile``) and for the reference's own
``src/DOPR853.cu`` compiled unmodified into ``oracle/_ref/libdopr853_ref.so``.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
import this module, and only as the reference results they compare against / the CPU baseline.  The shipped
``eobgpu_amd`` package never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libdopr853_ref.so")

_lib = None
_ref = None

P = ctypes.POINTER(ctypes.c_double)
PI32 = ctypes.POINTER(ctypes.c_int)
PU8 = ctypes.POINTER(ctypes.c_ubyte)

NADD = 80


def build(quiet: bool = True) -> None:
    """Compile the oracle (and oracle/_ref when the reference tree exists)."""
    out = subprocess.run(["make", "-C", HERE], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def ref_lib():
    """The reference's DOPR853.cu host build, or None if it was not built here."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        # lazy binding: the library's one unresolved symbol belongs to
        # dormandPrinceSteps_wrap, which is never called from here
        _ref = ctypes.CDLL(REF_PATH, mode=os.RTLD_LAZY)
    return _ref


def _p(a):
    return a.ctypes.data_as(P)


def _pi(a):
    return a.ctypes.data_as(PI32)


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def coeff_doubles() -> int:
    return lib().orc_coeff_doubles()


def tableau():
    out = np.zeros(12 + 144 + 12 + 12 + 3)
    lib().orc_tableau(_p(out))
    return out


# --------------------------------------------------------------- physics API
def evaluate_H(r, phi, pr, pphi, m1, m2, chi1, chi2, K=0.0, d5=0.0, dSO=0.0, dSS=0.0):
    r, phi, pr, pphi, m1, m2, chi1, chi2 = [_c(x) for x in np.broadcast_arrays(r, phi, pr, pphi, m1, m2, chi1, chi2)]
    n = r.size
    out = np.zeros(n)
    lib().orc_evaluate_H(_p(out), _p(r), _p(phi), _p(pr), _p(pphi), _p(m1), _p(m2), _p(chi1), _p(chi2),
                         ctypes.c_double(K), ctypes.c_double(d5), ctypes.c_double(dSO), ctypes.c_double(dSS), n)
    return out


def grad_H(arg, add):
    arg = _c(arg); add = _c(add)
    n = arg.shape[-1]
    g = np.zeros((4, n))
    lib().orc_grad_H(_p(arg), _p(g), _p(add), n)
    return g


def hessian_H(arg, add):
    arg = _c(arg); add = _c(add)
    n = arg.shape[-1]
    h = np.zeros((16, n))
    lib().orc_hessian_H(_p(arg), _p(h), _p(add), n)
    return h.reshape(4, 4, n)


def newton_prefixes(m1, m2):
    m1 = _c(np.atleast_1d(m1)); m2 = _c(np.atleast_1d(m2))
    n = m1.size
    out = np.zeros((35, n, 2))
    lib().orc_newton_prefixes(_p(out), _p(m1), _p(m2), n)
    return out[..., 0] + 1j * out[..., 1]


def make_additional_args(m1, m2, chi1, chi2, omega0=None, K=0.0, d5=0.0, dSO=0.0, dSS=0.0, h22_calib=0.0):
    """additionalArgs [80][n] as SEOBNRv4PHM.setup_class builds it (eob/seobnrv4phm.py:1922-1964)."""
    m1 = np.atleast_1d(np.asarray(m1, dtype=float)); m2 = np.atleast_1d(np.asarray(m2, dtype=float))
    n = m1.size
    M = m1 + m2
    m1s, m2s = m1 / M, m2 / M
    add = np.zeros((NADD, n))
    add[0] = m1s; add[1] = m2s; add[2] = chi1; add[3] = chi2
    add[4] = 0.0 if omega0 is None else omega0
    add[5] = K; add[6] = d5; add[7] = dSO; add[8] = dSS; add[9] = h22_calib
    pf = newton_prefixes(m1s, m2s)
    add[10::2] = pf.real
    add[11::2] = pf.imag
    return add


def fac_coeffs(add, use_hm=1):
    add = _c(add)
    n = add.shape[1]
    m1, m2, c1, c2 = add[0], add[1], add[2], add[3]
    eta = m1 * m2 / (m1 + m2) ** 2
    chiS = 0.5 * (c1 + c2)
    chiA = 0.5 * (c1 - c2)
    a = (1 - 2 * eta) * chiS + (m1 - m2) / (m1 + m2) * chiA
    out = np.zeros((n, coeff_doubles()))
    lib().orc_fac_coeffs(_p(out), use_hm, _p(_c(m1)), _p(_c(m2)), _p(_c(eta)), _p(_c(a)), _p(_c(chiS)), _p(_c(chiA)), n)
    return out


def rr_force(arg, add, coeffs):
    arg = _c(arg); add = _c(add); coeffs = _c(coeffs)
    n = arg.shape[-1]
    F = np.zeros((2, n))
    idx = np.arange(n, dtype=np.int32)
    lib().orc_rr_force(_p(F), _p(arg), _p(add), n, _p(coeffs), _pi(idx))
    return F


def rhs(arg, add, coeffs):
    arg = _c(arg); add = _c(add); coeffs = _c(coeffs)
    n = arg.shape[-1]
    k = np.zeros((4, n))
    idx = np.arange(n, dtype=np.int32)
    lib().orc_rhs(_p(arg), _p(k), _p(add), n, _p(coeffs), _pi(idx))
    return k


def initial_conditions(r0, add, coeffs, max_iter=1000, err=1e-12):
    """pphi0, pr0, omega0 for circular-orbit start at r0; add row 4 gets omega0."""
    r0 = _c(np.atleast_1d(r0)); coeffs = _c(coeffs)
    add = np.array(add, dtype=np.float64, order="C")
    n = r0.size
    pphi0 = np.zeros(n); pr0 = np.zeros(n); om0 = np.zeros(n)
    lib().orc_initial_conditions(_p(r0), _p(add), _p(coeffs), n, _p(pphi0), _p(pr0), _p(om0), max_iter, ctypes.c_double(err))
    return pphi0, pr0, om0, add


def root_find_all(x0, add, max_iter=1000, err=1e-12):
    """root_find_all: (r, pphi) [2][n] solving IC_cons = 0 from x0 [2][n] (add row 4 = omega0)."""
    x0 = _c(x0); add = _c(add)
    n = x0.shape[1]
    out = np.zeros((2, n))
    lib().orc_root_find_all(_p(out), _p(x0), _p(add), int(max_iter), ctypes.c_double(err), n)
    return out


def ic_diss(pr, args, add, coeffs):
    """IC_diss through its kernel: returns (out, args, add, grad, grad_temp_force, hess, force);
    args and add are copies updated as the reference updates its arrays."""
    pr = _c(np.atleast_1d(pr)); args = np.array(args, dtype=np.float64, order="C"); add = np.array(add, dtype=np.float64, order="C")
    coeffs = _c(coeffs)
    n = pr.size
    out = np.zeros(n); g = np.zeros((4, n)); gt = np.zeros((4, n)); hs = np.zeros((16, n)); F = np.zeros((2, n))
    lib().orc_ic_diss(_p(out), _p(pr), _p(args), _p(add), _p(g), _p(gt), _p(hs), _p(F), _p(coeffs), n)
    return out, args, add, g, gt, hs, F


def root_find_scalar_all(bounds, args, add, coeffs, max_iter=1000, err=1e-12):
    """root_find_scalar_all: pr [n] with IC_diss(pr) = 0 in bounds [2][n]."""
    bounds = _c(bounds); args = _c(args); add = _c(add); coeffs = _c(coeffs)
    n = args.shape[1]
    out = np.zeros(n)
    lib().orc_root_find_scalar_all(_p(out), _p(bounds), _p(args), _p(add), _p(coeffs), int(max_iter), ctypes.c_double(err), n)
    return out


# --------------------------------------------------------------- integrator
def integrate(condBound, add, coeffs, hInit=0.01, max_step=400, tmax=1e13, fix_step=False, rtol=1e-12, atol=1e-12,
              step_num=None, denseOutput=None, denseOutputLoc=None, batch_coupled=False):
    """DOPR853.integrate semantics (eob/new_Dormand_horiz.py:424-778).

    Returns (denseOutputLoc [max_step][n], denseOutput [max_step][4][n], step_num, stop_info, iterations).
    """
    condBound = _c(condBound); add = _c(add); coeffs = _c(coeffs)
    n = condBound.shape[1]
    restart = denseOutput is not None
    if restart:
        dO = np.array(denseOutput, dtype=np.float64, order="C")
        dL = np.array(denseOutputLoc, dtype=np.float64, order="C")
        sn = np.array(step_num, dtype=np.int32)
        max_step = dO.shape[0]
    else:
        dO = np.zeros((max_step, 4, n)); dL = np.zeros((max_step, n)); sn = np.zeros(n, dtype=np.int32)
    si = np.zeros(n, dtype=np.int32)
    lib().orc_integrate.restype = ctypes.c_long
    it = lib().orc_integrate(_p(condBound), _p(add), _p(coeffs), n, ctypes.c_double(hInit), max_step, ctypes.c_double(tmax),
                             int(fix_step), ctypes.c_double(rtol), ctypes.c_double(atol), _pi(sn), _p(dO), _p(dL),
                             int(restart), _pi(si), int(batch_coupled))
    return dL, dO, sn, si, it


def run_trajectory(condBound, add, coeffs, max_step=400, rtol=1e-12, atol=1e-12, hInit=0.01, step_back=10.0,
                   fine_step=0.05, batch_coupled=False, fix_step1=False):
    """SEOBNRv4PHM.run_trajectory two-pass semantics; t in units of M.

    Returns (t [n][max_step], traj [n][4][max_step], num_steps, stop_info).
    """
    condBound = _c(condBound); add = _c(add); coeffs = _c(coeffs)
    n = condBound.shape[1]
    t = np.zeros((max_step, n)); traj = np.zeros((max_step, 4, n))
    ns = np.zeros(n, dtype=np.int32); si = np.zeros(n, dtype=np.int32)
    lib().orc_run_trajectory(_p(condBound), _p(add), _p(coeffs), n, max_step, ctypes.c_double(rtol), ctypes.c_double(atol),
                             ctypes.c_double(hInit), ctypes.c_double(step_back), ctypes.c_double(fine_step),
                             _p(t), _p(traj), _pi(ns), _pi(si), int(batch_coupled), int(fix_step1))
    return t.T.copy(), traj.transpose(2, 1, 0).copy(), ns, si


def dopr_attempt(y, h, add, coeffs, rtol=1e-12, atol=1e-12):
    """One attempted DOP853 step per lane: (k [12][4][n], y_new [4][n], err [n])."""
    y = _c(y); h = _c(np.broadcast_to(np.asarray(h, dtype=np.float64), (y.shape[1],))); add = _c(add); coeffs = _c(coeffs)
    n = y.shape[1]
    k = np.zeros((12, 4, n)); yn = np.zeros((4, n)); err = np.zeros(n)
    lib().orc_dopr_attempt(_p(y), _p(h), _p(add), n, _p(coeffs), _p(k), _p(yn), _p(err), ctypes.c_double(rtol),
                           ctypes.c_double(atol))
    return k, yn, err


def hlm22(r, phi, v, Hreal, eta, vPhiIn, prefix, coeffs, h22_calib=0.0):
    """EOBGetSpinFactorizedWaveform for (2,2) on knots [numSys][L] (src/EOB.cu:1616-1677):
    prefix complex [numSys]; returns complex [numSys][L]."""
    r = _c(r); phi = _c(phi); v = _c(v); Hreal = _c(Hreal); vPhiIn = _c(vPhiIn); eta = _c(eta); coeffs = _c(coeffs)
    pf = _c(np.asarray(prefix, dtype=np.complex128).view(np.float64))
    nb, L = r.shape
    out = np.zeros((nb, L, 2))
    idx = np.arange(nb, dtype=np.int32)
    lib().orc_hlm22(_p(out), _p(r), _p(phi), _p(v), _p(Hreal), _p(eta), _p(vPhiIn), _p(pf), ctypes.c_double(h22_calib), nb, L,
                    _p(coeffs), _pi(idx))
    return out[..., 0] + 1j * out[..., 1]


def hlm(l, m, r, phi, pphi, v, Hreal, eta, vPhiIn, prefix, coeffs, h22_calib=0.0):
    """EOBGetSpinFactorizedWaveform for any (l, m), 2 <= l <= 8, 1 <= m <= l, on knots
    [numSys][L] (src/EOB.cu:1236-1613); prefix complex [numSys]; returns complex [numSys][L]."""
    r = _c(r); phi = _c(phi); pphi = _c(pphi); v = _c(v); Hreal = _c(Hreal); vPhiIn = _c(vPhiIn); eta = _c(eta); coeffs = _c(coeffs)
    pf = _c(np.asarray(prefix, dtype=np.complex128).view(np.float64))
    nb, L = r.shape
    out = np.zeros((nb, L, 2))
    idx = np.arange(nb, dtype=np.int32)
    lib().orc_hlm(_p(out), int(l), int(m), _p(r), _p(phi), _p(pphi), _p(v), _p(Hreal), _p(eta), _p(vPhiIn), _p(pf),
                  ctypes.c_double(h22_calib), nb, L, _p(coeffs), _pi(idx))
    return out[..., 0] + 1j * out[..., 1]


def error(solOld, h, ks, rtol=1e-12, atol=1e-12):
    """Restated error norm; ks = dict of stage arrays k1,k2,k3,k6..k10 ([4][n])."""
    solOld = _c(solOld); h = _c(h)
    n = solOld.shape[1]
    err = np.zeros(n); solNew = np.zeros_like(solOld)
    lib().orc_error(_p(err), _p(solOld), _p(solNew), _p(h), *[_p(_c(ks[k])) for k in ("k1", "k2", "k3", "k6", "k7", "k8", "k9", "k10")],
                    n, 4, ctypes.c_double(rtol), ctypes.c_double(atol))
    return err, solNew


def controller(err, errOld, prevReject, h):
    err = _c(err); errOld = _c(errOld).copy(); h = _c(h).copy()
    pr = np.array(prevReject, dtype=np.uint8)
    flag = np.zeros(err.size, dtype=np.uint8)
    lib().orc_controller(flag.ctypes.data_as(PU8), _p(err), _p(errOld), pr.ctypes.data_as(PU8), _p(h), err.size)
    return flag.astype(bool), errOld, pr.astype(bool), h


def ref_error(solOld, h, ks):
    """The reference's error_wrap (src/DOPR853.cu:620-668), rtol = atol = 1e-12."""
    r = ref_lib()
    solOld = _c(solOld); h = _c(h)
    n = solOld.shape[1]
    err = np.zeros(n); solNew = np.zeros_like(solOld)
    f = getattr(r, "_Z10error_wrapPdS_S_S_S_S_S_S_S_S_S_S_ii")
    f(_p(err), _p(solOld), _p(solNew), _p(h), *[_p(_c(ks[k])) for k in ("k1", "k2", "k3", "k6", "k7", "k8", "k9", "k10")], n, 4)
    return err, solNew


def ref_controller(err, errOld, prevReject, h):
    """The reference's controllerSuccess_wrap (src/DOPR853.cu:764-797)."""
    r = ref_lib()
    err = _c(err); errOld = _c(errOld).copy(); h = _c(h).copy()
    pr = np.array(prevReject, dtype=np.uint8)
    flag = np.zeros(err.size, dtype=np.uint8)
    x = np.zeros(err.size)
    f = getattr(r, "_Z22controllerSuccess_wrapPbPdS0_S_S0_S0_ii")
    f(flag.ctypes.data_as(PU8), _p(err), _p(errOld), pr.ctypes.data_as(PU8), _p(h), _p(x), err.size, 4)
    return flag.astype(bool), errOld, pr.astype(bool), h


# ------------------------------------------------------------------ waveform
def cgamma(z):
    out = np.zeros(2)
    lib().orc_cgamma(ctypes.c_double(z.real), ctypes.c_double(z.imag), _p(out))
    return out[0] + 1j * out[1]


def ylm_equator(l, m, phi):
    out = np.zeros(2)
    lib().orc_ylm_equator(l, m, ctypes.c_double(phi), _p(out))
    return out[0] + 1j * out[1]


def get_hlm22(traj, lengths, add, coeffs, h22_calib=0.0):
    """(2,2) mode on knots for traj [nbin][4][len] (eob/seobnrv4phm.py:2138-2221)."""
    traj = _c(traj); lengths = _c(lengths, np.int32); add = _c(add); coeffs = _c(coeffs)
    nbin, _, L = traj.shape
    out = np.zeros((nbin, L, 2))
    lib().orc_get_hlm22(_p(out), _p(traj), _pi(lengths), nbin, L, _p(add), _p(coeffs), ctypes.c_double(h22_calib))
    return out[..., 0] + 1j * out[..., 1]


def flux_mode_abs2(l, m, v, Hreal, pphi, vPhi2, prefix_abs, eta, coeff_row, h22_calib=0.0):
    """|h_lm|^2 as the flux evaluates it (src/EOB.cu:1706-1941), one binary's record."""
    v = _c(v); Hreal = _c(Hreal); pphi = _c(pphi); vPhi2 = _c(vPhi2); row = _c(coeff_row)
    out = np.zeros(v.size)
    lib().orc_flux_mode_abs2(_p(out), int(l), int(m), _p(v), _p(Hreal), _p(pphi), _p(vPhi2), int(v.size), ctypes.c_double(prefix_abs),
                             ctypes.c_double(eta), ctypes.c_double(h22_calib), _p(row))
    return out


def get_hlm(l, m, traj, lengths, add, coeffs, h22_calib=0.0):
    """Mode (l, m) on knots for traj [nbin][4][len] (eob/seobnrv4phm.py:2138-2221)."""
    traj = _c(traj); lengths = _c(lengths, np.int32); add = _c(add); coeffs = _c(coeffs)
    nbin, _, L = traj.shape
    out = np.zeros((nbin, L, 2))
    lib().orc_get_hlm(_p(out), int(l), int(m), _p(traj), _pi(lengths), nbin, L, _p(add), _p(coeffs), ctypes.c_double(h22_calib))
    return out[..., 0] + 1j * out[..., 1]


def td_interp2(dataTime, ts, propArrays, c1, c2, c3, old_length, old_lengths, data_length, ms, inds, lengths, max_length,
               legacy_stride=True):
    """TD2 (src/interpolate.cu:441-549) restated with numpy: per binary b and sample
    i < lengths[b], sum over modes of (re + i im) e^{-i m phi}, each series the
    cubic at the sample's interval inds[b][i]; series s of binary b starts at
    b (2M+1) old_length + s * stride with stride = old_lengths[b] (the reference,
    legacy_stride) or old_length (the padded layout).  Returns complex
    [numBinAll][data_length] (zero past lengths[b])."""
    nbin = len(old_lengths)
    M = len(ms)
    out = np.zeros((nbin, data_length), dtype=np.complex128)
    y = np.asarray(propArrays, dtype=np.float64); c1 = np.asarray(c1); c2 = np.asarray(c2); c3 = np.asarray(c3)
    ts = np.asarray(ts, dtype=np.float64).reshape(nbin, old_length)
    inds = np.asarray(inds).reshape(nbin, max_length)
    for b in range(nbin):
        n = int(lengths[b])
        if n == 0:
            continue
        t = np.asarray(dataTime[:n], dtype=np.float64)
        ind = inds[b, :n].astype(np.int64)
        x = t - ts[b, ind]
        stride = int(old_lengths[b]) if legacy_stride else old_length
        base = b * (2 * M + 1) * old_length

        def ev(sub):
            k = base + sub * stride + ind
            return y[k] + c1[k] * x + c2[k] * x * x + c3[k] * x * x * x
        phi = ev(2 * M)
        acc = np.zeros(n, dtype=np.complex128)
        for j, m in enumerate(ms):
            acc += (ev(j) + 1j * ev(M + j)) * np.exp(-1j * (m * phi))
        out[b, :n] = acc
    return out


# ------------------------------------------------------------------- splines
def interpolate_TD(t, y, lengths, nbin, nsubs):
    """interpolate_TD: t [maxlen][nbin], y [maxlen][nsubs*nbin] -> (B, c1, c2, c3)."""
    t = _c(t); y = _c(y); lengths = _c(lengths, np.int32)
    B = np.zeros_like(y); ud = np.zeros_like(y); dg = np.zeros_like(y); ld = np.zeros_like(y)
    lib().orc_interpolate_TD(_p(t), _p(y), _p(B), _p(ud), _p(dg), _p(ld), _pi(lengths), nbin, nsubs)
    return B, ud, dg, ld


def spline_eval(tnew, t, y, c1, c2, c3, length, deriv=0):
    """One interpolant: knots t[:length], coefficient arrays 1-D."""
    tnew = _c(tnew); t = _c(t); y = _c(y); c1 = _c(c1); c2 = _c(c2); c3 = _c(c3)
    out = np.zeros(tnew.size)
    lib().orc_spline_eval(_p(out), _p(tnew), tnew.size, _p(t), 1, _p(y), _p(c1), _p(c2), _p(c3), 1, int(length), int(deriv))
    return out


def likelihood_partials(hp, hc, data, psd, Fp, Fc, ts, df):
    """all_in_one_likelihood partial sums [nblk][nbin][nchan] (complex)."""
    hp = _c(np.asarray(hp, dtype=np.complex128).view(np.float64)); hc = _c(np.asarray(hc, dtype=np.complex128).view(np.float64))
    data = _c(np.asarray(data, dtype=np.complex128).view(np.float64))
    psd = _c(psd); Fp = _c(Fp); Fc = _c(Fc); ts = _c(ts)
    nchan, F = psd.shape
    nbin = Fp.shape[1]
    nblk = (F + 1023) // 1024
    out = np.zeros((nblk, nbin, nchan, 2))
    lib().orc_likelihood(_p(out), _p(hp), _p(hc), _p(data), _p(psd), _p(Fp), _p(Fc), _p(ts), ctypes.c_double(df), nbin, nchan, F)
    return out[..., 0] + 1j * out[..., 1]


# ------------------------------------------------------------ whole pass
MTSUN_SI = 4.925491025543575903411922162094833998e-6


def psd_weight(f, f0=200.0):
    """1/S(f) for S(f) = (f/f0)^-4 + 1 + (f/f0)^2, with 1/S(0) = 0."""
    x = np.asarray(f, dtype=np.float64) / f0
    with np.errstate(divide="ignore"):
        S = x ** -4.0 + 1.0 + x * x
    return np.where(np.asarray(f) > 0, 1.0 / S, 0.0)


def hplus_from_trajectory(t, traj, ns, add, coeffs, total_mass, dt, N, amp=1.0):
    """h+ = amp Re h22(t_j) on t_j = t_0 + j dt (seconds), j < N, for trajectories
    t [n][L] (units of M), traj [n][4][L], ns[n] points: the (2,2) mode on the
    knots (get_hlms, eob/seobnrv4phm.py:2138-2221), not-a-knot splines of phi and
    Re, Im of h22 e^{2 i phi} (interpolate_TD, src/interpolate.cu:11-308),
    evaluated on the uniform grid up to the last knot
    (eob/seobnrv4phm.py:2319-2372).  Returns h [n][N]."""
    n = len(ns)
    amp = np.broadcast_to(np.asarray(amp, dtype=np.float64), (n,))
    total_mass = np.broadcast_to(np.asarray(total_mass, dtype=np.float64), (n,))
    h22 = get_hlm22(traj, ns, add, coeffs)
    out = np.zeros((n, N))
    for i in range(n):
        L = int(ns[i])
        tk = t[i, :L]
        phi = traj[i, 1, :L]
        s22 = h22[i, :L] * np.exp(2j * phi)
        y = np.stack([phi, s22.real, s22.imag], axis=1)
        B, c1, c2, c3 = interpolate_TD(tk[:, None].copy(), y, np.array([L], dtype=np.int32), 1, 3)
        dT = dt / (total_mass[i] * MTSUN_SI)
        nsamp = min(N, int((tk[-1] - tk[0]) / dT) + 1)
        tj = tk[0] + np.arange(nsamp) * dT
        ev = [spline_eval(tj, tk, y[:, q].copy(), c1[:, q].copy(), c2[:, q].copy(), c3[:, q].copy(), L) for q in range(3)]
        out[i, :nsamp] = amp[i] * (ev[1] * np.cos(2 * ev[0]) + ev[2] * np.sin(2 * ev[0]))
    return out


def fd_logl(h, data, df, f0=200.0):
    """log L = -1/2 4 df sum_k |d_k - rfft(h)_k|^2 / S(f_k) per row of h (the
    reference's rfft without a dt factor and get_ll finish,
    eob/seobnrv4phm.py:742-807)."""
    ht = np.fft.rfft(h, axis=-1)
    w = psd_weight(np.arange(ht.shape[-1]) * df, f0)
    return -0.5 * 4 * df * np.sum(np.abs(data[None, :] - ht) ** 2 * w[None, :], axis=-1)


def pipeline_logl(m1, m2, chi1, chi2, r0, data, dt, N, amp=1.0, max_step=4096, rtol=1e-10, atol=1e-10, hInit=0.01,
                  fix_step1=False, f0=200.0):
    """The whole pass for a batch: tables, circular initial conditions, two-pass
    trajectory, (2,2) knots, splines, h+ on N samples, rfft and log L against
    data (complex [N/2+1]).  Returns (logl, (t, traj, ns, stop_info), h)."""
    m1 = np.atleast_1d(np.asarray(m1, dtype=np.float64))
    add = make_additional_args(m1, m2, chi1, chi2)
    co = fac_coeffs(add)
    r0 = np.broadcast_to(np.asarray(r0, dtype=np.float64), m1.shape).copy()
    pphi0, pr0, om0, add = initial_conditions(r0, add, co)
    cond = np.array([r0, np.zeros(m1.size), pr0, pphi0])
    t, traj, ns, si = run_trajectory(cond, add, co, max_step=max_step, rtol=rtol, atol=atol, hInit=hInit, fix_step1=fix_step1)
    M = m1 + np.asarray(m2, dtype=np.float64)
    h = hplus_from_trajectory(t, traj, ns, add, co, M, dt, N, amp)
    df = 1.0 / (N * dt)
    return fd_logl(h, data, df, f0), (t, traj, ns, si), h


def fac_coeff_names():
    """Field names of the oracle's FacCoeffs record in memory order (read from
    the struct in oracle_flux.hpp), so tests can address coefficients by name."""
    import re
    text = open(os.path.join(HERE, "oracle_flux.hpp")).read()
    body = text[text.index("struct FacCoeffs {"):]
    body = body[:body.index("};")]
    names = []
    for line in body.splitlines():
        line = line.split("//")[0].strip()
        if line.startswith("double "):
            names += [x.strip() for x in line[len("double "):].rstrip(";").split(",")]
    assert len(names) == coeff_doubles(), (len(names), coeff_doubles())
    return names


def fac_coeffs_named(add, use_hm=1):
    """fac_coeffs(add) as a list (one dict name -> value per binary)."""
    co = fac_coeffs(add, use_hm)
    names = fac_coeff_names()
    return [dict(zip(names, row)) for row in co]


# ---------------------------------------------------------- detector projection
def detector_information(ra, dec, gmst, psi, tensors, vertices):
    """(Fplus, Fcross, time_shift) [nifo][n]: get_detector_information
    (eob/seobnrv4phm.py:597-634) restated with numpy, gmst given."""
    phi = np.asarray(ra) - np.asarray(gmst)
    theta = np.pi / 2 - np.asarray(dec)
    psi = np.asarray(psi)
    u = np.array([np.cos(phi) * np.cos(theta), np.cos(theta) * np.sin(phi), -np.sin(theta)]).T
    v = np.array([-np.sin(phi), np.cos(phi), np.zeros_like(phi)]).T
    m = -u * np.sin(psi)[:, None] - v * np.cos(psi)[:, None]
    n = -u * np.cos(psi)[:, None] + v * np.sin(psi)[:, None]
    ep = np.einsum("...i,...j->...ij", m, m) - np.einsum("...i,...j->...ij", n, n)
    ec = np.einsum("...i,...j->...ij", m, n) + np.einsum("...i,...j->...ij", n, m)
    D = np.asarray(tensors)
    Fp = np.einsum("kij,lij->kl", D, ep)
    Fc = np.einsum("kij,lij->kl", D, ec)
    om = np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)]).T
    dt = np.einsum("ij,kj->ki", om, np.asarray(vertices)) / 299792458.0
    return Fp, Fc, dt


def detector_signal(hp_fd, hc_fd, Fp, Fc, dt, df, k_min=0):
    """[n][nifo][F] detector-frame signal (eob/seobnrv4phm.py:742-767)."""
    F = hp_fd.shape[-1]
    f = np.arange(F) * df
    hp = np.array(hp_fd, dtype=np.complex128)
    hc = np.array(hc_fd, dtype=np.complex128)
    hp[:, :k_min] = 0.0
    hc[:, :k_min] = 0.0
    return (Fp.T[:, :, None] * hp[:, None, :] + Fc.T[:, :, None] * hc[:, None, :]) * np.exp(-1j * 2.0 * np.pi * f[None, None, :]
                                                                                           * dt.T[:, :, None])
