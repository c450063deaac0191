"""Batched, device-resident SEOBNR pipeline on MI355X.

One ``EOBBatch`` holds the per-binary tables (Hamiltonian coefficients and the
35-mode flux table), the initial state and the dense trajectory output in HBM.
Every stage is a HIP kernel from libeobhip.so:

  setup      additionalArgs -> tables             (SEOBNRv4PHM.setup_class, eob/seobnrv4phm.py:1922-1964)
  initial    r0 -> (pphi0, pr0, omega0)            (circular start; IC_diss false position, src/EOB.cu:2666-2834)
  trajectory persistent or plain DOP853 batch     (run_trajectory, eob/seobnrv4phm.py:2072-2136)
  hlm22      (2,2) mode on the trajectory knots   (get_hlms, eob/seobnrv4phm.py:2138-2221)
"""
from __future__ import annotations

import collections
import ctypes
import math
import time

import numpy as np
import torch

from ._lib import call, lib
from .utils import current_stream

MTSUN_SI = 4.925491025543575903411922162094833998e-6
NADD = 80


def _dbl_factorial(n):
    out = 1
    for k in range(n, 1, -2):
        out *= k
    return out


def _ipow_seq(x, k):
    """x^k (k >= 1) by repeated multiplication, the order the device uses."""
    p = x
    for _ in range(k - 1):
        p = p * x
    return p


def newton_prefix_constants():
    """The parameter-independent factors n_lm of the 35 prefixes (complex [35]):
    (i m)^l 8 pi / (2l+1)!! sqrt((l+1)(l+2) / (l(l-1))) for even l + m and
    -(i m)^l i 16 pi / (2l+1)!! sqrt(m2v) for odd (Damour, Iyer & Nagar 2008
    Eqs. 5-7)."""
    out = np.zeros(35, dtype=np.complex128)
    q = 0
    for l in range(2, 9):
        for m in range(1, l + 1):
            df = _dbl_factorial(2 * l + 1)
            n = (1j * m) ** l
            if (l + m) % 2 == 0:
                n = n * (8.0 * math.pi / df) * math.sqrt((l + 1) * (l + 2) / (l * (l - 1)))
            else:
                m2v = (2.0 * l + 1) * (l + 2) * (l * l - m * m) / ((2.0 * l - 1) * (l + 1) * l * (l - 1))
                n = -n * (1j * 16.0 * math.pi / df) * math.sqrt(m2v)
            out[q] = n
            q += 1
    return out


def newton_prefixes(m1, m2):
    """Newtonian multipole prefixes for the 35 flux modes (l = 2..8, m = 1..l).

    Damour, Iyer & Nagar 2008 Eqs. 5-7, following EOBComputeNewtonMultipolePrefixes
    (eob/seobnrv4phm.py:1703-1776).  Returns complex [35][n].  The arithmetic is
    ordered as the device's (``eobhip_params_setup``), which gives the same bits.
    """
    m1 = np.atleast_1d(np.asarray(m1, dtype=np.float64))
    m2 = np.atleast_1d(np.asarray(m2, dtype=np.float64))
    M = m1 + m2
    x1, x2 = m1 / M, m2 / M
    eta = m1 * m2 / (M * M)
    nlm = newton_prefix_constants()
    out = np.zeros((35, m1.size), dtype=np.complex128)
    q = 0
    for l in range(2, 9):
        for m in range(1, l + 1):
            eps = (l + m) % 2
            sign = 1.0 if m % 2 == 0 else -1.0
            t1 = (m1 != m2) | (sign == 1.0)
            k = l + eps - 1
            cdeg = np.where(t1, _ipow_seq(x2, k) + sign * _ipow_seq(x1, k),
                            -1.0 if l in (2, 3) else (-0.5 if l in (4, 5) else 0.0))
            out[q].real = (nlm[q].real * eta) * cdeg
            out[q].imag = (nlm[q].imag * eta) * cdeg
            q += 1
    return out


def make_additional_args(m1, m2, chi1, chi2, omega0=None, K=0.0, d5=0.0, dSO=0.0, dSS=0.0, h22_calib=0.0):
    """additionalArgs [80][n] as SEOBNRv4PHM.setup_class lays it out (masses scaled to M = 1)."""
    m1 = np.atleast_1d(np.asarray(m1, dtype=np.float64))
    m2 = np.atleast_1d(np.asarray(m2, dtype=np.float64))
    M = m1 + m2
    m1s, m2s = m1 / M, m2 / M
    add = np.zeros((NADD, m1.size))
    add[0], add[1], add[2], add[3] = m1s, m2s, chi1, chi2
    add[4] = 0.0 if omega0 is None else omega0
    add[5], add[6], add[7], add[8], add[9] = K, d5, dSO, dSS, h22_calib
    pf = newton_prefixes(m1s, m2s)
    add[10::2] = pf.real
    add[11::2] = pf.imag
    return add


# the modes the reference's SEOBNRv4HM evaluates (eob/eobstuff.py:56)
REFERENCE_MODES = ((2, 2), (2, 1), (3, 3), (4, 4), (5, 5))


class EOBBatch:
    """Device-resident tables and trajectory buffers for n binaries."""

    def __init__(self, additional_args, max_step=4096, device="cuda", use_hm=1):
        if isinstance(additional_args, torch.Tensor):
            add = additional_args.to(device=device, dtype=torch.float64).contiguous()
        else:
            add = torch.as_tensor(np.ascontiguousarray(additional_args, dtype=np.float64), device=device)
        self._alloc(int(add.shape[1]), max_step, device)
        self.add = add
        self.use_hm = int(use_hm)
        call("eobhip_bin_setup", add.data_ptr(), self.n, int(use_hm), self.bindata.data_ptr(), self.modes.data_ptr(), current_stream())

    @classmethod
    def from_params(cls, m1, m2, chi1, chi2, max_step=4096, device="cuda", use_hm=1):
        """A batch whose additional args and tables are built on the device from
        the masses and spins (``set_params``)."""
        m1 = m1 if isinstance(m1, torch.Tensor) else np.atleast_1d(np.asarray(m1, dtype=np.float64))
        b = cls.__new__(cls)
        b._alloc(int(m1.shape[0]), max_step, device)
        b.add = torch.empty((NADD, b.n), dtype=torch.float64, device=device)
        b.use_hm = int(use_hm)
        b.set_params(m1, m2, chi1, chi2)
        return b

    def set_params(self, m1, m2, chi1, chi2):
        """New binaries for the same buffers: masses (any common unit) and aligned
        spins [n], as device tensors or host values; additional args and tables are
        built on the device in one launch (``eobhip_params_setup``, the device form
        of SEOBNRv4PHM.setup_class, eob/seobnrv4phm.py:1922-1964).  Host values go
        to the device in one copy through pinned memory."""
        n = self.n
        if all(isinstance(x, torch.Tensor) and x.device.type == "cuda" for x in (m1, m2, chi1, chi2)):
            cols = [x.to(dtype=torch.float64).reshape(n).contiguous() for x in (m1, m2, chi1, chi2)]
        else:
            st = getattr(self, "_param_stage", None)
            if st is None:
                self._param_stage = st = torch.empty((4, n), dtype=torch.float64, pin_memory=True)
                self._param_dev = torch.empty((4, n), dtype=torch.float64, device=self.device)
            # the previous copy out of the staging rows must be done before they are rewritten
            ev = getattr(self, "_param_copied", None)
            if ev is not None:
                ev.synchronize()
            a = st.numpy()
            for j, x in enumerate((m1, m2, chi1, chi2)):
                a[j] = x.cpu().numpy() if isinstance(x, torch.Tensor) else np.broadcast_to(np.asarray(x, dtype=np.float64), (n,))
            self._param_dev.copy_(st, non_blocking=True)
            self._param_copied = torch.cuda.Event()
            self._param_copied.record(torch.cuda.current_stream(self.device))
            cols = [self._param_dev[j] for j in range(4)]
        call("eobhip_params_setup", cols[0].data_ptr(), cols[1].data_ptr(), cols[2].data_ptr(), cols[3].data_ptr(), n, self.use_hm,
             0.0, 0.0, 0.0, 0.0, 0.0, self.add.data_ptr(), self.bindata.data_ptr(), self.modes.data_ptr(), current_stream())
        self._param_cols = cols   # keep device inputs alive while the launch is queued

    def _alloc(self, n, max_step, device):
        self.n = int(n)
        self.device = device
        self.max_step = int(max_step)
        L = lib()
        self.bindata = torch.empty(int(L.eobhip_bindata_bytes()) * self.n, dtype=torch.uint8, device=device)
        self.modes = torch.empty(int(L.eobhip_modes_bytes()) * self.n, dtype=torch.uint8, device=device)
        self.dense = torch.zeros((self.n, self.max_step, 5), dtype=torch.float64, device=device)
        self.num_steps = torch.zeros(self.n, dtype=torch.int32, device=device)
        self.stop_info = torch.zeros(self.n, dtype=torch.int32, device=device)
        self.pass1_steps = torch.zeros(self.n, dtype=torch.int32, device=device)
        self.attempts = torch.zeros(self.n, dtype=torch.int64, device=device)
        self.scratch = torch.empty(int(L.eobhip_integrate_scratch_bytes(self.n)), dtype=torch.uint8, device=device)
        self.cond = None

    # ------------------------------------------------------------------ setup
    def initial_conditions(self, r0, max_iter=1000, err=1e-12):
        """Circular start at r0 (device tensor or host values [n]); returns
        (pphi0, pr0, omega0) as views of the batch's state that the next call
        overwrites."""
        if isinstance(r0, torch.Tensor):   # a device tensor is used as is (no host round trip)
            r0 = r0.to(device=self.device, dtype=torch.float64).reshape(self.n)
        else:
            r0 = torch.as_tensor(np.broadcast_to(np.asarray(r0, dtype=np.float64), (self.n,)).copy(), device=self.device)
        # one launch writes the whole initial state and omega0 (additional-args row 4);
        # every host launch before the integrator delays the pass
        cond = torch.empty((4, self.n), dtype=torch.float64, device=self.device)
        call("eobhip_initial_state", r0.data_ptr(), self.bindata.data_ptr(), self.modes.data_ptr(), self.n, cond.data_ptr(),
             self.add[4].data_ptr(), int(max_iter), float(err), current_stream())
        self.cond = cond
        # views, not copies (a copy would cost a launch in every pass): pphi0 and
        # pr0 are rows of the batch's initial state and omega0 is additional-args
        # row 4, so the next call overwrites all three
        return cond[3], cond[2], self.add[4]

    def set_initial_state(self, cond):
        self.cond = torch.as_tensor(np.ascontiguousarray(cond, dtype=np.float64), device=self.device).reshape(4, self.n).contiguous()

    # ------------------------------------------------------------- trajectory
    def integrate(self, rtol=1e-12, atol=1e-12, hInit=0.01, step_back=10.0, fine_step=0.05, two_pass=True, fix_step=False,
                  mode="persistent", group=64, block=256, grid_groups=0, x0=None, step0=None, use_stop=True, tmax=1e13, done=None):
        """Run the batch.  Returns the driver iteration count (plain and coupled modes) or 0.

        mode: "persistent" (one work-stealing launch), "plain" (one launch per
        attempted step) -- both integrate each binary as a batch of one -- or
        "coupled" (one launch per attempted step with the reference's batch-coupled
        stopping criterion, ``eobhip_integrate_batch_coupled``; ``self.old_omegas``
        then holds the criterion's omega history of the last pass).
        x0/step0 (device tensors [n]) continue an integration from row step0 of self.dense.
        done (pinned host int32 tensor [n], zeroed; persistent mode): done[b] becomes 1
        when binary b's outputs are in memory (``eobhip_integrate_signal``)."""
        if self.cond is None:
            raise RuntimeError("initial state not set")
        if mode not in ("persistent", "plain", "coupled"):
            raise ValueError(f"unknown mode {mode!r}")
        if mode == "coupled":
            if done is not None:
                raise ValueError("completion flags need the persistent mode")
            om = getattr(self, "old_omegas", None)
            if om is None or tuple(om.shape) != (self.max_step, self.n):
                om = torch.zeros((self.max_step, self.n), dtype=torch.float64, device=self.device)
                self.old_omegas = om
            it = ctypes.c_longlong(0)
            call("eobhip_integrate_batch_coupled", self.cond.data_ptr(), self.bindata.data_ptr(), self.modes.data_ptr(), self.n,
                 self.max_step, float(rtol), float(atol), float(hInit), float(step_back), float(fine_step), float(tmax),
                 int(bool(two_pass)), int(bool(fix_step)), None if x0 is None else x0.data_ptr(),
                 None if step0 is None else step0.data_ptr(), int(bool(use_stop)), int(group), self.dense.data_ptr(),
                 self.num_steps.data_ptr(), self.stop_info.data_ptr(), self.pass1_steps.data_ptr(), self.attempts.data_ptr(),
                 om.data_ptr(), self.scratch.data_ptr(), ctypes.byref(it), current_stream())
            return int(it.value)
        if done is not None:
            if mode != "persistent":
                raise ValueError("completion flags need the persistent mode")
            if done.device.type != "cpu" or not done.is_pinned() or done.dtype != torch.int32 or done.numel() < self.n:
                raise ValueError("done must be a pinned host int32 tensor of n elements")
            call("eobhip_integrate_signal", self.cond.data_ptr(), self.bindata.data_ptr(), self.modes.data_ptr(), self.n,
                 self.max_step, float(rtol), float(atol), float(hInit), float(step_back), float(fine_step), float(tmax),
                 int(bool(two_pass)), int(bool(fix_step)), None if x0 is None else x0.data_ptr(),
                 None if step0 is None else step0.data_ptr(), int(bool(use_stop)), int(group), int(block), int(grid_groups),
                 self.dense.data_ptr(), self.num_steps.data_ptr(), self.stop_info.data_ptr(), self.pass1_steps.data_ptr(),
                 self.attempts.data_ptr(), self.scratch.data_ptr(), done.data_ptr(), current_stream())
            return 0
        it = ctypes.c_longlong(0)
        call("eobhip_integrate", self.cond.data_ptr(), self.bindata.data_ptr(), self.modes.data_ptr(), self.n, self.max_step,
             float(rtol), float(atol), float(hInit), float(step_back), float(fine_step), float(tmax), int(bool(two_pass)),
             int(bool(fix_step)),
             None if x0 is None else x0.data_ptr(), None if step0 is None else step0.data_ptr(), int(bool(use_stop)),
             0 if mode == "persistent" else 1, int(group), int(block), int(grid_groups), self.dense.data_ptr(),
             self.num_steps.data_ptr(), self.stop_info.data_ptr(), self.pass1_steps.data_ptr(), self.attempts.data_ptr(),
             self.scratch.data_ptr(), ctypes.byref(it), current_stream())
        return int(it.value)

    def trajectory_host(self):
        """(t [n][L], traj [n][4][L], num_steps) on the host, t in units of M, zero past num_steps."""
        ns = self.num_steps.cpu().numpy()
        Lmax = int(ns.max()) if self.n else 0
        d = self.dense[:, :Lmax].cpu().numpy().copy()
        for b in range(self.n):
            d[b, ns[b]:] = 0.0
        return d[:, :, 0], d[:, :, 1:].transpose(0, 2, 1), ns

    # ----------------------------------------------------------- (2,2) knots
    def hlm22_knots(self, out=None):
        if out is None:
            out = torch.zeros((self.n, self.max_step, 3), dtype=torch.float64, device=self.device)
        call("eobhip_hlm22_knots", self.dense.data_ptr(), self.num_steps.data_ptr(), self.bindata.data_ptr(), self.modes.data_ptr(),
             self.n, self.max_step, out.data_ptr(), current_stream())
        self.knots = out
        return out

    # ------------------------------------------------- splines and resample
    def build_splines(self):
        """Not-a-knot splines of phi, Re and Im of h22 e^{2 i phi} over each binary's knots."""
        if getattr(self, "coef", None) is None:
            self.coef = torch.empty((self.n, self.max_step, 12), dtype=torch.float64, device=self.device)
            self.spline_scratch = torch.empty(int(lib().eobhip_spline_scratch_doubles(self.n, self.max_step)), dtype=torch.float64,
                                              device=self.device)
        call("eobhip_spline_build_batched", self.dense.data_ptr(), self.knots.data_ptr(), self.num_steps.data_ptr(), self.n,
             self.max_step, self.coef.data_ptr(), self.spline_scratch.data_ptr(), current_stream())
        return self.coef

    def resample_hplus(self, total_mass_msun, dt, N, amp=None, out=None):
        """h+ = amp Re(h22) on t_j = t_0 + j dt (seconds), j < N (zero past the last knot)."""
        M = torch.as_tensor(np.broadcast_to(np.asarray(total_mass_msun, dtype=np.float64), (self.n,)).copy(), device=self.device)
        dT_M = (dt / (M * MTSUN_SI)).contiguous()
        if amp is None:
            amp = torch.ones(self.n, dtype=torch.float64, device=self.device)
        else:
            amp = torch.as_tensor(np.broadcast_to(np.asarray(amp, dtype=np.float64), (self.n,)).copy(), device=self.device)
        if out is None:
            out = torch.empty((self.n, N), dtype=torch.float64, device=self.device)
        self.nsamp = torch.zeros(self.n, dtype=torch.int32, device=self.device)
        call("eobhip_resample_hplus", self.dense.data_ptr(), self.coef.data_ptr(), self.num_steps.data_ptr(), amp.data_ptr(),
             dT_M.data_ptr(), self.n, self.max_step, int(N), out.data_ptr(), self.nsamp.data_ptr(), current_stream())
        self._keep = (dT_M, amp)
        return out


    # ------------------------------------------------------------- higher modes
    def hlm_modes_knots(self, modes=REFERENCE_MODES, out=None):
        """The modes (l, m) on every binary's knots (get_hlms, eob/seobnrv4phm.py:2138-2221,
        and the series compute_full_waveform splines, :2336-2356): knots
        [n][max_step][2 M + 1] = Re(h_j e^{i m_j phi}) for the M modes, their Im, phi."""
        ls, ms = _mode_arrays(modes)
        nm = len(ls)
        L = lib()
        key = (tuple(ls), tuple(ms))
        if getattr(self, "_wave_key", None) != key:
            self.wtab = torch.empty(int(L.eobhip_wave_mode_bytes()) * self.n * nm, dtype=torch.uint8, device=self.device)
            self._wave_key = key
        call("eobhip_wave_tables", self.add.data_ptr(), self.n, self.use_hm, self.bindata.data_ptr(), self.modes.data_ptr(), nm,
             _host_ints(ls), _host_ints(ms), self.wtab.data_ptr(), current_stream())
        if out is None:
            out = torch.zeros((self.n, self.max_step, 2 * nm + 1), dtype=torch.float64, device=self.device)
        call("eobhip_hlm_modes_knots", self.dense.data_ptr(), self.num_steps.data_ptr(), self.bindata.data_ptr(), self.wtab.data_ptr(),
             self.add.data_ptr(), self.n, self.max_step, nm, out.data_ptr(), current_stream())
        self.mode_knots, self.mode_ms = out, ms
        return out

    def build_mode_splines(self):
        """Not-a-knot splines of the 2 M + 1 series of ``hlm_modes_knots``:
        coef [n][max_step][4 (2 M + 1)]."""
        nser = self.mode_knots.shape[2]
        L = lib()
        self.mode_coef = torch.empty((self.n, self.max_step, 4 * nser), dtype=torch.float64, device=self.device)
        scratch = torch.empty(int(L.eobhip_spline_series_scratch_doubles(self.n, self.max_step, nser)), dtype=torch.float64,
                              device=self.device)
        call("eobhip_spline_build_series", self.dense.data_ptr(), self.mode_knots.data_ptr(), self.num_steps.data_ptr(), self.n,
             self.max_step, nser, self.mode_coef.data_ptr(), scratch.data_ptr(), current_stream())
        self._keep_scratch = scratch
        return self.mode_coef

    def resample_modes(self, total_mass_msun, dt, N, amp=None):
        """(h+, hx) [n][N]: the real and imaginary parts of amp * sum_j h_j(t_j) on
        t_j = t_0 + j dt (seconds), j < N, zero past the last knot -- the reference's
        template over its modes (eob/seobnrv4phm.py:2366-2372, 720, 742-744)."""
        M = torch.as_tensor(np.broadcast_to(np.asarray(total_mass_msun, dtype=np.float64), (self.n,)).copy(), device=self.device)
        dT_M = (dt / (M * MTSUN_SI)).contiguous()
        a = torch.as_tensor(np.broadcast_to(np.asarray(1.0 if amp is None else amp, dtype=np.float64), (self.n,)).copy(),
                            device=self.device)
        hp = torch.empty((self.n, N), dtype=torch.float64, device=self.device)
        hc = torch.empty_like(hp)
        self.nsamp = torch.zeros(self.n, dtype=torch.int32, device=self.device)
        ms = self.mode_ms
        call("eobhip_resample_modes", self.dense.data_ptr(), self.mode_coef.data_ptr(), self.num_steps.data_ptr(), a.data_ptr(),
             dT_M.data_ptr(), self.n, self.max_step, int(N), len(ms), _host_ints(ms), hp.data_ptr(), hc.data_ptr(),
             self.nsamp.data_ptr(), current_stream())
        self._keep = (dT_M, a)
        return hp, hc


    # ------------------------------------------------- merger-ringdown attachment
    def build_rp_splines(self):
        """Not-a-knot splines of r and pr over the knots, rp_coef [n][max_step][8]
        (compute_full_waveform's r_pr_spline, eob/seobnrv4phm.py:2327-2334)."""
        knots = self.dense[:, :, [1, 3]].contiguous()
        self.rp_coef = torch.empty((self.n, self.max_step, 8), dtype=torch.float64, device=self.device)
        scratch = torch.empty(int(lib().eobhip_spline_series_scratch_doubles(self.n, self.max_step, 2)), dtype=torch.float64,
                              device=self.device)
        call("eobhip_spline_build_series", self.dense.data_ptr(), knots.data_ptr(), self.num_steps.data_ptr(), self.n, self.max_step,
             2, self.rp_coef.data_ptr(), scratch.data_ptr(), current_stream())
        self._keep_rp = (knots, scratch)
        return self.rp_coef

    def imr_params(self, total_mass_msun, dt):
        """NQC coefficients, remnant, QNMs and ringdown coefficients of every binary
        and mode of ``hlm_modes_knots`` (which must hold only modes with ringdown
        fits: (2,2), (2,1), (3,3), (4,4), (5,5)), for the grid t_j = t_0 + j dt
        seconds (``eobhip_imr_params``).  Needs ``build_mode_splines`` first; builds
        the r, pr splines.  Returns a dict of device tensors."""
        ls, ms = (list(x) for x in self._wave_key)
        nm = len(ls)
        L = lib()
        m1, m2, c1, c2 = self.add[:4].cpu().numpy()
        same = (np.abs(m1 - m2) < 1e-4) & (((np.abs(c1) < 1e-4) & (np.abs(c2) < 1e-4)) | (np.abs(c1 - c2) < 1e-4))
        if same.any():
            # equal masses with equal spins: the odd-m modes vanish and the NQC
            # systems are singular; the reference stops there too (:2391-2395)
            raise NotImplementedError(f"binaries {np.flatnonzero(same).tolist()} have equal masses and spins")
        self.build_rp_splines()
        M = torch.as_tensor(np.broadcast_to(np.asarray(total_mass_msun, dtype=np.float64), (self.n,)).copy(), device=self.device)
        dT_M = (dt / (M * MTSUN_SI)).contiguous()
        dev = dict(dtype=torch.float64, device=self.device)
        out = dict(qnm=torch.empty((self.n, nm, 2), **dev), nqc=torch.empty((self.n, nm, 5), **dev),
                   rd=torch.empty((self.n, nm, int(L.eobhip_imr_rd_doubles())), **dev),
                   binf=torch.zeros((self.n, int(L.eobhip_imr_bin_doubles())), **dev),
                   bini=torch.zeros((self.n, int(L.eobhip_imr_bin_ints())), dtype=torch.int32, device=self.device), dT_M=dT_M)
        call("eobhip_imr_params", self.add[0].data_ptr(), self.add[1].data_ptr(), self.add[2].data_ptr(), self.add[3].data_ptr(),
             self.dense.data_ptr(), self.num_steps.data_ptr(), self.mode_coef.data_ptr(), self.rp_coef.data_ptr(), dT_M.data_ptr(), self.n,
             self.max_step, nm, _host_ints(ls), _host_ints(ms), out["qnm"].data_ptr(), out["nqc"].data_ptr(), out["rd"].data_ptr(),
             out["binf"].data_ptr(), out["bini"].data_ptr(), current_stream())
        self.imr = out
        return out

    def imr_template(self, total_mass_msun, dt, N, amp=None, batch_ringdown=False, params=None):
        """(h+, hx) [n][N]: the IMR template -- amp times the sum over the modes of
        the NQC-corrected modes up to the attachment sample and of the ringdown
        after it (compute_full_waveform, eob/seobnrv4phm.py:2308-2556, summed as
        BBHWaveformTD does, :720) on t_j = t_0 + j dt seconds, j < N.  Each binary's
        ringdown lasts its own 30 (2,2) damping times; batch_ringdown=True pads every
        binary to the batch's longest, as the reference does (:2514).  params: the
        result of ``imr_params`` for the same masses and dt (computed here if None)."""
        p = self.imr_params(total_mass_msun, dt) if params is None else params
        if batch_ringdown:
            p["bini"][:, 2] = p["bini"][:, 2].max()
        ls, ms = (list(x) for x in self._wave_key)
        a = torch.as_tensor(np.broadcast_to(np.asarray(1.0 if amp is None else amp, dtype=np.float64), (self.n,)).copy(),
                            device=self.device)
        hp = torch.empty((self.n, N), dtype=torch.float64, device=self.device)
        hc = torch.empty_like(hp)
        self.nsamp = torch.zeros(self.n, dtype=torch.int32, device=self.device)
        call("eobhip_imr_resample", self.dense.data_ptr(), self.num_steps.data_ptr(), self.mode_coef.data_ptr(), self.rp_coef.data_ptr(),
             a.data_ptr(), p["dT_M"].data_ptr(), self.n, self.max_step, int(N), len(ls), _host_ints(ls), _host_ints(ms),
             p["nqc"].data_ptr(), p["rd"].data_ptr(), p["binf"].data_ptr(), p["bini"].data_ptr(), hp.data_ptr(), hc.data_ptr(),
             self.nsamp.data_ptr(), current_stream())
        self._keep = (a,)
        return hp, hc


def _mode_arrays(modes):
    modes = [(int(l), int(m)) for l, m in modes]
    if not 1 <= len(modes) <= 15:
        raise ValueError("between 1 and 15 modes")
    for l, m in modes:
        if not (2 <= l <= 8 and 1 <= m <= l):
            raise ValueError(f"mode {(l, m)} outside 2 <= l <= 8, 1 <= m <= l")
    return [l for l, _ in modes], [m for _, m in modes]


def _host_ints(xs):
    """A host int32 array for a by-value pointer argument (kept alive by ctypes' call)."""
    import ctypes
    return (ctypes.c_int * len(xs))(*xs)


def rfft_batched(h, out=None):
    """Row-wise real-to-complex FFT (hipFFT D2Z), no dt factor (as the reference's rfft)."""
    n, N = h.shape
    if out is None:
        out = torch.empty((n, N // 2 + 1), dtype=torch.complex128, device=h.device)
    call("eobhip_rfft_batched", h.data_ptr(), out.data_ptr(), int(n), int(N), current_stream())
    return out


def fd_likelihood(htilde, dtilde, df, f0=200.0, partial=None, out=None):
    """log L_b = -1/2 4 df sum_k |d_k - h_k|^2 / S(f_k), analytic S(f) = (f/f0)^-4 + 1 + (f/f0)^2."""
    n, F = htilde.shape
    if partial is None:
        partial = torch.empty(int(lib().eobhip_fd_likelihood_partial_doubles(n, F)), dtype=torch.float64, device=htilde.device)
    if out is None:
        out = torch.empty(n, dtype=torch.float64, device=htilde.device)
    call("eobhip_fd_likelihood", htilde.data_ptr(), dtilde.data_ptr(), partial.data_ptr(), int(n), int(F), float(df), float(f0),
         out.data_ptr(), current_stream())
    return out


def _device_vector(x, n, dev):
    """x (scalar, array or device tensor) as a float64 device vector of n; a device
    tensor is used as is, so that no host copy (which waits for the stream) is made."""
    if isinstance(x, torch.Tensor) and x.device.type == "cuda":
        x = x.to(dtype=torch.float64).reshape(-1)
        return x.expand(n).contiguous() if x.numel() == 1 else x.reshape(n).contiguous()
    return torch.as_tensor(np.broadcast_to(np.asarray(x, dtype=np.float64), (n,)).copy(), device=dev)


class SupportLikelihood:
    """log L_b = -1/2 4 df sum_k |d~_k - rfft(h_b)_k|^2 / S(f_k) for many templates,
    computed from each template's support instead of its N-point transform.

    With DD = sum |d~|^2 w, g_j = Re sum_k d~_k w_k e^{2 pi i jk/N} and the lag
    kernel of w (all fixed by the data set and set up once), a template of L
    non-zero samples needs h on j < M (M = power of two >= 2L, at most N), one
    M-point real FFT and two dot products:  log L = -2 df (DD - 2 h.g + hh).
    The result equals the full-length path (``resample_hplus`` + ``rfft_batched``
    + ``fd_likelihood``) up to rounding; only the order of the sums differs.
    """

    def __init__(self, dtilde, N, df, f0=200.0, m_min_log2=10, device="cuda"):
        L = lib()
        self.N, self.df, self.f0, self.m_min = int(N), float(df), float(f0), int(m_min_log2)
        self.log2N = self.N.bit_length() - 1
        if (1 << self.log2N) != self.N:
            raise ValueError("N must be a power of two")
        self.device = device
        d = torch.as_tensor(dtilde, dtype=torch.complex128, device=device).contiguous()
        if d.numel() != self.N // 2 + 1:
            raise ValueError("dtilde must have N/2 + 1 bins")
        self.g = torch.empty(self.N, dtype=torch.float64, device=device)
        self.dd = torch.empty(1, dtype=torch.float64, device=device)
        self.bank = torch.empty(int(L.eobhip_lik_bank_doubles(self.N, self.m_min)), dtype=torch.float64, device=device)
        scratch = torch.empty(int(L.eobhip_lik_setup_scratch_bytes(self.N)), dtype=torch.uint8, device=device)
        call("eobhip_lik_setup", d.data_ptr(), self.N, self.df, self.f0, self.m_min, self.g.data_ptr(), self.dd.data_ptr(),
             self.bank.data_ptr(), scratch.data_ptr(), current_stream())
        self._keep = (d, scratch)
        self._work = {}
        self.grows = 0   # work-buffer reallocations so far (each may make the host wait for the device)

    def reserve(self, n):
        """Size the current stream's work buffers for calls of up to n templates of
        any support (M = N each), so that no call on this stream allocates."""
        gx, gx2 = ctypes.c_int(0), ctypes.c_int(0)
        call("eobhip_lik_partial_blocks", self.log2N, ctypes.byref(gx), ctypes.byref(gx2))
        self._buffers(n, n * self.N, n * (self.N // 2 + 1), gx.value, gx2.value)

    def _buffers(self, n, ph, px, gx, gx2):
        """Work buffers of the calls on the current stream, kept between calls and
        grown when a call needs more (calls on different streams never share
        them, so chunks of one pass can run concurrently)."""
        key = torch.cuda.current_stream(self.device).cuda_stream
        w = self._work.get(key)
        need = {"n": n, "h": max(ph, 1), "x": max(px, 1), "cross": n * gx, "hh": n * gx2}
        if w is None or any(w["cap"][k] < v for k, v in need.items()):
            # grow by half again at least, so that calls of varying sizes reallocate rarely
            cap = {k: max(v, 0 if w is None else w["cap"][k] + w["cap"][k] // 2) for k, v in need.items()}
            self.grows += 1
            dev = self.device
            w = {"cap": cap,
                 "nsamp": torch.empty(cap["n"], dtype=torch.int32, device=dev),
                 "mlog2": torch.empty(cap["n"], dtype=torch.int32, device=dev),
                 "mlog2_host": torch.empty(cap["n"], dtype=torch.int32, pin_memory=True),
                 "offs_host": torch.empty(2 * cap["n"], dtype=torch.int64, pin_memory=True),
                 "offs": torch.empty(2 * cap["n"], dtype=torch.int64, device=dev),
                 "hbuf": torch.empty(cap["h"], dtype=torch.float64, device=dev),
                 "X": torch.empty(cap["x"], dtype=torch.complex128, device=dev),
                 "cross": torch.empty(max(cap["cross"], 1), dtype=torch.float64, device=dev),
                 "hh": torch.empty(max(cap["hh"], 1), dtype=torch.float64, device=dev)}
            self._work[key] = w
        return w

    def __call__(self, batch, total_mass_msun, dt, amp=None, out=None, terms=None):
        return self.complete(self.begin(batch, total_mass_msun, dt, amp), out=out, terms=terms)

    def warm_transforms(self):
        """Create the transform plans of every size (2^m_min .. N) and row chunk
        (1, 2, 4, .. 64 rows) on the current stream -- plans are kept per stream
        and size -- so that no later call on this stream spends host time
        creating one (a call whose templates fall into new sizes or row counts
        would otherwise)."""
        dev, chunk = self.device, 64
        src = torch.zeros(chunk << self.log2N, dtype=torch.float64, device=dev)
        dst = torch.empty(chunk * ((self.N >> 1) + 1), dtype=torch.complex128, device=dev)
        for k in range(self.m_min, self.log2N + 1):
            rows = 1
            while rows <= chunk:
                call("eobhip_rfft_batched", src.data_ptr(), dst.data_ptr(), rows, 1 << k, current_stream())
                rows *= 2
        torch.cuda.current_stream(dev).synchronize()

    def begin(self, batch, total_mass_msun, dt, amp=None, dT_M=None):
        """First half of a call: support sizes on the device and their copy to the
        host, queued on the current stream.  ``complete`` waits for that copy only,
        so a caller can queue the first halves of several batches (on their own
        streams) before it waits for any of them.  dT_M (optional device tensor
        [n]): the sample spacing in units of each binary's mass, dt / (M MTSUN),
        when the caller has it (total_mass_msun is then not used)."""
        n, dev = batch.n, self.device
        if dT_M is None:
            M = _device_vector(total_mass_msun, n, dev)
            dT_M = (dt / (M * MTSUN_SI)).contiguous()
        amp_t = torch.ones(n, dtype=torch.float64, device=dev) if amp is None else _device_vector(amp, n, dev)
        w = self._buffers(n, 0, 0, 0, 0)
        nsamp, mlog2 = w["nsamp"][:n], w["mlog2"][:n]
        call("eobhip_lik_sizes", batch.dense.data_ptr(), batch.num_steps.data_ptr(), dT_M.data_ptr(), n, batch.max_step, self.N,
             self.m_min, nsamp.data_ptr(), mlog2.data_ptr(), current_stream())
        ml_host = w["mlog2_host"][:n]
        ml_host.copy_(mlog2, non_blocking=True)
        ready = torch.cuda.Event()
        ready.record(torch.cuda.current_stream(dev))
        return {"batch": batch, "n": n, "dT_M": dT_M, "amp": amp_t, "nsamp": nsamp, "mlog2": mlog2, "ml_host": ml_host,
                "ready": ready, "stream": torch.cuda.current_stream(dev)}

    def complete(self, h, out=None, terms=None):
        """Second half: waits for the sizes, lays the templates out by transform
        size, and queues the resample, the M-point transforms and the sums on the
        stream ``begin`` ran on.  Returns log L [n]."""
        batch, n, dev, st = h["batch"], h["n"], self.device, h["stream"]
        nsamp, mlog2 = h["nsamp"], h["mlog2"]
        h["ready"].synchronize()
        ml = h["ml_host"].numpy().copy()
        order = np.argsort(ml, kind="stable")
        rowoff = np.empty(n, dtype=np.int64)
        xoff = np.empty(n, dtype=np.int64)
        buckets = []
        ph = px = 0
        for k in np.unique(ml):
            idx = order[ml[order] == k]
            Mk = 1 << int(k)
            buckets.append((Mk, ph, px, idx.size))
            rowoff[idx] = ph + Mk * np.arange(idx.size)
            xoff[idx] = px + (Mk // 2 + 1) * np.arange(idx.size)
            ph += Mk * idx.size
            px += (Mk // 2 + 1) * idx.size
        kmax = int(ml.max()) if n else self.m_min
        gx, gx2 = ctypes.c_int(0), ctypes.c_int(0)
        call("eobhip_lik_partial_blocks", kmax, ctypes.byref(gx), ctypes.byref(gx2))
        with torch.cuda.stream(st):
            sp = current_stream()
            w = self._buffers(n, ph, px, gx.value, gx2.value)
            # row offsets then transform offsets, one contiguous copy
            oh = w["offs_host"][:2 * n]
            oh[:n].numpy()[:] = rowoff
            oh[n:].numpy()[:] = xoff
            offs = w["offs"][:2 * n]
            offs.copy_(oh, non_blocking=True)
            hbuf, X, cross, hh = w["hbuf"], w["X"], w["cross"], w["hh"]
            call("eobhip_lik_resample", batch.dense.data_ptr(), batch.coef.data_ptr(), batch.num_steps.data_ptr(),
                 h["amp"].data_ptr(), h["dT_M"].data_ptr(), nsamp.data_ptr(), mlog2.data_ptr(), offs[:n].data_ptr(), n,
                 batch.max_step, kmax, self.g.data_ptr(), hbuf.data_ptr(), cross.data_ptr(), sp)
            for Mk, h0, x0, cnt in buckets:
                call("eobhip_rfft_batched", hbuf.data_ptr() + 8 * h0, X.data_ptr() + 16 * x0, int(cnt), int(Mk), sp)
            if out is None:
                out = torch.empty(n, dtype=torch.float64, device=dev)
            call("eobhip_lik_finish", X.data_ptr(), offs[n:].data_ptr(), mlog2.data_ptr(), n, kmax, self.m_min, self.bank.data_ptr(),
                 cross.data_ptr(), hh.data_ptr(), self.dd.data_ptr(), self.df, out.data_ptr(),
                 None if terms is None else terms.data_ptr(), sp)
        self.nsamp, self.mlog2 = nsamp, mlog2
        self._tmp = (h["dT_M"], h["amp"])
        return out


def _runs_beside(a, b, cycles):
    """True if a short kernel on stream b finishes while a spin of `cycles` runs on a."""
    start = torch.cuda.Event(enable_timing=True)
    end_a = torch.cuda.Event(enable_timing=True)
    end_b = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(a)
    call("eobhip_spin", int(cycles), a.cuda_stream)
    end_a.record(a)
    b.wait_event(start)
    call("eobhip_spin", 0, b.cuda_stream)
    end_b.record(b)
    torch.cuda.synchronize()
    return start.elapsed_time(end_b) < 0.5 * start.elapsed_time(end_a)


def concurrent_streams(k, device="cuda", cycles=2_000_000, candidates=32):
    """k streams (default priority, from torch's pool) whose kernels run
    concurrently with each other's.  A process gets a few hardware queues
    (GPU_MAX_HW_QUEUES); further streams share them, and a kernel queued behind
    a long one in a shared queue waits for it -- which would serialise chunks
    meant to overlap.  Each candidate is tested against the streams already
    taken, in both directions, with a bounded spin kernel.  Returns fewer than
    k streams if the candidates run out."""
    chosen = []
    for _ in range(candidates):
        if len(chosen) == k:
            break
        s = torch.cuda.Stream(device=device)
        call("eobhip_spin", 0, s.cuda_stream)   # first use of the stream
        if all(_runs_beside(c, s, cycles) and _runs_beside(s, c, cycles) for c in chosen):
            chosen.append(s)
    torch.cuda.synchronize()
    return chosen


def expected_length(r0, m1, m2, chi1=0.0, chi2=0.0):
    """Expected attempted steps of a binary's two-pass integration, up to a
    factor: r0^1.165 nu^-0.292 exp(0.150 chi_eff), a least-squares fit of
    log(attempts) to log r0, log nu and chi_eff over the 1024-binary uneven
    batch (rank correlation 0.98; orbits to merger, r0^(5/2)/nu, rank 0.96,
    puts more of the longest binaries among shorter ones).  Used only to group
    binaries of similar length; it changes no result."""
    m1 = np.asarray(m1, dtype=np.float64)
    m2 = np.asarray(m2, dtype=np.float64)
    nu = m1 * m2 / (m1 + m2) ** 2
    chi_eff = (m1 * np.asarray(chi1, dtype=np.float64) + m2 * np.asarray(chi2, dtype=np.float64)) / (m1 + m2)
    return np.exp(1.165 * np.log(np.asarray(r0, dtype=np.float64)) - 0.292 * np.log(nu) + 0.150 * chi_eff)


class ChunkedPass:
    """One pass over n binaries -- tables, initial conditions, the two-pass
    trajectory, the (2,2) knots, the splines and the support-length likelihood
    (``SupportLikelihood``) -- with the batch split into groups of binaries of
    similar expected length (``fractions`` of the batch, longest first), each
    with its own buffers and stream.  Three chunks by default: with the
    caller's stream they use four hardware queues, the number a process gets
    (GPU_MAX_HW_QUEUES); streams beyond that share a queue, and work queued
    behind a long integrator in a shared queue waits for it.

    Every chunk's persistent integrator starts at once, so together they occupy
    the SIMDs as one launch would.  The chunks of shorter binaries finish first
    and run their knots, splines, transforms and sums while the longer chunks
    are still integrating: that downstream work fills the integrator's tail,
    in which only the longest binaries' wavefronts still run.  The chunk of
    the longest binaries is the smallest, since its downstream work is the
    part left after the last trajectory ends.  Each binary's
    numbers are those of the one-batch pass (the kernels treat binaries
    independently; the transforms run on other batch sizes, so log L may
    differ in the last bits).
    """

    def __init__(self, m1, m2, chi1, chi2, r0, lik, total_mass, dt, amp=None, max_step=4096, fractions=(0.1, 0.2, 0.7),
                 rtol=1e-10, device="cuda", order=None):
        m1, m2, chi1, chi2, r0 = (np.atleast_1d(np.asarray(x, dtype=np.float64)) for x in (m1, m2, chi1, chi2, r0))
        n = m1.size
        fr = np.asarray(fractions, dtype=np.float64)
        if fr.ndim != 1 or fr.size < 1 or fr.size > n or np.any(fr <= 0):
            raise ValueError("fractions: 1..n positive shares of the batch, longest binaries first")
        if order is None:
            order = np.argsort(-expected_length(r0, m1, m2, chi1, chi2), kind="stable")   # longest first
        self.n, self.lik, self.dt, self.rtol, self.device = n, lik, float(dt), float(rtol), device
        M = np.broadcast_to(np.asarray(total_mass, dtype=np.float64), (n,))
        A = np.broadcast_to(np.asarray(1.0 if amp is None else amp, dtype=np.float64), (n,))
        # chunk sizes from the shares, each at least one binary
        cuts = np.round(np.cumsum(fr / fr.sum()) * n).astype(np.int64)[:-1]
        cuts = np.clip(np.maximum(cuts, np.arange(1, fr.size)), None, n - (fr.size - 1) + np.arange(fr.size - 1))
        streams = concurrent_streams(fr.size, device=device)
        if len(streams) < fr.size:
            raise RuntimeError(f"found {len(streams)} streams that run concurrently, {fr.size} chunks need as many")
        self.groups = [np.asarray(g) for g in np.split(np.asarray(order), cuts)]
        self.streams = streams
        self.batches, self.r0, self.M, self.amp, self.idx, self.outs = [], [], [], [], [], []
        for g in self.groups:
            self.batches.append(EOBBatch(make_additional_args(m1[g], m2[g], chi1[g], chi2[g]), max_step=max_step, device=device))
            # per-chunk inputs live on the device: a host array would be copied
            # (and its stream waited for) inside the pass
            self.r0.append(torch.as_tensor(r0[g].copy(), device=device))
            self.M.append(torch.as_tensor(M[g].copy(), device=device))
            self.amp.append(torch.as_tensor(A[g].copy(), device=device))
            self.idx.append(torch.as_tensor(g, dtype=torch.int64, device=device))
            self.outs.append(torch.empty(g.size, dtype=torch.float64, device=device))
        self.out = torch.empty(n, dtype=torch.float64, device=device)
        torch.cuda.synchronize(device)

    def run(self, group=64, block=64, grid_groups=0, events=None):
        """Launch the pass; returns log L [n] in the caller's order (valid on the
        current stream once it returns).  ``events`` (optional) gets
        'traj_end_longest' recorded when the longest chunk's trajectory ends."""
        cur = torch.cuda.current_stream(self.device)
        for st in self.streams:
            st.wait_stream(cur)
        for k, (b, st) in enumerate(zip(self.batches, self.streams)):   # longest first
            with torch.cuda.stream(st):
                call("eobhip_bin_setup", b.add.data_ptr(), b.n, 1, b.bindata.data_ptr(), b.modes.data_ptr(), current_stream())
                b.initial_conditions(self.r0[k])
                b.integrate(rtol=self.rtol, atol=self.rtol, two_pass=True, group=group, block=block, grid_groups=grid_groups)
                if k == 0 and events is not None:
                    events["traj_end_longest"].record(st)
        pending = []
        for k in reversed(range(len(self.batches))):   # shortest first: they finish first
            b, st = self.batches[k], self.streams[k]
            with torch.cuda.stream(st):
                b.hlm22_knots(out=getattr(b, "knots", None))
                b.build_splines()
                pending.append((k, self.lik.begin(b, self.M[k], self.dt, amp=self.amp[k])))
        for k, h in pending:
            self.lik.complete(h, out=self.outs[k])
            with torch.cuda.stream(self.streams[k]):
                self.out.index_copy_(0, self.idx[k], self.outs[k])
        for st in self.streams:
            cur.wait_stream(st)
        return self.out


class _Rows:
    """Downstream buffers for up to ``cap`` binaries gathered from a batch: the
    rows, tables and per-binary inputs of the binaries being processed, the
    knots and spline coefficients, and their log L.  ``view(k)`` is the first k
    binaries in the shape ``EOBBatch``'s knot and spline calls and
    ``SupportLikelihood`` expect."""

    def __init__(self, cap, max_step, device):
        L = lib()
        self.cap, self.max_step, self.device = cap, max_step, device
        self.bd_bytes, self.md_bytes = int(L.eobhip_bindata_bytes()), int(L.eobhip_modes_bytes())
        self.dense = torch.empty((cap, max_step, 5), dtype=torch.float64, device=device)
        self.num_steps = torch.empty(cap, dtype=torch.int32, device=device)
        self.bindata = torch.empty((cap, self.bd_bytes), dtype=torch.uint8, device=device)
        self.modes = torch.empty((cap, self.md_bytes), dtype=torch.uint8, device=device)
        self.knots = torch.empty((cap, max_step, 3), dtype=torch.float64, device=device)
        self.coef = torch.empty((cap, max_step, 12), dtype=torch.float64, device=device)
        self.spline_scratch = torch.empty(int(L.eobhip_spline_scratch_doubles(cap, max_step)), dtype=torch.float64, device=device)
        self.dT_M = torch.empty(cap, dtype=torch.float64, device=device)
        self.amp = torch.empty(cap, dtype=torch.float64, device=device)
        self.pos = torch.empty(cap, dtype=torch.int64, device=device)
        self.out = torch.empty(cap, dtype=torch.float64, device=device)
        self.idx = torch.empty(cap, dtype=torch.int64, device=device)
        self.idx_host = torch.empty(cap, dtype=torch.int64, pin_memory=True)

    class _View:
        hlm22_knots = EOBBatch.hlm22_knots
        build_splines = EOBBatch.build_splines

    def view(self, k):
        v = self._View()
        v.n, v.max_step, v.device = k, self.max_step, self.device
        v.dense, v.num_steps = self.dense[:k], self.num_steps[:k]
        v.bindata, v.modes = self.bindata[:k], self.modes[:k]
        v.knots, v.coef, v.spline_scratch = self.knots[:k], self.coef[:k], self.spline_scratch
        return v


def drain_finished(done, nslot, launch, complete, is_ready, integrator_ended, poll_s=2e-5, spin_below=32, max_group=None):
    """The host's schedule in a streamed pass.

    done: host int array the integrator fills (nonzero = finished).
    launch(new, slot) queues the first half of the group of newly finished
    binaries ``new`` on downstream slot ``slot`` and returns a handle;
    is_ready(handle) says whether that group's support sizes are on the host;
    complete(handle) queues its second half.  Second halves are queued oldest
    group first; at most ``nslot`` groups are in flight and a slot is given
    to a new group only once its previous group's second half is queued; the
    host never waits for a group's sizes while binaries may still finish.
    While more than ``spin_below`` binaries remain, polls sleep ``poll_s`` and
    ready second halves go before new groups; for the last ones, new groups
    go first.
    A group takes at most ``max_group`` binaries (all newly finished ones if None).
    Raises if ``integrator_ended()`` while flags are missing.  Returns the
    groups (index arrays into done) in launch order."""
    n = done.size
    taken = np.zeros(n, dtype=bool)
    left = n
    pending = collections.deque()   # (handle, slot) of groups whose second half is not queued yet, oldest first
    groups = []
    while True:
        # while many binaries remain, queue ready second halves first; for the
        # last few, start a newly finished group first (its first half then runs
        # beside the previous group's second half instead of after its launches)
        last_few = left <= spin_below
        if pending and not last_few and is_ready(pending[0][0]):
            complete(pending.popleft()[0])
            continue
        if len(pending) < nslot:
            new = np.flatnonzero((done != 0) & ~taken)
            if max_group is not None:
                new = new[:max_group]
            if new.size:
                taken[new] = True
                left -= new.size
                groups.append(new)
                busy = {sl for _, sl in pending}
                slot = next(i for i in range(nslot) if i not in busy)
                pending.append((launch(new, slot), slot))
                continue
        if pending and last_few and is_ready(pending[0][0]):
            complete(pending.popleft()[0])
            continue
        if left == 0:
            if not pending:
                break
            complete(pending.popleft()[0])   # nothing else left to do: wait for its sizes
            continue
        if integrator_ended() and not (done != 0).all():
            raise RuntimeError("the integrator ended without flagging every binary")
        if left > spin_below:
            time.sleep(poll_s)   # the last few binaries are polled without sleeping (a sleep lasts 50 us or more)
    return groups


class StreamedPass:
    """One pass over n binaries, like ``ChunkedPass``, with the downstream work
    started per group of finished binaries instead of per chunk.

    The whole batch (ordered longest expected first by ``set_params``; in the
    caller's order after ``set_params_device``) integrates in one
    persistent launch on its own stream (``eobhip_integrate_signal``), which
    raises a pinned host flag per binary once the binary's rows are in memory.
    The host polls the flags; whenever binaries have finished it gathers their
    rows and tables into a downstream buffer (``_Rows``) and queues knots,
    splines and the first half of the likelihood for them on a free one of
    the downstream streams (three by default); a group's second half (which needs its support
    sizes on the host) is queued once they have arrived, oldest group first,
    and the host never blocks on them while binaries may still finish.
    Downstream groups therefore size themselves to the rate at which binaries
    finish, and what remains after the last trajectory ends is the downstream
    work of the few binaries that finished last.

    Each binary's numbers are those of the one-batch pass (the integrator and
    the knot, spline and likelihood kernels treat binaries independently; the
    transforms run on other batch sizes, so log L may differ in the last
    bits).  ``run`` blocks the calling thread until every binary's work is
    queued.  The constructor runs ``warm_passes`` passes (one by default):
    the first pass of a new object allocates work buffers and loads kernels.
    """

    def __init__(self, m1, m2, chi1, chi2, r0, lik, total_mass, dt, amp=None, max_step=4096, rtol=1e-10, device="cuda",
                 order=None, poll_s=2e-5, down_streams=3, warm_passes=1, max_group=128):
        m1, m2, chi1, chi2, r0 = (np.atleast_1d(np.asarray(x, dtype=np.float64)) for x in (m1, m2, chi1, chi2, r0))
        n = m1.size
        if n < 1:
            raise ValueError("empty batch")
        if down_streams < 1:
            raise ValueError("down_streams must be >= 1")
        if order is None:
            order = np.argsort(-expected_length(r0, m1, m2, chi1, chi2), kind="stable")   # longest first
        order = np.asarray(order, dtype=np.int64)
        if order.shape != (n,) or not np.array_equal(np.sort(order), np.arange(n)):
            raise ValueError("order must be a permutation of range(n)")
        self.n, self.lik, self.dt, self.rtol, self.device, self.poll_s = n, lik, float(dt), float(rtol), device, float(poll_s)
        # one stream for the integrator and down_streams for the downstream groups
        # (three by default: with the integrator's that makes the four hardware
        # queues a process gets; the caller's stream, idle during the pass, shares
        # one); fewer if fewer run concurrently, at least one
        streams = concurrent_streams(1 + down_streams, device=device)
        if len(streams) < 2:
            raise RuntimeError(f"found {len(streams)} streams that run concurrently, the pass needs at least 2")
        self.integ_stream, self.down_streams = streams[0], streams[1:]
        # the binaries' parameters in pass order: host staging (pinned) and device rows
        # m1, m2, chi1, chi2, r0, total mass, amplitude, caller position
        self._stage = torch.empty((8, n), dtype=torch.float64, pin_memory=True)
        self._par = torch.empty((8, n), dtype=torch.float64, device=device)
        self._staged = None
        self.r0, self.M, self.amp = self._par[4], self._par[5], self._par[6]
        self.pos = torch.empty(n, dtype=torch.int64, device=device)
        self.batch = EOBBatch.from_params(m1, m2, chi1, chi2, max_step=max_step, device=device)
        self.set_params(m1, m2, chi1, chi2, r0, total_mass=total_mass, amp=amp, order=order)
        self.done = torch.zeros(n, dtype=torch.int32, pin_memory=True)
        self.max_group = max(1, min(int(max_group), n))
        self.rows = [_Rows(self.max_group, self.batch.max_step, device) for _ in self.down_streams]
        for st in self.down_streams:
            with torch.cuda.stream(st):
                lik.warm_transforms()
                # work buffers for the largest group, allocated now: an allocation in
                # the middle of a pass can make the host wait for the integrator
                lik.reserve(self.max_group)
        self.out = torch.empty(n, dtype=torch.float64, device=device)
        self.groups = []   # caller indices of each downstream group of the last run
        torch.cuda.synchronize(device)
        # the first pass allocates the likelihood's work buffers, loads the
        # downstream kernels and fills the allocator's caches; run it here
        for _ in range(int(warm_passes)):
            self.run()
            torch.cuda.synchronize(device)

    def set_params(self, m1, m2, chi1, chi2, r0, total_mass=None, amp=None, order=None):
        """New binaries (same count) for the next ``run``: masses, aligned spins,
        initial separations [n]; total_mass (solar masses, for the sample spacing)
        defaults to m1 + m2, amp to 1.  The pass order (longest expected first,
        ``expected_length``) is computed here on the host; the parameters go to
        the device in one copy on the current stream, and ``run`` builds the
        additional args, tables and initial state from them on the device."""
        n = self.n
        m1, m2, chi1, chi2, r0 = (np.broadcast_to(np.asarray(x, dtype=np.float64), (n,)) for x in (m1, m2, chi1, chi2, r0))
        if order is None:
            order = np.argsort(-expected_length(r0, m1, m2, chi1, chi2), kind="stable")   # longest first
        order = np.asarray(order, dtype=np.int64)
        if order.shape != (n,) or not np.array_equal(np.sort(order), np.arange(n)):
            raise ValueError("order must be a permutation of range(n)")
        M = m1 + m2 if total_mass is None else np.broadcast_to(np.asarray(total_mass, dtype=np.float64), (n,))
        A = np.broadcast_to(np.asarray(1.0 if amp is None else amp, dtype=np.float64), (n,))
        if self._staged is not None:
            self._staged.synchronize()   # the previous copy has left the staging rows
        st = self._stage.numpy()
        for j, x in enumerate((m1, m2, chi1, chi2, r0, M, A)):
            st[j] = x[order]
        st[7] = order
        self._par.copy_(self._stage, non_blocking=True)
        self.pos.copy_(self._par[7].to(torch.int64))
        self._staged = torch.cuda.Event()
        self._staged.record(torch.cuda.current_stream(self.device))
        self.order = order

    def set_params_device(self, par):
        """New binaries whose parameters are already in device memory: par is a
        float64 tensor [7, n] on this pass's device with rows m1, m2, chi1,
        chi2, r0, total mass (solar masses) and amplitude.  The binaries keep
        the caller's order (no host work and no host synchronisation: one
        device copy on the current stream); ``run`` builds the additional
        args, tables and initial state from them on the device, as after
        ``set_params``.  With at most one binary per SIMD (n <= 4 x CUs) every
        binary starts at once, so the order does not change the pass time."""
        n = self.n
        if not isinstance(par, torch.Tensor) or par.dtype != torch.float64 or tuple(par.shape) != (7, n):
            raise ValueError(f"par must be a float64 tensor of shape (7, {n})")
        if par.device != self._par.device:
            raise ValueError(f"par must be on {self._par.device}")
        if getattr(self, "_iota", None) is None:
            self._iota = torch.arange(n, dtype=torch.int64, device=self._par.device)
            self._iota_host = np.arange(n, dtype=np.int64)
        self._par[:7].copy_(par)
        self._par[7].copy_(self._iota)
        self.pos.copy_(self._iota)
        self.order = self._iota_host

    def _launch(self, new, slot):
        rows, st, k = self.rows[slot], self.down_streams[slot], new.size
        rows.idx_host[:k].numpy()[:] = new
        with torch.cuda.stream(st):
            idx = rows.idx[:k]
            idx.copy_(rows.idx_host[:k], non_blocking=True)
            b = self.batch
            # one launch (each host launch costs as much as a small group's kernels)
            call("eobhip_gather_binaries", idx.data_ptr(), k, self.n, b.max_step, b.dense.data_ptr(), b.num_steps.data_ptr(),
                 b.bindata.data_ptr(), rows.bd_bytes, b.modes.data_ptr(), rows.md_bytes, self.M.data_ptr(), self.amp.data_ptr(),
                 self.pos.data_ptr(), self.dt, MTSUN_SI, rows.dense.data_ptr(), rows.num_steps.data_ptr(), rows.bindata.data_ptr(),
                 rows.modes.data_ptr(), rows.dT_M.data_ptr(), rows.amp.data_ptr(), rows.pos.data_ptr(), current_stream())
            v = rows.view(k)
            # the support sizes need only the rows: queued before the knots and
            # splines, they reach the host while those run, and the second half
            # (queued after them on this stream) is ready to go when they end
            h = self.lik.begin(v, None, self.dt, amp=rows.amp[:k], dT_M=rows.dT_M[:k])
            v.hlm22_knots(out=v.knots)
            v.build_splines()
        return h, slot, k

    def _complete(self, pend):
        h, slot, k = pend
        rows, st = self.rows[slot], self.down_streams[slot]
        self.lik.complete(h, out=rows.out[:k])
        with torch.cuda.stream(st):
            self.out.index_copy_(0, rows.pos[:k], rows.out[:k])

    def run(self, group=64, block=256, grid_groups=0, events=None):
        """Launch the pass; returns log L [n] in the caller's order (valid on the
        current stream once it returns).  ``events`` (optional) gets
        'traj_end_longest' recorded when the integration ends."""
        cur = torch.cuda.current_stream(self.device)
        for st in (self.integ_stream, *self.down_streams):
            st.wait_stream(cur)
        done = self.done.numpy()
        done[:] = 0   # the previous run's integrator raised every flag, so it writes none of these any more
        b = self.batch
        with torch.cuda.stream(self.integ_stream):
            # additional args and tables from this pass's parameters (set_params)
            b.set_params(self._par[0], self._par[1], self._par[2], self._par[3])
            b.initial_conditions(self.r0)
            b.integrate(rtol=self.rtol, atol=self.rtol, two_pass=True, group=group, block=block, grid_groups=grid_groups,
                        done=self.done)
            ended = torch.cuda.Event()
            ended.record(self.integ_stream)
            if events is not None:
                events["traj_end_longest"].record(self.integ_stream)
        groups = drain_finished(done, len(self.down_streams), self._launch, self._complete, lambda h: h[0]["ready"].query(),
                                ended.query, self.poll_s, max_group=self.max_group)
        self.groups = [self.order[g] for g in groups]
        for st in (self.integ_stream, *self.down_streams):
            cur.wait_stream(st)
        return self.out


class GraphPass:
    """The whole pass for a fixed batch of n binaries captured once as a HIP graph
    and replayed for every new batch of parameters.

    The pass is the one-batch pass with the full-length likelihood: additional
    args and tables from the parameters (``eobhip_params_setup``), circular
    initial state, the two-pass trajectory in one persistent launch, the (2,2)
    knots, the splines, h+ on N samples, the N-point transforms and the sum
    against the data.  Every buffer is allocated when the object is built,
    every input is read from device memory, and nothing in the pass waits for
    the host, so the launches (about 20, the transforms included) are recorded
    once and a pass is one graph launch; ``set_params`` copies the next batch's
    parameters into the graph's input rows beforehand.  (The support-length
    likelihood reads template sizes back to the host between its halves and
    is not captured.)  A pass gives the numbers of the same kernels launched
    one by one."""

    def __init__(self, n, dtilde, N, dt, df=None, max_step=4096, rtol=1e-10, group=64, block=256, f0=200.0, device="cuda"):
        L = lib()
        self.n, self.N, self.dt, self.rtol, self.group, self.block, self.f0, self.device = (int(n), int(N), float(dt), float(rtol),
                                                                                          int(group), int(block), float(f0), device)
        self.F = self.N // 2 + 1
        self.df = 1.0 / (self.N * self.dt) if df is None else float(df)
        self.data = torch.as_tensor(dtilde, dtype=torch.complex128, device=device).contiguous()
        if self.data.numel() != self.F:
            raise ValueError("dtilde must have N/2 + 1 bins")
        n = self.n
        self.stream = torch.cuda.Stream(device=device)
        # inputs: m1, m2, chi1, chi2, r0, dT_M, amp (device rows and their pinned staging)
        self._stage = torch.empty((7, n), dtype=torch.float64, pin_memory=True)
        self._staged = None
        self.par = torch.zeros((7, n), dtype=torch.float64, device=device)
        self.batch = EOBBatch.from_params(np.full(n, 0.6), np.full(n, 0.4), np.zeros(n), np.zeros(n), max_step=max_step,
                                          device=device)
        b = self.batch
        b.knots = torch.zeros((n, b.max_step, 3), dtype=torch.float64, device=device)
        b.coef = torch.zeros((n, b.max_step, 12), dtype=torch.float64, device=device)
        b.spline_scratch = torch.zeros(int(L.eobhip_spline_scratch_doubles(n, b.max_step)), dtype=torch.float64, device=device)
        self.cond = torch.zeros((4, n), dtype=torch.float64, device=device)
        self.nsamp = torch.zeros(n, dtype=torch.int32, device=device)
        self.hp = torch.zeros((n, self.N), dtype=torch.float64, device=device)
        self.ht = torch.zeros((n, self.F), dtype=torch.complex128, device=device)
        self.partial = torch.zeros(int(L.eobhip_fd_likelihood_partial_doubles(n, self.F)), dtype=torch.float64, device=device)
        self.logl = torch.zeros(n, dtype=torch.float64, device=device)
        self.graph = None

    def set_params(self, m1, m2, chi1, chi2, r0, amp=1.0):
        """The next batch: masses (solar masses), aligned spins, r0 [n], amplitude
        scale; copied to the graph's input rows on the current stream."""
        n = self.n
        m1, m2, chi1, chi2, r0, amp = (np.broadcast_to(np.asarray(x, dtype=np.float64), (n,)) for x in (m1, m2, chi1, chi2, r0, amp))
        if self._staged is not None:
            self._staged.synchronize()
        st = self._stage.numpy()
        st[0], st[1], st[2], st[3], st[4] = m1, m2, chi1, chi2, r0
        st[5] = self.dt / ((m1 + m2) * MTSUN_SI)
        st[6] = amp
        self.par.copy_(self._stage, non_blocking=True)
        self._staged = torch.cuda.Event()
        self._staged.record(torch.cuda.current_stream(self.device))

    def _body(self):
        b, n, s, par = self.batch, self.n, current_stream(), self.par
        call("eobhip_params_setup", par[0].data_ptr(), par[1].data_ptr(), par[2].data_ptr(), par[3].data_ptr(), n, 1, 0.0, 0.0, 0.0,
             0.0, 0.0, b.add.data_ptr(), b.bindata.data_ptr(), b.modes.data_ptr(), s)
        call("eobhip_initial_state", par[4].data_ptr(), b.bindata.data_ptr(), b.modes.data_ptr(), n, self.cond.data_ptr(),
             b.add[4].data_ptr(), 1000, 1e-12, s)
        it = ctypes.c_longlong(0)
        call("eobhip_integrate", self.cond.data_ptr(), b.bindata.data_ptr(), b.modes.data_ptr(), n, b.max_step, self.rtol, self.rtol,
             0.01, 10.0, 0.05, 1e13, 1, 0, None, None, 1, 0, self.group, self.block, 0, b.dense.data_ptr(), b.num_steps.data_ptr(),
             b.stop_info.data_ptr(), b.pass1_steps.data_ptr(), b.attempts.data_ptr(), b.scratch.data_ptr(), ctypes.byref(it), s)
        call("eobhip_hlm22_knots", b.dense.data_ptr(), b.num_steps.data_ptr(), b.bindata.data_ptr(), b.modes.data_ptr(), n,
             b.max_step, b.knots.data_ptr(), s)
        call("eobhip_spline_build_batched", b.dense.data_ptr(), b.knots.data_ptr(), b.num_steps.data_ptr(), n, b.max_step,
             b.coef.data_ptr(), b.spline_scratch.data_ptr(), s)
        call("eobhip_resample_hplus", b.dense.data_ptr(), b.coef.data_ptr(), b.num_steps.data_ptr(), par[6].data_ptr(),
             par[5].data_ptr(), n, b.max_step, self.N, self.hp.data_ptr(), self.nsamp.data_ptr(), s)
        call("eobhip_rfft_batched", self.hp.data_ptr(), self.ht.data_ptr(), n, self.N, s)
        call("eobhip_fd_likelihood", self.ht.data_ptr(), self.data.data_ptr(), self.partial.data_ptr(), n, self.F, self.df, self.f0,
             self.logl.data_ptr(), s)

    def run_eager(self):
        """The same launches one by one on the pass's stream (the reference for the graph)."""
        cur = torch.cuda.current_stream(self.device)
        self.stream.wait_stream(cur)
        with torch.cuda.stream(self.stream):
            self._body()
        cur.wait_stream(self.stream)
        return self.logl

    def capture(self):
        """Record the pass (after one eager pass on the same stream, which creates
        the transform plans the pass uses)."""
        self.run_eager()
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=self.stream):
            self._body()
        torch.cuda.synchronize(self.device)
        self.graph = g
        return g

    def run(self):
        """One pass as a graph launch; returns log L [n] (valid on the current stream)."""
        if self.graph is None:
            self.capture()
        cur = torch.cuda.current_stream(self.device)
        self.stream.wait_stream(cur)
        with torch.cuda.stream(self.stream):
            self.graph.replay()
        cur.wait_stream(self.stream)
        return self.logl


def analytic_psd(f, f0=200.0):
    """S(f) = (f/f0)^-4 + 1 + (f/f0)^2 (S(0) = inf)."""
    x = np.asarray(f, dtype=np.float64) / f0
    with np.errstate(divide="ignore"):
        return x ** -4.0 + 1.0 + x * x
