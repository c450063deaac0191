"""Drop-in for the reference's ``eob.pyEOB`` Cython module (src/eob.pyx).

Same function names and positional arguments; array arguments are GPU arrays
(torch tensors or objects with ``__cuda_array_interface__``) or raw device
addresses.  Calls are synchronous, as the reference's wraps are
(cudaDeviceSynchronize after each launch).  Errors raise ``EobHipError``.
"""
from __future__ import annotations

import ctypes

from ._lib import call, lib
from .utils import current_stream, pointer_adjust, sync

__all__ = ["grad_Ham_align_AD", "hessian_Ham_align_AD", "evaluate_Ham_align_AD", "IC_cons", "root_find_all", "SEOBNRv5Class"]


@pointer_adjust
def grad_Ham_align_AD(arg, grad_out, additionalArgs, numSys):
    """src/eob.pyx:72-79 -> grad_Ham_align_AD_wrap (include/EOB.hh:8)."""
    call("eobhip_grad_Ham_align_AD", arg, grad_out, additionalArgs, int(numSys), current_stream())
    sync()


@pointer_adjust
def hessian_Ham_align_AD(arg, hessian_out, additionalArgs, numSys):
    """src/eob.pyx:81-89 -> hessian_Ham_align_AD_wrap (include/EOB.hh:9)."""
    call("eobhip_hessian_Ham_align_AD", arg, hessian_out, additionalArgs, int(numSys), current_stream())
    sync()


@pointer_adjust
def evaluate_Ham_align_AD(out, r, phi, pr, pphi, m_1, m_2, chi_1, chi_2, K, d5, dSO, dSS, numSys):
    """src/eob.pyx:92-102 -> evaluate_Ham_align_AD_wrap (include/EOB.hh:11).  K..dSS are scalars."""
    call("eobhip_evaluate_Ham_align_AD", out, r, phi, pr, pphi, m_1, m_2, chi_1, chi_2, float(K), float(d5), float(dSO),
         float(dSS), int(numSys), current_stream())
    sync()


@pointer_adjust
def IC_cons(res, x, args, additionalArgs, grad_out, numSys):
    """src/eob.pyx:94-101 -> IC_cons_wrap (include/EOB.hh:13)."""
    call("eobhip_IC_cons", res, x, args, additionalArgs, grad_out, int(numSys), current_stream())
    sync()


@pointer_adjust
def root_find_all(xOut, x0In, argsIn, additionalArgsIn, max_iter, err, numBinAll, n, num_args, num_add_args):
    """src/eob.pyx:51-58 -> root_find_all_wrap (include/EOB.hh:6): (r, pphi) [2][numBinAll] from
    the discrete Newton iteration on IC_cons, one binary per thread."""
    call("eobhip_root_find_all", xOut, x0In, argsIn, additionalArgsIn, int(max_iter), float(err), int(numBinAll), int(n),
         int(num_args), int(num_add_args), current_stream())
    sync()


class SEOBNRv5Class:
    """src/eob.pyx:106-256: owns per-binary waveform-coefficient tables on the GPU."""

    def __init__(self):
        h = ctypes.c_void_p()
        call("eobhip_seobnr_create", ctypes.byref(h))
        self._h = h
        self._prefix_ref = None

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and self._h.value:
                lib().eobhip_seobnr_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def conceiled_ptr(self):
        """Address of the native handle (src/eob.pyx:254-256)."""
        return int(self._h.value)

    def update_information(self, m1, m2, eta, tplspin, chi_S, chi_A, use_hm, numSys, newtonian_prefixes_hlms):
        self._prefix_ref = newtonian_prefixes_hlms   # the handle keeps the pointer, as the reference does
        self._update(m1, m2, eta, tplspin, chi_S, chi_A, int(use_hm), int(numSys), newtonian_prefixes_hlms)

    @pointer_adjust
    def _update(self, m1, m2, eta, tplspin, chi_S, chi_A, use_hm, numSys, pref):
        call("eobhip_seobnr_update_information", self._h, m1, m2, eta, tplspin, chi_S, chi_A, use_hm, numSys, pref,
             current_stream())
        sync()

    def deallocate_information(self):
        call("eobhip_seobnr_deallocate_information", self._h)

    @pointer_adjust
    def RR_force(self, force_out, grad_out, args, additionalArgs, numSys, hCoeffs_index):
        call("eobhip_seobnr_RR_force", self._h, force_out, grad_out, args, additionalArgs, int(numSys), hCoeffs_index,
             current_stream())
        sync()

    @pointer_adjust
    def ODE_Ham_align_AD(self, x, arg, k, additionalArgs, numSys, hCoeffs_index):
        call("eobhip_seobnr_ODE_Ham_align_AD", self._h, x, arg, k, additionalArgs, int(numSys), hCoeffs_index, current_stream())
        sync()

    @pointer_adjust
    def EOBGetSpinFactorizedWaveform(self, out, r, phi, pr, pphi, v, Hreal, eta, vPhiInput, l, m, newtonian_prefix, h22_calib,
                                     numSys, traj_length, hCoeffs_index):
        call("eobhip_seobnr_EOBGetSpinFactorizedWaveform", self._h, out, r, phi, pr, pphi, v, Hreal, eta, vPhiInput, int(l), int(m),
             newtonian_prefix, float(h22_calib), int(numSys), int(traj_length), hCoeffs_index, current_stream())
        sync()

    @pointer_adjust
    def IC_diss(self, out, pr, args, additionalArgs, grad_out, grad_temp_force, hess_out, force_out, numSys):
        """src/eob.pyx:133-144 -> IC_diss_wrap (include/EOB.hh:381)."""
        call("eobhip_seobnr_IC_diss", self._h, out, pr, args, additionalArgs, grad_out, grad_temp_force, hess_out, force_out,
             int(numSys), current_stream())
        sync()

    @pointer_adjust
    def root_find_scalar_all(self, pr_res, start_bounds, argsIn, additionalArgsIn, max_iter, err, numBinAll, num_args, num_add_args):
        """src/eob.pyx:169-176 -> root_find_scalar_all_wrap (include/EOB.hh:379)."""
        call("eobhip_seobnr_root_find_scalar_all", self._h, pr_res, start_bounds, argsIn, additionalArgsIn, int(max_iter),
             float(err), int(numBinAll), int(num_args), int(num_add_args), current_stream())
        sync()

    def compute_hlms(self, *args, **kwargs):
        raise NotImplementedError("compute_hlms is dead code in the reference (src/EOB.cu:3319) and is not provided")
