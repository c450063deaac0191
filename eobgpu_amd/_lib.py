"""ctypes binding of libeobhip.so (the HIP kernels behind the C ABI in include/eobhip.h).

The library is the only compute path: there is no CPU fallback.  If it is
missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libeobhip.so")

_lock = threading.Lock()
_lib = None

c_int = ctypes.c_int
c_double = ctypes.c_double
c_vp = ctypes.c_void_p
c_size = ctypes.c_size_t
c_ll_p = ctypes.POINTER(ctypes.c_longlong)

# name -> argtypes (restype int unless listed in _RESTYPES)
_SIGS = {
    "eobhip_last_error": [],
    "eobhip_version": [],
    "eobhip_dop853_tableau": [c_vp],
    "eobhip_spin": [ctypes.c_longlong, c_vp],
    "eobhip_gather_binaries": [c_vp, c_int, c_int, c_int, c_vp, c_vp, c_vp, ctypes.c_size_t, c_vp, ctypes.c_size_t, c_vp, c_vp, c_vp,
                               c_double, c_double, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "eobhip_evaluate_Ham_align_AD": [c_vp] * 9 + [c_double] * 4 + [c_int, c_vp],
    "eobhip_grad_Ham_align_AD": [c_vp, c_vp, c_vp, c_int, c_vp],
    "eobhip_hessian_Ham_align_AD": [c_vp, c_vp, c_vp, c_int, c_vp],
    "eobhip_IC_cons": [c_vp] * 5 + [c_int, c_vp],
    "eobhip_root_find_all": [c_vp, c_vp, c_vp, c_vp, c_int, c_double, c_int, c_int, c_int, c_int, c_vp],
    "eobhip_seobnr_IC_diss": [c_vp] * 9 + [c_int, c_vp],
    "eobhip_seobnr_root_find_scalar_all": [c_vp] * 5 + [c_int, c_double, c_int, c_int, c_int, c_vp],
    "eobhip_seobnr_create": [c_vp],
    "eobhip_seobnr_destroy": [c_vp],
    "eobhip_seobnr_update_information": [c_vp] * 7 + [c_int, c_int, c_vp, c_vp],
    "eobhip_seobnr_deallocate_information": [c_vp],
    "eobhip_seobnr_RR_force": [c_vp] * 5 + [c_int, c_vp, c_vp],
    "eobhip_seobnr_ODE_Ham_align_AD": [c_vp] * 5 + [c_int, c_vp, c_vp],
    "eobhip_seobnr_EOBGetSpinFactorizedWaveform": [c_vp] * 10 + [c_int, c_int, c_vp, c_double, c_int, c_int, c_vp, c_vp],
    "eobhip_dormandPrinceSteps": [c_vp] * 17 + [c_int, c_int, c_int, c_vp, c_vp, c_vp],
    "eobhip_error": [c_vp] * 12 + [c_int, c_int, c_vp],
    "eobhip_error_tol": [c_vp] * 12 + [c_int, c_int, c_double, c_double, c_vp],
    "eobhip_controllerSuccess": [c_vp] * 6 + [c_int, c_int, c_vp],
    "eobhip_bindata_bytes": [],
    "eobhip_modes_bytes": [],
    "eobhip_bin_setup": [c_vp, c_int, c_int, c_vp, c_vp, c_vp],
    "eobhip_params_setup": [c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_double, c_double, c_double, c_double, c_double, c_vp, c_vp,
                            c_vp, c_vp],
    "eobhip_newton_prefix_constants": [c_vp],
    "eobhip_initial_conditions": [c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_int, c_double, c_vp],
    "eobhip_initial_state": [c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_int, c_double, c_vp],
    "eobhip_integrate_scratch_bytes": [c_int],
    "eobhip_integrate": [c_vp, c_vp, c_vp, c_int, c_int, c_double, c_double, c_double, c_double, c_double, c_double,
                         c_int, c_int, c_vp, c_vp, c_int, c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                         c_ll_p, c_vp],
    "eobhip_integrate_signal": [c_vp, c_vp, c_vp, c_int, c_int, c_double, c_double, c_double, c_double, c_double, c_double,
                                c_int, c_int, c_vp, c_vp, c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                c_vp, c_vp],
    "eobhip_integrate_batch_coupled": [c_vp, c_vp, c_vp, c_int, c_int, c_double, c_double, c_double, c_double, c_double, c_double,
                                       c_int, c_int, c_vp, c_vp, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_ll_p,
                                       c_vp],
    "eobhip_hlm22_knots": [c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_vp, c_vp],
    "eobhip_wave_mode_bytes": [],
    "eobhip_wave_tables": [c_vp, c_int, c_int, c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp],
    "eobhip_hlm_modes_knots": [c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_int, c_vp, c_vp],
}
_RESTYPES = {
    "eobhip_last_error": ctypes.c_char_p,
    "eobhip_bindata_bytes": c_size,
    "eobhip_modes_bytes": c_size,
    "eobhip_integrate_scratch_bytes": c_size,
    "eobhip_wave_mode_bytes": c_size,
}
# optional groups: present once the corresponding sources are built
_SIGS.update({
    "eobhip_interpolate_TD": [c_vp] * 7 + [c_int, c_int, c_vp],
    "eobhip_all_in_one_likelihood": [c_vp] * 8 + [c_double, c_int, c_int, c_int, c_vp],
    "eobhip_spline_build_batched": [c_vp, c_vp, c_vp, c_int, c_int, c_vp, c_vp, c_vp],
    "eobhip_spline_scratch_doubles": [c_int, c_int],
    "eobhip_resample_hplus": [c_vp] * 5 + [c_int, c_int, c_int, c_vp, c_vp, c_vp],
    "eobhip_resample_hpc": [c_vp] * 5 + [c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp],
    "eobhip_spline_series_scratch_doubles": [c_int, c_int, c_int],
    "eobhip_spline_build_series": [c_vp, c_vp, c_vp, c_int, c_int, c_int, c_vp, c_vp, c_vp],
    "eobhip_resample_modes": [c_vp] * 5 + [c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp],
    "eobhip_TDInterp2": [c_vp] * 7 + [c_int, c_vp, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_vp],
    "eobhip_detector_response": [c_vp] * 4 + [c_int, c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp],
    "eobhip_detector_signal": [c_vp] * 5 + [c_int, c_int, c_int, c_double, c_int, c_vp, c_vp],
    "eobhip_rfft_batched": [c_vp, c_vp, c_int, c_int, c_vp],
    "eobhip_fd_likelihood": [c_vp, c_vp, c_vp, c_int, c_int, c_double, c_double, c_vp, c_vp],
    "eobhip_fd_likelihood_partial_doubles": [c_int, c_int],
    "eobhip_fd_likelihood_static_max_partials": [],
    "eobhip_fd_likelihood_static": [c_vp, c_vp, c_int, c_int, c_double, c_double, c_vp, c_vp],
})
_RESTYPES["eobhip_fd_likelihood_static_max_partials"] = c_size
_SIGS.update({
    "eobhip_imr_fit": [c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_int, c_vp, c_vp],
    "eobhip_kerr_qnm": [c_int, c_int, c_vp, c_int, c_vp, c_vp],
    "eobhip_imr_params": [c_vp] * 9 + [c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "eobhip_imr_resample": [c_vp] * 6 + [c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "eobhip_imr_rd_doubles": [],
    "eobhip_imr_bin_doubles": [],
    "eobhip_imr_bin_ints": [],
})
for _k in ("eobhip_imr_rd_doubles", "eobhip_imr_bin_doubles", "eobhip_imr_bin_ints"):
    _RESTYPES[_k] = c_size
_RESTYPES["eobhip_fd_likelihood_partial_doubles"] = c_size
_RESTYPES["eobhip_spline_scratch_doubles"] = c_size
_RESTYPES["eobhip_spline_series_scratch_doubles"] = c_size
_SIGS.update({
    "eobhip_lik_bank_doubles": [c_int, c_int],
    "eobhip_lik_setup_scratch_bytes": [c_int],
    "eobhip_lik_setup": [c_vp, c_int, c_double, c_double, c_int, c_vp, c_vp, c_vp, c_vp, c_vp],
    "eobhip_lik_sizes": [c_vp, c_vp, c_vp, c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp],
    "eobhip_lik_partial_blocks": [c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int)],
    "eobhip_lik_resample": [c_vp] * 8 + [c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp],
    "eobhip_lik_finish": [c_vp, c_vp, c_vp, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_double, c_vp, c_vp, c_vp],
})
_RESTYPES["eobhip_lik_bank_doubles"] = c_size
_RESTYPES["eobhip_lik_setup_scratch_bytes"] = c_size
_OPTIONAL = {}


class EobHipError(RuntimeError):
    pass


def lib():
    """The loaded library (raises EobHipError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise EobHipError(
                f"{LIB_PATH} is missing: build it with `python -m eobgpu_amd.build` "
                "(there is no CPU fallback for the HIP path)")
        L = ctypes.CDLL(LIB_PATH)
        for name, argt in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = argt
            f.restype = _RESTYPES.get(name, c_int)
        for name, argt in _OPTIONAL.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.argtypes = argt
                f.restype = _RESTYPES.get(name, c_int)
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS) + [n for n in _OPTIONAL if hasattr(lib(), n)]


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().eobhip_last_error().decode(errors="replace")
        raise EobHipError(f"{what or 'eobhip'} failed (code {rc}): {msg}")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)
