"""Drop-in for the reference's ``eob.pydopr853`` Cython module (src/dopr853.pyx).

``ode_class`` is an ``eobgpu_amd.pyEOB.SEOBNRv5Class`` (anything exposing
``conceiled_ptr`` to a native handle).  Flags are 1-byte bool arrays.
"""
from __future__ import annotations

from ._lib import call
from .utils import current_stream, device_ptr, pointer_adjust, sync

__all__ = ["dormandPrinceSteps", "error", "controllerSuccess"]


@pointer_adjust
def _steps(x, xCurrent, solOld, h, arg, additionalArgs, k1, k2, k3, k4, k5, k6, k7, k8, k9, k10, ak_term_buffer, nargs, numEq,
           num_add_args, handle, hCoeffs_index):
    call("eobhip_dormandPrinceSteps", x, xCurrent, solOld, h, arg, additionalArgs, k1, k2, k3, k4, k5, k6, k7, k8, k9, k10,
         ak_term_buffer, int(nargs), int(numEq), int(num_add_args), handle, hCoeffs_index, current_stream())


def dormandPrinceSteps(x, xCurrent, solOld, h, arg, additionalArgs, k1, k2, k3, k4, k5, k6, k7, k8, k9, k10, ak_term_buffer,
                       nargs, numEq, num_add_args, ode_class, hCoeffs_index):
    """src/dopr853.pyx:66-137 -> dormandPrinceSteps_wrap (src/DOPR853.cu:357-517)."""
    _steps(x, xCurrent, solOld, h, arg, additionalArgs, k1, k2, k3, k4, k5, k6, k7, k8, k9, k10, ak_term_buffer, nargs, numEq,
           num_add_args, int(ode_class.conceiled_ptr), hCoeffs_index)
    sync()


@pointer_adjust
def error(err, solOld, solNew, h, k1, k2, k3, k6, k7, k8, k9, k10, numEq, nargs):
    """src/dopr853.pyx:139-184 -> error_wrap (src/DOPR853.cu:539-618); 13th argument numEq, 14th nargs."""
    call("eobhip_error", err, solOld, solNew, h, k1, k2, k3, k6, k7, k8, k9, k10, int(numEq), int(nargs), current_stream())
    sync()


@pointer_adjust
def controllerSuccess(flagSuccess, err, errOld, previousReject, h, x, numEq, nargs):
    """src/dopr853.pyx:186-214 -> controllerSuccess_wrap (src/DOPR853.cu:675-762)."""
    call("eobhip_controllerSuccess", flagSuccess, err, errOld, previousReject, h, x, int(numEq), int(nargs), current_stream())
    sync()
