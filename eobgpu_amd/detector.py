"""Detector projection of the geocentre polarizations (the step between the
template and the likelihood for interferometer data).

Reference: BBHWaveformTD.get_detector_information (eob/seobnrv4phm.py:597-634)
and the "detector_fd" branch of BBHWaveformTD.__call__ (:742-767).  The
reference reads the detector tensors and vertices from bilby's interferometer
definitions and the Greenwich mean sidereal time from lal; here the caller
passes them (``tensors`` [nifo][3][3], ``vertices`` [nifo][3] in metres,
``gmst`` [n] in radians).
"""
from __future__ import annotations

import numpy as np
import torch

from ._lib import call
from .utils import current_stream

__all__ = ["get_detector_information", "detector_signal", "resample_hpc"]


def _dev(x, n, device):
    if isinstance(x, torch.Tensor) and x.is_cuda:
        return x.to(dtype=torch.float64).reshape(-1).expand(n).contiguous() if x.numel() == 1 else x.to(dtype=torch.float64).contiguous()
    return torch.as_tensor(np.broadcast_to(np.asarray(x, dtype=np.float64), (n,)).copy(), device=device)


def get_detector_information(ra, dec, gmst, psi, tensors, vertices, device="cuda"):
    """(Fplus, Fcross, time_shift), each [nifo][n], for n sky positions and
    polarization angles (eob/seobnrv4phm.py:597-634)."""
    n = int(np.size(ra)) if not isinstance(ra, torch.Tensor) else int(ra.numel())
    D = torch.as_tensor(np.ascontiguousarray(tensors, dtype=np.float64) if not isinstance(tensors, torch.Tensor) else tensors,
                        dtype=torch.float64, device=device).reshape(-1, 9).contiguous()
    V = torch.as_tensor(np.ascontiguousarray(vertices, dtype=np.float64) if not isinstance(vertices, torch.Tensor) else vertices,
                        dtype=torch.float64, device=device).reshape(-1, 3).contiguous()
    nifo = int(D.shape[0])
    if V.shape[0] != nifo:
        raise ValueError("one vertex per detector tensor")
    ra_t, dec_t, g_t, psi_t = (_dev(x, n, device) for x in (ra, dec, gmst, psi))
    Fp = torch.empty((nifo, n), dtype=torch.float64, device=device)
    Fc = torch.empty_like(Fp)
    dt = torch.empty_like(Fp)
    call("eobhip_detector_response", ra_t.data_ptr(), dec_t.data_ptr(), g_t.data_ptr(), psi_t.data_ptr(), n, D.data_ptr(), V.data_ptr(),
         nifo, Fp.data_ptr(), Fc.data_ptr(), dt.data_ptr(), current_stream())
    return Fp, Fc, dt


def detector_signal(hp_fd, hc_fd, Fplus, Fcross, time_shift, df, k_min=0, out=None):
    """signal [n][nifo][F] = (F+ hp~ + Fx hc~) exp(-2 pi i f dt), zero below bin k_min
    (eob/seobnrv4phm.py:742-767)."""
    n, F = hp_fd.shape
    nifo = int(Fplus.shape[0])
    if out is None:
        out = torch.empty((n, nifo, F), dtype=torch.complex128, device=hp_fd.device)
    call("eobhip_detector_signal", hp_fd.data_ptr(), hc_fd.data_ptr(), Fplus.data_ptr(), Fcross.data_ptr(), time_shift.data_ptr(), int(n),
         nifo, int(F), float(df), int(k_min), out.data_ptr(), current_stream())
    return out


def resample_hpc(batch, total_mass_msun, dt, N, amp=None):
    """h+ and hx [n][N] of a batch whose splines are built: the real and imaginary
    parts of amp h22(t_j), the two series the reference transforms
    (template.real, template.imag, eob/seobnrv4phm.py:742-744)."""
    from .pipeline import MTSUN_SI
    dev = batch.device
    n = batch.n
    M = _dev(total_mass_msun, n, dev)
    dT_M = (dt / (M * MTSUN_SI)).contiguous()
    a = torch.ones(n, dtype=torch.float64, device=dev) if amp is None else _dev(amp, n, dev)
    hp = torch.empty((n, N), dtype=torch.float64, device=dev)
    hc = torch.empty_like(hp)
    nsamp = torch.zeros(n, dtype=torch.int32, device=dev)
    call("eobhip_resample_hpc", batch.dense.data_ptr(), batch.coef.data_ptr(), batch.num_steps.data_ptr(), a.data_ptr(), dT_M.data_ptr(),
         n, batch.max_step, int(N), hp.data_ptr(), hc.data_ptr(), nsamp.data_ptr(), current_stream())
    return hp, hc
