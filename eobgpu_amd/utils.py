"""Argument plumbing for the drop-in modules.

``pointer_adjust`` plays the role of ``bbhx.utils.utility.pointer_adjust`` that
the reference's Cython modules use (src/eob.pyx:4, src/dopr853.pyx, src/tdinterp.pyx):
array arguments become integer device addresses, scalars pass through.  Device
arrays are torch tensors on the GPU or anything exposing
``__cuda_array_interface__``; raw integers are taken as addresses.  Host
(numpy) arrays are rejected: every entry point runs on the GPU.
"""
from __future__ import annotations

import functools

import numpy as np
import torch


def device_ptr(a):
    if a is None:
        return None
    if isinstance(a, torch.Tensor):
        if not a.is_cuda:
            raise TypeError("host tensor passed to a device entry point; move it to the GPU first")
        if not a.is_contiguous():
            raise ValueError("non-contiguous tensor passed to a device entry point")
        return a.data_ptr()
    if hasattr(a, "__cuda_array_interface__"):
        return int(a.__cuda_array_interface__["data"][0])
    if isinstance(a, np.ndarray):
        raise TypeError("numpy (host) array passed to a device entry point; use a GPU tensor")
    if isinstance(a, (int, np.integer)):
        return int(a)
    if hasattr(a, "conceiled_ptr"):
        return int(a.conceiled_ptr)
    raise TypeError(f"cannot take a device address of {type(a)!r}")


def _is_array(a):
    return isinstance(a, torch.Tensor) or hasattr(a, "__cuda_array_interface__") or isinstance(a, np.ndarray)


def pointer_adjust(func):
    """Replace array arguments by device addresses before calling ``func``."""

    @functools.wraps(func)
    def wrapper(*args, **kwargs):
        args = [device_ptr(a) if _is_array(a) else a for a in args]
        kwargs = {k: (device_ptr(v) if _is_array(v) else v) for k, v in kwargs.items()}
        return func(*args, **kwargs)

    return wrapper


def current_stream():
    return torch.cuda.current_stream().cuda_stream


def sync():
    torch.cuda.current_stream().synchronize()
