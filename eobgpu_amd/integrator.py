"""Drop-in for the reference's batched DOPR853 driver and its EOB stopping criterion.

``DOPR853.integrate`` keeps the reference's call and return convention
(eob/new_Dormand_horiz.py:219-243, 424-778): ``condBound [4][n]`` and
``argsData [80][n]`` in, ``(denseOutputLoc [max_step][n], denseOutput
[max_step][4][n], step_num [n])`` out, optional restart from row ``step_num``
of caller-provided dense arrays, and ``fix_step`` (step reset to ``hInit``
after every attempt, ``hInit`` halved for good on a reject).

The reference drives one attempted step per Python iteration (about 88 kernel
launches and a host sync each).  Here the whole integration of every binary is
one launch of the persistent integrator kernel (``eobhip_integrate``): each
lane group owns one binary from first step to stop, evaluates the fused EOB
right-hand side (Hamiltonian gradient + 35-mode flux), the DOP853 stages,
error norm and controller, and the stopping criterion in registers, and
writes accepted rows straight to HBM.

Semantics kept from the reference:
  * the first attempted step of a lane that is rejected ends that lane (the
    stop flag starts True, eob/new_Dormand_horiz.py:583, 761);
  * a lane ends when ``x >= tmax``, ``step_num >= max_step - 1`` or the
    stopping criterion returns a non-zero code (:761);
  * stop codes 2/3/4/5 (eob/seobnrv4phm.py:113-133);
  * a restart row index of -1 addresses the last row, as numpy indexing does.

The reference's criterion couples the binaries of a batch: its third check
reads the omega history at a binary's position among the binaries accepted in
the iteration, and while any binary is past step 5 only those store their
omega (eob/seobnrv4phm.py:121-140), so a binary's stop code and step count
depend on its batch mates.  ``DOPR853(batch_coupled=True)`` (the default when a
stopping criterion is given) reproduces that with the coupled driver
(``eobhip_integrate_batch_coupled``: one launch per attempted step plus a
one-block criterion launch); ``batch_coupled=False`` integrates every binary as
a batch of one in the persistent kernel, whose results do not depend on the
batch (the batched pipeline and the server use that form).
"""
from __future__ import annotations

import numpy as np
import torch

from ._lib import lib
from .pipeline import EOBBatch
from .pyEOB import SEOBNRv5Class
from .utils import sync

__all__ = ["DOPR853", "StoppingCriterion"]


def _as_device(a, device, dtype=torch.float64):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=dtype)
    if hasattr(a, "__cuda_array_interface__"):
        return torch.as_tensor(a, device=device).to(dtype)
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=device)


class StoppingCriterion:
    """eob/seobnrv4phm.py:61-150.

    Passing an instance to ``DOPR853`` switches the EOB stop codes on inside the
    integrator kernels, where they are evaluated on the k1 stage of the step
    after each accepted one (that stage is f(x, y_new), the same right-hand
    side the reference evaluates again for the criterion, :94); the coupled
    driver keeps its omega history in ``old_omegas``.  ``__call__`` is the
    reference's decision as a standalone batched call for callers that drive
    their own stepping: it evaluates the right-hand side through
    ``eob_c_class.ODE_Ham_align_AD`` and applies codes 2-5 with the
    reference's indexing of the omega history (third check at the position
    inside ``index_update``; omegas stored only for the binaries past step 5
    while any is, eob/seobnrv4phm.py:113-140).
    """

    def __init__(self, eob_c_class, max_step, use_gpu=False, read_out_to_cpu=False):
        self.eob_c_class = eob_c_class
        self.max_step = int(max_step)
        self.use_gpu = use_gpu
        self.read_out_to_cpu = read_out_to_cpu
        self.old_omegas = None

    def setup(self, numSys):
        self.numSys = int(numSys)
        self.old_omegas = torch.zeros((self.max_step, self.numSys), dtype=torch.float64, device="cuda")

    def reset(self):
        self.old_omegas = None

    def __call__(self, step_num, denseOutput, additionalArgs, index_update):
        if self.read_out_to_cpu:
            raise NotImplementedError("read_out_to_cpu is not supported by the stopping criterion (as in the reference)")
        if self.old_omegas is None:
            raise ValueError("Must use `setup` attribute before using this method.")
        dev = self.old_omegas.device
        y = _as_device(denseOutput, dev).reshape(4, -1).contiguous()
        n_here = y.shape[1]
        add = _as_device(additionalArgs, dev).reshape(-1, n_here).contiguous()
        idx = _as_device(index_update, dev, torch.int64).reshape(-1)
        sn_all = _as_device(step_num, dev, torch.int64).reshape(-1)
        derivs = torch.zeros((4, n_here), dtype=torch.float64, device=dev)
        x = torch.zeros(1, dtype=torch.float64, device=dev)
        self.eob_c_class.ODE_Ham_align_AD(x, y.reshape(-1), derivs.reshape(-1), add.reshape(-1), n_here,
                                          idx.to(torch.int32))
        r, drdt, omega, dprdt = y[0], derivs[0], derivs[1], derivs[2]
        stop = torch.zeros(n_here, dtype=torch.int64, device=dev)
        stop = torch.where(r < 1.0, 2, stop)
        stop = torch.where((stop == 0) & (r < 6.0) & (drdt > 0.0), 3, stop)
        stop = torch.where((stop == 0) & (r < 6.0) & (dprdt > 0.0), 4, stop)
        sn = sn_all[idx]
        here = (stop == 0) & (sn > 5)
        if bool(here.any()):
            still = torch.nonzero(here).reshape(-1)          # positions inside index_update
            s_st = sn[still]
            second = omega[still] < self.old_omegas[s_st - 1, idx[still]]
            rows = s_st[:, None] + torch.arange(-4, 0, device=dev)[None, :]
            hist = self.old_omegas[rows, still[:, None].expand(-1, 4)]   # column = position, as the reference
            third = (hist[:, 1:] - hist[:, :-1] < 0.0).all(dim=1)
            c5 = (r[still] < 6.0) & second & third
            stop[still] = stop[still] + 5 * c5.to(torch.int64)
            self.old_omegas[s_st, idx[still]] = omega[still]
        else:
            self.old_omegas[sn, idx] = omega
        return stop


class DOPR853:
    """eob/new_Dormand_horiz.py:219-243.

    ``ode`` is the ``pyEOB.SEOBNRv5Class`` whose right-hand side is integrated
    (the EOB equations of motion); the kernel derives the Hamiltonian and flux
    tables from ``argsData`` the way SEOBNRv4PHM.setup_class fills both
    (eob/seobnrv4phm.py:1922-1964), so the object need not hold tables.
    Generic right-hand sides (the reference's pendulum harness) are not on
    this path and are rejected.
    """

    def __init__(self, ode, stopping_criterion=None, tmax=1e300, max_step=int(1e6), use_gpu=True, read_out_to_cpu=False,
                 group=64, block=256, batch_coupled=True):
        if ode is not None and not isinstance(ode, SEOBNRv5Class):
            raise TypeError("DOPR853 integrates the EOB equations of motion only: pass a pyEOB.SEOBNRv5Class (or None)")
        if not use_gpu:
            raise ValueError("there is no CPU path: DOPR853 runs on the GPU only")
        lib()  # fail here, not at the first step, if the HIP library is missing
        self.ode = ode
        self.stopping_criterion = stopping_criterion
        self.tmax = float(tmax)
        self.max_step = int(max_step)
        self.use_gpu = True
        self.read_out_to_cpu = bool(read_out_to_cpu)
        self.group = int(group)
        self.block = int(block)
        self.batch_coupled = bool(batch_coupled)
        self.stop_info = None
        self.attempts = None

    def integrate(self, condBound, argsData, hInit=None, step_num=None, denseOutput=None, denseOutputLoc=None,
                  fix_step=False):
        """eob/new_Dormand_horiz.py:424-778; returns (denseOutputLoc, denseOutput, step_num)."""
        dev = torch.device("cuda")
        add = _as_device(argsData, dev)
        n = int(add.shape[-1])
        add = add.reshape(-1, n).contiguous()
        max_step = self.max_step
        if hInit is None:
            hInit = 0.01
        batch = EOBBatch(add, max_step=max_step, device=dev)
        x0 = step0 = None
        if denseOutput is None:
            cond = _as_device(condBound, dev).reshape(4, n).contiguous()
        else:
            dout = _as_device(denseOutput, dev)
            dloc = _as_device(denseOutputLoc, dev)
            if dout.ndim != 3 or tuple(dout.shape) != (max_step, 4, n):
                raise ValueError(f"denseOutput must have shape ({max_step}, 4, {n})")
            if tuple(dloc.shape) != (max_step, n):
                raise ValueError(f"denseOutputLoc must have shape ({max_step}, {n})")
            if step_num is None:
                raise ValueError("step_num is required with denseOutput")
            sn = _as_device(step_num, dev, torch.int64).reshape(-1)
            if sn.numel() != n:
                raise ValueError("step_num must have one entry per binary")
            if bool(((sn < -max_step) | (sn >= max_step)).any()):
                raise IndexError("step_num out of range of denseOutput")
            if bool((sn >= max_step - 1).any()):
                # the reference's first accepted step would index denseOutput[max_step]
                raise IndexError("step_num leaves no row of denseOutput for the next step")
            row = torch.where(sn < 0, sn + max_step, sn)
            ar = torch.arange(n, device=dev)
            x0 = dloc[row, ar].contiguous()
            cond = dout[row, :, ar].T.contiguous()
            # existing rows travel with the run; rows the run accepts are overwritten
            batch.dense[:, :, 0] = dloc.T
            batch.dense[:, :, 1:] = dout.permute(2, 0, 1)
            step0 = sn.to(torch.int32).contiguous()
        batch.cond = cond
        use_stop = self.stopping_criterion is not None
        sc = self.stopping_criterion
        if use_stop and hasattr(sc, "setup"):
            sc.setup(n)
        if use_stop and self.batch_coupled:
            # the reference's batch semantics; the criterion's history array is the driver's
            om = getattr(sc, "old_omegas", None)
            if isinstance(om, torch.Tensor) and om.is_cuda and om.dtype == torch.float64 and tuple(om.shape) == (max_step, n) \
                    and om.is_contiguous():
                batch.old_omegas = om
            batch.integrate(hInit=float(hInit), two_pass=False, fix_step=bool(fix_step), tmax=self.tmax, x0=x0, step0=step0,
                            use_stop=True, mode="coupled", group=self.group)
        else:
            # each binary as a batch of one: the persistent kernel keeps the last
            # four omegas in registers
            batch.integrate(hInit=float(hInit), two_pass=False, fix_step=bool(fix_step), tmax=self.tmax, x0=x0, step0=step0,
                            use_stop=use_stop, group=self.group, block=self.block)
        sync()
        loc = batch.dense[:, :, 0].T.contiguous()
        traj = batch.dense[:, :, 1:].permute(1, 2, 0).contiguous()
        steps = batch.num_steps.to(torch.int64) - 1
        self.stop_info = batch.stop_info
        self.attempts = batch.attempts
        if use_stop and hasattr(self.stopping_criterion, "reset"):
            self.stopping_criterion.reset()
        if self.read_out_to_cpu:
            return loc.cpu().numpy(), traj.cpu().numpy(), steps.cpu().numpy()
        return loc, traj, steps
