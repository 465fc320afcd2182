"""PyTorch autograd bridge of the finite-source light curves.

    from lcbinint_amd.torch_ops import FiniteSourceLightCurves
    lc = FiniteSourceLightCurves(offsets, times, nbin=None, options=None)
    mag = lc(params)          # params: (n_curves, 8) float64 on cuda:
                              # t0, tE, u0, alpha, s, q, rho, u_LD per curve
    loss(mag).backward()      # params.grad via the device derivatives

The forward pass is the plain finite-source method (the reference's values);
the backward pass contracts the incoming gradient with the per-epoch
derivatives d mag / d params of lcbg_batch_finite_source_grad_device and sums
them per curve on the device.
"""
from __future__ import annotations

import numpy as np
import torch

from . import capi

PARAM_NAMES = capi.GRAD_PARAMS  # t0, tE, u0, alpha, s, q, rho, ld_c


def _curves_from(params: torch.Tensor):
    host = params.detach().to("cpu", torch.float64).numpy()
    return [capi.Curve(t0=r[0], tE=r[1], u0=r[2], alpha=r[3], s=r[4], q=r[5], rho=r[6],
                       ld_c=r[7], ld_d=0.0) for r in host]


class _LightCurveFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, owner):
        n = owner.n_epochs
        dev = params.device
        stream = torch.cuda.current_stream(dev).cuda_stream
        # the curve upload and the run share the caller's stream, so that the
        # kernels read the new curves under any torch.cuda.stream context
        owner.batch.update_curves(_curves_from(params), stream)
        mag = torch.empty(n, dtype=torch.float64, device=dev)
        method = torch.empty(n, dtype=torch.int32, device=dev)
        dmag = torch.empty((n, 8), dtype=torch.float64, device=dev)
        capi.finite_source_grad(owner.batch, owner.d_off.data_ptr(), owner.d_t.data_ptr(), n,
                                mag.data_ptr(), method.data_ptr(), 0, dmag.data_ptr(), stream)
        ctx.save_for_backward(dmag)
        ctx.owner = owner
        owner.last_method = method
        return mag

    @staticmethod
    def backward(ctx, grad_mag):
        (dmag,) = ctx.saved_tensors
        owner = ctx.owner
        contrib = grad_mag.to(torch.float64)[:, None] * dmag
        grad = torch.zeros((owner.n_curves, 8), dtype=torch.float64, device=dmag.device)
        grad.index_add_(0, owner.curve_of, contrib)
        return grad, None


class FiniteSourceLightCurves:
    """Light curves of n_curves curves with per-curve epochs
    times[offsets[c]:offsets[c+1]] as a differentiable function of the
    (n_curves, 8) parameter tensor (VBM convention, as LightCurve)."""

    def __init__(self, offsets, times, nbin=None, options: capi.Options | None = None,
                 device: str = "cuda"):
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        times = np.ascontiguousarray(times, dtype=np.float64)
        self.n_curves = len(offsets) - 1
        self.n_epochs = int(offsets[-1])
        if len(times) != self.n_epochs:
            raise ValueError(f"times has {len(times)} entries, offsets end at {self.n_epochs}")
        placeholder = [capi.Curve(0.0, 1.0, 0.1, 0.0, 1.0, 1.0, 1e-3, 0.0, 0.0)] * self.n_curves
        self.batch = capi.Batch(placeholder, options)
        if nbin is not None:
            self.batch.set_source_bins(np.broadcast_to(np.asarray(nbin, dtype=np.int32),
                                                       (self.n_curves,)))
        self.d_off = torch.from_numpy(offsets).to(device)
        self.d_t = torch.from_numpy(times).to(device)
        self.curve_of = torch.repeat_interleave(
            torch.arange(self.n_curves, device=device), torch.from_numpy(np.diff(offsets)).to(device))
        self.last_method = None
        self._dev_epochs = None

    def device_epochs(self, dev):
        """The offsets and times as device tensors (uploaded once per device)."""
        if self._dev_epochs is None or self._dev_epochs[0].device != dev:
            self._dev_epochs = (torch.from_numpy(self.offsets).to(dev), torch.from_numpy(self.times).to(dev))
        return self._dev_epochs

    def __call__(self, params: torch.Tensor) -> torch.Tensor:
        if params.shape != (self.n_curves, 8):
            raise ValueError(f"params must be ({self.n_curves}, 8), got {tuple(params.shape)}")
        return _LightCurveFn.apply(params, self)

    def close(self):
        self.batch.close()


DYNAMIC_PARAM_NAMES = capi.DYN_GRAD_PARAMS  # t0, tE, u0, alpha, s, q, rho, ld_c, piEN, piEE, g1, g2


class _DynamicLightCurveFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, owner):
        host = params.detach().to("cpu", torch.float64).numpy()
        curves = [capi.Curve(t0=r[0], tE=r[1], u0=r[2], alpha=r[3], s=r[4], q=r[5], rho=r[6], ld_c=r[7],
                             ld_d=0.0) for r in host]
        dyn = [capi.Dynamics(piEN=r[8], piEE=r[9], g1=r[10], g2=r[11], **fixed)
               for r, fixed in zip(host, owner.fixed)]
        dev = params.device if params.is_cuda else torch.device("cuda")
        stream = torch.cuda.current_stream(dev).cuda_stream
        owner.batch.update_curves(curves, stream)
        d_off, d_t = owner.device_epochs(dev)
        n = len(owner.times)
        mag = torch.empty(n, dtype=torch.float64, device=dev)
        method = torch.empty(n, dtype=torch.int32, device=dev)
        d = torch.empty((n, len(DYNAMIC_PARAM_NAMES)), dtype=torch.float64, device=dev)
        dyn_arr = (capi.Dynamics * len(dyn))(*dyn)
        # trajectories, magnifications and derivatives on the device, in order
        # on the caller's stream
        capi.check(capi.lcbg_batch_light_curves_dynamic_device(
            owner.batch.handle, dyn_arr, owner.orbit_mode, d_off.data_ptr(), d_t.data_ptr(), n,
            mag.data_ptr(), method.data_ptr(), None, d.data_ptr(), stream),
            "lcbg_batch_light_curves_dynamic_device")
        if not bool(torch.isfinite(mag).all()):
            # a non-finite magnification is a numerical error (lens_model.cpp:228-231)
            raise capi.LcbgError(2, "lcbg_batch_light_curves_dynamic_device (non-finite magnification)")
        owner.last_method = method.cpu().numpy()
        ctx.save_for_backward(d)
        ctx.owner = owner
        return mag.to(params.device)

    @staticmethod
    def backward(ctx, grad_mag):
        (dmag,) = ctx.saved_tensors
        owner = ctx.owner
        contrib = grad_mag.to(torch.float64)[:, None] * dmag
        grad = torch.zeros((owner.n_curves, len(DYNAMIC_PARAM_NAMES)), dtype=torch.float64,
                           device=dmag.device)
        grad.index_add_(0, owner.curve_of.to(dmag.device), contrib)
        return grad, None


class DynamicLightCurves:
    """Light curves with annual parallax and/or orbital motion as a
    differentiable function of the (n_curves, 12) parameter tensor (columns
    DYNAMIC_PARAM_NAMES: t0, tE, u0, alpha, s, q, rho, u_LD, piEN, piEE, g1, g2).
    `fixed` holds per curve the parameters that are not differentiated (ra,
    dec, tfix, g3, lom_szs, lom_ar).  Values and derivatives come from
    lcbg_batch_light_curves_dynamic_device (trajectories, magnifications and
    their derivatives on the device, on the caller's current stream); parallax
    needs the Earth table (capi.load_earth_ephemeris)."""

    def __init__(self, offsets, times, fixed, orbit_mode=capi.ORBIT_CIRCULAR, nbin=None,
                 options: capi.Options | None = None):
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.times = np.ascontiguousarray(times, dtype=np.float64)
        self.n_curves = len(self.offsets) - 1
        if len(fixed) != self.n_curves:
            raise ValueError(f"fixed needs {self.n_curves} entries, got {len(fixed)}")
        keys = ("ra", "dec", "tfix", "g3", "lom_szs", "lom_ar")
        self.fixed = [{k: float(f.get(k, 1.0 if k == "lom_ar" else 0.0)) for k in keys} for f in fixed]
        self.orbit_mode = int(orbit_mode)
        placeholder = [capi.Curve(0.0, 1.0, 0.1, 0.0, 1.0, 1.0, 1e-3, 0.0, 0.0)] * self.n_curves
        self.batch = capi.Batch(placeholder, options)
        if nbin is not None:
            self.batch.set_source_bins(np.broadcast_to(np.asarray(nbin, dtype=np.int32), (self.n_curves,)))
        self.curve_of = torch.repeat_interleave(torch.arange(self.n_curves),
                                                torch.from_numpy(np.diff(self.offsets)))
        self.last_method = None
        self._dev_epochs = None

    def device_epochs(self, dev):
        """The offsets and times as device tensors (uploaded once per device)."""
        if self._dev_epochs is None or self._dev_epochs[0].device != dev:
            self._dev_epochs = (torch.from_numpy(self.offsets).to(dev), torch.from_numpy(self.times).to(dev))
        return self._dev_epochs

    def __call__(self, params: torch.Tensor) -> torch.Tensor:
        if params.shape != (self.n_curves, len(DYNAMIC_PARAM_NAMES)):
            raise ValueError(f"params must be ({self.n_curves}, {len(DYNAMIC_PARAM_NAMES)}), "
                             f"got {tuple(params.shape)}")
        return _DynamicLightCurveFn.apply(params, self)

    def close(self):
        self.batch.close()
