"""Differentiable batched Poisson NLL for PSF calibration (no reference counterpart).

``PoissonNLL.apply(estimator, params, shared, samples)`` returns the per-spot
negative log-likelihood sum_px mu' - x' log mu' (the negated
ComputeLogLikelihood of the reference, Estimation.h:138-156) of the
estimator's PSF model.  ``params`` [N,K] are per-spot parameters, ``shared``
the PSF parameters shared by all spots (``Estimator.shared_parameters()``):
sigma_x/sigma_y (fixed-sigma Gaussian), {s0, gamma, d, A, B} per axis
(astigmatic Gaussian) or the [Nz,Ny,Nx,64] coefficient table (cubic spline).
The backward pass is a hand-written HIP adjoint kernel that recomputes the
expected images and Jacobians from the parameters instead of storing them.
"""
from __future__ import annotations

import torch


class PoissonNLL(torch.autograd.Function):
    @staticmethod
    def forward(ctx, estimator, params, shared, samples):
        n = params.shape[0]
        K = estimator.numparams
        if params.dtype != torch.float32 or not params.is_cuda:
            raise ValueError("params must be a float32 HIP tensor")
        params = params.contiguous()
        samples = samples.contiguous()
        if tuple(samples.shape) != (n, *estimator.sampleshape):
            raise ValueError(f"samples must have shape {(n, *estimator.sampleshape)}")
        if shared is not None:
            shared = shared.contiguous()
            if shared.numel() != estimator.num_shared():
                raise ValueError(f"shared must hold {estimator.num_shared()} values")
        nll = torch.empty(n, dtype=torch.float32, device=params.device)
        estimator._nll_grad(params, samples, shared, None, nll, None, None)
        ctx.estimator = estimator
        ctx.save_for_backward(params, shared if shared is not None else torch.empty(0, device=params.device), samples)
        ctx.has_shared = shared is not None
        return nll

    @staticmethod
    def backward(ctx, gnll):
        params, shared, samples = ctx.saved_tensors
        est = ctx.estimator
        shared = shared if ctx.has_shared else None
        gnll = gnll.contiguous().float()
        gp = torch.empty_like(params) if ctx.needs_input_grad[1] else None
        gs = None
        if shared is not None and ctx.needs_input_grad[2]:
            gs = torch.zeros_like(shared)
        est._nll_grad(params, samples, shared, gnll, None, gp, gs)
        return None, gp, gs, None


def poisson_nll(estimator, params, samples, shared=None):
    """Per-spot Poisson NLL, differentiable w.r.t. params and shared."""
    return PoissonNLL.apply(estimator, params, shared, samples)
