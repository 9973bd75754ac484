"""Differentiable batched Poisson NLL for PSF calibration (no reference counterpart).

``PoissonNLL.apply(estimator, params, shared, samples)`` returns the per-spot
negative log-likelihood sum_px mu' - x' log mu' (the negated
ComputeLogLikelihood of the reference, Estimation.h:138-156) of the
estimator's PSF model.  ``params`` [N,K] are per-spot parameters, ``shared``
the PSF parameters shared by all spots (``Estimator.shared_parameters()``):
sigma_x/sigma_y (fixed-sigma Gaussian), {s0, gamma, d, A, B} per axis
(astigmatic Gaussian) or the [Nz,Ny,Nx,64] coefficient table (cubic spline).
The adjoint is a hand-written HIP kernel that recomputes the expected images and
Jacobians from the parameters instead of storing them.  When an input needs a
gradient, the forward pass runs it once with unit upstream weights, which also
yields the NLL, and keeps the resulting gradients (n x K, and the shared
gradient); the backward pass scales them by the upstream gradient.  A spot's
parameter gradient is then the kernel's own g * sums (bit for bit where the kernel
sums in FP32, the spline forms; within one FP32 rounding where it sums in FP64).
The shared gradient is scaled as a whole when the upstream gradient is one value
broadcast over the spots (a stride-0 tensor, what ``.sum()`` passes back), and
recomputed with the per-spot weights otherwise (no host synchronisation either
way).
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
        need_p = ctx.needs_input_grad[1]
        need_s = shared is not None and ctx.needs_input_grad[2]
        # gradients for unit upstream weights, from the same pass as the NLL
        gp1 = torch.empty_like(params) if need_p else None
        gs1 = torch.zeros_like(shared) if need_s else None
        estimator._nll_grad(params, samples, shared, None, nll, gp1, gs1)
        empty = torch.empty(0, device=params.device)
        ctx.estimator = estimator
        ctx.save_for_backward(params, shared if shared is not None else empty, samples,
                              gp1 if need_p else empty, gs1 if need_s else empty)
        ctx.has_shared = shared is not None
        return nll

    @staticmethod
    def backward(ctx, gnll):
        params, shared, samples, gp1, gs1 = ctx.saved_tensors
        est = ctx.estimator
        shared = shared if ctx.has_shared else None
        gnll = gnll.float()
        gp = gs = None
        if ctx.needs_input_grad[1]:
            gp = gp1 * gnll.reshape(-1, 1)
        if shared is not None and ctx.needs_input_grad[2]:
            if gnll.dim() == 1 and gnll.stride(0) == 0:  # one weight for every spot
                gs = gs1 * gnll[0]
            else:
                gs = torch.zeros_like(shared)
                est._nll_grad(params, samples, shared, gnll.contiguous(), None, None, gs)
        return None, gp, gs, None


def poisson_nll(estimator, params, samples, shared=None):
    """Per-spot Poisson NLL, differentiable w.r.t. params and shared."""
    return PoissonNLL.apply(estimator, params, shared, samples)
