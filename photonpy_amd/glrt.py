"""Generalized likelihood-ratio test of spot vs background (photonpy/cpp/glrt.py).

``CreateEstimator(psf, ctx)`` wraps an estimator (GLRT_CreatePSF): fits run on the
wrapped estimator, and ``Estimate`` also returns three diagnostics per spot,
{ll_on, ll_off, mean}: the Poisson log-likelihood of the fit and of a flat
background at ``mean``.  As in the reference, ``mean`` is the mean sample of the
first spot of each call; ``meanOfFirstSpot=False`` uses every spot's own mean.
"""
from __future__ import annotations

import numpy as np
from scipy.stats import norm

from .context import Context
from .estimator import Estimator


def CreateEstimator(psf: Estimator, ctx: Context, meanOfFirstSpot=True) -> Estimator:
    lib = ctx.smlm.lib
    inst = lib.GLRT_CreatePSF(psf.inst, ctx.inst if ctx else None)
    est = Estimator(ctx, inst)
    est.glrt_model = psf  # the wrapper computes with psf: keep it alive
    if not meanOfFirstSpot:
        lib.GLRT_SetMeanOfFirstSpot(inst, 0)
    return est


def ComputeLogPFA(diag):
    """2 log Phi(-sqrt(T_g)) with T_g = max(0, 2 (ll_on - ll_off)), as glrt.py:23-34
    computes it (its comment names the false-alarm probability 2 Phi(-sqrt(T_g)))."""
    diag = np.asarray(diag)
    Tg = np.maximum(0, 2 * (diag[:, 0] - diag[:, 1]))
    return 2 * norm.logcdf(-np.sqrt(Tg))
