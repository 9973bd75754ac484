"""Pins the oracle's GLRT diagnostics (GLRT_PSF.cu:103-129) and the Python ComputeLogPFA.

ll_on = sum max(x,1e-8) log(max(mu,1e-8)) - max(mu,1e-8) with mu from the closed-form
model; ll_off the same with mu = mean; mean = the first spot's mean sample (the
reference's indexing) or the spot's own.  A model whose mu is the flat mean gives
ll_on == ll_off.
"""
import numpy as np

from oracle import oracle as O
from photonpy_amd import glrt
from tests import _closed as C

RNG = np.random.default_rng(99)


def _np_glrt(mu, smp, mean):
    mu = np.maximum(mu, 1e-8).reshape(len(mu), -1)
    x = np.maximum(smp, 1e-8).reshape(len(smp), -1)
    on = np.sum(x * np.log(mu) - mu, axis=1)
    off = np.sum(x * np.log(mean)[:, None] - mean[:, None], axis=1)
    return np.stack([on, off, mean], 1)


def test_glrt_closed_form_and_first_spot_mean():
    R = 10
    th = C.synth_params("fixed", R, 30, RNG)
    smp = RNG.poisson(C.mu_fixed(R, th, 1.5)).astype(np.float64)
    m = O.gauss2d(R, 1.5)
    mu = C.mu_fixed(R, th, 1.5)
    d0 = m.glrt(th, smp, True, dtype=np.float64)
    mean0 = np.full(30, smp[0].mean())
    assert np.allclose(d0, _np_glrt(mu, smp, mean0), rtol=1e-12, atol=1e-9)
    assert np.all(d0[:, 2] == d0[0, 2])  # every spot takes the first spot's mean
    d1 = m.glrt(th, smp, False, dtype=np.float64)
    assert np.allclose(d1, _np_glrt(mu, smp, smp.reshape(30, -1).mean(1)), rtol=1e-12, atol=1e-9)
    d32 = m.glrt(th, smp, False, dtype=np.float32)
    assert np.allclose(d32, d1, rtol=1e-5)


def test_glrt_flat_model_equals_background_hypothesis():
    R = 10
    smp = RNG.poisson(8.0, (20, R, R)).astype(np.float64)
    mean = smp.reshape(20, -1).mean(1)
    th = np.stack([np.full(20, 4.5), np.full(20, 4.5), np.zeros(20), mean], 1)  # I = 0: mu = mean
    d = O.gauss2d(R, 1.5).glrt(th, smp, False, dtype=np.float64)
    assert np.allclose(d[:, 0], d[:, 1], rtol=1e-14)


def test_glrt_with_constants():
    R = 11
    th, cst = C.synth_variant("const_sigma", R, 10, RNG)
    smp = RNG.poisson(C.mu_variant("const_sigma", R, th, cst)).astype(np.float64)
    d = O.gauss2d_const_sigma(R).glrt(th, smp, False, dtype=np.float64, constants=cst)
    ref = _np_glrt(C.mu_variant("const_sigma", R, th, cst), smp, smp.reshape(10, -1).mean(1))
    assert np.allclose(d, ref, rtol=1e-12, atol=1e-9)


def test_log_pfa():
    diag = np.array([[10.0, 8.0, 1.0], [5.0, 5.0, 1.0], [3.0, 7.0, 1.0], [100.0, 10.0, 1.0]])
    lp = glrt.ComputeLogPFA(diag)
    from scipy.special import ndtr
    # T_g = 4, 0, max(0, -8) = 0, 180
    assert np.isclose(lp[0], 2 * np.log(ndtr(-2.0)))
    assert np.isclose(lp[1], 2 * np.log(0.5)) and np.isclose(lp[2], 2 * np.log(0.5))
    assert lp[3] < lp[0] < lp[1]
