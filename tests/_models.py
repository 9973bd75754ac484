"""Matching (HIP estimator, oracle model, closed-form mu) triples for the tests."""
import numpy as np

import photonpy_amd as pa
from oracle import oracle as O
from tests import _closed as C

ROI = {"fixed": 10, "sigma": 11, "astig": 16, "spline": 20}

_spline_cache = {}


def spline_coefs():
    if "c" not in _spline_cache:
        stack, zs = C.spline_zstack()
        coefs = O.spline_compute(stack)
        _spline_cache["c"] = (coefs, zs[1], zs[1] + 61 * 13.3)
    return _spline_cache["c"]


def make(kind, ctx, R=None, B=True):
    """Returns (estimator, oracle_model, mu_fn(theta) in FP64)."""
    R = R or ROI[kind]
    g = pa.GaussianPSFMethods(ctx)
    if kind == "fixed":
        return (g.CreatePSF_XYIBg(R, 1.5, True), O.gauss2d(R, 1.5), lambda t: C.mu_fixed(R, t, 1.5))
    if kind == "sigma":
        return (g.CreatePSF_XYIBgSigma(R, 1.6, True), O.gauss2d_sigma(R, 1.6), lambda t: C.mu_sigma(R, t))
    if kind == "astig":
        b = C.ASTIG_B if B else 0.0
        cal = pa.Gauss3D_Calibration(C.ASTIG_X + [b], C.ASTIG_Y + [b], C.ASTIG_ZRANGE)
        om = O.astig(R, C.ASTIG_X, C.ASTIG_Y, C.ASTIG_ZRANGE, B=(b, b))
        return (g.CreatePSF_XYZIBg(R, cal, True), om, lambda t: C.mu_astig(R, t, C.ASTIG_X, C.ASTIG_Y, b, b))
    if kind == "spline":
        coefs, zmin, zmax = spline_coefs()
        m = pa.CSplineMethods(ctx)
        om = O.spline(R, coefs, zmin, zmax)
        return (m.CreatePSF_XYZIBg(R, pa.CSplineCalibration(zmin, zmax, coefs), m.FlatBg, True), om,
                lambda t: om.expected_value(t, np.float64))
    raise ValueError(kind)


def samples(kind, n, seed, R=None, B=True):
    """Seeded true params and Poisson samples (mu from the FP64 closed form / oracle)."""
    rng = np.random.default_rng(seed)
    R = R or ROI[kind]
    th = C.synth_params(kind, R, n, rng)
    with pa.Context() as ctx:
        _, om, fn = make(kind, ctx, R, B=B)
        mu = fn(th)
    return th, rng.poisson(mu).astype(np.float32)
