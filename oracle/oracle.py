"""ORACLE — test infrastructure only.

ctypes front end of the CPU restatement in ``oracle/smlm_oracle.cpp``.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
may import this module; the product package ``photonpy_amd`` never does.

The float32 entry points reproduce the reference CPU estimator
(``cpp/SMLMLib/Estimators/EstimatorImpl.h:108-185`` driving
``Estimation.h:194-251``); the float64 entry points evaluate the same
mathematics in FP64.  Model descriptors:

* ``gauss2d(R, sx, sy)``            Gauss2D_Model_XYIBg      (GaussianPSFModels.h:35-152)
* ``gauss2d_sigma(R, initialSigma)``  Gauss2D_Model_XYIBgSigma (GaussianPSFModels.h:308-387)
* ``astig(R, x4, y4, zrange, B)``     Gauss2D_Model_XYZIBg     (GaussianPSFModels.h:636-751),
  with an optional quartic term ``B`` (per axis) inside the square root
* ``spline(R, coefs, zmin, zmax)``    CSpline_Model_XYZIBg     (CubicSplinePSFModels.h:14-154)
* ``gauss2d_const_sigma(R)``          Gauss2D_Model_XYIBgConstSigma (GaussianPSFModels.h:155-229),
  constants [sigma_x, sigma_y] per spot
* ``gauss2d_tilted_bg(R)``            Gauss2D_Model_XYITiltedBg (GaussianPSFModels.h:232-304),
  constants [sigma_x, sigma_y] per spot
* ``gauss2d_sigma_xy(R, isx, isy)``   Gauss2D_Model_XYIBgSigmaXY (GaussianPSFModels.h:388-467)
* ``gauss2d_tilted_bg_sigma_xy(R, isx, isy)``  Gauss2D_Model_XYITiltedBgSigmaXY (:471-553)
* ``spline(R, coefs, zmin, zmax, mode)``  mode 1: CSpline_Model_XYZITiltedBg (CubicSplinePSFModels.h:157-298),
  mode 2: CSpline_Model_XYZIBgImage (:300-443), constants = an R x R background image per spot
* ``intensity_bg(R)``                 IntensityBgModel (GaussianPSFModels.h:557-623), constants
  [x, y, sigma] per spot; CRLB from |diag| (Estimation.h:318-333)
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libsmlm_oracle.so")


class OrModel(ct.Structure):
    _fields_ = [
        ("kind", ct.c_int32),
        ("roisize", ct.c_int32),
        ("p", ct.c_double * 16),
        ("shape", ct.c_int32 * 3),
        ("coefs", ct.c_void_p),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE], stdout=subprocess.DEVNULL)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        _lib = ct.CDLL(_SO)
        F, D, I, P = ct.POINTER(ct.c_float), ct.POINTER(ct.c_double), ct.POINTER(ct.c_int32), ct.c_void_p
        M = ct.POINTER(OrModel)
        for sfx, R in (("f32", F), ("f64", D)):
            getattr(_lib, f"or_expected_value_{sfx}").argtypes = [M, ct.c_int, R, R, R, ct.c_int]
            getattr(_lib, f"or_derivatives_{sfx}").argtypes = [M, ct.c_int, R, R, R, R, ct.c_int]
            scal = ct.c_float if sfx == "f32" else ct.c_double
            getattr(_lib, f"or_estimate_{sfx}").argtypes = [M, ct.c_int, R, R, R, R, I, R, ct.c_int, ct.c_int, scal,
                                                            ct.c_int]
            getattr(_lib, f"or_crlb_{sfx}").argtypes = [M, ct.c_int, R, R, R, R, R, ct.c_int]
            getattr(_lib, f"or_loglik_{sfx}").argtypes = [M, ct.c_int, R, R, R, R, ct.c_int]
        for sfx, R in (("f32", F), ("f64", D)):
            getattr(_lib, f"or_glrt_{sfx}").argtypes = [M, ct.c_int, R, R, R, ct.c_int, R, ct.c_int]
        _lib.or_num_params.argtypes = [ct.c_int]
        _lib.or_num_consts.argtypes = [ct.c_int, ct.c_int]
        _lib.or_spline_compute.argtypes = [I, F, F, ct.c_int]
    return _lib


def _nthreads(n):
    if n is None:
        return os.cpu_count() or 1
    return int(n)


class Model:
    """A PSF model descriptor plus convenience wrappers around the oracle."""

    def __init__(self, kind, roisize, p=(), shape=(0, 0, 0), coefs=None):
        self.desc = OrModel()
        self.desc.kind = kind
        self.desc.roisize = roisize
        for i, v in enumerate(p):
            self.desc.p[i] = v
        for i, v in enumerate(shape):
            self.desc.shape[i] = v
        self._coefs = None
        if coefs is not None:
            self._coefs = np.ascontiguousarray(coefs, dtype=np.float32)
            self.desc.coefs = self._coefs.ctypes.data
        self.kind = kind
        self.roisize = roisize
        self.K = lib().or_num_params(kind)
        self.NC = lib().or_num_consts(kind, roisize)

    # -- helpers -----------------------------------------------------------
    def _arr(self, a, dt, shape=None):
        a = np.ascontiguousarray(a, dtype=dt)
        if shape is not None:
            a = a.reshape(shape)
        return a

    def _ptr(self, a, dt):
        if a is None:
            return None
        return a.ctypes.data_as(ct.POINTER(ct.c_float if dt == np.float32 else ct.c_double))

    def _cst(self, constants, n, dt):
        if self.NC == 0:
            return None
        if constants is None:
            raise ValueError(f"model kind {self.kind} needs constants of shape ({n}, {self.NC})")
        return self._arr(constants, dt, (n, self.NC))

    # -- API ---------------------------------------------------------------
    def expected_value(self, params, dtype=np.float32, nthreads=None, constants=None):
        p = self._arr(params, dtype, (-1, self.K))
        n, R = len(p), self.roisize
        c = self._cst(constants, n, dtype)
        ev = np.zeros((n, R, R), dtype)
        f = lib().or_expected_value_f32 if dtype == np.float32 else lib().or_expected_value_f64
        f(ct.byref(self.desc), n, self._ptr(p, dtype), self._ptr(c, dtype), self._ptr(ev, dtype), _nthreads(nthreads))
        return ev

    def derivatives(self, params, dtype=np.float32, nthreads=None, constants=None):
        p = self._arr(params, dtype, (-1, self.K))
        n, R = len(p), self.roisize
        c = self._cst(constants, n, dtype)
        ev = np.zeros((n, R, R), dtype)
        dv = np.zeros((n, self.K, R, R), dtype)
        f = lib().or_derivatives_f32 if dtype == np.float32 else lib().or_derivatives_f64
        f(ct.byref(self.desc), n, self._ptr(p, dtype), self._ptr(c, dtype), self._ptr(dv, dtype), self._ptr(ev, dtype),
          _nthreads(nthreads))
        return dv, ev

    def estimate(self, samples, initial=None, iterations=50, lambda_=1e-18, trace_len=0, dtype=np.float32, nthreads=None,
                 constants=None):
        R = self.roisize
        s = self._arr(samples, dtype, (-1, R, R))
        n = len(s)
        c = self._cst(constants, n, dtype)
        ini = None if initial is None else self._arr(initial, dtype, (n, self.K))
        params = np.zeros((n, self.K), dtype)
        iters = np.zeros(n, np.int32)
        trace = np.zeros((n, max(trace_len, 1), self.K), dtype) if trace_len > 0 else None
        f = lib().or_estimate_f32 if dtype == np.float32 else lib().or_estimate_f64
        f(ct.byref(self.desc), n, self._ptr(s, dtype), self._ptr(c, dtype), self._ptr(ini, dtype), self._ptr(params, dtype),
          iters.ctypes.data_as(ct.POINTER(ct.c_int32)), self._ptr(trace, dtype), trace_len, iterations,
          lambda_, _nthreads(nthreads))
        if trace_len > 0:
            return params, iters, trace
        return params, iters

    def crlb(self, params, samples=None, dtype=np.float32, nthreads=None, constants=None):
        p = self._arr(params, dtype, (-1, self.K))
        n, R = len(p), self.roisize
        c = self._cst(constants, n, dtype)
        s = None if samples is None else self._arr(samples, dtype, (n, R, R))
        out = np.zeros((n, self.K), dtype)
        chi = np.zeros(n, dtype)
        f = lib().or_crlb_f32 if dtype == np.float32 else lib().or_crlb_f64
        f(ct.byref(self.desc), n, self._ptr(p, dtype), self._ptr(c, dtype), self._ptr(s, dtype), self._ptr(out, dtype),
          self._ptr(chi, dtype), _nthreads(nthreads))
        return out, chi

    def glrt(self, params, samples, mean_of_first_spot=True, dtype=np.float32, nthreads=None, constants=None):
        """GLRT diagnostics [n, 3] = {ll_on, ll_off, mean} of fitted params (GLRT_PSF.cu:103-129)."""
        p = self._arr(params, dtype, (-1, self.K))
        n, R = len(p), self.roisize
        c = self._cst(constants, n, dtype)
        s = self._arr(samples, dtype, (n, R, R))
        out = np.zeros((n, 3), dtype)
        f = lib().or_glrt_f32 if dtype == np.float32 else lib().or_glrt_f64
        f(ct.byref(self.desc), n, self._ptr(p, dtype), self._ptr(c, dtype), self._ptr(s, dtype),
          int(bool(mean_of_first_spot)), self._ptr(out, dtype), _nthreads(nthreads))
        return out

    def loglik(self, params, samples, dtype=np.float32, nthreads=None, constants=None):
        p = self._arr(params, dtype, (-1, self.K))
        n, R = len(p), self.roisize
        c = self._cst(constants, n, dtype)
        s = self._arr(samples, dtype, (n, R, R))
        ll = np.zeros(n, dtype)
        f = lib().or_loglik_f32 if dtype == np.float32 else lib().or_loglik_f64
        f(ct.byref(self.desc), n, self._ptr(p, dtype), self._ptr(c, dtype), self._ptr(s, dtype), self._ptr(ll, dtype),
          _nthreads(nthreads))
        return ll


def gauss2d(roisize, sx, sy=None):
    return Model(0, roisize, (sx, sx if sy is None else sy))


def gauss2d_sigma(roisize, initial_sigma):
    return Model(1, roisize, (initial_sigma,))


def astig(roisize, x, y, zrange, B=(0.0, 0.0)):
    """x, y: (s0, gamma, d, A) per axis in the reference's units; B: quartic coefficient per axis."""
    return Model(2, roisize, tuple(x[:4]) + tuple(y[:4]) + (zrange[0], zrange[1], B[0], B[1]))


def spline(roisize, coefs, zmin, zmax, mode=0):
    """mode: 0 flat background, 1 tilted background, 2 background image (CSpline_CreatePSF modes)."""
    coefs = np.ascontiguousarray(coefs, dtype=np.float32)
    assert coefs.ndim == 4 and coefs.shape[3] == 64
    return Model({0: 3, 1: 9, 2: 10}[mode], roisize, (zmin, zmax), shape=coefs.shape[:3], coefs=coefs)


def gauss2d_const_sigma(roisize):
    return Model(4, roisize)


def gauss2d_tilted_bg(roisize):
    return Model(5, roisize)


def gauss2d_sigma_xy(roisize, initial_sigma_x, initial_sigma_y):
    return Model(6, roisize, (initial_sigma_x, initial_sigma_y))


def gauss2d_tilted_bg_sigma_xy(roisize, initial_sigma_x, initial_sigma_y):
    return Model(7, roisize, (initial_sigma_x, initial_sigma_y))


def intensity_bg(roisize):
    return Model(8, roisize)


def spline_compute(data, border=0):
    """CSpline_Compute restatement: data[Nz,Ny,Nx] -> coefs[Nz-3+2b, Ny-3+2b, Nx-3+2b, 64]."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    shape = np.array(data.shape, dtype=np.int32)
    out = np.zeros((*(shape - 3 + 2 * border), 64), np.float32)
    lib().or_spline_compute(shape.ctypes.data_as(ct.POINTER(ct.c_int32)),
                            data.ctypes.data_as(ct.POINTER(ct.c_float)),
                            out.ctypes.data_as(ct.POINTER(ct.c_float)), border)
    return out
