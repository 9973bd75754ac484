"""Gaussian PSF factories, mirroring photonpy/cpp/gaussian.py:15-237.

``Gauss3D_Calibration(x, y, zrange)`` takes 4 coefficients per axis
``{s0, gamma, d, A}`` as in the reference (GaussianPSF.h:27-33); a fifth
coefficient ``B`` per axis adds the quartic term B*t^4/d^2 under the square
root of sigma(z) and selects the extended factory.
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from .context import Context
from .estimator import Estimator


class Gauss3D_Calibration(ctypes.Structure):
    """Same memory layout as the reference struct (x[4], y[4], zrange[2])."""
    _fields_ = [("x", ctypes.c_float * 4), ("y", ctypes.c_float * 4), ("zrange", ctypes.c_float * 2)]

    def __init__(self, x=[1.3, 2, 3, 0], y=[1.3, -2, 3, 0], zrange=[-3, 3]):
        super().__init__()
        self.x = (ctypes.c_float * 4)(*x[:4])
        self.y = (ctypes.c_float * 4)(*y[:4])
        self.zrange = (ctypes.c_float * 2)(*zrange)
        self.B = (float(x[4]) if len(x) > 4 else 0.0, float(y[4]) if len(y) > 4 else 0.0)

    def sigma(self):
        return self.x[0], self.y[0]

    @classmethod
    def from_astig_params(cls, s0, c, d, A, B, zrange):
        """sigma_{x,y}(z) = s0 sqrt(1 + u^2 + A u^3 + B u^4), u = (z -+ c)/d, mapped to the
        reference parametrisation gamma = +-c, A_ref = A/d, B_ref = B/d^2."""
        x = [s0, c, d, A / d, B / d ** 2]
        y = [s0, -c, d, A / d, B / d ** 2]
        return cls(x, y, zrange)

    @classmethod
    def from_file(cls, filename):
        """YAML or JSON with keys x, y (and optionally range).  Pickled .npy files are not read."""
        ext = os.path.splitext(filename)[1]
        with open(filename, "r") as f:
            if ext == ".json":
                calibration = json.load(f)
            else:
                import yaml
                calibration = yaml.safe_load(f)
        zr = calibration.get("range", [-3, 3])
        return cls(calibration.get("x"), calibration.get("y"), zr)

    def save(self, filename):
        ext = os.path.splitext(filename)[1]
        with open(filename, "w") as f:
            if ext == ".json":
                json.dump(self.__dict__(), f)
            else:
                import yaml
                yaml.safe_dump(self.__dict__(), f)

    def __dict__(self):
        x = np.array(self.x, dtype=float).tolist()
        y = np.array(self.y, dtype=float).tolist()
        if any(self.B):
            x.append(self.B[0])
            y.append(self.B[1])
        return {'x': x, 'y': y, 'range': np.array(self.zrange, dtype=float).tolist()}


class _CalibQ(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float * 4), ("y", ctypes.c_float * 4), ("minz", ctypes.c_float),
                ("maxz", ctypes.c_float), ("Bx", ctypes.c_float), ("By", ctypes.c_float)]


class GaussianPSFMethods:
    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.lib = ctx.smlm
        self._lib = ctx.smlm.lib

    def _make(self, inst, calib):
        if not inst:
            self.ctx.smlm.check()
            raise RuntimeError("photonpy_amd: estimator creation failed")
        return Estimator(self.ctx, inst, calib)

    def CreatePSF_XYIBg(self, roisize, sigma, cuda=True) -> Estimator:
        """Parameters [x,y,I,bg].  sigma may be a scalar or (sigma_x, sigma_y); None gives
        the model whose sigma x, y are per-spot constants (gaussian.py:175-191)."""
        if sigma is not None:
            if np.isscalar(sigma):
                sigma_x, sigma_y = sigma, sigma
            else:
                sigma_x, sigma_y = sigma
            inst = self._lib.Gauss2D_CreatePSF_XYIBg(roisize, sigma_x, sigma_y, int(cuda), self.ctx.inst)
        else:
            inst = self._lib.Gauss2D_CreatePSF_XYIBgConstSigma(roisize, int(cuda), self.ctx.inst)
        return self._make(inst, sigma)

    def CreatePSF_XYITiltedBg(self, roisize, cuda=True) -> Estimator:
        """Parameters [x,y,I,bg,bgx,bgy], mu = bg + bgx (x - roisize/2) + bgy (y - roisize/2)
        + I Gaussian(x, y); sigma x, y are per-spot constants (gaussian.py:193-200)."""
        inst = self._lib.Gauss2D_CreatePSF_XYITiltedBg(roisize, int(cuda), self.ctx.inst)
        return self._make(inst, None)

    def CreatePSF_XYIBgSigmaXY(self, roisize, initialSigma, cuda=True) -> Estimator:
        """Parameters [x,y,I,bg,sx,sy] (gaussian.py:212-221)."""
        if np.isscalar(initialSigma):
            initialSigma = [initialSigma, initialSigma]
        inst = self._lib.Gauss2D_CreatePSF_XYIBgSigmaXY(roisize, initialSigma[0], initialSigma[1], int(cuda),
                                                        self.ctx.inst)
        return self._make(inst, initialSigma)

    def CreatePSF_XYITiltedBgSigmaXY(self, roisize, initialSigma, cuda=True) -> Estimator:
        """Parameters [x,y,I,bg,bgx,bgy,sx,sy] (gaussian.py:223-232)."""
        if np.isscalar(initialSigma):
            initialSigma = [initialSigma, initialSigma]
        inst = self._lib.Gauss2D_CreatePSF_XYITiltedBgSigmaXY(roisize, initialSigma[0], initialSigma[1], int(cuda),
                                                              self.ctx.inst)
        return self._make(inst, initialSigma)

    def EstimateIBg(self, images, sigma, xy, roipos=None, useCuda=False):
        """Intensity and background of spots at known positions (gaussian.py:154-172):
        returns (IBg [n,2], crlb [n,2]).  sigma is broadcast to [n,2] as the reference
        does, and the native side reads its i-th element for spot i."""
        images = np.ascontiguousarray(images, dtype=np.float32)
        xy = np.ascontiguousarray(xy, dtype=np.float32)
        if images.ndim != 3:
            raise ValueError("images must be [n, imgw, imgw]")
        numspots = len(images)
        if xy.shape != (numspots, 2):
            raise ValueError(f"xy must have shape {(numspots, 2)}")
        imgw = images.shape[1]
        result = np.zeros((numspots, 2), dtype=np.float32)
        crlb = np.zeros((numspots, 2), dtype=np.float32)
        sigma_ = np.zeros((numspots, 2), dtype=np.float32)
        sigma_[:] = sigma
        if roipos is None:
            roipos = np.zeros((numspots, 2))
        roipos = np.ascontiguousarray(roipos, dtype=np.int32)
        if roipos.shape != (numspots, 2):
            raise ValueError(f"roipos must have shape {(numspots, 2)}")
        if numspots > 0:
            self._lib.Gauss2D_EstimateIntensityBg(images, result, crlb, numspots, xy, roipos, sigma_, imgw, 100,
                                                  bool(useCuda))
            self.ctx.smlm.check()
        return result, crlb

    def CreatePSF_XYIBgSigma(self, roisize, initialSigma, cuda=True) -> Estimator:
        """Parameters [x,y,I,bg,sigma]."""
        inst = self._lib.Gauss2D_CreatePSF_XYIBgSigma(roisize, initialSigma, int(cuda), self.ctx.inst)
        return self._make(inst, initialSigma)

    def CreatePSF_XYZIBg(self, roisize, calib: Gauss3D_Calibration, cuda=True) -> Estimator:
        """Parameters [x,y,z,I,bg]."""
        B = getattr(calib, "B", (0.0, 0.0))
        if B[0] != 0.0 or B[1] != 0.0:
            q = _CalibQ()
            q.x = (ctypes.c_float * 4)(*calib.x)
            q.y = (ctypes.c_float * 4)(*calib.y)
            q.minz, q.maxz = calib.zrange[0], calib.zrange[1]
            q.Bx, q.By = B
            inst = self._lib.Gauss2D_CreatePSF_XYZIBgQ(roisize, ctypes.byref(q), int(cuda), self.ctx.inst)
        else:
            inst = self._lib.Gauss2D_CreatePSF_XYZIBg(roisize, ctypes.byref(calib), int(cuda), self.ctx.inst)
        return self._make(inst, calib)


Gaussian = GaussianPSFMethods
