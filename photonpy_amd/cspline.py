"""Cubic-spline PSF factories, mirroring photonpy/cpp/cspline.py:16-213."""
from __future__ import annotations

import ctypes

import numpy as np
import numpy.ctypeslib as ctl

from . import estimator
from .context import Context


class CSplineCalibration(ctypes.Structure):
    """Same layout as CSpline_Calibration (CubicSplinePSF.h:12-23)."""
    _fields_ = [
        ("n_voxels_z", ctypes.c_int32),
        ("n_voxels_y", ctypes.c_int32),
        ("n_voxels_x", ctypes.c_int32),
        ("z_min", ctypes.c_float),
        ("z_max", ctypes.c_float),
        ("coefs", ctypes.c_void_p),
    ]

    def __init__(self, z_min, z_max, coefs):
        super().__init__()
        coefs = np.ascontiguousarray(coefs, dtype=np.float32)
        if coefs.ndim != 4 or coefs.shape[3] != 64:
            raise ValueError(f"spline coefficients must have shape [Nz,Ny,Nx,64], got {coefs.shape}")
        self.coefsx = coefs  # keeps the array alive while the struct points at it
        self.n_voxels_x = coefs.shape[2]
        self.n_voxels_y = coefs.shape[1]
        self.n_voxels_z = coefs.shape[0]
        self.coefs = coefs.ctypes.data
        self.z_min = z_min
        self.z_max = z_max
        self.zrange = [z_min, z_max]


class CSplinePSF(estimator.Estimator):
    def __init__(self, ctx: Context, inst, calib):
        super().__init__(ctx, inst, calib)


class CSplineMethods:
    FlatBg = 0
    TiltedBg = 1
    BgImage = 2

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self._lib = ctx.smlm.lib

    def CreatePSF_XYZIBg(self, roisize, calib: CSplineCalibration, fitMode=FlatBg, cuda=True) -> CSplinePSF:
        inst = self._lib.CSpline_CreatePSF(roisize, ctypes.byref(calib), None, fitMode, cuda, self.ctx.inst)
        if not inst:
            self.ctx.smlm.check()
            raise RuntimeError("photonpy_amd: spline estimator creation failed")
        return CSplinePSF(self.ctx, inst, calib)

    def ComputeSplineCoeff(self, data, border=0):
        """Spline coefficients [*(shape-3+2*border), 64] of a 3D stack (cspline.py:159-174)."""
        data = np.ascontiguousarray(data, dtype=np.float32)
        if data.ndim != 3:
            raise ValueError("ComputeSplineCoeff expects a 3D array [z,y,x]")
        outshape = np.array(data.shape) - 3 + border * 2
        spline = np.zeros((*outshape, 64), dtype=np.float32)
        shape = np.ascontiguousarray(data.shape, dtype=np.int32)
        self._lib.CSpline_Compute(shape, data, spline, border)
        self.ctx.smlm.check()
        return spline

    def CreatePSFFromZStack(self, roisize, zstack, zrange, fitMode=FlatBg, cuda=True, border=0) -> CSplinePSF:
        spline_coeff = self.ComputeSplineCoeff(zstack, border)
        calib = CSplineCalibration(zrange[0], zrange[-1], spline_coeff)
        calib.zstack = zstack
        return self.CreatePSF_XYZIBg(roisize, calib, fitMode, cuda)
