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

    @classmethod
    def from_file(cls, filename):
        """A calibration file (cspline.py:42-56): ``.mat`` (from_file_nmeth) or a numpy
        ``.npz`` holding ``zmin``, ``zmax`` (nm) and ``coeff``.  The reference reads a pickled
        ``.npy`` here; pickles are not loaded."""
        if filename.endswith(".mat"):
            return cls.from_file_nmeth(filename)
        with np.load(filename, allow_pickle=False) as d:
            # the file has z in nm, the estimator um
            return cls(float(d["zmin"]) * 1e-3, float(d["zmax"]) * 1e-3, d["coeff"])

    @classmethod
    def from_file_nmeth(cls, filename):
        """Spline of the Nature Methods cubic-spline calibration format (cspline.py:58-92):
        a [a, b, nz, 64] table whose per-cell coefficients are ordered (z, x, y) powers becomes
        [nz, a, b, 64] with (z, y, x) powers -- the reference's transpose (2, 0, 1, 3, 5, 4),
        which puts the file's first voxel axis on y; the z range is centred,
        z_min = -nz dz / 2, z_max = (nz - 1) dz / 2 (um, dz in nm)."""
        import scipy.io
        mat = scipy.io.loadmat(filename)
        try:
            spline = mat["SXY"]["cspline"].item()
            coefs = spline["coeff"].item()
        except (KeyError, ValueError):
            spline = mat["cspline"].item()
            coefs = spline[0]
        try:
            dz = float(np.squeeze(mat["parameters"]["dz"][0][0]))
        except (KeyError, ValueError, IndexError):
            raise ValueError("Required parameters for C-spline can't be found in the .mat file")
        coefs = np.ascontiguousarray(coefs, dtype=np.float32)
        a, b, nz = coefs.shape[:-1]
        coefs = coefs.reshape((a, b, nz, 4, 4, 4))
        coefs = np.ascontiguousarray(np.transpose(coefs, (2, 0, 1, 3, 5, 4))).reshape((nz, a, b, 64))
        nzf = float(coefs.shape[0])
        z_min = -nzf * dz * 1e-3 / 2
        z_max = (nzf - 1) * dz * 1e-3 / 2
        return cls(z_min, z_max, coefs)


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

    def Evaluate(self, splineCoefs, shift, border=0):
        """Value, d/dx, d/dy, d/dz of the spline at every grid voxel shifted by `shift`
        (z, y, x): [*(shape[:3] - 2*border), 4] (cspline.py:176-186, CSpline_Evaluate)."""
        splineCoefs = np.ascontiguousarray(splineCoefs, dtype=np.float32)
        if splineCoefs.ndim != 4 or splineCoefs.shape[-1] != 64:
            raise ValueError("spline coefficients must have shape [Nz,Ny,Nx,64]")
        shift = np.ascontiguousarray(shift, dtype=np.float32)
        outshape = np.array(splineCoefs.shape[:3]) - border * 2
        values = np.zeros((*outshape, 4), dtype=np.float32)
        self._lib.CSpline_Evaluate(np.array(splineCoefs.shape, dtype=np.int32), splineCoefs, shift, border, values)
        self.ctx.smlm.check()
        return values

    def CreatePSFFromFile(self, roisize, filename, fitMode=FlatBg, cuda=True) -> CSplinePSF:
        """``.zstack`` (z-stack file) or ``.mat`` (Nature Methods calibration), as
        cspline.py:188-207; ``.pickle`` files are not loaded."""
        if filename.endswith(".mat"):
            return self.CreatePSF_XYZIBg(roisize, CSplineCalibration.from_file_nmeth(filename), fitMode, cuda)
        if filename.endswith(".zstack"):
            from .psf_io import load_zstack
            zstack, zrange = load_zstack(filename)
            return self.CreatePSFFromZStack(roisize, zstack, zrange, fitMode, cuda)
        raise RuntimeError(f"Unknown psf extension: {filename}")

    def CreatePSFFromZStack(self, roisize, zstack, zrange, fitMode=FlatBg, cuda=True, border=0) -> CSplinePSF:
        spline_coeff = self.ComputeSplineCoeff(zstack, border)
        calib = CSplineCalibration(zrange[0], zrange[-1], spline_coeff)
        calib.zstack = zstack
        return self.CreatePSF_XYZIBg(roisize, calib, fitMode, cuda)
