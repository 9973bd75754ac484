"""Z-stack files and PSF rendering helpers (photonpy/smlm/psf.py:217-276).

``.zstack``: a 24-byte little-endian header {int32 version = 1, int32 shape[3] (z, y, x),
float32 zmin, zmax} followed by the float32 samples, the format CSpline_CreatePSF_FromFile
reads (CubicSplinePSF.cu:109-133).
"""
from __future__ import annotations

import numpy as np

HEADER_DTYPE = np.dtype([("version", "<i4"), ("dims", "<i4", 3), ("zrange", "<f4", 2)])


def save_zstack(zstack, zrange, fn):
    zstack = np.ascontiguousarray(zstack, dtype=np.float32)
    if zstack.ndim != 3:
        raise ValueError("a z-stack is a 3D array [z, y, x]")
    with open(fn, "wb") as f:
        np.array([(1, zstack.shape, (zrange[0], zrange[-1]))], dtype=HEADER_DTYPE).tofile(f)
        zstack.tofile(f)


def load_zstack(fn):
    """Returns (zstack [z, y, x] float32, [zmin, zmax])."""
    with open(fn, "rb") as f:
        h = np.fromfile(f, dtype=HEADER_DTYPE, count=1)
        if len(h) != 1:
            raise ValueError(f"{fn}: truncated z-stack header")
        version, shape, zrange = h[0]
        if version != 1:
            raise ValueError("Wrong ZStack file version (should be 1)")
        data = np.fromfile(f, dtype="<f4")
    shape = tuple(int(v) for v in shape)
    if data.size != np.prod(shape):
        raise ValueError(f"{fn}: expected {np.prod(shape)} samples, found {data.size}")
    return data.reshape(shape), np.asarray(zrange, dtype=np.float32)


def psf_to_zstack(psf, zrange, intensity=1, bg=0):
    """Expected images of a 3D PSF at the ROI centre for each z in zrange (psf.py:217-237)."""
    if psf.numparams not in (5, 7):
        raise ValueError("psf_to_zstack needs a 3D PSF (x, y, z, I, bg[, tilts])")
    params = np.zeros((len(zrange), psf.numparams), dtype=np.float32)
    params[:, 0] = psf.sampleshape[0] / 2
    params[:, 1] = psf.sampleshape[1] / 2
    params[:, 2] = zrange
    params[:, 3] = intensity
    params[:, 4] = bg
    return psf.ExpectedValue(params)
