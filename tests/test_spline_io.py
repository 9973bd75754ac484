"""Spline calibration IO on the CPU: .zstack files (smlm/psf.py:256-276,
CubicSplinePSF.cu:109-133) and the Nature Methods .mat reordering (cspline.py:58-92)."""
import os
import struct

import numpy as np
import pytest

from photonpy_amd.psf_io import load_zstack, save_zstack
from photonpy_amd.cspline import CSplineCalibration

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "tetrapod-psf.zstack")


def test_fixture_header_and_slices():
    with open(FIXTURE, "rb") as f:
        version, s0, s1, s2, zmin, zmax = struct.unpack("<4i2f", f.read(24))
    assert (version, s0, s1, s2) == (1, 51, 40, 40)
    zs, zr = load_zstack(FIXTURE)
    assert zs.shape == (51, 40, 40) and zs.dtype == np.float32
    assert np.allclose(zr, [zmin, zmax]) and np.allclose(zr, [-1.375, 1.375])
    # each slice is a normalised PSF image
    assert np.allclose(zs.sum(axis=(1, 2)), 1.0, atol=1e-4)
    assert os.path.getsize(FIXTURE) == 24 + zs.size * 4


def test_save_load_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    zs = rng.random((7, 9, 11)).astype(np.float32)
    fn = str(tmp_path / "a.zstack")
    save_zstack(zs, [-0.5, 0.25, 0.5], fn)
    with open(fn, "rb") as f:
        assert struct.unpack("<4i2f", f.read(24)) == (1, 7, 9, 11, -0.5, 0.5)
    z2, zr = load_zstack(fn)
    assert np.array_equal(z2, zs) and np.allclose(zr, [-0.5, 0.5])
    with open(fn, "r+b") as f:
        f.write(struct.pack("<i", 2))
    with pytest.raises(ValueError, match="version"):
        load_zstack(fn)


def test_nature_methods_mat_reordering(tmp_path):
    """A known [nz, ny, nx] table (per-cell powers z, y, x) written in the .mat layout -- voxel
    axes (y, x, z), per-cell powers (z, x, y) -- reads back as that table, with the centred
    z range.  Also pins that a coefficient moves with the power it multiplies: the table is
    the spline of a polynomial whose x and y powers differ, evaluated at one sub-voxel point."""
    import scipy.io
    rng = np.random.default_rng(1)
    nz, ny, nx = 5, 4, 3
    coefs = rng.random((nz, ny, nx, 4, 4, 4)).astype(np.float32)  # [z,y,x][m(z),n(y),o(x)]
    # .mat cell: voxel axes x, y, z and coefficient order (z-power, x-power, y-power)
    mat = np.ascontiguousarray(np.transpose(coefs, (1, 2, 0, 3, 5, 4))).reshape(ny, nx, nz, 64)
    dz = 50.0
    fn = str(tmp_path / "cal.mat")
    scipy.io.savemat(fn, {"cspline": np.array([(mat, dz)], dtype=[("coeff", "O"), ("dz", "O")]),
                          "parameters": {"dz": dz}})
    cal = CSplineCalibration.from_file_nmeth(fn)
    assert (cal.n_voxels_z, cal.n_voxels_y, cal.n_voxels_x) == (nz, ny, nx)
    assert np.array_equal(cal.coefsx.reshape(nz, ny, nx, 4, 4, 4), coefs)
    assert np.isclose(cal.z_min, -nz * dz * 1e-3 / 2) and np.isclose(cal.z_max, (nz - 1) * dz * 1e-3 / 2)
    # value of cell (0, 0, 0) at (w, v, u) = (0.5, 0.25, 0.75): sum c[m,n,o] w^m v^n u^o, with
    # the .mat powers (z, x, y) read as mat[..., m, o, n]
    w, v, u = 0.5, 0.25, 0.75
    P = lambda t: t ** np.arange(4)
    ours = np.einsum("mno,m,n,o->", coefs[0, 0, 0], P(w), P(v), P(u))
    m0 = mat.reshape(ny, nx, nz, 4, 4, 4)[0, 0, 0]  # (z, x, y) powers
    theirs = np.einsum("mox,m,o,x->", m0, P(w), P(u), P(v))
    assert np.isclose(ours, theirs)


def test_npz_calibration(tmp_path):
    coefs = np.ones((3, 2, 2, 64), np.float32)
    fn = str(tmp_path / "c.npz")
    np.savez(fn, zmin=-400.0, zmax=400.0, coeff=coefs)
    cal = CSplineCalibration.from_file(fn)
    assert np.isclose(cal.z_min, -0.4) and np.isclose(cal.z_max, 0.4)
    assert np.array_equal(cal.coefsx, coefs)
