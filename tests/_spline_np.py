"""FP64 numpy evaluation of the cubic-spline PSF (CubicSplinePSFModels.h:64-153) with the
cell index of every pixel, for analytic checks of the coefficient gradient."""
import numpy as np


def frame(th, shape, zmin, zmax):
    Nz, Ny, Nx = shape
    du = (zmax - zmin) / Nz
    tx = th[:, 0] - Nx * 0.5
    ty = th[:, 1] - Ny * 0.5
    spx = -tx - np.trunc(-tx)
    spy = -ty - np.trunc(-ty)
    spz = (th[:, 2] - zmin) / du
    k = np.clip(np.trunc(spz).astype(int), 0, Nz - 1)
    spz = np.clip(spz - k, 0, 1)
    return tx, ty, spx, spy, spz, k, du


def basis(spx, spy, spz):
    """[N, 64] with q = 16 m + 4 n + o -> z^m y^n x^o."""
    m, n, o = np.meshgrid(np.arange(4), np.arange(4), np.arange(4), indexing="ij")
    return (spz[:, None] ** m.ravel()) * (spy[:, None] ** n.ravel()) * (spx[:, None] ** o.ravel())


def cells(th, R, shape, zmin, zmax):
    """flat cell index [N, R, R] of every pixel and the per-spot basis [N, 64]."""
    Nz, Ny, Nx = shape
    tx, ty, spx, spy, spz, k, du = frame(th, shape, zmin, zmax)
    y = np.arange(R)[None, :, None]
    x = np.arange(R)[None, None, :]
    j = np.clip(np.trunc(y - ty[:, None, None]).astype(int), 0, Ny - 1)
    i = np.clip(np.trunc(x - tx[:, None, None]).astype(int), 0, Nx - 1)
    cell = (k[:, None, None] * Ny + j) * Nx + i
    return cell, basis(spx, spy, spz)
