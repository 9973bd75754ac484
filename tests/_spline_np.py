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


def lagrange(t):
    """[N, 4] weights of the cubic through the nodes -1, 0, 1, 2 at t, and their derivatives."""
    t = np.asarray(t, np.float64)[:, None]
    nodes = np.array([-1.0, 0.0, 1.0, 2.0])
    w = np.ones((t.shape[0], 4))
    dw = np.zeros((t.shape[0], 4))
    for m in range(4):
        others = [nodes[j] for j in range(4) if j != m]
        den = np.prod([nodes[m] - o for o in others])
        w[:, m] = np.prod([t[:, 0] - o for o in others], axis=0) / den
        dw[:, m] = sum(np.prod([t[:, 0] - others[k] for k in range(3) if k != i], axis=0) for i in range(3)) / den
    return w, dw


def grid_from_coefs(coefs):
    """Samples [Nz+3, Ny+3, Nx+3] read off the cell polynomials: sample d along an axis is the
    value of cell clamp(d-1, 0, N-1) at its local node d - cell - 1."""
    Nz, Ny, Nx = coefs.shape[:3]
    C = coefs.astype(np.float64).reshape(Nz, Ny, Nx, 4, 4, 4)

    def axis(N):
        d = np.arange(N + 3)
        c = np.clip(d - 1, 0, N - 1)
        return c, (d - c - 1).astype(np.float64)

    cz, tz = axis(Nz)
    cy, ty = axis(Ny)
    cx, tx = axis(Nx)
    P = lambda t: t[:, None] ** np.arange(4)[None, :]
    sub = C[cz][:, cy][:, :, cx]  # [Dz, Dy, Dx, 4, 4, 4]
    return np.einsum("zyxmno,zm,yn,xo->zyx", sub, P(tz), P(ty), P(tx))


def grid_eval(th, R, grid, shape, zmin, zmax):
    """mu / psf of the spline at every pixel from the sample grid (FP64)."""
    Nz, Ny, Nx = shape
    tx, ty, spx, spy, spz, k, du = frame(th, shape, zmin, zmax)
    wx, _ = lagrange(spx)
    wy, _ = lagrange(spy)
    wz, _ = lagrange(spz)
    y = np.arange(R)[None, :, None]
    x = np.arange(R)[None, None, :]
    j = np.clip(np.trunc(y - ty[:, None, None]).astype(int), 0, Ny - 1)
    i = np.clip(np.trunc(x - tx[:, None, None]).astype(int), 0, Nx - 1)
    f = np.zeros((len(th), R, R))
    for m in range(4):
        for n in range(4):
            for q in range(4):
                g = grid[(k[:, None, None] + m), j + n, i + q]
                f += g * (wz[:, m, None, None] * wy[:, n, None, None] * wx[:, q, None, None])
    return f
