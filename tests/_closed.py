"""Independent FP64 closed forms used to pin the oracle and to make test data.

These are written from the mathematics of the pixel-integrated Gaussian
(mu = bg + I * DeltaX * DeltaY, DeltaX = 1/2 [erf((x - tx + 1/2)/(sqrt2 s)) -
erf((x - tx - 1/2)/(sqrt2 s))]) with scipy's erf, not from the oracle's code.
"""
import numpy as np
from scipy.special import erf


def _delta(coord, t, s):
    """[N, R] integrated Gaussian over pixels coord-1/2..coord+1/2."""
    c = coord[None, :] - t[:, None]
    s = np.asarray(s, dtype=np.float64)
    s = s[:, None] if s.ndim else s
    return 0.5 * (erf((c + 0.5) / (np.sqrt(2) * s)) - erf((c - 0.5) / (np.sqrt(2) * s)))


def gauss_mu(R, x, y, I, bg, sx, sy):
    x, y, I, bg = (np.asarray(v, np.float64) for v in (x, y, I, bg))
    sx = np.broadcast_to(np.asarray(sx, np.float64), x.shape)
    sy = np.broadcast_to(np.asarray(sy, np.float64), x.shape)
    coord = np.arange(R, dtype=np.float64)
    DX = _delta(coord, x, sx)
    DY = _delta(coord, y, sy)
    return bg[:, None, None] + I[:, None, None] * DY[:, :, None] * DX[:, None, :]


def astig_sigma(z, s0, gamma, d, A, B=0.0):
    """Reference parametrisation: s0 sqrt(1 + t^2/d^2 + A t^3/d^2 + B t^4/d^2), t = z - gamma."""
    t = np.asarray(z, np.float64) - gamma
    return s0 * np.sqrt(1 + t ** 2 / d ** 2 + A * t ** 3 / d ** 2 + B * t ** 4 / d ** 2)


def mu_fixed(R, theta, sx, sy=None):
    th = np.asarray(theta, np.float64)
    return gauss_mu(R, th[:, 0], th[:, 1], th[:, 2], th[:, 3], sx, sx if sy is None else sy)


def mu_sigma(R, theta):
    th = np.asarray(theta, np.float64)
    return gauss_mu(R, th[:, 0], th[:, 1], th[:, 2], th[:, 3], th[:, 4], th[:, 4])


def mu_astig(R, theta, cx, cy, Bx=0.0, By=0.0):
    th = np.asarray(theta, np.float64)
    sx = astig_sigma(th[:, 2], *cx[:4], B=Bx)
    sy = astig_sigma(th[:, 2], *cy[:4], B=By)
    return gauss_mu(R, th[:, 0], th[:, 1], th[:, 3], th[:, 4], sx, sy)


def fd_jacobian(mu_fn, theta, eps):
    """Central differences of mu_fn (FP64) w.r.t. each column of theta: [N, K, R, R]."""
    th = np.asarray(theta, np.float64)
    K = th.shape[1]
    eps = np.broadcast_to(np.asarray(eps, np.float64), (K,))
    cols = []
    for k in range(K):
        tp, tm = th.copy(), th.copy()
        tp[:, k] += eps[k]
        tm[:, k] -= eps[k]
        cols.append((mu_fn(tp) - mu_fn(tm)) / (2 * eps[k]))
    return np.stack(cols, 1)


# C3 astigmatism (sigma(z) = 1.3 sqrt(1 + u^2 + 0.1 u^3 + 0.05 u^4), u = (z -+ 200)/400 nm),
# in the reference's parametrisation {s0, gamma, d, A_ref = A/d} plus B_ref = B/d^2.
ASTIG_X = [1.3, 200.0, 400.0, 0.1 / 400.0]
ASTIG_Y = [1.3, -200.0, 400.0, 0.1 / 400.0]
ASTIG_B = 0.05 / 400.0 ** 2
ASTIG_ZRANGE = [-600.0, 600.0]


def synth_params(kind, R, n, rng):
    """Seeded true parameters following the benchmark workload definitions."""
    c = (R - 1) / 2
    x = c + rng.uniform(-1.5, 1.5, n)
    y = c + rng.uniform(-1.5, 1.5, n)
    if kind == "fixed":
        return np.stack([x, y, rng.uniform(500, 5000, n), rng.uniform(1, 20, n)], 1)
    if kind == "sigma":
        I = np.exp(rng.uniform(np.log(200), np.log(20000), n))
        return np.stack([x, y, I, rng.uniform(0.5, 30, n), rng.uniform(1.2, 2.0, n)], 1)
    if kind in ("astig", "spline"):
        z = rng.uniform(-400, 400, n)
        return np.stack([x, y, z, rng.uniform(500, 8000, n), rng.uniform(2, 20, n)], 1)
    raise ValueError(kind)


def spline_zstack(nz=64, ny=29, nx=29, dz=13.3, z0=-418.95, emitter=14.0):
    """Astigmatic Gaussian (I=1, bg=0) rendered as a [nz, ny, nx] z-stack with the
    emitter at pixel `emitter`, z(s) = z0 + dz*s (the spline workload's PSF)."""
    z = z0 + dz * np.arange(nz)
    th = np.zeros((nz, 5))
    th[:, 0] = emitter
    th[:, 1] = emitter
    th[:, 2] = z
    th[:, 3] = 1.0
    th[:, 4] = 0.0
    assert ny == nx
    return mu_astig(nx, th, ASTIG_X, ASTIG_Y, ASTIG_B, ASTIG_B), z


def mu_tilted(R, theta, sx, sy, tilt=True):
    """mu = bg + bgx (x - R/2) + bgy (y - R/2) + I DeltaX DeltaY for theta [x, y, I, bg, bgx, bgy, ...];
    with tilt=False the tilt terms are left out."""
    th = np.asarray(theta, np.float64)
    mu = gauss_mu(R, th[:, 0], th[:, 1], th[:, 2], th[:, 3], sx, sy)
    if tilt:
        c = np.arange(R, dtype=np.float64) - R / 2
        mu = mu + th[:, 4, None, None] * c[None, None, :] + th[:, 5, None, None] * c[None, :, None]
    return mu


def synth_variant(kind, R, n, rng):
    """Seeded true parameters and per-spot constants of the model variants:
    const_sigma, tilted_bg ([sx, sy] constants), sigma_xy, tilted_bg_sigma_xy, intensity_bg
    ([x, y, sigma] constants)."""
    c = (R - 1) / 2
    x = c + rng.uniform(-1.5, 1.5, n)
    y = c + rng.uniform(-1.5, 1.5, n)
    I = rng.uniform(500, 5000, n)
    bg = rng.uniform(2, 20, n)
    sx, sy = rng.uniform(1.2, 2.0, n), rng.uniform(1.2, 2.0, n)
    if kind == "const_sigma":
        return np.stack([x, y, I, bg], 1), np.stack([sx, sy], 1)
    if kind == "tilted_bg":
        # tilts small enough that mu stays positive over the ROI
        bgx, bgy = rng.uniform(-0.2, 0.2, n), rng.uniform(-0.2, 0.2, n)
        return np.stack([x, y, I, bg, bgx, bgy], 1), np.stack([sx, sy], 1)
    if kind == "sigma_xy":
        return np.stack([x, y, I, bg, sx, sy], 1), None
    if kind == "tilted_bg_sigma_xy":
        return np.stack([x, y, I, bg, np.zeros(n), np.zeros(n), sx, sy], 1), None
    if kind == "intensity_bg":
        return np.stack([I, bg], 1), np.stack([x, y, sx], 1)
    raise ValueError(kind)


def mu_variant(kind, R, theta, cst=None):
    th = np.asarray(theta, np.float64)
    if kind == "const_sigma":
        return gauss_mu(R, th[:, 0], th[:, 1], th[:, 2], th[:, 3], cst[:, 0], cst[:, 1])
    if kind == "tilted_bg":
        return mu_tilted(R, th, cst[:, 0], cst[:, 1])
    if kind == "sigma_xy":
        return gauss_mu(R, th[:, 0], th[:, 1], th[:, 2], th[:, 3], th[:, 4], th[:, 5])
    if kind == "tilted_bg_sigma_xy":
        return mu_tilted(R, th, th[:, 6], th[:, 7], tilt=False)
    if kind == "intensity_bg":
        return gauss_mu(R, cst[:, 0], cst[:, 1], th[:, 0], th[:, 1], cst[:, 2], cst[:, 2])
    raise ValueError(kind)
