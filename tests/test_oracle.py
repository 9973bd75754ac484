"""Pins the CPU oracle (oracle/smlm_oracle.cpp) against what the mathematics fixes.

The reference ships no golden vectors or known-answer tests for this path
(its tests/ are plotting scripts; the one assert, tests/psf_queue.py:51, is
never called), so every oracle routine is pinned here by an independent
check chosen so that a plausible mistake (dropped term, wrong sign or index,
transposed operand) fails:

* expected values   vs scipy-erf closed forms (tests/_closed.py)
* Jacobians         vs central differences of those closed forms (the free-sigma
                    column is sigma x the derivative, as GaussianPSFModels.h:359,368 computes it)
* LM optimizer      noise-free fixed point; FP64 fit vs scipy minimisation of the Poisson NLL;
                    iteration-count semantics of Estimation.h:194-251
* CRLB / chi^2      vs numpy FP64 inverse of the closed-form Fisher matrix; continuous limit
* spline            CSpline_Compute vs numpy.linalg.solve of the 64x64 system; cubic
                    reproduction; model interpolates the z-stack at grid points
* log-likelihood    vs closed form
"""
import numpy as np
import pytest
from scipy.optimize import minimize

from oracle import oracle as O
from tests import _closed as C

RNG = np.random.default_rng(1234)


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30)


# ------------------------------------------------------------ expected values
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-6)])
def test_ev_fixed_closed_form(dtype, tol):
    R = 10
    th = C.synth_params("fixed", R, 50, RNG)
    m = O.gauss2d(R, 1.5, 1.7)
    ev = m.expected_value(th, dtype)
    ref = C.mu_fixed(R, th, 1.5, 1.7)
    assert _rel(ev, ref) < tol


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-6)])
def test_ev_sigma_closed_form(dtype, tol):
    R = 11
    th = C.synth_params("sigma", R, 50, RNG)
    ev = O.gauss2d_sigma(R, 1.6).expected_value(th, dtype)
    assert _rel(ev, C.mu_sigma(R, th)) < tol


@pytest.mark.parametrize("B", [0.0, C.ASTIG_B])
def test_ev_astig_closed_form(B):
    R = 16
    th = C.synth_params("astig", R, 50, RNG)
    m = O.astig(R, C.ASTIG_X, C.ASTIG_Y, C.ASTIG_ZRANGE, B=(B, B))
    ref = C.mu_astig(R, th, C.ASTIG_X, C.ASTIG_Y, B, B)
    assert _rel(m.expected_value(th, np.float64), ref) < 1e-12
    assert _rel(m.expected_value(th, np.float32), ref) < 5e-6


def test_ev_sums_to_photon_count():
    # a large ROI collects (almost) every photon: sum(mu - bg) -> I
    R = 32
    th = np.array([[15.5, 15.5, 1000.0, 0.0]])
    ev = O.gauss2d(R, 1.5).expected_value(th, np.float64)
    assert abs(ev.sum() - 1000.0) < 1e-6


# ------------------------------------------------------------- derivatives
def test_deriv_fixed_fd():
    R = 10
    th = C.synth_params("fixed", R, 20, RNG)
    dv, _ = O.gauss2d(R, 1.5).derivatives(th, np.float64)
    fd = C.fd_jacobian(lambda t: C.mu_fixed(R, t, 1.5), th, [1e-6, 1e-6, 1e-3, 1e-4])
    assert _rel(dv, fd) < 1e-6


def test_deriv_sigma_fd_and_reference_scaling():
    R = 11
    th = C.synth_params("sigma", R, 20, RNG)
    dv, _ = O.gauss2d_sigma(R, 1.6).derivatives(th, np.float64)
    fd = C.fd_jacobian(lambda t: C.mu_sigma(R, t), th, [1e-6, 1e-6, 1e-3, 1e-4, 1e-7])
    assert _rel(dv[:, :4], fd[:, :4]) < 1e-6
    # the reference's sigma column equals sigma * dmu/dsigma
    assert _rel(dv[:, 4], th[:, 4, None, None] * fd[:, 4]) < 1e-6


@pytest.mark.parametrize("B", [0.0, C.ASTIG_B])
def test_deriv_astig_fd(B):
    R = 16
    th = C.synth_params("astig", R, 20, RNG)
    m = O.astig(R, C.ASTIG_X, C.ASTIG_Y, C.ASTIG_ZRANGE, B=(B, B))
    dv, _ = m.derivatives(th, np.float64)
    fd = C.fd_jacobian(lambda t: C.mu_astig(R, t, C.ASTIG_X, C.ASTIG_Y, B, B), th, [1e-6, 1e-6, 1e-3, 1e-3, 1e-4])
    assert _rel(dv, fd) < 1e-6


# ------------------------------------------------------------ LM optimizer
@pytest.mark.parametrize("kind", ["fixed", "sigma", "astig"])
def test_lm_noise_free_fixed_point(kind):
    R = {"fixed": 10, "sigma": 11, "astig": 16}[kind]
    th = C.synth_params(kind, R, 200, RNG)
    if kind == "fixed":
        m, mu = O.gauss2d(R, 1.5), C.mu_fixed(R, th, 1.5)
    elif kind == "sigma":
        m, mu = O.gauss2d_sigma(R, 1.6), C.mu_sigma(R, th)
    else:
        m, mu = O.astig(R, C.ASTIG_X, C.ASTIG_Y, C.ASTIG_ZRANGE), C.mu_astig(R, th, C.ASTIG_X, C.ASTIG_Y)
    # start 1% (and 0.2 px) away from the truth; the noise-free optimum is the truth
    start = th * (1 + 0.01 * RNG.standard_normal(th.shape))
    start[:, :2] = th[:, :2] + 0.2
    p, it = m.estimate(mu, initial=start, iterations=200, lambda_=1e-15, dtype=np.float64)
    assert np.all(it <= 200)
    scale = np.ones(th.shape[1])
    if kind == "astig":
        scale[2] = 100.0  # z in nm
    err = np.abs(p - th) / np.maximum(np.abs(th), 1) / scale
    assert np.max(err) < 1e-4


def _nll(theta, mu_fn, smp):
    mu = np.maximum(mu_fn(theta[None])[0], 1e-8)
    x = np.maximum(smp, 1e-8)
    return np.sum(mu - x * np.log(mu))


@pytest.mark.parametrize("kind", ["fixed", "sigma"])
def test_lm_matches_scipy_mle(kind):
    R = 9
    rng = np.random.default_rng(7)
    th = C.synth_params(kind, R, 6, rng)
    if kind == "fixed":
        m, fn = O.gauss2d(R, 1.5), (lambda t: C.mu_fixed(R, t, 1.5))
    else:
        m, fn = O.gauss2d_sigma(R, 1.6), (lambda t: C.mu_sigma(R, t))
    smp = rng.poisson(fn(th)).astype(np.float64)
    p, it = m.estimate(smp, iterations=200, lambda_=1e-15, dtype=np.float64)
    for i in range(len(th)):
        res = minimize(_nll, p[i] * (1 + 1e-3), args=(fn, smp[i]), method="Nelder-Mead",
                       options={"xatol": 1e-10, "fatol": 1e-12, "maxiter": 40000, "maxfev": 80000})
        # the oracle's optimum is at least as good as scipy's and the two coincide
        assert _nll(p[i], fn, smp[i]) <= res.fun + 1e-5
        assert np.allclose(p[i], res.x, rtol=2e-4, atol=2e-4)


def test_lm_iteration_semantics():
    R = 10
    th = C.synth_params("fixed", R, 50, RNG)
    m = O.gauss2d(R, 1.5)
    smp = RNG.poisson(C.mu_fixed(R, th, 1.5)).astype(np.float32)
    # zero iterations: params are the initial value, iterations = 1 (i + 1 with i = 0)
    p0, it0 = m.estimate(smp, initial=th, iterations=0)
    assert np.all(it0 == 1) and np.allclose(p0, th.astype(np.float32))
    # max+1 when the stop never fires within the budget
    p2, it2 = m.estimate(smp, iterations=2)
    assert np.all(it2 <= 3) and np.any(it2 == 3)
    # the stop never fires at i = 0
    p1, it1 = m.estimate(smp, initial=th, iterations=30)
    assert np.all(it1 >= 2)
    # the trace holds the parameters before each step
    p3, it3, tr = m.estimate(smp, initial=th, iterations=30, trace_len=30)
    assert np.allclose(tr[:, 0], th.astype(np.float32))
    assert np.array_equal(p3, p1) and np.array_equal(it3, it1)


def test_lm_clamps():
    R = 10
    m = O.gauss2d(R, 1.5)
    smp = np.zeros((1, R, R), np.float32)
    smp[0, 0, 0] = 50  # all light in a corner: x,y run into the clamp box [1.5, R-2.5]
    p, _ = m.estimate(smp, iterations=30)
    # the clamp box holds even when the fit degenerates (fmaxf/fminf map NaN to the bound);
    # I and bg use `if (v < lo)` clamps, which let a NaN through as in the reference
    assert 1.5 <= p[0, 0] <= R - 2.5 and 1.5 <= p[0, 1] <= R - 2.5


# ---------------------------------------------------------------- CRLB
def test_crlb_vs_numpy_fisher():
    R = 11
    th = C.synth_params("sigma", R, 20, RNG)
    fn = lambda t: C.mu_sigma(R, t)
    fd = C.fd_jacobian(fn, th, [1e-6, 1e-6, 1e-3, 1e-4, 1e-7])
    fd[:, 4] *= th[:, 4, None, None]  # the reference's sigma column
    mu = fn(th)
    J = fd.reshape(len(th), 5, -1)
    FI = np.einsum("nks,nls->nkl", J / np.maximum(mu.reshape(len(th), 1, -1), 1e-8), J)
    ref = np.sqrt(np.diagonal(np.linalg.inv(FI), axis1=1, axis2=2))
    crlb, chi = O.gauss2d_sigma(R, 1.6).crlb(th, mu, dtype=np.float64)
    assert _rel(crlb, ref) < 1e-6
    assert np.max(np.abs(chi)) < 1e-18  # sample == mu


def test_crlb_continuous_limit():
    # bg = 0, large ROI: CRLB_x -> sqrt(sigma^2 + 1/12) / sqrt(I) for a pixelated Gaussian
    R, s, I = 41, 3.0, 1e4
    th = np.array([[20.0, 20.0, I, 0.0]])
    crlb, _ = O.gauss2d(R, s).crlb(th, dtype=np.float64)
    approx = np.sqrt(s ** 2 + 1 / 12) / np.sqrt(I)
    assert abs(crlb[0, 0] / approx - 1) < 0.01
    assert abs(crlb[0, 2] / np.sqrt(I) - 1) < 0.01  # Poisson count of I photons


def test_crlb_singular_is_inf():
    # zero intensity and zero background: the Fisher matrix is degenerate
    th = np.array([[4.5, 4.5, 0.0, 0.0]])
    crlb, _ = O.gauss2d(10, 1.5).crlb(th)
    assert np.all(np.isinf(crlb))


def test_chisq_definition():
    R = 10
    th = C.synth_params("fixed", R, 10, RNG)
    mu = C.mu_fixed(R, th, 1.5)
    smp = RNG.poisson(mu).astype(np.float64)
    _, chi = O.gauss2d(R, 1.5).crlb(th, smp, dtype=np.float64)
    ref = np.sum((smp - mu) ** 2 / mu, axis=(1, 2))
    assert _rel(chi, ref) < 1e-10


def test_loglik_closed_form():
    R = 10
    th = C.synth_params("fixed", R, 10, RNG)
    mu = C.mu_fixed(R, th, 1.5)
    smp = RNG.poisson(mu).astype(np.float64)
    ll = O.gauss2d(R, 1.5).loglik(th, smp, dtype=np.float64)
    ref = np.sum(np.maximum(smp, 1e-8) * np.log(mu) - mu, axis=(1, 2))
    assert _rel(ll, ref) < 1e-10


# ---------------------------------------------------------------- spline
def _spline_system():
    t = np.arange(4) - 1.0
    V = t[:, None] ** np.arange(4)[None, :]
    return np.kron(np.kron(V, V), V)  # rows (z,y,x) offsets, cols (z,y,x) powers


def test_spline_compute_vs_numpy_solve():
    data = RNG.normal(size=(6, 7, 8)).astype(np.float32)
    coefs = O.spline_compute(data)
    A = _spline_system()
    for (i, j, k) in [(0, 0, 0), (2, 3, 4), (2, 3, 0)]:
        b = data[i:i + 4, j:j + 4, k:k + 4].astype(np.float64).ravel()
        c = np.linalg.solve(A, b)
        assert np.allclose(coefs[i, j, k], c, rtol=1e-5, atol=1e-5)


def test_spline_reproduces_cubic_and_grid_values():
    z, y, x = np.meshgrid(np.arange(7.0), np.arange(6.0), np.arange(8.0), indexing="ij")
    f = 0.5 + 0.1 * x - 0.02 * y ** 2 + 0.003 * z ** 3 + 0.01 * x * y * z
    coefs = O.spline_compute(f.astype(np.float32))
    # cell (i,j,k) at fractional offset 0 equals data[i+1, j+1, k+1]
    assert np.allclose(coefs[..., 0], f[1:-2, 1:-2, 1:-2], atol=2e-5)
    # the cubic is reproduced exactly: evaluate every cell at offset (0.3, 0.6, 0.2)
    fz, fy, fx = 0.3, 0.6, 0.2
    basis = np.array([fz ** m * fy ** n * fx ** o for m in range(4) for n in range(4) for o in range(4)])
    zz, yy, xx = np.meshgrid(np.arange(4.0), np.arange(3.0), np.arange(5.0), indexing="ij")
    Z, Y, X = zz + 1 + fz, yy + 1 + fy, xx + 1 + fx
    ref = 0.5 + 0.1 * X - 0.02 * Y ** 2 + 0.003 * Z ** 3 + 0.01 * X * Y * Z
    assert np.allclose(coefs @ basis, ref, atol=5e-5)


def test_spline_model_interpolates_zstack():
    stack, zs = C.spline_zstack()
    coefs = O.spline_compute(stack)
    zmin, zmax = zs[1], zs[1] + 61 * 13.3
    R = 20
    m = O.spline(R, coefs, zmin, zmax)
    # integer sub-pixel offsets: pixel x samples data column x - tx + 14
    th = np.array([[9.0, 10.0, zs[30], 1.0, 0.0], [8.0, 9.0, zs[45], 1.0, 0.0]])
    ev = m.expected_value(th, np.float64)
    for n, s in enumerate([30, 45]):
        tx, ty = th[n, 0], th[n, 1]
        ref = stack[s, int(14 - ty):int(14 - ty) + R, int(14 - tx):int(14 - tx) + R]
        assert np.allclose(ev[n], np.maximum(ref, 0), atol=2e-6)
    # off-grid it tracks the rendered Gaussian closely
    th2 = np.array([[9.3, 9.6, 123.0, 1000.0, 5.0]])
    ref2 = C.mu_astig(R, th2 + np.array([[0, 0, 0, 0, 0]]), C.ASTIG_X, C.ASTIG_Y, C.ASTIG_B, C.ASTIG_B)
    assert np.max(np.abs(m.expected_value(th2, np.float64) - ref2)) < 2.0  # photons, of ~100 peak (cubic interpolation)


def test_spline_deriv_fd():
    stack, zs = C.spline_zstack()
    coefs = O.spline_compute(stack)
    m = O.spline(20, coefs, zs[1], zs[1] + 61 * 13.3)
    th = np.array([[9.3, 9.6, 123.0, 1000.0, 5.0], [8.7, 10.2, -211.0, 3000.0, 2.0]])
    dv, _ = m.derivatives(th, np.float64)
    fd = C.fd_jacobian(lambda t: m.expected_value(t, np.float64), th, [1e-5, 1e-5, 1e-3, 1e-3, 1e-3])
    # where the spline value is negative mu is clamped to bg, yet the reference keeps the
    # x/y/z derivatives (CubicSplinePSFModels.h:143-150): compare where the PSF is positive
    unit = th.copy()
    unit[:, 3], unit[:, 4] = 1.0, 0.0
    pos = (m.expected_value(unit, np.float64) > 0)[:, None]
    assert _rel(np.where(pos, dv, 0), np.where(pos, fd, 0)) < 1e-5
    assert np.all(dv[:, 3][~pos[:, 0]] == 0)
