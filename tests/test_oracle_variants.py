"""Pins the oracle's model variants (GaussianPSFModels.h:155-623) against the mathematics.

* expected values  vs the scipy-erf closed forms of tests/_closed.py, with per-spot
                   constants (sigma x/y; x, y, sigma for the intensity/background model)
* Jacobians        vs central differences; the sigma columns of the SigmaXY models are
                   sigma x dmu/dsigma (GaussianPSFModels.h:459-460), and the tilted-background
                   SigmaXY model's mu leaves the tilt out while its Jacobian carries the tilt
                   columns (x - R/2, y - R/2) (GaussianPSFModels.h:542-548)
* LM optimizer     noise-free fixed point; FP64 fit vs scipy minimisation of the Poisson NLL
* CRLB             vs the numpy inverse of the FD Fisher matrix; the intensity/background
                   CRLB vs the closed-form 2x2 inverse
"""
import numpy as np
import pytest
from scipy.optimize import minimize

from oracle import oracle as O
from tests import _closed as C

RNG = np.random.default_rng(4321)
R = 11
ISX, ISY = 1.5, 1.7  # initial sigma of the SigmaXY models
KINDS = ["const_sigma", "tilted_bg", "sigma_xy", "tilted_bg_sigma_xy", "intensity_bg"]


def model(kind):
    return {
        "const_sigma": lambda: O.gauss2d_const_sigma(R),
        "tilted_bg": lambda: O.gauss2d_tilted_bg(R),
        "sigma_xy": lambda: O.gauss2d_sigma_xy(R, ISX, ISY),
        "tilted_bg_sigma_xy": lambda: O.gauss2d_tilted_bg_sigma_xy(R, ISX, ISY),
        "intensity_bg": lambda: O.intensity_bg(R),
    }[kind]()


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30)


EPS = {
    "const_sigma": [1e-6, 1e-6, 1e-3, 1e-4],
    "tilted_bg": [1e-6, 1e-6, 1e-3, 1e-4, 1e-6, 1e-6],
    "sigma_xy": [1e-6, 1e-6, 1e-3, 1e-4, 1e-7, 1e-7],
    "tilted_bg_sigma_xy": [1e-6, 1e-6, 1e-3, 1e-4, 1e-6, 1e-6, 1e-7, 1e-7],
    "intensity_bg": [1e-3, 1e-4],
}


def test_param_and_constant_counts():
    want = {"const_sigma": (4, 2), "tilted_bg": (6, 2), "sigma_xy": (6, 0), "tilted_bg_sigma_xy": (8, 0),
            "intensity_bg": (2, 3)}
    for k, (K, NC) in want.items():
        m = model(k)
        assert (m.K, m.NC) == (K, NC), k


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 3e-6)])
def test_ev_closed_form(kind, dtype, tol):
    th, cst = C.synth_variant(kind, R, 40, RNG)
    if kind == "tilted_bg_sigma_xy":
        th[:, 4:6] = RNG.uniform(-0.2, 0.2, (len(th), 2))  # the tilt does not enter mu
    ev = model(kind).expected_value(th, dtype, constants=cst)
    assert _rel(ev, C.mu_variant(kind, R, th, cst)) < tol


def test_constants_are_per_spot():
    th, cst = C.synth_variant("const_sigma", R, 2, RNG)
    th[1] = th[0]
    cst[1] = cst[0] * 1.2
    ev = model("const_sigma").expected_value(th, np.float64, constants=cst)
    assert not np.allclose(ev[0], ev[1])
    with pytest.raises(ValueError):
        model("const_sigma").expected_value(th)


@pytest.mark.parametrize("kind", KINDS)
def test_deriv_fd(kind):
    th, cst = C.synth_variant(kind, R, 20, RNG)
    dv, ev = model(kind).derivatives(th, np.float64, constants=cst)
    fd = C.fd_jacobian(lambda t: C.mu_variant(kind, R, t, cst), th, EPS[kind])
    if kind == "sigma_xy":
        fd[:, 4:6] *= th[:, 4:6, None, None]  # the reference's sigma columns
    if kind == "tilted_bg_sigma_xy":
        fd[:, 6:8] *= th[:, 6:8, None, None]
        # mu has no tilt, so the FD tilt columns vanish; the Jacobian carries them anyway
        assert np.max(np.abs(fd[:, 4:6])) < 1e-9
        c = np.arange(R) - R / 2
        assert np.array_equal(dv[:, 4], np.broadcast_to(c[None, :], (len(th), R, R)))
        assert np.array_equal(dv[:, 5], np.broadcast_to(c[:, None], (len(th), R, R)))
        dv, fd = dv[:, [0, 1, 2, 3, 6, 7]], fd[:, [0, 1, 2, 3, 6, 7]]
    assert _rel(dv, fd) < 1e-6
    assert _rel(ev, C.mu_variant(kind, R, th, cst)) < 1e-12


@pytest.mark.parametrize("kind", KINDS)
def test_lm_noise_free_fixed_point(kind):
    th, cst = C.synth_variant(kind, R, 100, RNG)
    mu = C.mu_variant(kind, R, th, cst)
    start = th * (1 + 0.01 * RNG.standard_normal(th.shape))
    if kind != "intensity_bg":
        start[:, :2] = th[:, :2] + 0.2
    if kind == "tilted_bg":
        start[:, 4:6] = 0.0
    p, it = model(kind).estimate(mu, initial=start, iterations=300, lambda_=1e-15, dtype=np.float64, constants=cst)
    err = np.abs(p - th) / np.maximum(np.abs(th), 1)
    if kind == "tilted_bg_sigma_xy":
        err = err[:, [0, 1, 2, 3, 6, 7]]  # the tilt does not enter mu and is left free
    assert np.max(err) < 1e-4


def test_intensity_bg_default_start():
    # Gauss2D_EstimateIntensityBg starts every spot at I = 1, bg = 0 (GaussianPSF.cu:230)
    th, cst = C.synth_variant("intensity_bg", R, 50, RNG)
    mu = C.mu_variant("intensity_bg", R, th, cst)
    p, _ = model("intensity_bg").estimate(mu, iterations=200, lambda_=1e-15, dtype=np.float64, constants=cst)
    assert np.max(np.abs(p - th) / th) < 1e-4
    _, it0 = model("intensity_bg").estimate(mu, iterations=0, dtype=np.float64, constants=cst)
    p0, _ = model("intensity_bg").estimate(mu, iterations=0, dtype=np.float64, constants=cst)
    # the start is not clamped before the first step (Estimation.h:200), so bg stays 0
    assert np.all(it0 == 1) and np.all(p0[:, 0] == 1.0) and np.all(p0[:, 1] == 0.0)


def _nll(theta, mu_fn, smp):
    mu = np.maximum(mu_fn(theta[None])[0], 1e-8)
    x = np.maximum(smp, 1e-8)
    return np.sum(mu - x * np.log(mu))


@pytest.mark.parametrize("kind", ["tilted_bg", "sigma_xy", "intensity_bg"])
def test_lm_matches_scipy_mle(kind):
    rng = np.random.default_rng(11)
    th, cst = C.synth_variant(kind, R, 4, rng)
    smp = rng.poisson(C.mu_variant(kind, R, th, cst)).astype(np.float64)
    p, _ = model(kind).estimate(smp, iterations=300, lambda_=1e-15, dtype=np.float64, constants=cst)
    for i in range(len(th)):
        ci = None if cst is None else cst[i:i + 1]
        fn = lambda t: C.mu_variant(kind, R, t, ci)
        if kind == "sigma_xy":
            # the reference's sigma column is sigma x dmu/dsigma, so its LM does not find
            # the MLE in sigma exactly; compare with the MLE over x, y, I, bg at its sigma
            f4 = lambda t4: fn(np.concatenate([t4[0], p[i, 4:]])[None])
            res = minimize(lambda t4: _nll(t4, lambda a: f4(a), smp[i]), p[i, :4] * (1 + 1e-3), method="Nelder-Mead",
                           options={"xatol": 1e-10, "fatol": 1e-12, "maxiter": 40000, "maxfev": 80000})
            assert np.allclose(p[i, :4], res.x, rtol=2e-4, atol=2e-4)
            continue
        res = minimize(_nll, p[i] * (1 + 1e-3), args=(fn, smp[i]), method="Nelder-Mead",
                       options={"xatol": 1e-10, "fatol": 1e-12, "maxiter": 40000, "maxfev": 80000})
        assert _nll(p[i], fn, smp[i]) <= res.fun + 1e-5
        assert np.allclose(p[i], res.x, rtol=2e-4, atol=2e-4)


@pytest.mark.parametrize("kind", ["const_sigma", "tilted_bg", "sigma_xy"])
def test_crlb_vs_numpy_fisher(kind):
    th, cst = C.synth_variant(kind, R, 20, RNG)
    mu = C.mu_variant(kind, R, th, cst)
    fd = C.fd_jacobian(lambda t: C.mu_variant(kind, R, t, cst), th, EPS[kind])
    if kind == "sigma_xy":
        fd[:, 4:6] *= th[:, 4:6, None, None]
    K = th.shape[1]
    J = fd.reshape(len(th), K, -1)
    FI = np.einsum("nks,nls->nkl", J / np.maximum(mu.reshape(len(th), 1, -1), 1e-8), J)
    ref = np.sqrt(np.diagonal(np.linalg.inv(FI), axis1=1, axis2=2))
    crlb, chi = model(kind).crlb(th, mu, dtype=np.float64, constants=cst)
    assert _rel(crlb, ref) < 1e-6
    assert np.max(np.abs(chi)) < 1e-18


def test_intensity_bg_crlb_closed_form():
    th, cst = C.synth_variant("intensity_bg", R, 20, RNG)
    psf = C.mu_variant("intensity_bg", R, np.stack([np.ones(20), np.zeros(20)], 1), cst).reshape(20, -1)
    mu = th[:, 0:1] * psf + th[:, 1:2]
    a, b, d = (psf * psf / mu).sum(1), (psf / mu).sum(1), (1 / mu).sum(1)
    det = a * d - b * b
    ref = np.stack([np.sqrt(d / det), np.sqrt(a / det)], 1)
    crlb, _ = model("intensity_bg").crlb(th, dtype=np.float64, constants=cst)
    assert _rel(crlb, ref) < 1e-9


# ------------------------------------------------------- spline background modes
# CSpline_Model_XYZITiltedBg (CubicSplinePSFModels.h:157-298) is the flat-background
# spline plus bgx (x - R/2) + bgy (y - R/2); CSpline_Model_XYZIBgImage (:300-443)
# places the spline one pixel further along x and y ((t - R/2) + (R - N + 1)/2 + 1/2
# = t - N/2 + 1) and scales a per-spot background image by bg.  Both reduce to the
# flat model, which the spline tests of test_oracle.py pin against the z-stack.
from tests import _models as MD  # noqa: E402

RS = 20


def _spline_models():
    coefs, zmin, zmax = MD.spline_coefs()
    return [O.spline(RS, coefs, zmin, zmax, mode=m) for m in (0, 1, 2)]


def _spline_theta(n, rng, K=5):
    th = np.zeros((n, K))
    th[:, :2] = 9.5 + rng.uniform(-1.5, 1.5, (n, 2))
    th[:, 2] = rng.uniform(-300, 300, n)
    th[:, 3] = rng.uniform(1000, 5000, n)
    th[:, 4] = rng.uniform(2, 10, n)
    if K == 7:
        th[:, 5:7] = rng.uniform(-0.1, 0.1, (n, 2))
    return th


def test_spline_tilted_is_flat_plus_tilt():
    flat, tilted, _ = _spline_models()
    assert (tilted.K, tilted.NC) == (7, 0)
    th = _spline_theta(30, RNG, 7)
    dv, ev = tilted.derivatives(th, np.float64)
    dvf, evf = flat.derivatives(th[:, :5], np.float64)
    c = np.arange(RS) - RS / 2
    tilt = th[:, 5, None, None] * c[None, None, :] + th[:, 6, None, None] * c[None, :, None]
    assert _rel(ev, evf + tilt) < 1e-12
    assert np.array_equal(dv[:, :5], dvf)
    assert np.array_equal(dv[:, 5], np.broadcast_to(c[None, :], (30, RS, RS)))
    assert np.array_equal(dv[:, 6], np.broadcast_to(c[:, None], (30, RS, RS)))


def test_spline_bgimage_is_shifted_flat_times_image():
    flat, _, image = _spline_models()
    assert (image.K, image.NC) == (5, RS * RS)
    th = _spline_theta(30, RNG)
    b = RNG.uniform(0.5, 1.5, (30, RS * RS))
    dv, ev = image.derivatives(th, np.float64, constants=b)
    sh = th.copy()
    sh[:, :2] += 1.0  # one pixel further
    sh[:, 4] = 0.0
    dvf, evf = flat.derivatives(sh, np.float64)
    bimg = b.reshape(30, RS, RS)
    assert _rel(ev, evf + th[:, 4, None, None] * bimg) < 1e-10
    assert _rel(dv[:, :4], dvf[:, :4]) < 1e-10
    assert np.array_equal(dv[:, 4], bimg)
    with pytest.raises(ValueError):
        image.expected_value(th)


@pytest.mark.parametrize("mode", [1, 2])
def test_spline_variant_noise_free_fixed_point(mode):
    m = _spline_models()[mode]
    rng = np.random.default_rng(17)
    th = _spline_theta(40, rng, 7 if mode == 1 else 5)
    cst = rng.uniform(0.5, 1.5, (40, RS * RS)) if mode == 2 else None
    mu = m.expected_value(th, np.float64, constants=cst)
    start = th.copy()
    start[:, :2] += 0.2
    start[:, 2] += 30
    start[:, 3] *= 1.05
    if mode == 1:
        start[:, 5:7] = 0
    p, _ = m.estimate(mu, initial=start, iterations=200, lambda_=1e-15, dtype=np.float64, constants=cst)
    err = np.abs(p - th) / np.maximum(np.abs(th), 1)
    assert np.max(err) < 1e-4


@pytest.mark.parametrize("mode", [1, 2])
def test_spline_variant_crlb_vs_numpy_fisher(mode):
    m = _spline_models()[mode]
    rng = np.random.default_rng(19)
    th = _spline_theta(10, rng, 7 if mode == 1 else 5)
    cst = rng.uniform(0.5, 1.5, (10, RS * RS)) if mode == 2 else None
    dv, mu = m.derivatives(th, np.float64, constants=cst)
    K = th.shape[1]
    J = dv.reshape(10, K, -1)
    FI = np.einsum("nks,nls->nkl", J / np.maximum(mu.reshape(10, 1, -1), 1e-8), J)
    ref = np.sqrt(np.diagonal(np.linalg.inv(FI), axis1=1, axis2=2))
    crlb, _ = m.crlb(th, dtype=np.float64, constants=cst)
    assert _rel(crlb, ref) < 1e-8
