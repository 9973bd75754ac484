"""Spline file IO and CSpline_Evaluate on the GPU, with the tetrapod z-stack fixture
(photonpy/examples/tetrapod-psf.zstack)."""
import ctypes as ct
import os

import numpy as np
import pytest

import photonpy_amd as pa
from photonpy_amd.psf_io import load_zstack, psf_to_zstack

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "tetrapod-psf.zstack")


@pytest.fixture(scope="module")
def ctx():
    with pa.Context() as c:
        yield c


def test_psf_from_file_matches_psf_from_zstack(ctx):
    m = pa.CSplineMethods(ctx)
    zs, zr = load_zstack(FIXTURE)
    a = m.CreatePSFFromFile(20, FIXTURE)
    b = m.CreatePSFFromZStack(20, zs, zr)
    assert a.spline_form() == "grid"
    lib = ctx.smlm.lib
    c_inst = lib.CSpline_CreatePSF_FromFile(20, FIXTURE.encode(), 0, True, ctx.inst)
    assert c_inst
    c = pa.CSplinePSF(ctx, c_inst, None)
    rng = np.random.default_rng(2)
    th = np.zeros((50, 5), np.float32)
    th[:, :2] = 9.5 + rng.uniform(-1.5, 1.5, (50, 2))
    th[:, 2] = rng.uniform(-1.2, 1.2, 50)
    th[:, 3] = 1000
    th[:, 4] = 2
    ev = b.ExpectedValue(th)
    assert np.array_equal(a.ExpectedValue(th), ev)
    assert np.array_equal(c.ExpectedValue(th), ev)
    # the limits of a file PSF come from the file's z range
    assert np.allclose(c.ParamLimits()[0][2], zr[0]) and np.allclose(c.ParamLimits()[1][2], zr[1])


def test_missing_file_raises(ctx):
    lib = ctx.smlm.lib
    assert not lib.CSpline_CreatePSF_FromFile(20, b"/nonexistent.zstack", 0, True, ctx.inst)
    assert "Can't open" in lib.SMLM_GetLastError().decode()
    # the error stays set until cleared, as documented for SMLM_GetLastError
    lib.SMLM_ClearError()
    assert lib.SMLM_GetLastError() == b""


def test_evaluate_reproduces_samples_and_matches_numpy(ctx):
    m = pa.CSplineMethods(ctx)
    zs, _ = load_zstack(FIXTURE)
    coefs = m.ComputeSplineCoeff(zs)
    v0 = m.Evaluate(coefs, [0, 0, 0])
    # zero shift: every voxel is its cell's node 0, i.e. sample (z+1, y+1, x+1)
    assert v0.shape == (*coefs.shape[:3], 4)
    assert np.max(np.abs(v0[..., 0] - zs[1:-2, 1:-2, 1:-2])) <= 1e-6 * zs.max()
    # a shifted evaluation against an FP64 numpy restatement of CSpline_Evaluate
    shift = np.array([0.3, -0.45, 1.7], np.float32)
    v = m.Evaluate(coefs, shift, border=1)
    d = -shift - np.trunc(-shift)
    P = lambda t: np.array([1, t, t * t, t ** 3], np.float64)
    dP = lambda t: np.array([0, 1, 2 * t, 3 * t * t], np.float64)
    C = coefs.astype(np.float64).reshape(*coefs.shape[:3], 4, 4, 4)
    S = np.array(coefs.shape[:3])
    ref = np.zeros_like(v, dtype=np.float64)
    for z in range(1, S[0] - 1):
        k = min(max(int(np.float32(z) - shift[0]), 0), S[0] - 1)
        for y in range(1, S[1] - 1):
            j = min(max(int(np.float32(y) - shift[1]), 0), S[1] - 1)
            for x in range(1, S[2] - 1):
                i = min(max(int(np.float32(x) - shift[2]), 0), S[2] - 1)
                c = C[k, j, i]
                o = ref[z - 1, y - 1, x - 1]
                o[0] = np.einsum("mno,m,n,o->", c, P(d[0]), P(d[1]), P(d[2]))
                o[1] = np.einsum("mno,m,n,o->", c, P(d[0]), P(d[1]), dP(d[2]))
                o[2] = np.einsum("mno,m,n,o->", c, P(d[0]), dP(d[1]), P(d[2]))
                o[3] = np.einsum("mno,m,n,o->", c, dP(d[0]), P(d[1]), P(d[2]))
    scale = np.abs(ref).max(axis=(0, 1, 2))
    assert np.all(np.abs(v - ref).max(axis=(0, 1, 2)) <= 2e-6 * scale)


def test_fit_on_file_psf_recovers_noise_free_truth(ctx):
    psf = pa.CSplineMethods(ctx).CreatePSFFromFile(24, FIXTURE)
    rng = np.random.default_rng(5)
    n = 2000
    th = np.zeros((n, 5), np.float32)
    th[:, :2] = 11.5 + rng.uniform(-1.5, 1.5, (n, 2))
    th[:, 2] = rng.uniform(-1.0, 1.0, n)
    th[:, 3] = rng.uniform(1000, 5000, n)
    th[:, 4] = rng.uniform(1, 10, n)
    mu = psf.ExpectedValue(th)
    start = th.copy()
    start[:, :2] += 0.2
    start[:, 2] += 0.05
    start[:, 3] *= 1.05
    psf.SetLevMarParams(1e-18, 100)
    p, _, _ = psf.Estimate(mu, initial=start)
    err = np.abs(p - th)
    ok = (err[:, 0] < 1e-3) & (err[:, 1] < 1e-3) & (err[:, 2] < 1e-3) & (err[:, 3] < 1e-3 * th[:, 3])
    assert ok.mean() > 0.99, ok.mean()
    # psf_to_zstack renders the PSF back on its own ROI
    zst = psf_to_zstack(psf, np.linspace(-1, 1, 5))
    assert zst.shape == (5, 24, 24) and np.all(np.isfinite(zst))
