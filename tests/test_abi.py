"""CPU checks of the C ABI and the Python host layer (no kernel launches).

The shapes, names, limits and ValueError behaviour follow the reference's
photonpy/cpp/estimator.py:27-396, gaussian.py:175-237, cspline.py:114-213 and
the factories in GaussianPSF.cu / CubicSplinePSF.cu.
"""
import ctypes as ct
import os
import subprocess

import numpy as np
import pytest

import photonpy_amd as pa
from photonpy_amd import lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    names = L.exported_symbols()
    assert "Estimator_Estimate" in names and "CSpline_Compute" in names
    out = subprocess.check_output(["nm", "-D", "--defined-only", L.LIB_PATH], text=True)
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    so = ct.CDLL(L.LIB_PATH)
    for n in names:
        getattr(so, n)


def test_reference_binding_layouts():
    # the reference's ctypes structs must match ours byte for byte
    assert ct.sizeof(pa.Gauss3D_Calibration) == 40
    assert ct.sizeof(pa.CSplineCalibration) == 32
    assert pa.CSplineCalibration.coefs.offset == 24
    assert ct.sizeof(L.EstimatorProperties) == 20 and ct.sizeof(L.LevMarParams) == 12


@pytest.fixture(scope="module")
def ctx():
    with pa.Context() as c:
        yield c


def test_fixed_properties(ctx):
    psf = pa.GaussianPSFMethods(ctx).CreatePSF_XYIBg(10, 1.5, True)
    assert psf.numparams == 4 and psf.samplecount == 100 and psf.numdiag == 0 and psf.numconst == 0
    assert list(psf.sampleshape) == [10, 10]
    assert psf.param_names == ["x", "y", "I", "bg"]
    lo, hi = psf.ParamLimits()
    assert np.allclose(lo, [1.5, 1.5, 10, 1e-6]) and np.allclose(hi, [7.5, 7.5, 1e9, 1e9])
    lm = psf.GetLevMarParams()
    assert lm.iterations == 50 and lm.lambda_ == np.float32(1e-18)
    psf.SetLevMarParams(1e-15, 30)
    assert psf.GetLevMarParams().iterations == 30
    psf.Destroy()


def test_sigma_and_astig_properties(ctx):
    g = pa.GaussianPSFMethods(ctx)
    s = g.CreatePSF_XYIBgSigma(11, 1.6, True)
    assert s.fmt == "x,y,\tI,bg,sigma" and s.param_names == ["x", "y", "I", "bg", "sigma"]
    assert np.allclose(s.ParamLimits()[1], [8.5, 8.5, 1e9, 1e9, 22])
    cal = pa.Gauss3D_Calibration([1.3, 200, 400, 2.5e-4], [1.3, -200, 400, 2.5e-4], [-600, 600])
    a = g.CreatePSF_XYZIBg(16, cal, True)
    assert a.param_names == ["x", "y", "z", "I", "bg"]
    assert np.allclose(a.ParamLimits()[0], [1.5, 1.5, -600, 10, 1e-6])
    q = pa.Gauss3D_Calibration([1.3, 200, 400, 2.5e-4, 3e-7], [1.3, -200, 400, 2.5e-4, 3e-7], [-600, 600])
    assert q.B == (3e-7, 3e-7)
    aq = g.CreatePSF_XYZIBg(16, q, True)
    assert aq.numparams == 5


def test_spline_properties_and_modes(ctx):
    coefs = np.zeros((5, 6, 6, 64), np.float32)
    m = pa.CSplineMethods(ctx)
    psf = m.CreatePSF_XYZIBg(12, pa.CSplineCalibration(-1.0, 1.0, coefs), m.FlatBg, True)
    assert psf.param_names == ["x", "y", "z", "I", "bg"]
    assert np.allclose(psf.ParamLimits()[0], [1.5, 1.5, -1, 10, 1e-6])
    with pytest.raises(pa.NativeError):
        m.CreatePSF_XYZIBg(12, pa.CSplineCalibration(-1.0, 1.0, coefs), 7, True)


def test_cuda_false_keeps_the_cpu_estimator_interface(ctx):
    # EstimatorImpl.h:120: the reference's CPU estimators report one diagnostics column
    g = pa.GaussianPSFMethods(ctx)
    cpu, gpu = g.CreatePSF_XYIBgSigma(11, 1.6, False), g.CreatePSF_XYIBgSigma(11, 1.6, True)
    assert cpu.numdiag == 1 and gpu.numdiag == 0
    assert cpu.numparams == gpu.numparams and cpu.fmt == gpu.fmt
    assert np.array_equal(cpu.limits, gpu.limits)


def test_shape_validation(ctx):
    psf = pa.GaussianPSFMethods(ctx).CreatePSF_XYIBg(10, 1.5, True)
    with pytest.raises(ValueError):
        psf.Estimate(np.zeros((3, 9, 10), np.float32))
    with pytest.raises(ValueError):
        psf.ExpectedValue(np.zeros((3, 5), np.float32))
    with pytest.raises(ValueError):
        psf.Estimate(np.zeros((3, 10, 10), np.float32), initial=np.zeros((3, 5), np.float32))
    with pytest.raises(ValueError):
        psf.CRLBAndChiSquare(np.zeros((3, 4)), samples=np.zeros((2, 10, 10)))
    with pytest.raises(ValueError):
        psf.ExpectedValue(np.zeros((3, 4)), roipos=np.zeros((3, 3), np.int32))
    # zero spots are a no-op, as in the reference (Estimator.cpp:105-106)
    p, diag, traces = psf.Estimate(np.zeros((0, 10, 10), np.float32))
    assert p.shape == (0, 4) and diag.shape == (0, 0) and traces == []
    assert psf.ExpectedValue(np.zeros((0, 4))).shape == (0, 10, 10)


def test_group_size_control(ctx):
    psf = pa.GaussianPSFMethods(ctx).CreatePSF_XYIBg(11, 1.5, True)
    assert psf.GetGroupSize() == 2
    psf.SetGroupSize(4)
    assert psf.GetGroupSize() == 4
    with pytest.raises(ValueError):
        psf.SetGroupSize(3)
    psf.SetGroupSize(0)
    assert psf.GetGroupSize() == 2


def test_context_destroys_estimators():
    c = pa.Context()
    g = pa.GaussianPSFMethods(c)
    for _ in range(3):
        g.CreatePSF_XYIBg(10, 1.5, True)
    c.Destroy()


def test_estimator_result_layout():
    from photonpy_amd.lib import EstimResultDType
    assert EstimResultDType.itemsize == 12
    assert [EstimResultDType.fields[n][1] for n in ("id", "chisq", "iterations")] == [0, 4, 8]


def test_queue_results_filtering():
    from photonpy_amd.estim_queue import EstimQueue_Results
    n = 6
    ids = np.array([3, 1, 5, 0, 4, 2], np.int32)
    est = np.arange(n * 4, dtype=np.float32).reshape(n, 4)
    r = EstimQueue_Results(["x", "y", "I", "bg"], (10, 10), est.copy(), np.zeros((n, 0), np.float32),
                           est.copy() * 2, np.arange(n), np.arange(n) * 0.5, np.zeros((n, 2), np.int32), None, ids)
    r.SortByID(isUnique=True)
    assert np.array_equal(r.ids, np.arange(n))
    assert np.array_equal(r.estim, est[np.argsort(ids)])
    r.Filter(r.estim[:, 0] > 8)
    assert np.all(r.estim[:, 0] > 8) and len(r.ids) == len(r.chisq) == len(r.crlb)
    assert r.ColIdx("I") == 2
