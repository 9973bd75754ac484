"""ROI sizes with no compiled per-size fit kernel (roisize < 3 or > 32).

The reference's models take any roisize (GaussianPSFModels.h:76-97, EstimatorImpl.h:97-104);
these sizes run on the any-size fit kernel (one wavefront per spot).  Checked against the
oracle with the parity rules of tests/_parity.py, plus the CRLB kernel and a mixed batch
that holds such a bin.
"""
import numpy as np
import pytest

import photonpy_amd as pa
from photonpy_amd.batch import MixedBatch
from tests import _models as MD
from tests import _parity as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    with pa.Context() as c:
        yield c


@pytest.mark.parametrize("kind,R", [("fixed", 40), ("sigma", 36), ("astig", 48), ("spline", 34), ("fixed", 33)])
def test_large_roi_estimate_matches_oracle(ctx, kind, R):
    n = 400
    # a large ROI starts far from the optimum (I from 0.7 x the ROI sum, background
    # included): fixed sigma gets 100 iterations so that most spots converge
    lam, iters = (1e-15, 100) if kind == "fixed" else (1e-18, 50)
    th, smp = MD.samples(kind, n, 60 + R, R=R)
    est, om, _ = MD.make(kind, ctx, R=R)
    est.SetLevMarParams(lam, iters)
    p_gpu, _, traces = est.Estimate(smp)
    it_gpu = est.last_iterations
    assert np.array_equal([len(t) for t in traces], np.minimum(it_gpu, iters))
    P.check_fit_parity(kind, om, smp, p_gpu, it_gpu, iters, lam, est.ParamLimits(), tag=f"{kind} R={R}",
                       min_witnesses=40)
    # fused CRLB equals the stand-alone CRLB call
    import torch
    out = est.estimate_device(torch.from_numpy(smp).cuda(), crlb=True)
    torch.cuda.synchronize()
    p = out["params"].cpu().numpy()
    crlb, chi = est.CRLBAndChiSquare(p, smp)
    assert np.array_equal(out["crlb"].cpu().numpy(), crlb, equal_nan=True)
    assert np.array_equal(out["chisq"].cpu().numpy(), chi, equal_nan=True)


def test_tiny_roi_follows_the_clamp_semantics(ctx):
    """roisize 2: the x/y clamp interval [1.5, R - 2.5] is empty, so every step puts x and
    y at 1.5, as the reference's fmaxf(1.5, fminf(v, -0.5)) does.  With 4 pixels for 4
    parameters the normal equations are near-singular: where an FP32 Cholesky meets a zero
    pivot the fit stops at the centre-of-mass start, and rounding decides which spots that
    is, so only the clamp of the spots that stepped is compared."""
    R, n = 2, 200
    rng = np.random.default_rng(5)
    smp = rng.poisson(50.0, (n, R, R)).astype(np.float32)
    est, om, _ = MD.make("fixed", ctx, R=R)
    est.SetLevMarParams(1e-15, 30)
    p_gpu, _, _ = est.Estimate(smp)
    p_ref, it_ref = om.estimate(smp, iterations=30, lambda_=1e-15)
    stepped = it_ref > 1
    assert stepped.any()
    assert np.all(p_ref[stepped, :2] == 1.5)
    assert np.all(est.last_iterations >= 1)
    stepped_gpu = est.last_iterations > 1
    assert stepped_gpu.any()
    assert np.all(p_gpu[stepped_gpu, :2] == 1.5)


def test_mixed_batch_with_a_large_bin(ctx):
    import torch
    ests, samples, counts = [], [], []
    for kind, R in (("sigma", 11), ("fixed", 40), ("spline", 34)):
        th, smp = MD.samples(kind, 300, 70 + R, R=R)
        est, _, _ = MD.make(kind, ctx, R=R)
        est.SetLevMarParams(1e-18, 30)
        ests.append(est)
        samples.append(torch.from_numpy(smp).cuda())
        counts.append(300)
    with MixedBatch(ests, counts, streams=2, tune=False) as mb:
        ref = mb.allocate_outputs()
        mb.run_plain(samples, ref)
        outs = mb.allocate_outputs()
        mb.run(samples, outs, use_graph=True)
        torch.cuda.synchronize()
    for a, b in zip(ref, outs):
        for k in a:
            assert np.array_equal(a[k].cpu().numpy(), b[k].cpu().numpy(), equal_nan=True), k
