"""EstimationQueue (estim_queue.py / EstimationQueue.cpp) on the GPU.

A queued spot is fitted by the same fused fit + CRLB kernels as
Estimator.estimate_device, so its estimate, iterations, CRLB and chi^2 must equal
the direct call's bit for bit, whatever batch and stream it landed in; the queue's
results are also checked against the oracle with the parity rules of tests/_parity.py.
"""
import numpy as np
import pytest

import photonpy_amd as pa
from tests import _models as MD
from tests import _parity as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    with pa.Context() as c:
        yield c


def _direct(est, smp, initial=None):
    import torch
    d = torch.from_numpy(smp).cuda()
    ini = None if initial is None else torch.from_numpy(initial).cuda()
    out = est.estimate_device(d, initial=ini, crlb=True)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


@pytest.mark.parametrize("kind", ["sigma", "spline"])
def test_queue_matches_direct_fit(ctx, kind):
    n = 5000 if kind == "sigma" else 1500
    th, smp = MD.samples(kind, n, 41)
    est, _, _ = MD.make(kind, ctx)
    est.SetLevMarParams(1e-18, 50)
    rng = np.random.default_rng(0)
    ids = rng.permutation(n).astype(np.int32)
    roipos = np.stack([ids, 2 * ids], 1).astype(np.int32)
    with pa.EstimQueue(est, batchSize=512, numStreams=3) as q:
        # small batches fit with more lanes per spot where compiled (R=11: 4, not 2)
        G = q.GroupSize()
        assert G == (4 if kind == "sigma" else est.GetGroupSize())
        est.SetGroupSize(G)
        ref = _direct(est, smp)
        pos = 0
        while pos < n:  # uneven chunks, so batches straddle Schedule calls
            m = int(rng.integers(1, 1200))
            q.Schedule(smp[pos:pos + m], roipos=roipos[pos:pos + m], ids=ids[pos:pos + m])
            pos += m
        q.WaitUntilDone()
        assert q.GetResultCount() == n
        r = q.GetResults()
        assert q.GetResultCount() == 0
    assert len(r.ids) == n
    r.SortByID(isUnique=True)
    order = np.argsort(ids)  # row of each id in the input
    assert np.array_equal(r.ids, np.arange(n))
    assert np.array_equal(r.roipos, roipos[order])
    assert np.array_equal(r.estim, ref["params"][order], equal_nan=True)
    assert np.array_equal(r.iterations, ref["iterations"][order])
    assert np.array_equal(r.crlb, ref["crlb"][order], equal_nan=True)
    assert np.array_equal(r.chisq, ref["chisq"][order], equal_nan=True)
    assert r.diagnostics.shape == (n, 0)
    # and the queue's results, back in input order, match the oracle (the reference's
    # queue fits each batch with the estimator's Estimate, EstimationQueue.cpp:277-299)
    p_in = np.empty_like(r.estim)
    it_in = np.empty_like(r.iterations)
    p_in[order] = r.estim
    it_in[order] = r.iterations
    _, om, _ = MD.make(kind, ctx)
    P.check_fit_parity(kind, om, smp, p_in, it_in, 50, 1e-18, est.ParamLimits(), tag=f"queue {kind}")


def test_queue_initial_values_and_samples(ctx):
    th, smp = MD.samples("fixed", 700, 43)
    est, _, _ = MD.make("fixed", ctx)
    initial = (th * 1.02).astype(np.float32)
    with pa.EstimQueue(est, batchSize=128, numStreams=2, keepSamples=True) as q:
        est.SetGroupSize(q.GroupSize())
        ref = _direct(est, smp, initial)
        q.Schedule(smp, ids=np.arange(700), initial=initial)
        q.WaitUntilDone()
        r = q.GetResults(getSampleData=True)
    r.SortByID(isUnique=True)
    assert np.array_equal(r.estim, ref["params"])
    assert np.array_equal(r.samples, smp)


def test_queue_partial_batches_idle_and_lengths(ctx):
    th, smp = MD.samples("sigma", 5, 45)
    est, _, _ = MD.make("sigma", ctx)
    with pa.EstimQueue(est, batchSize=4, numStreams=3) as q:
        q.Schedule(smp, ids=np.arange(5))
        # 4 spots form a full batch; the fifth waits in the batch being filled
        for _ in range(400):
            if q.IsIdle() and q.GetResultCount() == 4:
                break
            import time
            time.sleep(0.005)
        assert q.GetResultCount() == 4 and q.IsIdle()
        q.WaitUntilDone()
        assert q.GetResultCount() == 5
        # whole batches only: room for 3 spots takes the 1-spot batch (newest first), not the 4
        r1 = q.GetResults(maxResults=3)
        assert len(r1.ids) == 1 and r1.ids[0] == 4
        r2 = q.GetResults()
        assert sorted(r2.ids.tolist()) == [0, 1, 2, 3]
        assert q.GetQueueLength() == 0


def test_queue_deleted_with_its_context():
    with pa.Context() as c:
        est = pa.GaussianPSFMethods(c).CreatePSF_XYIBg(10, 1.5, True)
        q = pa.EstimQueue(est, batchSize=64)
        q.Schedule(np.ones((10, 10, 10), np.float32))
        q.inst = None  # the context owns it now
    # leaving the context deleted the queue before its estimator
