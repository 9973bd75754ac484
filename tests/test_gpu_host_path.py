"""The host-pointer Estimator_Estimate on calls large enough for its chunked upload
(samples copied through pinned staging buffers in chunks of up to 32 MB, each chunk
fitted while the next is copied): parameters and iteration counts equal, bit for
bit, those of one device-pointer fit over the same spots, with and without per-spot
constants and initial values."""
import numpy as np
import pytest
import torch

import photonpy_amd as pa
from tests import _models as MD

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    with pa.Context() as c:
        yield c


def host_estimate(est, smp, constants=None, initial=None):
    """Estimator_Estimate with host buffers and no trace (the C-ABI call a binding makes)."""
    n = len(smp)
    K = est.numparams
    params = np.zeros((n, K), np.float32)
    iters = np.zeros(n, np.int32)
    diag = np.zeros((n, 0), np.float32)
    cons = np.zeros((n, 0), np.float32) if constants is None else np.ascontiguousarray(constants, np.float32)
    rp = np.zeros((n, 2), np.int32)
    init = None if initial is None else np.ascontiguousarray(initial, np.float32)
    est._lib.Estimator_Estimate(est.inst, n, np.ascontiguousarray(smp, np.float32), cons, rp, init, params, diag,
                                iters, None, 0)
    est._check()  # raises the library's error, if the call set one
    return params, iters


def test_chunked_host_call_equals_device_fit(ctx):
    # 150,000 11x11 ROIs: 3 chunks of up to 69,327 spots (32 MB each)
    _, smp = MD.samples("sigma", 150_000, 17)
    est, _, _ = MD.make("sigma", ctx)
    p_host, it_host = host_estimate(est, smp)
    out = est.estimate_device(torch.from_numpy(smp).cuda(), crlb=False)
    p_dev = out["params"].cpu().numpy()
    it_dev = out["iterations"].cpu().numpy()
    assert np.array_equal(p_host, p_dev, equal_nan=True)
    assert np.array_equal(it_host, it_dev)


def test_chunked_host_call_with_constants_and_initial(ctx):
    # per-spot sigma constants and caller-supplied starts travel with their chunk
    R, n = 11, 150_000
    rng = np.random.default_rng(5)
    _, smp = MD.samples("fixed", n, 19, R=R)
    est = pa.GaussianPSFMethods(ctx).CreatePSF_XYIBg(R, None, True)
    cons = np.stack([rng.uniform(1.2, 1.8, n), rng.uniform(1.2, 1.8, n)], axis=1).astype(np.float32)
    init = np.stack([rng.uniform(4, 6, n), rng.uniform(4, 6, n), rng.uniform(100, 400, n),
                     rng.uniform(1, 5, n)], axis=1).astype(np.float32)
    p_host, it_host = host_estimate(est, smp, constants=cons, initial=init)
    out = est.estimate_device(torch.from_numpy(smp).cuda(), initial=torch.from_numpy(init).cuda(), crlb=False,
                              constants=torch.from_numpy(cons).cuda())
    assert np.array_equal(p_host, out["params"].cpu().numpy(), equal_nan=True)
    assert np.array_equal(it_host, out["iterations"].cpu().numpy())
