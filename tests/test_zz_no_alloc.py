"""Estimator_EstimateDeviceNoAlloc: the fused fit + CRLB with no allocation.

Its work counters live in a fixed device array of the library instead of memory the
estimator allocates, so its results must equal Estimator_EstimateDevice's bit for bit
(every spot's fit is independent of which lane group claims it), on two calls in a
row, the second with fewer spots, and the library's allocation tallies must not move.
"""
import ctypes as ct

import numpy as np
import pytest

import photonpy_amd as pa
from tests import _models as MD

pytestmark = pytest.mark.gpu

KEYS = ("params", "iterations", "crlb", "chisq")


@pytest.fixture(scope="module")
def ctx():
    with pa.Context() as c:
        yield c


def _outputs(n, K):
    import torch
    return {"params": torch.full((n, K), -1.0, device="cuda"),
            "iterations": torch.full((n,), -1, dtype=torch.int32, device="cuda"),
            "crlb": torch.full((n, K), -1.0, device="cuda"),
            "chisq": torch.full((n,), -1.0, device="cuda")}


def _no_alloc(est, smp, out):
    import torch
    stream = torch.cuda.current_stream().cuda_stream
    return est._lib.Estimator_EstimateDeviceNoAlloc(
        est.inst, smp.shape[0], smp.data_ptr(), None, None, out["params"].data_ptr(), out["iterations"].data_ptr(),
        None, 0, out["crlb"].data_ptr(), out["chisq"].data_ptr(), stream)


def _device_bytes(est):
    vals = [ct.c_uint64() for _ in range(4)]
    est._lib.CudaGetMemoryUse(*[ct.byref(v) for v in vals])
    return vals[0].value, vals[2].value  # pinned, device


def _same(a, b):
    for k in KEYS:
        x, y = a[k].cpu().numpy(), b[k].cpu().numpy()
        assert np.array_equal(x, y, equal_nan=True), k


@pytest.mark.parametrize("kind,sizes", [("sigma", (3000, 1000)), ("astig", (1500, 400)), ("spline", (700, 250))])
def test_no_alloc_matches_estimate_device(ctx, kind, sizes):
    import torch
    th, smp = MD.samples(kind, sizes[0], 17)
    est, _, _ = MD.make(kind, ctx)
    est.SetLevMarParams(1e-18, 50)
    d_all = torch.from_numpy(smp).cuda()
    K = est.numparams
    if kind == "spline":
        # a fresh spline estimator has no device tables yet, and this entry point may not
        # upload them: it refuses with a clear error
        out = _outputs(10, K)
        assert _no_alloc(est, d_all[:10], out) != 0
        with pytest.raises(pa.NativeError, match="not on the device yet"):
            est._check()
    for n in sizes:  # two calls in a row, the second smaller
        d = d_all[:n].contiguous()
        ref = est.estimate_device(d, crlb=True)
        torch.cuda.synchronize()
        out = _outputs(n, K)
        torch.cuda.synchronize()
        before = _device_bytes(est)
        assert _no_alloc(est, d, out) == 0
        torch.cuda.synchronize()
        assert _device_bytes(est) == before
        _same(ref, out)
        assert (out["iterations"].cpu().numpy() >= 1).all()
