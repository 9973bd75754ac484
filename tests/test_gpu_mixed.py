"""Mixed-model, mixed-size batch vs the plain per-bin launches.

Every spot's fit is independent of scheduling (which stream, which wavefront, when), so
the batch (eager or replayed from its HIP graph) must reproduce the plain launches bit for
bit; the plain launches themselves are parity-tested against the oracle in
test_gpu_parity.
"""
import numpy as np
import pytest

import photonpy_amd as pa
from photonpy_amd.batch import MixedBatch
from tests import _models as MD

pytestmark = pytest.mark.gpu

SIZES = [7, 9, 11, 13, 16, 20, 25]
KEYS = ("params", "iterations", "crlb", "chisq")


@pytest.fixture(scope="module")
def ctx():
    with pa.Context() as c:
        yield c


def _bins(ctx, kinds, sizes, n, seed=1):
    import torch
    ests, samples, counts = [], [], []
    rng = np.random.default_rng(seed)
    for kind in kinds:
        for R in sizes:
            m = n + int(rng.integers(0, n))
            th, smp = MD.samples(kind, m, seed + R, R=R)
            est, _, _ = MD.make(kind, ctx, R=R)
            est.SetLevMarParams(1e-18, 30)
            ests.append(est)
            samples.append(torch.from_numpy(smp).cuda())
            counts.append(m)
    return ests, samples, counts


def _cat(outs, key):
    return [o[key].cpu().numpy() for o in outs]


def _same(ref, outs, tag):
    for key in KEYS:
        for b, (a, c) in enumerate(zip(_cat(ref, key), _cat(outs, key))):
            assert np.array_equal(a, c, equal_nan=True), (tag, key, b)


def test_batch_matches_plain(ctx):
    import torch
    ests, samples, counts = _bins(ctx, ["fixed", "sigma", "astig", "spline"], SIZES, 300)
    with MixedBatch(ests, counts, streams=4, tune=False) as mb:
        assert mb.num_streams() == 4
        ref = mb.allocate_outputs()
        mb.run_plain(samples, ref)
        eager = mb.allocate_outputs()
        mb.run(samples, eager, use_graph=False)
        torch.cuda.synchronize()
        _same(ref, eager, "eager")
        # graph replays, with eager work and synchronisation between them, each checked
        outs = mb.allocate_outputs()
        for rep in range(3):
            for o in outs:
                for v in o.values():
                    v.fill_(-1)
            mb.run(samples, outs, use_graph=True)
            torch.cuda.synchronize()
            _ = (samples[0] * 2).sum().item()
            _same(ref, outs, f"replay {rep}")
        # other output buffers than the captured ones: refused (a new batch captures them)
        outs2 = mb.allocate_outputs()
        with pytest.raises(pa.NativeError, match="create a new MixedBatch"):
            mb.run(samples, outs2, use_graph=True)
    # a new batch captures them
    with MixedBatch(ests, counts, streams=4, tune=False) as mb2:
        mb2.run(samples, outs2, use_graph=True)
        torch.cuda.synchronize()
        _same(ref, outs2, "second batch")


def test_stream_count_does_not_change_results(ctx):
    import torch
    ests, samples, counts = _bins(ctx, ["sigma", "astig"], [9, 16], 500, seed=4)
    results = []
    for ns in (1, 2, 8):
        with MixedBatch(ests, counts, streams=ns, tune=False) as mb:
            outs = mb.allocate_outputs()
            mb.run(samples, outs)
            torch.cuda.synchronize()
            results.append(outs)
    for r in results[1:]:
        _same(results[0], r, "streams")
