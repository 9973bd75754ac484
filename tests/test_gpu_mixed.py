"""Persistent mixed-model / mixed-size batch vs the plain per-bin launches.

Every spot's fit is independent of scheduling, so the persistent launch (eager or
replayed from its HIP graph) must reproduce the plain launches bit for bit; the
plain launches themselves are parity-tested against the oracle in test_gpu_parity.
"""
import numpy as np
import pytest

import photonpy_amd as pa
from photonpy_amd.batch import MixedBatch
from tests import _models as MD

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="the persistent mixed-batch launch hung once on the GPU (build with per-variant "
                            "__noinline__ functions); its cause has not been shown, so the path is switched off"),
]

SIZES = [7, 9, 11, 13, 16, 20, 25]


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


def test_persistent_matches_plain(ctx):
    import torch
    ests, samples, counts = _bins(ctx, ["fixed", "sigma", "astig", "spline"], SIZES, 300)
    with MixedBatch(ests, counts, [4] * len(ests)) as mb:
        ref = mb.allocate_outputs()
        mb.run_plain(samples, ref)
        eager = mb.allocate_outputs()
        mb.run(samples, eager, use_graph=False)
        torch.cuda.synchronize()
        for key in ("params", "iterations", "crlb", "chisq"):
            for a, b in zip(_cat(ref, key), _cat(eager, key)):
                assert np.array_equal(a, b, equal_nan=True), key
        # graph replays, with eager work and synchronisation between them, each checked
        for rep in range(3):
            outs = mb.allocate_outputs()
            mb.run(samples, outs, use_graph=True)
            torch.cuda.synchronize()
            _ = (samples[0] * 2).sum().item()
            for key in ("params", "iterations", "crlb", "chisq"):
                for a, b in zip(_cat(ref, key), _cat(outs, key)):
                    assert np.array_equal(a, b, equal_nan=True), (rep, key)


def test_unsupported_size_rejected(ctx):
    est = pa.GaussianPSFMethods(ctx).CreatePSF_XYIBg(12, 1.5, True)
    with pytest.raises(pa.NativeError, match="roisize 12"):
        MixedBatch([est], [10], [8])


def test_tile_lengths_do_not_change_results(ctx):
    import torch
    ests, samples, counts = _bins(ctx, ["sigma", "astig"], [9, 16], 500, seed=4)
    results = []
    for spg in (1, 8, 32):
        with MixedBatch(ests, counts, [spg] * len(ests)) as mb:
            outs = mb.allocate_outputs()
            mb.run(samples, outs)
            torch.cuda.synchronize()
            results.append(_cat(outs, "params"))
    for r in results[1:]:
        for a, b in zip(results[0], r):
            assert np.array_equal(a, b, equal_nan=True)
