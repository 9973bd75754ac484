"""The mixed batch (the C5 workload's form) against the oracle, bin by bin.

Every (model, ROI size) bin of the mixed workload -- sizes {7, 9, 11, 13, 16, 20, 25}
x {fixed sigma, free sigma, astigmatic, spline} -- is fitted through
MixedBatch.run(use_graph=True) with the tuned lane-group sizes and refill batches the
benchmark uses, and each bin's results are checked against the oracle's FP32
estimator with the parity rules of tests/_parity.py.  At R = 25 the spline's pixels
run past the 26-cell grid, where the reference clamps the cell index to the grid's
edge (CubicSplinePSFModels.h:122-128); the oracle clamps the same way, so parity
holds there too.  A full-size (4,000,000 ROI) noise-free batch checks the fixed
point property on every bin.
"""
import numpy as np
import pytest

import photonpy_amd as pa
from photonpy_amd import synthetic
from photonpy_amd.batch import MixedBatch
from tests import _closed as C
from tests import _models as MD
from tests import _parity as P

pytestmark = pytest.mark.gpu

SIZES = synthetic.MIXED["sizes"]
KINDS = synthetic.MIXED["kinds"]


@pytest.fixture(scope="module")
def ctx():
    with pa.Context() as c:
        yield c


def _lm(kind):
    return synthetic.MODEL_LM[kind]  # (lambda, iterations) of the single-model workloads


def test_every_mixed_bin_matches_the_oracle(ctx):
    import torch
    ests, oms, samples, host = [], [], [], []
    for ki, kind in enumerate(KINDS):
        for R in SIZES:
            n = 500 + 40 * R
            th, smp = MD.samples(kind, n, 100 + 10 * ki + R, R=R)
            est, om, _ = MD.make(kind, ctx, R=R)
            lam, iters = _lm(kind)
            est.SetLevMarParams(lam, iters)
            ests.append(est)
            oms.append(om)
            host.append(smp)
            samples.append(torch.from_numpy(smp).cuda())
    counts = [len(s) for s in host]
    print(f"mixed parity: {len(ests)} bins, {sum(counts)} spots", flush=True)
    with MixedBatch(ests, counts, streams=4, tune=True) as mb:
        outs = mb.allocate_outputs()
        mb.run(samples, outs, use_graph=True)
        torch.cuda.synchronize()
        groups = [e.GetGroupSize() for e in ests]
    summary, failures = [], []
    b = 0
    for kind in KINDS:
        lam, iters = _lm(kind)
        for R in SIZES:
            est, om = ests[b], oms[b]
            p = outs[b]["params"].cpu().numpy()
            it = outs[b]["iterations"].cpu().numpy()
            tag = f"{kind} R={R} G={groups[b]}"
            try:
                w, bad = P.check_fit_parity(kind, om, host[b], p, it, iters, lam, est.ParamLimits(), tag=tag,
                                            min_witnesses=100)
                summary.append(f"{tag}: {w} witnesses of {len(p)}, {bad} differ")
            except AssertionError as ex:
                failures.append(str(ex).split("\n")[0])
            b += 1
    print("\n".join(summary))
    assert not failures, "\n".join(failures)


def test_full_size_mixed_batch_noise_free_fixed_point(ctx):
    """4,000,000 ROIs in the workload's bins, samples = the expected image of the true
    parameters (no noise), so the truth is the fit's fixed point:
    - started 0.2 px and 5% in I away from the truth, every bin returns the truth (the
      spline bins on spots that stay on one side of the grid centre, see below);
    - from the centre-of-mass start, every spot whose stop fired (off the bounds) is at
      the truth (spots that use up the iteration budget of the reference's LM -- 30 for
      fixed sigma -- are not counted: large ROIs start far off in I and bg);
    and the CRLBs are finite."""
    import torch
    rng = np.random.default_rng(7)
    n = synthetic.MIXED["n"]
    size_idx = rng.choice(len(SIZES), size=n, p=synthetic.MIXED["weights"])
    kind_idx = rng.integers(0, len(KINDS), size=n)
    ests, samples, truth, starts, kinds = [], [], [], [], []
    for ki, kind in enumerate(KINDS):
        for si, R in enumerate(SIZES):
            m = int(np.sum((kind_idx == ki) & (size_idx == si)))
            th = C.synth_params(kind, R, m, rng).astype(np.float32)
            est, _, _ = MD.make(kind, ctx, R=R)
            lam, iters = _lm(kind)
            est.SetLevMarParams(lam, iters)
            samples.append(est.expected_value_device(torch.from_numpy(th).cuda()))
            start = th.copy()
            start[:, :2] += 0.2
            start[:, 3 if P.zcol(kind) else 2] *= 1.05
            starts.append(torch.from_numpy(start).cuda())
            ests.append(est)
            truth.append(th)
            kinds.append((kind, R))
    torch.cuda.synchronize()
    print(f"full-size mixed batch: {sum(len(t) for t in truth)} ROIs in {len(ests)} bins", flush=True)
    counts = [len(t) for t in truth]
    with MixedBatch(ests, counts, streams=4, tune=True) as mb:
        near = mb.allocate_outputs()
        mb.run(samples, near, use_graph=True, initial=starts)
        torch.cuda.synchronize()
    with MixedBatch(ests, counts, streams=4, tune=True) as mb:
        com = mb.allocate_outputs()
        mb.run(samples, com, use_graph=True)
        torch.cuda.synchronize()
    rows, failures = [], []
    for (kind, R), th, o1, o2, est in zip(kinds, truth, near, com, ests):
        lam, iters = _lm(kind)
        lo, hi = est.ParamLimits()

        def at_truth(p):
            iI = 3 if P.zcol(kind) else 2
            err = np.abs(p - th)
            ok = (err[:, 0] < 1e-3) & (err[:, 1] < 1e-3) & (err[:, iI] < 1e-3 * th[:, iI])
            if P.zcol(kind):
                ok &= err[:, 2] < 1.0
            return ok

        # The spline's sub-pixel offset truncates toward zero (CubicSplinePSFModels.h:89-92):
        # past x or y = Nx/2 (13 px for the 26-cell grid) the model jumps by one pixel, so
        # the truth of a spot on that side is no fixed point of a start on the other.  Only
        # the R = 25 bin (centre 12 +- 1.5 px, start + 0.2) reaches it; those spots are
        # left out of the near-start check and counted.
        clean = np.ones(len(th), bool)
        if kind == "spline":
            half = 26 / 2
            clean = np.all(th[:, :2] + 0.2 < half, axis=1)
        f_near = at_truth(o1["params"].cpu().numpy())[clean].mean()
        p2 = o2["params"].cpu().numpy()
        stopped = (o2["iterations"].cpu().numpy() <= iters) & ~P.on_bounds(p2, lo, hi)
        f_com = at_truth(p2)[stopped].mean() if stopped.any() else 1.0
        crlb = o1["crlb"].cpu().numpy()
        fin = np.all(np.isfinite(crlb) & (crlb > 0), axis=1).mean()
        rows.append(f"{kind} R={R} n={len(th)}: near-start fixed point {f_near:.5f} "
                    f"({(~clean).sum()} past the grid centre left out); COM start: "
                    f"{stopped.mean():.3f} stopped, {f_com:.5f} of them at the truth; finite CRLB {fin:.5f}")
        if not (f_near > 0.999 and f_com > 0.999 and fin > 0.999):
            failures.append(rows[-1])
    print("\n".join(rows))
    assert not failures, "\n".join(failures)
