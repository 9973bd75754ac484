"""Autotuner for the fit kernels and the mixed batch (no reference counterpart).

Knobs tuned per (model, ROI size) and stored in ``tuning.json`` next to this file:

* ``group``    lanes per spot of the fit kernel (a compiled power of two); with the
               block size it sets the spots per block and, for the grid-form spline,
               the LDS window tables per block
* ``refill``   stopped lane groups of a wavefront that are refilled together
* ``segs``     work counters of a launch larger than the resident grid: 1, or 8
               contiguous spot segments with a counter each

and for a mixed batch ``streams``, the number of concurrent branches.
``tune_group`` times the compiled group sizes x refill batches of an estimator on
device-resident samples and keeps the fastest; ``tune_streams`` times branch counts
of a MixedBatch.  ``apply`` sets an estimator's tuned values.
"""
from __future__ import annotations

import json
import os
import time

_HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(_HERE, "tuning.json")

_KIND = {"x,y,I,bg": "fixed", "x,y,\tI,bg,sigma": "sigma"}
# group sizes compiled for each ROI size (csrc/fit_sizes.h)
COMPILED_GROUPS = {3: [1], 4: [1], 5: [2], 6: [2], 7: [1, 2, 4], 8: [2], 9: [2, 4], 10: [2, 4, 8, 16], 11: [2, 4],
                   12: [4], 13: [4, 8], 14: [4], 15: [4], 16: [4, 8], 17: [8], 18: [8], 19: [8], 20: [8, 16],
                   21: [8], 22: [8], 23: [8], 24: [8], 25: [8, 16], 26: [16], 27: [16], 28: [16], 29: [16],
                   30: [16], 31: [16], 32: [16]}
# grid-form spline (PPA_GRID_SIZES)
GRID_GROUPS = {3: [4], 4: [4], 5: [4], 6: [4], 7: [4, 8], 8: [8], 9: [8, 16], 10: [8], 11: [8, 16], 12: [8],
               13: [8, 16], 14: [16], 15: [16], 16: [16, 32], 17: [16], 18: [32], 19: [32], 20: [16, 32, 64],
               21: [32], 22: [32], 23: [64], 24: [64], 25: [16, 32, 64], 26: [64], 27: [64], 28: [64], 29: [64],
               30: [64], 31: [64], 32: [64]}


def key(est):
    kind = _KIND.get(est.fmt)
    if est.numconst > 0 or est.fmt not in _KIND and est.numparams != 5:
        return f"{est.fmt}/{est.numconst}:{int(est.sampleshape[0])}"
    if kind is None:
        kind = "astig"
        if est.num_shared() > 10:
            kind = "splinegrid" if est.spline_form() == "grid" else "spline"
    return f"{kind}:{int(est.sampleshape[0])}"


def compiled_groups(est):
    R = int(est.sampleshape[0])
    k = key(est)
    if "/" in k:  # the other Gaussian variants: default group only
        return []
    return (GRID_GROUPS if k.startswith("splinegrid") else COMPILED_GROUPS).get(R, [])


def load():
    if os.path.exists(TABLE):
        with open(TABLE) as f:
            return json.load(f)
    return {}


def save(table):
    with open(TABLE, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def apply(est):
    """Set the estimator's fit-kernel group size from the tuning table (if tuned)."""
    entry = load().get(key(est), {})
    if entry.get("group"):
        est.SetGroupSize(int(entry["group"]))
    if entry.get("refill") is not None:
        est.SetRefill(int(entry["refill"]))
    if entry.get("segs"):
        est.SetWorkSegments(int(entry["segs"]))
    return est.GetGroupSize()


def _time(fn, reps=3):
    import torch
    fn()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t)
    return best


def tune_group(est, samples, refills=(0, 1, 2, 4, 8), segments=(1, 8), persist=True):
    """Time every compiled group size x refill batch x work-segment count for this
    estimator's ROI size on device samples [N,R,R]; keep the fastest."""
    results = {}
    for g in compiled_groups(est):
        est.SetGroupSize(g)
        for rf in refills:
            if rf > 64 // g:
                continue
            est.SetRefill(rf)
            for sg in segments:
                est.SetWorkSegments(sg)
                results[(g, rf, sg)] = _time(lambda: est.estimate_device(samples, crlb=True))
    best = min(results, key=results.get)
    est.SetGroupSize(best[0])
    est.SetRefill(best[1])
    est.SetWorkSegments(best[2])
    if persist:
        t = load()
        e = t.setdefault(key(est), {})
        e["group"], e["refill"], e["segs"] = int(best[0]), int(best[1]), int(best[2])
        e["loc_per_s"] = {f"G{g}_refill{rf}_segs{sg}": samples.shape[0] / v for (g, rf, sg), v in results.items()}
        save(t)
    return best, results


def tune_streams(estimators, samples, candidates=(1, 2, 4, 8), persist=True):
    """Time branch counts of a MixedBatch over these bins (graph replay); keep the fastest."""
    from .batch import MixedBatch
    counts = [s.shape[0] for s in samples]
    results = {}
    for ns in candidates:
        with MixedBatch(estimators, counts, streams=ns, tune=False) as mb:
            outs = mb.allocate_outputs()
            results[ns] = _time(lambda: mb.run(samples, outs))
    best = min(results, key=results.get)
    if persist:
        t = load()
        t["mixed_batch"] = {"streams": int(best), "seconds": {str(k): v for k, v in results.items()}}
        save(t)
    return best, results


def batch_streams(default=4):
    return int(load().get("mixed_batch", {}).get("streams", default))


def main():
    """python -m photonpy_amd.autotune [--configs c1,c2,c3,c4] [--n 0] [--bins] [--batch-n 400000]"""
    import argparse
    import torch
    from . import Context
    from . import synthetic
    ap = argparse.ArgumentParser()
    ap.add_argument("--configs", default="c1,c2,c3,c4")
    ap.add_argument("--n", type=int, default=0, help="spots per config run (0 = the workload's own count)")
    ap.add_argument("--batch-n", type=int, default=0, help="also tune the mixed batch's stream count on this many ROIs")
    ap.add_argument("--bins", action="store_true",
                    help="tune every (model, ROI size) bin of the mixed workload not yet in the table, on its own ROIs")
    args = ap.parse_args()
    with Context() as ctx:
        for c in [c for c in args.configs.split(",") if c]:
            w = synthetic.WORKLOADS[c]
            est = synthetic.make_estimator(ctx, w["kind"], w["R"])
            est.SetLevMarParams(w["lam"], w["iters"])
            _, smp = synthetic.make_samples(est, w["kind"], min(args.n, w["n"]) if args.n > 0 else w["n"])
            best, res = tune_group(est, smp)
            print(c, key(est), "best", best,
                  {f"{g}/{r}/{s}": round(smp.shape[0] / v / 1e6, 2) for (g, r, s), v in res.items()}, flush=True)
            del smp
            torch.cuda.empty_cache()
        if args.bins:
            # the keys of the configs above keep their own workload's choice
            tuned = set(load())
            for b in synthetic.mixed_bins(ctx):
                k = key(b["est"])
                if k in tuned or not compiled_groups(b["est"]):
                    continue
                best, res = tune_group(b["est"], b["samples"])
                tuned.add(k)
                print("bin", k, b["n"], "best", best,
                      {f"{g}/{r}/{s}": round(b["n"] / v / 1e6, 2) for (g, r, s), v in res.items()}, flush=True)
                del b
                torch.cuda.empty_cache()
        if args.batch_n:
            bins = synthetic.mixed_bins(ctx, n=args.batch_n)
            for b in bins:
                apply(b["est"])
            best, res = tune_streams([b["est"] for b in bins], [b["samples"] for b in bins])
            print("mixed batch streams best", best, {k: round(args.batch_n / v / 1e6, 2) for k, v in res.items()},
                  flush=True)


if __name__ == "__main__":
    main()
