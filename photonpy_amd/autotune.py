"""Autotuner for the fit kernels (no reference counterpart).

Two knobs are tuned per (model, ROI size) and stored in ``tuning.json`` next
to this file:

* ``group``            lanes per spot of the fit kernel (a compiled power of two)
* ``spots_per_group``  tile length of the persistent mixed batch, in spots per lane group

``tune_group`` times the compiled group sizes of an estimator on device-resident
samples and keeps the fastest; ``tune_batch`` times candidate tile lengths of a
MixedBatch.  ``apply`` sets an estimator's tuned group size.
"""
from __future__ import annotations

import json
import os
import time

_HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(_HERE, "tuning.json")

_KIND = {"x,y,I,bg": "fixed", "x,y,\tI,bg,sigma": "sigma"}
# group sizes compiled for each ROI size (csrc/fit_sizes.h)
COMPILED_GROUPS = {3: [1], 4: [1], 5: [2], 6: [2], 7: [1, 2, 4], 8: [2], 9: [2, 4], 10: [2, 4], 11: [2, 4],
                   12: [4], 13: [4, 8], 14: [4], 15: [4], 16: [4, 8], 17: [8], 18: [8], 19: [8], 20: [8, 16],
                   21: [8], 22: [8], 23: [8], 24: [8], 25: [8, 16], 26: [16], 27: [16], 28: [16], 29: [16],
                   30: [16], 31: [16], 32: [16]}


def key(est):
    kind = _KIND.get(est.fmt)
    if kind is None:
        kind = "spline" if est.num_shared() > 10 else "astig"
    return f"{kind}:{int(est.sampleshape[0])}"


def load():
    if os.path.exists(TABLE):
        with open(TABLE) as f:
            return json.load(f)
    return {}


def save(table):
    with open(TABLE, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def lookup_spots_per_group(est, default=8):
    return int(load().get(key(est), {}).get("spots_per_group", default))


def apply(est):
    """Set the estimator's fit-kernel group size from the tuning table (if tuned)."""
    entry = load().get(key(est), {})
    if entry.get("group"):
        est.SetGroupSize(int(entry["group"]))
    if entry.get("refill") is not None:
        est.SetRefill(int(entry["refill"]))
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


def tune_group(est, samples, refills=(0, 1, 2, 4, 8), persist=True):
    """Time every compiled group size x refill batch for this estimator's ROI size on
    device samples [N,R,R]; keep the fastest pair."""
    R = int(est.sampleshape[0])
    results = {}
    for g in COMPILED_GROUPS.get(R, []):
        est.SetGroupSize(g)
        for rf in refills:
            if rf > 64 // g:
                continue
            est.SetRefill(rf)
            results[(g, rf)] = _time(lambda: est.estimate_device(samples, crlb=True))
    best = min(results, key=results.get)
    est.SetGroupSize(best[0])
    est.SetRefill(best[1])
    if persist:
        t = load()
        e = t.setdefault(key(est), {})
        e["group"], e["refill"] = int(best[0]), int(best[1])
        e["loc_per_s"] = {f"G{g}_refill{rf}": samples.shape[0] / v for (g, rf), v in results.items()}
        save(t)
    return best, results


def tune_batch(estimators, samples, candidates=(2, 4, 8, 16, 32), persist=True):
    """Time tile lengths (spots per lane group) of the persistent batch; one length for all bins."""
    from .batch import MixedBatch
    counts = [s.shape[0] for s in samples]
    results = {}
    for spg in candidates:
        with MixedBatch(estimators, counts, [spg] * len(estimators)) as mb:
            outs = mb.allocate_outputs()
            results[spg] = _time(lambda: mb.run(samples, outs))
    best = min(results, key=results.get)
    if persist:
        t = load()
        for e in estimators:
            t.setdefault(key(e), {})["spots_per_group"] = int(best)
        save(t)
    return best, results


def main():
    """python -m photonpy_amd.autotune [--configs c1,c2,c3,c4] [--n 200000]"""
    import argparse
    import torch
    from . import Context
    from . import synthetic
    ap = argparse.ArgumentParser()
    ap.add_argument("--configs", default="c1,c2,c3,c4")
    ap.add_argument("--n", type=int, default=200_000)
    args = ap.parse_args()
    with Context() as ctx:
        for c in args.configs.split(","):
            w = synthetic.WORKLOADS[c]
            est = synthetic.make_estimator(ctx, w["kind"], w["R"])
            est.SetLevMarParams(w["lam"], w["iters"])
            _, smp = synthetic.make_samples(est, w["kind"], min(args.n, w["n"]) if c != "c1" else w["n"])
            best, res = tune_group(est, smp)
            print(c, key(est), "best", best, {f"{g}/{r}": round(smp.shape[0] / v / 1e6, 2) for (g, r), v in res.items()},
                  flush=True)
            del smp
            torch.cuda.empty_cache()


if __name__ == "__main__":
    main()
