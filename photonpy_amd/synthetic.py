"""Synthetic ROI batches of the benchmark workloads (for bench.py and the autotuner).

True parameters are drawn with numpy PCG64 as each workload states; expected
images come from the library's own expected-value kernel and Poisson noise
from a seeded torch generator on the GPU, so the batch is created in HBM.
"""
from __future__ import annotations

import numpy as np

WORKLOADS = {
    "c1": dict(kind="fixed", R=10, n=10_000, iters=30, lam=1e-15,
               workload="2D Gaussian MLE fixed sigma=1.5, 10x10 px, 10,000 spots, 30 iterations + CRLB"),
    "c2": dict(kind="sigma", R=11, n=1_000_000, iters=50, lam=1e-18,
               workload="2D Gaussian MLE free sigma, 11x11 px, 1,000,000 spots, 50 iterations + CRLB"),
    "c3": dict(kind="astig", R=16, n=2_000_000, iters=50, lam=1e-18,
               workload="astigmatic 3D Gaussian MLE (A=0.1, B=0.05), 16x16 px, 2,000,000 spots, 50 iterations + CRLB"),
    "c4": dict(kind="spline", R=20, n=500_000, iters=50, lam=1e-18,
               workload="cubic-spline 3D PSF MLE (61x26x26x64 coefficients), 20x20 px, 500,000 spots, 50 iterations + CRLB"),
}

ASTIG = dict(s0=1.3, c=200.0, d=400.0, A=0.1, B=0.05, zrange=[-600.0, 600.0])


def astig_calibration():
    from .gaussian import Gauss3D_Calibration
    a = ASTIG
    return Gauss3D_Calibration.from_astig_params(a["s0"], a["c"], a["d"], a["A"], a["B"], a["zrange"])


def spline_table(ctx):
    """61x26x26x64 spline of the astigmatic PSF rendered on 29x29 px at 13.3 nm z steps
    (emitter at pixel 14), computed on the GPU.  Returns (coefs, z_min, z_max)."""
    from .cspline import CSplineMethods
    from .gaussian import GaussianPSFMethods
    ast = GaussianPSFMethods(ctx).CreatePSF_XYZIBg(29, astig_calibration(), True)
    zs = -418.95 + 13.3 * np.arange(64)
    th = np.zeros((64, 5), np.float32)
    th[:, 0] = th[:, 1] = 14.0
    th[:, 2] = zs
    th[:, 3] = 1.0
    stack = ast.ExpectedValue(th)
    ast.Destroy()
    coefs = CSplineMethods(ctx).ComputeSplineCoeff(stack)
    return coefs, float(zs[1]), float(zs[1] + 61 * 13.3)


def make_estimator(ctx, kind, R, spline=None):
    from .cspline import CSplineCalibration, CSplineMethods
    from .gaussian import GaussianPSFMethods
    g = GaussianPSFMethods(ctx)
    if kind == "fixed":
        return g.CreatePSF_XYIBg(R, 1.5, True)
    if kind == "sigma":
        return g.CreatePSF_XYIBgSigma(R, 1.6, True)
    if kind == "astig":
        return g.CreatePSF_XYZIBg(R, astig_calibration(), True)
    if kind == "spline":
        coefs, zmin, zmax = spline if spline is not None else spline_table(ctx)
        m = CSplineMethods(ctx)
        return m.CreatePSF_XYZIBg(R, CSplineCalibration(zmin, zmax, coefs), m.FlatBg, True)
    raise ValueError(kind)


def true_params(kind, R, n, seed=1):
    rng = np.random.default_rng(seed)
    c = (R - 1) / 2
    x = c + rng.uniform(-1.5, 1.5, n)
    y = c + rng.uniform(-1.5, 1.5, n)
    if kind == "fixed":
        cols = [x, y, rng.uniform(500, 5000, n), rng.uniform(1, 20, n)]
    elif kind == "sigma":
        cols = [x, y, np.exp(rng.uniform(np.log(200), np.log(20000), n)), rng.uniform(0.5, 30, n),
                rng.uniform(1.2, 2.0, n)]
    else:
        cols = [x, y, rng.uniform(-400, 400, n), rng.uniform(500, 8000, n), rng.uniform(2, 20, n)]
    return np.stack(cols, 1).astype(np.float32)


def make_samples(est, kind, n, seed=1, device="cuda"):
    """(true params [n,K] numpy, Poisson samples [n,R,R] on the device)."""
    import torch
    R = int(est.sampleshape[0])
    th = true_params(kind, R, n, seed)
    mu = est.expected_value_device(torch.from_numpy(th).to(device))
    gen = torch.Generator(device=device).manual_seed(seed)
    smp = torch.poisson(mu, generator=gen)
    del mu
    return th, smp


MIXED = dict(n=4_000_000, sizes=[7, 9, 11, 13, 16, 20, 25], weights=[0.1, 0.2, 0.25, 0.15, 0.15, 0.1, 0.05],
             kinds=["fixed", "sigma", "astig", "spline"],
             workload="mixed batch: 4,000,000 ROIs, sizes {7,9,11,13,16,20,25} weighted "
                      "{.1,.2,.25,.15,.15,.1,.05}, models uniform over fixed/free-sigma/astig/spline, + CRLB")
# per-model LM settings of the single-model workloads
MODEL_LM = {"fixed": (1e-15, 30), "sigma": (1e-18, 50), "astig": (1e-18, 50), "spline": (1e-18, 50)}


def mixed_bins(ctx, n=None, seed=1):
    """The mixed workload as (model, ROI size) bins: each ROI draws its size from the
    weighted list and its model uniformly; bins are filled with that many ROIs of the
    model's own parameter distribution.  Returns a list of dicts (kind, R, n, est,
    samples [n,R,R] on the device, true params)."""
    n = n or MIXED["n"]
    rng = np.random.default_rng(seed)
    size_idx = rng.choice(len(MIXED["sizes"]), size=n, p=MIXED["weights"])
    kind_idx = rng.integers(0, len(MIXED["kinds"]), size=n)
    counts = np.zeros((len(MIXED["kinds"]), len(MIXED["sizes"])), np.int64)
    np.add.at(counts, (kind_idx, size_idx), 1)
    table = spline_table(ctx)
    bins = []
    for ki, kind in enumerate(MIXED["kinds"]):
        for si, R in enumerate(MIXED["sizes"]):
            m = int(counts[ki, si])
            if m == 0:
                continue
            est = make_estimator(ctx, kind, R, spline=table)
            lam, iters = MODEL_LM[kind]
            est.SetLevMarParams(lam, iters)
            th, smp = make_samples(est, kind, m, seed=seed + 1000 * ki + R)
            bins.append(dict(kind=kind, R=R, n=m, est=est, samples=smp, params=th, iters=iters))
    return bins
