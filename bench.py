"""Throughput benchmark of the MI355X spot fitter.

`python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2]`

One step = one fused launch (Levenberg-Marquardt fit + CRLB + chi^2) over the
whole batch of synthetic ROIs, already resident in HBM.  Prints one JSON line.

Workloads (the default is c2):
  c1  2D Gaussian fixed sigma=1.5, 10x10 px, 10,000 spots, 30 iterations + CRLB
  c2  2D Gaussian free sigma, 11x11 px, 1,000,000 spots, 50 iterations + CRLB
  c3  astigmatic 3D Gaussian (quartic sigma(z)), 16x16 px, 2,000,000 spots, 50 iterations + CRLB
  c4  cubic-spline 3D PSF (61x26x26 cells), 20x20 px, 500,000 spots, 50 iterations + CRLB
  c4adj  spline PSF adjoint (forward NLL + backward to params and all coefficients), 65,536 ROIs
  c5  mixed batch: 4,000,000 ROIs over 7 sizes x 4 models, batch form vs plain per-bin launches
  queue  C2's ROIs from host memory through the EstimQueue (the reference's streaming path)
Synthetic data (photonpy_amd/synthetic.py): true parameters from numpy
PCG64(seed 1) as the workload states, expected images from the library's own
expected-value kernel, Poisson noise from a seeded torch generator on the GPU.
Tuned lane-group sizes are read from photonpy_amd/tuning.json when present.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLE localizations/s per model (vs CPU + plain GPU); PSF-adjoint grad evals/s"
PEAK_FP32 = 157.3  # TFLOP/s, MI355X FP32 (vector and matrix)
# The roofline's "bound" names the compute roofline ("mfma" in the output contract);
# the fit issues FP32 vector instructions, no matrix instructions.
COMPUTE_NOTE = "compute-bound at the FP32 peak; FP32 VALU instructions (no MFMA), packed FP32 disabled"

# ---------------------------------------------------------------------------
# Algorithmic FLOP count per localization (FP32 model + normal equations, FP64
# solve and inversion counted as FLOPs too; erf/exp evaluations not counted).
#   per pixel-iteration   K^2 + 4K + 4 (normal equations) + model terms
#   per row/column-iter   row/column terms of the separable Gaussians
#   per iteration         FP64 damping + Cholesky + solve: 4K^2 + 4K + 2K^3/3
#   initial estimate      3S (centre of mass)
#   CRLB kernel           S (model + K^2 + 2K + 3) + (W+H) rowcol + 8K^3/3
MODEL_TERMS = {"fixed": 5, "sigma": 8, "astig": 8, "spline": 422, "splinegrid": 136}
ROWCOL_TERMS = {"fixed": 11, "sigma": 15, "astig": 16, "spline": 0, "splinegrid": 0}
# The grid-form spline (tables made by CSpline_Compute) evaluates a pixel from 4x4 entries of
# two z-contracted window tables: 64 FMA + 8 (finish); per spot-iteration the window costs
# (R+3)^2 x 8 FMA of z-contraction, counted per row below as (R+3)^2 * 16 / (2R).


def fit_flops_per_loc(kind, R, executed_iters):
    """Fit kernel: centre of mass + executed iterations."""
    K = 4 if kind == "fixed" else 5
    S = R * R
    window = (R + 3) ** 2 * 16 if kind == "splinegrid" else 0
    per_iter = S * (K * K + 4 * K + 4 + MODEL_TERMS[kind]) + 2 * R * ROWCOL_TERMS[kind] + window + \
        (4 * K * K + 4 * K + 2 * K ** 3 / 3)
    return per_iter * executed_iters + 3 * S


def crlb_flops_per_loc(kind, R):
    """CRLB kernel: Fisher matrix + chi^2 + FP64 LU inversion."""
    K = 4 if kind == "fixed" else 5
    S = R * R
    window = (R + 3) ** 2 * 16 if kind == "splinegrid" else 0
    return S * (MODEL_TERMS[kind] + K * K + 2 * K + 3) + 2 * R * ROWCOL_TERMS[kind] + window + 8 * K ** 3 / 3


def hbm_traffic(workload):
    """HBM bytes per fit-kernel launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE scaled by a plain copy's readings), from profiles/traffic_<workload>.txt
    (`key = value` lines), or None when that file is absent."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if "=" in line and not line.lstrip().startswith("#"):
                k, v = line.split("=", 1)
                vals[k.strip()] = v.strip()
    try:
        return float(vals["fit_kernel_hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def hbm_bytes_per_loc(kind, R):
    K = 4 if kind == "fixed" else 5
    return 4 * R * R + 4 * (2 * K + 2)  # ROI in; params, CRLB, chi^2, iterations out


def cpu_baseline(cfg, smp_host, est, target_s=12.0):
    """The oracle (FP32 restatement of the reference estimator) on host cores, over a
    bounded prefix of the same ROIs: Estimate + CRLB + chi^2 per spot."""
    from oracle import oracle as O
    kind, R = cfg["kind"], cfg["R"]
    if kind == "fixed":
        om = O.gauss2d(R, 1.5)
    elif kind == "sigma":
        om = O.gauss2d_sigma(R, 1.6)
    elif kind == "astig":
        om = O.astig(R, [1.3, 200.0, 400.0, 0.1 / 400], [1.3, -200.0, 400.0, 0.1 / 400], [-600.0, 600.0],
                     B=(0.05 / 400 ** 2, 0.05 / 400 ** 2))
    else:
        coefs = est.shared_parameters("cpu").numpy().reshape(61, 26, 26, 64)
        lo, hi = est.ParamLimits()
        om = O.spline(R, coefs, float(lo[2]), float(hi[2]))
    threads = min(16, os.cpu_count() or 1)

    def run(m):
        s = smp_host[:m]
        t = time.perf_counter()
        p, _ = om.estimate(s, iterations=cfg["iters"], lambda_=cfg["lam"], nthreads=threads)
        om.crlb(p, s, nthreads=threads)
        return time.perf_counter() - t

    m = min(len(smp_host), 4096)
    dt = run(m)
    m2 = int(min(len(smp_host), max(m, m * target_s / max(dt, 1e-3))))
    dt2 = run(m2)
    return {"value": m2 / dt2, "unit": "localizations/s", "cores": threads, "kind": "port",
            "sample": f"first {m2} of the same ROIs, oracle Estimate+CRLB+chi^2, {threads} threads, {dt2:.1f} s"}


def host_api_rate(est, smp_host):
    """Localizations/s through Estimator_Estimate with host buffers (H2D + fit + D2H)."""
    m = len(smp_host)
    K = est.numparams
    params = np.zeros((m, K), np.float32)
    iters = np.zeros(m, np.int32)
    diag = np.zeros((m, 0), np.float32)
    cons = np.zeros((m, 0), np.float32)
    rp = np.zeros((m, 2), np.int32)
    est._lib.Estimator_Estimate(est.inst, m, smp_host, cons, rp, None, params, diag, iters, None, 0)
    t1 = time.perf_counter()
    est._lib.Estimator_Estimate(est.inst, m, smp_host, cons, rp, None, params, diag, iters, None, 0)
    return m / (time.perf_counter() - t1)


def bench_mixed(args):
    """Mixed batch (4,000,000 ROIs over 28 model x size bins): the batch form (bins on
    concurrent streams, one HIP graph launch per step) against the plain per-bin launches
    one after another on one stream; both compute fit + CRLB + chi^2 for every spot."""
    import torch
    import photonpy_amd as pa
    from photonpy_amd import autotune, synthetic
    from photonpy_amd.batch import MixedBatch
    with pa.Context() as ctx:
        bins = synthetic.mixed_bins(ctx, n=args.n or None)
        ests = [b["est"] for b in bins]
        counts = [b["n"] for b in bins]
        samples = [b["samples"] for b in bins]
        n = sum(counts)
        stream = torch.cuda.current_stream()
        with MixedBatch(ests, counts, streams=autotune.batch_streams(), tune=True) as mb:
            outs = mb.allocate_outputs()
            ref = mb.allocate_outputs()

            def timed(fn):
                for _ in range(args.warmup):
                    fn()
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                t0 = time.perf_counter()
                e0.record(stream)
                for _ in range(args.steps):
                    fn()
                e1.record(stream)
                torch.cuda.synchronize()
                return (time.perf_counter() - t0) * 1e3 / args.steps, e0.elapsed_time(e1) / args.steps

            plain_ms, plain_gpu = timed(lambda: mb.run_plain(samples, ref))
            batch_ms, batch_gpu = timed(lambda: mb.run(samples, outs, use_graph=True))
            same = all(bool(((a[k] == b[k]) | (a[k].isnan() & b[k].isnan())).all()) if a[k].is_floating_point()
                       else torch.equal(a[k], b[k]) for a, b in zip(outs, ref) for k in a)
            flops = 0.0
            for b, o in zip(bins, outs):
                it = np.minimum(o["iterations"].cpu().numpy(), b["iters"]).astype(np.float64)
                fk = "splinegrid" if b["est"].spline_form() == "grid" else b["kind"]
                per0 = fit_flops_per_loc(fk, b["R"], 0)
                per1 = fit_flops_per_loc(fk, b["R"], 1) - per0
                flops += b["n"] * (per0 + crlb_flops_per_loc(fk, b["R"])) + per1 * it.sum()
            streams = mb.num_streams()
            forms = sorted({b["est"].spline_form() for b in bins if b["kind"] == "spline"})
    achieved = flops / (batch_gpu * 1e-3) / 1e12
    print(json.dumps({
        "metric": METRIC, "value": n / (batch_ms * 1e-3), "unit": "localizations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": batch_ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": synthetic.MIXED["workload"], "spots": n, "bins": len(bins), "streams": streams,
                   "spline_form": forms},
        "roofline": {"kernel": "all fit and CRLB kernels of the batch", "bound": "mfma", "compute": COMPUTE_NOTE,
                     "achieved": achieved,
                     "peak": PEAK_FP32, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32, "traffic": None},
        "cpu_baseline": None,
        "detail": {"plain_ms_per_step": plain_ms, "plain_loc_per_s": n / (plain_ms * 1e-3),
                   "batch_over_plain": plain_ms / batch_ms, "batch_gpu_ms": batch_gpu, "plain_gpu_ms": plain_gpu,
                   "results_identical": bool(same),
                   "bins": [{"kind": b["kind"], "R": b["R"], "n": b["n"]} for b in bins]}}))


def bench_queue(args):
    """The reference's own throughput path (photonpy/tests/psf_queue.py): C2's ROIs in
    host memory streamed through an EstimQueue (batches of 4096 on 4 streams, scheduled
    10,000 at a time, results collected as they finish).  Host-to-host rate, PCIe
    included; the queue's batches use the same fused kernels as the device path."""
    import torch
    import photonpy_amd as pa
    from photonpy_amd import autotune, synthetic
    cfg = synthetic.WORKLOADS["c2"]
    n = args.n or cfg["n"]
    with pa.Context() as ctx:
        est = synthetic.make_estimator(ctx, cfg["kind"], cfg["R"])
        est.SetLevMarParams(cfg["lam"], cfg["iters"])
        autotune.apply(est)
        _, smp = synthetic.make_samples(est, cfg["kind"], n, seed=1)
        host = smp.cpu().numpy()
        del smp
        torch.cuda.empty_cache()
        ids = np.arange(n, dtype=np.int32)
        chunk = 10_000

        def run_once(q):
            # as photonpy/tests/psf_queue.py: schedule, collect what has finished, flush, drain
            got = 0
            for s0 in range(0, n, chunk):
                q.Schedule(host[s0:s0 + chunk], ids=ids[s0:s0 + chunk])
                if q.GetResultCount() > 0:
                    got += len(q.GetResults().ids)
            q.Flush()
            while not q.IsIdle():
                time.sleep(0.0005)
            got += len(q.GetResults().ids)
            return got

        def timed(batch, steps, warmup):
            # the queue is created outside the timed region, as the reference's benchmark does
            with pa.EstimQueue(est, batchSize=batch, numStreams=4) as q:
                for _ in range(warmup):
                    run_once(q)
                t0 = time.perf_counter()
                for _ in range(steps):
                    got = run_once(q)
                return got, (time.perf_counter() - t0) / steps

        got, dt = timed(4096, args.steps, args.warmup)
        got_big, dt_big = timed(65536, 1, 1)
    assert got == n and got_big == n, (got, got_big, n)
    print(json.dumps({
        "metric": METRIC, "value": n / dt, "unit": "localizations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "EstimQueue from host memory: " + cfg["workload"], "spots": n, "batch": 4096,
                   "streams": 4, "schedule_chunk": chunk},
        "roofline": None, "cpu_baseline": None,
        "detail": {"note": "host-to-host including PCIe transfers; compare detail.pcie_inclusive_loc_per_s of c2",
                   "batch65536_loc_per_s": n / dt_big}}))


def bench_adjoint(args):
    """Spline PSF calibration step: forward Poisson NLL + hand-written adjoint on 65,536
    20x20 ROIs, gradients w.r.t. per-spot parameters and all 2,638,976 spline coefficients
    (mu and J recomputed in the backward pass, nothing stored)."""
    import torch
    import photonpy_amd as pa
    from photonpy_amd import synthetic
    from photonpy_amd.autograd import poisson_nll
    n = args.n or 65_536
    with pa.Context() as ctx:
        est = synthetic.make_estimator(ctx, "spline", 20)
        th, smp = synthetic.make_samples(est, "spline", n, seed=1)
        params = torch.from_numpy(th).cuda().requires_grad_(True)
        coefs = est.shared_parameters().requires_grad_(True)
        stream = torch.cuda.current_stream()

        def step():
            params.grad = None
            coefs.grad = None
            loss = poisson_nll(est, params, smp, coefs).sum()
            loss.backward()

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        ev0 = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
        ev1 = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
        t0 = time.perf_counter()
        for k in range(args.steps):
            ev0[k].record(stream)
            step()
            ev1[k].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gpu_ms = float(np.mean([a.elapsed_time(b) for a, b in zip(ev0, ev1)]))
    ms = wall * 1e3 / args.steps
    atomics = n * 400 * 64
    print(json.dumps({
        "metric": METRIC, "value": n / (ms * 1e-3), "unit": "ROI gradient evaluations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "cubic-spline PSF adjoint: forward NLL + backward w.r.t. params and 61x26x26x64 "
                               "coefficients, 20x20 px, 65,536 ROIs", "roi": 20, "spots": n},
        "detail": {"gpu_ms_per_step": gpu_ms, "coefficient_atomics_per_step": atomics,
                   "atomic_adds_per_s": atomics / (gpu_ms * 1e-3)}}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c4adj", "c5", "queue"])
    ap.add_argument("--n", type=int, default=0, help="override the number of spots")
    ap.add_argument("--group", type=int, default=0, help="lane-group size override (0 = tuned/default)")
    ap.add_argument("--refill", type=int, default=0, help="refill batch override (0 = tuned/default)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-e2e", action="store_true", help="skip the PCIe-inclusive host-API leg")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB 5 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(spots shard independently across GPUs, see the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch
    import photonpy_amd as pa

    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes, "launches": 5}))
        return

    from photonpy_amd import autotune, synthetic

    if args.config == "c4adj":
        return bench_adjoint(args)
    if args.config == "c5":
        return bench_mixed(args)
    if args.config == "queue":
        return bench_queue(args)
    cfg = synthetic.WORKLOADS[args.config]
    n = args.n or cfg["n"]
    torch.cuda.set_device(0)
    with pa.Context() as ctx:
        est = synthetic.make_estimator(ctx, cfg["kind"], cfg["R"])
        est.SetLevMarParams(cfg["lam"], cfg["iters"])
        autotune.apply(est)
        if args.group:
            est.SetGroupSize(args.group)
        if args.refill:
            est.SetRefill(args.refill)
        th, smp = synthetic.make_samples(est, cfg["kind"], n, seed=1)
        torch.cuda.synchronize()

        for _ in range(args.warmup):
            out = est.estimate_device(smp, crlb=True)
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        ev0 = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
        ev1 = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
        # HIP events on the launch stream around each kernel (fit, then CRLB)
        est.set_phase_timing(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            ev0[k].record(stream)
            out = est.estimate_device(smp, crlb=True)
            ev1[k].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        launches, fit_ms, crlb_ms = est.phase_times()
        est.set_phase_timing(False)
        assert launches == args.steps, (launches, args.steps)
        step_gpu_ms = float(np.mean([a.elapsed_time(b) for a, b in zip(ev0, ev1)]))
        ms_per_step = wall * 1e3 / args.steps

        it = out["iterations"].cpu().numpy()
        p = out["params"].cpu().numpy()
        executed = np.minimum(it, cfg["iters"]).astype(np.float64)
        fkind = "splinegrid" if est.spline_form() == "grid" else cfg["kind"]
        per0 = fit_flops_per_loc(fkind, cfg["R"], 0)
        per1 = fit_flops_per_loc(fkind, cfg["R"], 1) - per0
        fit_flops = n * per0 + per1 * executed.sum()
        crlb_flops = n * crlb_flops_per_loc(fkind, cfg["R"])
        achieved = fit_flops / (fit_ms * 1e-3) / 1e12
        traffic = hbm_traffic(args.config) if n == cfg["n"] else None

        # PCIe-inclusive rate through the host-pointer API (samples H2D, results D2H)
        m = min(n, 200_000)
        smp_host = smp[:m].cpu().numpy()
        e2e = None if args.no_e2e else host_api_rate(est, smp_host)

        cpu = None if args.no_cpu else cpu_baseline(cfg, smp.cpu().numpy(), est)
        group = est.GetGroupSize()
        form = est.spline_form()

    value = n / (ms_per_step * 1e-3)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "localizations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": cfg["workload"], "roi": cfg["R"], "spots": n, "iterations": cfg["iters"],
                   "lanes_per_spot": group, **({"spline_form": form} if form else {})},
        "roofline": {"kernel": "fit_kernel", "bound": "mfma", "compute": COMPUTE_NOTE, "achieved": achieved,
                     "peak": PEAK_FP32,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32, "traffic": traffic},
        "cpu_baseline": cpu,
        "detail": {
            "step_gpu_ms": step_gpu_ms,
            "fit_kernel_ms": fit_ms,
            "crlb_kernel_ms": crlb_ms,
            "fit_kernel_algorithmic_flop": fit_flops,
            "crlb_kernel_algorithmic_flop": crlb_flops,
            "crlb_kernel_tflops": crlb_flops / (crlb_ms * 1e-3) / 1e12 if crlb_ms > 0 else None,
            "mean_iterations": float(executed.mean()),
            "stopped_fraction": float(np.mean(it <= cfg["iters"])),
            "hbm_bytes_per_launch": n * hbm_bytes_per_loc(cfg["kind"], cfg["R"]),
            "hbm_GBps": n * hbm_bytes_per_loc(cfg["kind"], cfg["R"]) / (step_gpu_ms * 1e-3) / 1e9,
            "pcie_inclusive_loc_per_s": e2e,
            "finite_fraction": float(np.mean(np.all(np.isfinite(p), axis=1))),
        },
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
