"""Throughput benchmark of the MI355X spot fitter.

`python bench.py [--gpus 1] [--steps K] [--warmup W] [--config all|c1|c2|c3|c4|c4adj|c5|queue]`

One step = one fused launch (Levenberg-Marquardt fit + CRLB + chi^2) over the
whole batch of synthetic ROIs, already resident in HBM.  Prints one JSON line.

The default run (`--config all`) reports the headline workload c2 in the
top-level fields (K timed steps after W warm-up steps) and times every other
workload of the metric under detail.configs, each with its own roofline and
(c1-c4) CPU baseline:
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
PEAK_FP32 = 157.3  # TFLOP/s, MI355X FP32 vector peak
# FP32 FMA rate without packed FP32 instructions, measured on one MI355X with a kernel of
# independent v_fma_f32 chains (2.36 wave instructions per CU per ns).  Packed FP32 is
# disabled in every kernel here, so this, not PEAK_FP32, is what they can reach.
ATTAINABLE_FP32 = 77.4  # TFLOP/s
# These kernels issue FP32 vector (VALU) instructions and no matrix instructions: the
# bound is the FP32 VALU roofline, labelled "valu".
BOUND = "valu"
COMPUTE_NOTE = "FP32 VALU roofline (no MFMA is issued); packed FP32 disabled"
# Algorithmic FLOP per ROI of the spline adjoint pass (NLL and gradients), per pixel of each spot:
# the spline value and x/y/z derivatives from the cell's 64 coefficients (139 FMA), mu, the
# NLL weight and the 5 parameter-gradient terms (~22), and the spot's share of its cell's
# coefficient gradient, sum_s W_s B_s[q] over the 64 coefficients (64 FMA): about 435 FLOP
# per pixel, 174,000 per 20x20 ROI.
ADJ_FLOP_PER_ROI = 174_000
# Coefficient atomics per ROI of the direct (untiled) form: 20x20 px x 64 coefficients.
ADJ_DIRECT_ATOMICS_PER_ROI = 400 * 64

# ---------------------------------------------------------------------------
# Algorithmic FLOP count per localization (FP32 model, normal equations and damped
# solve, the FP64 CRLB inversion counted as FLOPs too; erf/exp evaluations not counted).
#   per pixel-iteration   K^2 + 4K + 4 (normal equations) + model terms
#   per row/column-iter   row/column terms of the separable Gaussians
#   per iteration         damping + Cholesky + solve: 4K^2 + 4K + 2K^3/3
#   initial estimate      3S (centre of mass)
#   CRLB kernel           S (model + K^2 + 2K + 3) + (W+H) rowcol + 8K^3/3
MODEL_TERMS = {"fixed": 5, "sigma": 8, "astig": 8, "spline": 422, "splinegrid": 136}
ROWCOL_TERMS = {"fixed": 11, "sigma": 15, "astig": 16, "spline": 0, "splinegrid": 0}
# The grid-form spline (tables made by CSpline_Compute) contracts, per spot-iteration, the
# window's four z-slices into a (value, z-derivative) table: (R+3)^2 x 8 FMA (the window
# term below).  Each pixel then needs its 4 x 4 block of that table.  Flat layout: 64 FMA
# per pixel + 8 (finish) = 136 FLOP.  Strip layout of length L (each lane walks vertical
# strips; a window row's x-contraction, 12 FMA, serves the four pixels below it): per
# pixel 16 FMA of y-contraction + 8 (finish), plus, per column piece of l pixels, l + 3
# row contractions, i.e. 24 (R^2 + 3 R P) FLOP per spot with P = ceil(R / L) pieces per
# column.


def spline_pixel_terms(R, strip):
    """Model FLOP per pixel of the grid-form spline at strip length `strip` (0: flat)."""
    if not strip:
        return MODEL_TERMS["splinegrid"]
    P = -(-R // strip)
    return 40 + 24 * (1 + 3 * P / R)


def fit_flops_per_loc(kind, R, executed_iters, strip=0):
    """Fit kernel: centre of mass + executed iterations."""
    K = 4 if kind == "fixed" else 5
    S = R * R
    window = (R + 3) ** 2 * 16 if kind == "splinegrid" else 0
    model = spline_pixel_terms(R, strip) if kind == "splinegrid" else MODEL_TERMS[kind]
    per_iter = S * (K * K + 4 * K + 4 + model) + 2 * R * ROWCOL_TERMS[kind] + window + \
        (4 * K * K + 4 * K + 2 * K ** 3 / 3)
    return per_iter * executed_iters + 3 * S


def crlb_flops_per_loc(kind, R, strip=0):
    """CRLB kernel: Fisher matrix + chi^2 + FP64 LU inversion."""
    K = 4 if kind == "fixed" else 5
    S = R * R
    window = (R + 3) ** 2 * 16 if kind == "splinegrid" else 0
    model = spline_pixel_terms(R, strip) if kind == "splinegrid" else MODEL_TERMS[kind]
    return S * (model + K * K + 2 * K + 3) + 2 * R * ROWCOL_TERMS[kind] + window + 8 * K ** 3 / 3


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


def host_info():
    """Host the CPU baseline runs on: logical CPUs, the CPUs this process may use, the model."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "model": model}


# A one-GPU job gets one GPU's share of the host's CPUs: worker pools stay within 16 threads.
CPU_SHARE = 16


def oracle_model(cfg, est):
    from oracle import oracle as O
    kind, R = cfg["kind"], cfg["R"]
    if kind == "fixed":
        return O.gauss2d(R, 1.5)
    if kind == "sigma":
        return O.gauss2d_sigma(R, 1.6)
    if kind == "astig":
        a = synthetic_astig()
        return O.astig(R, a[0], a[1], [-600.0, 600.0], B=(a[2], a[2]))
    coefs = est.shared_parameters("cpu").numpy().reshape(61, 26, 26, 64)
    lo, hi = est.ParamLimits()
    return O.spline(R, coefs, float(lo[2]), float(hi[2]))


def synthetic_astig():
    """Reference parametrisation of the astigmatic workload: {s0, gamma, d, A/d} per axis, B/d^2."""
    return [1.3, 200.0, 400.0, 0.1 / 400], [1.3, -200.0, 400.0, 0.1 / 400], 0.05 / 400 ** 2


# ROIs of each workload kept on the host for its CPU baseline
CPU_MAX_ROIS = 131_072


def cpu_job(cfg, smp, est):
    """The CPU baseline of one workload, prepared now (oracle model, host copy of the
    first ROIs) and run later by the caller, after every GPU measurement, so that the
    host load of the oracle's threads cannot slow a GPU leg that follows it."""
    om = oracle_model(cfg, est)
    host_rois = smp[:CPU_MAX_ROIS].cpu().numpy()
    return lambda: cpu_baseline(cfg, host_rois, om)


def cpu_baseline(cfg, smp_host, om, run_s=2.0, repeats=5):
    """The oracle (FP32 restatement of the reference estimator) on the host: Estimate + CRLB
    + chi^2 over the first m ROIs of the same batch (the batch repeated when it is smaller),
    m sized from a warm run so that one run takes about run_s seconds; one warm-up run,
    then the median of `repeats` timed runs."""
    host = host_info()
    threads = max(1, min(CPU_SHARE, host["usable_cpus"]))
    total = len(smp_host)

    def rois(m):
        if m <= total:
            return smp_host[:m]
        return np.concatenate([smp_host] * (m // total) + [smp_host[:m % total]])

    def run(s):
        t = time.perf_counter()
        p, _ = om.estimate(s, iterations=cfg["iters"], lambda_=cfg["lam"], nthreads=threads)
        om.crlb(p, s, nthreads=threads)
        return time.perf_counter() - t

    m = min(total, 1024)
    run(rois(m))  # first run: thread start-up, page faults
    dt = run(rois(m))
    for _ in range(2):  # per-call overhead makes small runs slow per ROI: rescale once more
        m = max(threads, int(m * run_s / max(dt, 1e-6)))
        s = rois(m)
        dt = run(s)  # also the warm-up of the timed runs
        if dt >= 0.6 * run_s:
            break
    times = sorted(run(s) for _ in range(repeats))
    med = times[len(times) // 2]
    print(f"bench: cpu baseline {cfg['kind']} R={cfg['R']}: {m} ROIs, median {med:.3f} s on {threads} threads",
          file=sys.stderr, flush=True)
    what = f"first {m} ROIs of the same batch" if m <= total else f"the first {total} ROIs of the batch, {m} in all"
    return {"value": m / med, "unit": "localizations/s", "cores": threads, "kind": "port",
            "sample": f"{what}; oracle Estimate + CRLB + chi^2 on {threads} threads (one GPU's {CPU_SHARE}-core "
                      f"share of the host); median of {repeats} runs after one warm-up, run after all GPU legs",
            "runs_s": [round(t, 4) for t in times], "host": host}


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


def line(metric_value, unit, steps, warmup, ms, workload, extra_config, roofline, cpu, detail):
    return {"metric": METRIC, "value": metric_value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic", "config": {"workload": workload, **extra_config}, "roofline": roofline,
            "cpu_baseline": cpu, "detail": detail}


def bench_fit(name, steps, warmup, n=0, group=0, refill=0, cpu=True, e2e=True, segs=0):
    """One single-model workload: `steps` timed fused launches after `warmup` untimed ones.
    Returns (bench line, CPU-baseline job or None); finish_cpu runs the job into the line."""
    import torch
    import photonpy_amd as pa
    from photonpy_amd import autotune, synthetic
    cfg = synthetic.WORKLOADS[name]
    n = n or cfg["n"]
    torch.cuda.set_device(0)
    with pa.Context() as ctx:
        est = synthetic.make_estimator(ctx, cfg["kind"], cfg["R"])
        est.SetLevMarParams(cfg["lam"], cfg["iters"])
        autotune.apply(est)
        if group:
            est.SetGroupSize(group)
        if refill:
            est.SetRefill(refill)
        if segs:
            est.SetWorkSegments(segs)
        th, smp = synthetic.make_samples(est, cfg["kind"], n, seed=1)
        torch.cuda.synchronize()
        for _ in range(warmup):
            out = est.estimate_device(smp, crlb=True)
        torch.cuda.synchronize()
        # the timed steps: nothing but the fits (no events between or inside them)
        t0 = time.perf_counter()
        for k in range(steps):
            out = est.estimate_device(smp, crlb=True)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms_per_step = wall * 1e3 / steps
        # the same steps again, instrumented: HIP events on the launch stream around each
        # step and around each kernel (fit, then CRLB) for the kernel times of the roofline
        stream = torch.cuda.current_stream()
        ev0 = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
        ev1 = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
        est.set_phase_timing(True)
        torch.cuda.synchronize()
        for k in range(steps):
            ev0[k].record(stream)
            out = est.estimate_device(smp, crlb=True)
            ev1[k].record(stream)
        torch.cuda.synchronize()
        launches, fit_ms, crlb_ms = est.phase_times()
        est.set_phase_timing(False)
        assert launches == steps, (launches, steps)
        step_gpu_ms = float(np.mean([a.elapsed_time(b) for a, b in zip(ev0, ev1)]))

        it = out["iterations"].cpu().numpy()
        p = out["params"].cpu().numpy()
        executed = np.minimum(it, cfg["iters"]).astype(np.float64)
        fkind = "splinegrid" if est.spline_form() == "grid" else cfg["kind"]
        strip = est.spline_strip_length() or 0
        per0 = fit_flops_per_loc(fkind, cfg["R"], 0, strip)
        per1 = fit_flops_per_loc(fkind, cfg["R"], 1, strip) - per0
        fit_flops = n * per0 + per1 * executed.sum()
        crlb_flops = n * crlb_flops_per_loc(fkind, cfg["R"], strip)
        achieved = fit_flops / (fit_ms * 1e-3) / 1e12
        traffic = hbm_traffic(name) if n == cfg["n"] else None
        e2e_rate = host_api_rate(est, smp[:min(n, 200_000)].cpu().numpy()) if e2e else None
        job = cpu_job(cfg, smp, est) if cpu else None
        group = est.GetGroupSize()
        form = est.spline_form()
        del smp, out
    torch.cuda.empty_cache()
    roof = {"kernel": "fit_kernel", "bound": BOUND, "compute": COMPUTE_NOTE, "achieved": achieved,
            "peak": PEAK_FP32, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32, "attainable_peak": ATTAINABLE_FP32,
            "frac_of_attainable": achieved / ATTAINABLE_FP32, "traffic": traffic}
    detail = {
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
        "pcie_inclusive_loc_per_s": e2e_rate,
        "finite_fraction": float(np.mean(np.all(np.isfinite(p), axis=1))),
    }
    extra = {"roi": cfg["R"], "spots": n, "iterations": cfg["iters"], "lanes_per_spot": group,
             **({"spline_form": form} if form else {}),
             **({"spline_strip_length": strip} if form == "grid" else {})}
    return line(n / (ms_per_step * 1e-3), "localizations/s", steps, warmup, ms_per_step, cfg["workload"], extra,
                roof, None, detail), job


def finish_cpu(result, job):
    """Runs a workload's CPU-baseline job into its bench line."""
    if job is None:
        return
    cpu_line = job()
    result["cpu_baseline"] = cpu_line
    result["detail"]["gpu_over_cpu"] = result["value"] / cpu_line["value"]


def bench_mixed(steps, warmup, n=0):
    """Mixed batch (4,000,000 ROIs over 28 model x size bins): the batch form (bins on
    concurrent streams, one HIP graph launch per step) against the plain per-bin launches
    one after another on one stream; both compute fit + CRLB + chi^2 for every spot.  The
    plain launches are also timed per bin (HIP events around each bin's fit and CRLB
    kernels) for the time shares."""
    import torch
    import photonpy_amd as pa
    from photonpy_amd import autotune, synthetic
    from photonpy_amd.batch import MixedBatch
    with pa.Context() as ctx:
        bins = synthetic.mixed_bins(ctx, n=n or None)
        ests = [b["est"] for b in bins]
        counts = [b["n"] for b in bins]
        samples = [b["samples"] for b in bins]
        n = sum(counts)
        stream = torch.cuda.current_stream()
        with MixedBatch(ests, counts, streams=autotune.batch_streams(), tune=True) as mb:
            outs = mb.allocate_outputs()
            ref = mb.allocate_outputs()

            def timed(fn):
                for _ in range(warmup):
                    fn()
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                t0 = time.perf_counter()
                e0.record(stream)
                for _ in range(steps):
                    fn()
                e1.record(stream)
                torch.cuda.synchronize()
                return (time.perf_counter() - t0) * 1e3 / steps, e0.elapsed_time(e1) / steps

            plain_ms, plain_gpu = timed(lambda: mb.run_plain(samples, ref))
            batch_ms, batch_gpu = timed(lambda: mb.run(samples, outs, use_graph=True))
            # Per-bin kernel times: each bin's plain fit + CRLB launch on its own (t1), and
            # the same launch over the bin's ROIs twice (t2).  t2 - t1 is the bin's marginal
            # work at full occupancy; 2 t1 - t2 its fixed cost (launch, ramp-up, tail), the
            # most any schedule of the bins can win back.
            def kernel_ms(e, smp):
                e.set_phase_timing(True)
                for _ in range(2):
                    e.estimate_device(smp, crlb=True)
                torch.cuda.synchronize()
                _, f_ms, c_ms = e.phase_times()
                e.set_phase_timing(False)
                return f_ms + c_ms

            bin_ms, bin_ms2 = [], []
            for e, smp in zip(ests, samples):
                bin_ms.append(kernel_ms(e, smp))
                doubled = torch.cat([smp, smp])
                bin_ms2.append(kernel_ms(e, doubled))
                del doubled
            same = all(bool(((a[k] == b[k]) | (a[k].isnan() & b[k].isnan())).all()) if a[k].is_floating_point()
                       else torch.equal(a[k], b[k]) for a, b in zip(outs, ref) for k in a)
            flops = 0.0
            for b, o in zip(bins, outs):
                it = np.minimum(o["iterations"].cpu().numpy(), b["iters"]).astype(np.float64)
                fk = "splinegrid" if b["est"].spline_form() == "grid" else b["kind"]
                strip = b["est"].spline_strip_length() or 0
                per0 = fit_flops_per_loc(fk, b["R"], 0, strip)
                per1 = fit_flops_per_loc(fk, b["R"], 1, strip) - per0
                flops += b["n"] * (per0 + crlb_flops_per_loc(fk, b["R"], strip)) + per1 * it.sum()
            streams = mb.num_streams()
            groups = [e.GetGroupSize() for e in ests]
            forms = sorted({b["est"].spline_form() for b in bins if b["kind"] == "spline"})
            meta = [(b["kind"], b["R"]) for b in bins]
        del samples, outs, ref, bins
    torch.cuda.empty_cache()
    total = max(sum(bin_ms), 1e-9)
    shares = sorted(({"kind": k, "R": R, "n": c, "lanes": g, "ms": round(t, 4), "share": round(t / total, 4),
                      "fixed_ms": round(2 * t - t2, 4)}
                     for (k, R), c, g, t, t2 in zip(meta, counts, groups, bin_ms, bin_ms2)), key=lambda d: -d["ms"])
    marginal = sum(t2 - t for t, t2 in zip(bin_ms, bin_ms2))
    achieved = flops / (batch_gpu * 1e-3) / 1e12
    roof = {"kernel": "all fit and CRLB kernels of the batch", "bound": BOUND, "compute": COMPUTE_NOTE,
            "achieved": achieved, "peak": PEAK_FP32, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32,
            "attainable_peak": ATTAINABLE_FP32, "frac_of_attainable": achieved / ATTAINABLE_FP32, "traffic": None}
    detail = {"plain_ms_per_step": plain_ms, "plain_loc_per_s": n / (plain_ms * 1e-3),
              "batch_over_plain": plain_ms / batch_ms, "batch_gpu_ms": batch_gpu, "plain_gpu_ms": plain_gpu,
              "results_identical": bool(same), "bin_time_shares": shares,
              "sum_of_bin_ms": total, "sum_of_marginal_bin_ms": marginal,
              "schedule_bound_over_plain": total / max(marginal, 1e-9)}
    return line(n / (batch_ms * 1e-3), "localizations/s", steps, warmup, batch_ms, synthetic.MIXED["workload"],
                {"spots": n, "bins": len(counts), "streams": streams, "spline_form": forms}, roof, None, detail)


def bench_queue(steps, warmup, n=0):
    """The reference's own throughput path (photonpy/tests/psf_queue.py): C2's ROIs in
    host memory streamed through an EstimQueue (batches of 4096 on 4 streams, scheduled
    10,000 at a time, results collected as they finish).  Host-to-host rate, PCIe
    included; the queue's batches use the same fused kernels as the device path."""
    import torch
    import photonpy_amd as pa
    from photonpy_amd import autotune, synthetic
    cfg = synthetic.WORKLOADS["c2"]
    n = n or cfg["n"]
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

        def timed(batch, nsteps, nwarm):
            # the queue is created outside the timed region, as the reference's benchmark does
            with pa.EstimQueue(est, batchSize=batch, numStreams=4) as q:
                for _ in range(nwarm):
                    run_once(q)
                t0 = time.perf_counter()
                for _ in range(nsteps):
                    got = run_once(q)
                return got, (time.perf_counter() - t0) / nsteps

        got, dt = timed(4096, steps, warmup)
        got_big, dt_big = timed(65536, 1, 1)
    assert got == n and got_big == n, (got, got_big, n)
    return line(n / dt, "localizations/s", steps, warmup, dt * 1e3, "EstimQueue from host memory: " + cfg["workload"],
                {"spots": n, "batch": 4096, "streams": 4, "schedule_chunk": chunk}, None, None,
                {"note": "host-to-host including PCIe transfers; compare detail.pcie_inclusive_loc_per_s of c2",
                 "batch65536_loc_per_s": n / dt_big})


def spline_adjoint_rounds(th, R, shape, zmin, zmax):
    """Rounds of the spline adjoint for spot parameters th: up to 64 spots of one window
    (z-slab, first pixel's cell row and column, FP32 as on the GPU) per round."""
    Nz, Ny, Nx = shape
    f32 = np.float32
    tx = (th[:, 0].astype(f32) - f32(Nx * 0.5)).astype(f32)
    ty = (th[:, 1].astype(f32) - f32(Ny * 0.5)).astype(f32)
    du = f32((f32(zmax) - f32(zmin)) / f32(Nz))
    spz = ((th[:, 2].astype(f32) - f32(zmin)) / du).astype(f32)
    k = np.clip(np.trunc(spz).astype(np.int64), 0, Nz - 1)
    i0 = np.clip(np.trunc((f32(0) - tx).astype(f32)).astype(np.int64), 0, Nx - 1)
    j0 = np.clip(np.trunc((f32(0) - ty).astype(f32)).astype(np.int64), 0, Ny - 1)
    _, cnt = np.unique((k * Ny + j0) * Nx + i0, return_counts=True)
    return int(np.sum((cnt + 63) // 64)), len(cnt)


def bench_adjoint(steps, warmup, n=0):
    """Spline PSF calibration step: forward Poisson NLL + hand-written adjoint on 65,536
    20x20 ROIs, gradients w.r.t. per-spot parameters and all 2,638,976 spline coefficients
    (mu and J recomputed from the parameters, nothing stored).  With inputs that need
    gradients the forward runs the adjoint once for unit upstream weights, yielding the
    NLL too, and the backward scales its gradients by the .sum()'s upstream gradient.
    The timed steps carry no events; a second pass of the same steps times forward and
    backward with HIP events on the launch stream."""
    import torch
    import photonpy_amd as pa
    from photonpy_amd import synthetic
    from photonpy_amd.autograd import poisson_nll
    n = n or 65_536
    R = 20
    with pa.Context() as ctx:
        est = synthetic.make_estimator(ctx, "spline", R)
        th, smp = synthetic.make_samples(est, "spline", n, seed=1)
        params = torch.from_numpy(th).cuda().requires_grad_(True)
        coefs = est.shared_parameters().requires_grad_(True)
        lo, hi = est.ParamLimits()
        rounds, windows = spline_adjoint_rounds(th, R, (61, 26, 26), float(lo[2]), float(hi[2]))
        stream = torch.cuda.current_stream()
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]

        def step(e=None):
            params.grad = None
            coefs.grad = None
            if e:
                e[0].record(stream)
            loss = poisson_nll(est, params, smp, coefs).sum()
            if e:
                e[1].record(stream)
            loss.backward()
            if e:
                e[2].record(stream)

        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            step()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        # the same steps again with HIP events around the forward and the backward
        for k in range(steps):
            step(ev[k])
        torch.cuda.synchronize()
        fwd_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev]))
        bwd_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in ev]))
        finite = bool(torch.isfinite(coefs.grad).all().item() and torch.isfinite(params.grad).all().item())
        del smp, params, coefs
    torch.cuda.empty_cache()
    ms = wall * 1e3 / steps
    # the adjoint pass (NLL, parameter and coefficient gradients for unit upstream weights)
    # runs in the forward; the backward scales its results
    achieved = n * ADJ_FLOP_PER_ROI / (fwd_ms * 1e-3) / 1e12
    atomics_tiled = rounds * R * R * 64  # at most one 64-lane atomic per pixel cell and round
    roof = {"kernel": "forward pass (window sort + spline_adjoint_round_kernel<true>: NLL and gradients)",
            "bound": BOUND,
            "compute": COMPUTE_NOTE, "achieved": achieved, "peak": PEAK_FP32, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32, "attainable_peak": ATTAINABLE_FP32,
            "frac_of_attainable": achieved / ATTAINABLE_FP32, "traffic": None}
    detail = {"forward_ms": fwd_ms, "backward_ms": bwd_ms, "adjoint_algorithmic_flop": n * ADJ_FLOP_PER_ROI,
              "windows": windows, "rounds": rounds,
              "coefficient_atomics_per_step": atomics_tiled,
              "coefficient_atomics_per_step_direct_form": n * ADJ_DIRECT_ATOMICS_PER_ROI,
              "gradients_finite": finite}
    return line(n / (ms * 1e-3), "ROI gradient evaluations/s", steps, warmup, ms,
                "cubic-spline PSF adjoint: forward NLL + backward w.r.t. params and 61x26x26x64 coefficients, "
                "20x20 px, 65,536 ROIs", {"roi": R, "spots": n}, roof, None, detail)


# timed steps / warm-up steps of the workloads reported under detail.configs of the
# default run (sized so each takes well under a second of GPU time).  The GPU runs its
# first few large launches about 4% slower than later ones (measured on C2: 3.44 ms per
# step over the 10 steps after 2 warm-up steps, 3.31-3.32 ms over every later block), so
# each workload warms up for at least about 30 ms of GPU time before its timed steps.
SUB_STEPS = {"c1": (200, 600), "c3": (10, 4), "c4": (5, 2), "c4adj": (10, 60), "c5": (5, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="all", choices=["all", "c1", "c2", "c3", "c4", "c4adj", "c5", "queue"])
    ap.add_argument("--n", type=int, default=0, help="override the number of spots")
    ap.add_argument("--group", type=int, default=0, help="lane-group size override (0 = tuned/default)")
    ap.add_argument("--refill", type=int, default=0, help="refill batch override (0 = tuned/default)")
    ap.add_argument("--segs", type=int, default=0, help="work-segment override: 1 or 8 (0 = tuned/default)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-e2e", action="store_true", help="skip the PCIe-inclusive host-API leg")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB 5 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(spots shard independently across GPUs, see the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes, "launches": 5}))
        return

    cpu = not args.no_cpu
    if args.config == "c4adj":
        print(json.dumps(bench_adjoint(args.steps, args.warmup, args.n)))
        return
    if args.config == "c5":
        print(json.dumps(bench_mixed(args.steps, args.warmup, args.n)))
        return
    if args.config == "queue":
        print(json.dumps(bench_queue(args.steps, args.warmup, args.n)))
        return
    if args.config != "all":
        res, job = bench_fit(args.config, args.steps, args.warmup, args.n, args.group, args.refill, cpu=cpu,
                             e2e=not args.no_e2e, segs=args.segs)
        finish_cpu(res, job)
        print(json.dumps(res))
        return
    # headline: c2 with the driver's K and W; then every other workload of the metric
    t_all = time.perf_counter()
    head, head_job = bench_fit("c2", args.steps, args.warmup, cpu=cpu, e2e=not args.no_e2e)
    print(f"bench: c2 {head['value']:.4g} loc/s", file=sys.stderr, flush=True)
    configs, jobs = {}, {}
    for name in ("c1", "c3", "c4"):
        k, w = SUB_STEPS[name]
        configs[name], jobs[name] = bench_fit(name, k, w, cpu=cpu, e2e=False)
        print(f"bench: {name} {configs[name]['value']:.4g} loc/s", file=sys.stderr, flush=True)
    k, w = SUB_STEPS["c4adj"]
    configs["c4adj"] = bench_adjoint(k, w)
    print(f"bench: c4adj {configs['c4adj']['value']:.4g} ROI/s", file=sys.stderr, flush=True)
    k, w = SUB_STEPS["c5"]
    configs["c5"] = bench_mixed(k, w)
    print(f"bench: c5 {configs['c5']['value']:.4g} loc/s", file=sys.stderr, flush=True)
    # CPU baselines last: no GPU leg runs after the host has been loaded
    finish_cpu(head, head_job)
    for name, job in jobs.items():
        finish_cpu(configs[name], job)
    for c in configs.values():
        for key in ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "dtype", "data"):
            c.pop(key, None)
    head["detail"]["configs"] = configs
    head["detail"]["bench_wall_s"] = time.perf_counter() - t_all
    print(json.dumps(head))


if __name__ == "__main__":
    main()
