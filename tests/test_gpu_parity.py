"""GPU parity of the HIP estimator with the CPU oracle (reference restatement).

Bar (spot-fitting contract): on converged spots the fitted x, y agree within
1e-3 px, z within 1 nm, and I, bg within 1e-4 relative (floor 1 photon);
expected values and Jacobians agree to FP32 rounding.  Sizes are small enough
for the oracle to finish in seconds; the full-size runs use size-independent
properties (noise-free fixed point, CRLB-normalised error statistics).
"""
import numpy as np
import pytest

import photonpy_amd as pa
from oracle import oracle as O
from tests import _closed as C
from tests import _models as MD
from tests import _parity as P

pytestmark = pytest.mark.gpu

KINDS = ["fixed", "sigma", "astig", "spline"]
TOL = P.TOL
_zcol, _on_bounds, _compare = P.zcol, P.on_bounds, P.compare


@pytest.fixture(scope="module")
def ctx():
    with pa.Context() as c:
        yield c


@pytest.mark.parametrize("kind", KINDS)
def test_expected_value_and_derivatives(ctx, kind):
    th, _ = MD.samples(kind, 200, 11)
    est, om, fn = MD.make(kind, ctx)
    ev = est.ExpectedValue(th)
    dv, ev2 = est.Derivatives(th)
    ev32 = om.expected_value(th, np.float32)
    dv32, _ = om.derivatives(th, np.float32)
    mu64 = fn(th)
    scale = np.abs(mu64).max(axis=(1, 2))[:, None, None]
    assert np.max(np.abs(ev - mu64) / scale) < 2e-6
    assert np.max(np.abs(ev - ev32) / scale) < 2e-6
    assert np.array_equal(ev, ev2)
    dscale = np.maximum(np.abs(dv32).max(axis=(2, 3), keepdims=True), 1e-6)
    assert np.max(np.abs(dv - dv32) / dscale) < 5e-5
    if kind == "spline":
        # the GPU contracts the 64 coefficients in a different order: same as FP64 oracle to FP32 rounding
        dv64, _ = om.derivatives(th, np.float64)
        assert np.max(np.abs(dv - dv64) / dscale) < 5e-5


@pytest.mark.parametrize("kind,B", [("fixed", True), ("sigma", True), ("astig", True), ("astig", False),
                                    ("spline", True)])
def test_estimate_parity(ctx, kind, B):
    """GPU Estimate vs the oracle at each model's workload ROI size.  ("astig", B=False)
    is the reference's own astigmatic model (no quartic term, GaussianPSFModels.h:636-751)."""
    n = {"fixed": 4000, "sigma": 4000, "astig": 1500, "spline": 600}[kind]
    iters = 30 if kind == "fixed" else 50
    lam = 1e-15 if kind == "fixed" else 1e-18
    th, smp = MD.samples(kind, n, 3, B=B)
    est, om, _ = MD.make(kind, ctx, B=B)
    est.SetLevMarParams(lam, iters)
    p_gpu, _, traces = est.Estimate(smp)
    it_gpu = est.last_iterations
    assert np.array_equal([len(t) for t in traces], np.minimum(it_gpu, iters))
    P.check_fit_parity(kind, om, smp, p_gpu, it_gpu, iters, lam, est.ParamLimits(), tag=f"{kind} B={B}")


@pytest.mark.parametrize("kind", KINDS)
def test_crlb_and_chisq(ctx, kind):
    th, smp = MD.samples(kind, 500, 5)
    est, om, _ = MD.make(kind, ctx)
    crlb, chi = est.CRLBAndChiSquare(th, smp)
    c64, x64 = om.crlb(th, smp, dtype=np.float64)
    c32, x32 = om.crlb(th, smp, dtype=np.float32)
    assert np.max(np.abs(crlb - c64) / c64) < 1e-4
    assert np.max(np.abs(crlb - c32) / c32) < 2e-3
    assert np.max(np.abs(chi - x64) / np.maximum(x64, 1)) < 1e-4
    # without samples chi^2 is zero and the CRLB is unchanged
    crlb2 = est.CRLBAndChiSquare(th)
    assert np.allclose(crlb2, crlb, rtol=1e-6)


@pytest.mark.parametrize("kind", KINDS)
def test_fused_fit_crlb_matches_separate_calls(ctx, kind):
    import torch
    th, smp = MD.samples(kind, 700, 9)
    est, _, _ = MD.make(kind, ctx)
    d_smp = torch.from_numpy(smp).cuda()
    out = est.estimate_device(d_smp, crlb=True)
    torch.cuda.synchronize()
    p = out["params"].cpu().numpy()
    p_host, _, _ = est.Estimate(smp)
    assert np.array_equal(p, p_host)
    assert np.array_equal(out["iterations"].cpu().numpy(), est.last_iterations)
    crlb, chi = est.CRLBAndChiSquare(p, smp)
    # well-conditioned spots: the two kernels sum the Fisher matrix in different orders,
    # which its inverse amplifies by up to its condition number (taken scale-free, on
    # the correlation form D^-1/2 FI D^-1/2)
    fi = est.FisherMatrix(p).astype(np.float64)
    d = np.sqrt(np.abs(np.diagonal(fi, axis1=1, axis2=2)))
    with np.errstate(all="ignore"):
        cond = np.linalg.cond(fi / (d[:, :, None] * d[:, None, :]))
    ok = np.all(np.isfinite(crlb), axis=1) & (crlb[:, 0] < 0.5) & (cond < 1e3)
    assert ok.mean() > 0.5
    assert np.allclose(out["crlb"].cpu().numpy()[ok], crlb[ok], rtol=1e-4)
    assert np.allclose(out["chisq"].cpu().numpy()[ok], chi[ok], rtol=1e-4, atol=1e-3)


def test_iteration_and_trace_semantics(ctx):
    th, smp = MD.samples("fixed", 300, 13)
    est, om, _ = MD.make("fixed", ctx)
    est.SetLevMarParams(1e-15, 0)
    p0, _, tr0 = est.Estimate(smp, initial=th.astype(np.float32))
    assert np.all(est.last_iterations == 1) and np.array_equal(p0, th.astype(np.float32))
    est.SetLevMarParams(1e-15, 2)
    est.Estimate(smp)
    _, it_ref = om.estimate(smp, iterations=2, lambda_=1e-15)
    assert np.array_equal(est.last_iterations, it_ref)
    est.SetLevMarParams(1e-15, 30)
    p, _, tr = est.Estimate(smp, initial=th.astype(np.float32))
    assert all(np.array_equal(t[0], th[i].astype(np.float32)) for i, t in enumerate(tr))
    assert np.all(est.last_iterations >= 2)


@pytest.mark.parametrize("kind", KINDS)
def test_degenerate_rois(ctx, kind):
    R = MD.ROI[kind]
    est, om, _ = MD.make(kind, ctx)
    est.SetLevMarParams(1e-18, 20)
    smp = np.zeros((4, R, R), np.float32)
    smp[1] = 7.0                      # flat
    smp[2, 0, 0] = 50.0               # single corner pixel
    smp[3] = 5.0
    smp[3, R // 2, R // 2] = np.nan   # NaN pixel
    p, _, _ = est.Estimate(smp)
    pr, _ = om.estimate(smp, iterations=20)
    # the all-zero ROI (0/0 centre of mass) and the NaN pixel give the oracle's NaN pattern;
    # flat and single-pixel ROIs are singular fits whose outcome depends on the solve's
    # precision (FP64 here, FP32 in the reference): they only have to stay in the clamp box
    for i in (0, 3):
        assert np.array_equal(np.isfinite(p[i]), np.isfinite(pr[i])), (i, p[i], pr[i])
    fin = np.isfinite(p[:, 0])
    assert np.all((p[fin, 0] >= 1.5) & (p[fin, 0] <= R - 2.5))
    assert np.all((p[fin, 1] >= 1.5) & (p[fin, 1] <= R - 2.5))


def test_invalid_roisize_raises(ctx):
    # any roisize >= 1 is fitted (sizes without a per-size kernel run on the any-size
    # kernel, tests/test_gpu_large_roi.py); a non-positive one is refused at creation
    with pytest.raises(pa.NativeError, match="roisize"):
        pa.GaussianPSFMethods(ctx).CreatePSF_XYIBg(0, 1.5, True)
    est = pa.GaussianPSFMethods(ctx).CreatePSF_XYIBg(40, 1.5, True)
    assert est.ExpectedValue(np.array([[20.0, 20.0, 100.0, 1.0]])).shape == (1, 40, 40)


@pytest.mark.parametrize("R,G", [(10, 4), (10, 16), (11, 4), (16, 8), (20, 16), (7, 2)])
def test_group_sizes_agree(ctx, R, G):
    th, smp = MD.samples("sigma", 800, 21, R=R)
    est, _, _ = MD.make("sigma", ctx, R=R)
    p1, _, _ = est.Estimate(smp)
    it1 = est.last_iterations.copy()
    est.SetGroupSize(G)
    p2, _, _ = est.Estimate(smp)
    conv = (it1 <= 50) & (est.last_iterations <= 50)
    # the group size changes the summation order, and a chaotic low-SNR path may then
    # stop elsewhere: allow 1 spot in 100 (a wrong pixel split would move most of them)
    bad = ~_compare("sigma", p2[conv], p1[conv])
    assert bad.sum() <= max(2, 1e-2 * conv.sum()), f"{bad.sum()} of {conv.sum()} differ"


@pytest.mark.parametrize("kind,R,Gs,Gw", [("fixed", 9, 2, 4), ("astig", 9, 2, 4), ("fixed", 10, 4, 2),
                                           ("astig", 10, 4, 2)])
def test_split_middle_row_agrees_with_row_layout(ctx, kind, R, Gs, Gw):
    # Gs lanes per spot split the middle row of each lane pair's band of 2R/Gs rows
    # between the pair (9x9 at 2 lanes, 10x10 at 4); at Gw lanes every lane owns whole
    # rows.  Both pixel partitions must give the same fits (astig: the z column's row
    # terms go through the split row too).
    th, smp = MD.samples(kind, 800, 29, R=R)
    est, _, _ = MD.make(kind, ctx, R=R)
    lo, hi = est.ParamLimits()
    est.SetGroupSize(Gw)
    pw, _, _ = est.Estimate(smp)
    est.SetGroupSize(Gs)
    ps, _, _ = est.Estimate(smp)
    ok = np.all(np.isfinite(pw), axis=1) & ~_on_bounds(pw, lo, hi)
    assert ok.mean() > 0.5
    bad = ~_compare(kind, ps[ok], pw[ok])
    assert bad.sum() <= max(2, 1e-2 * ok.sum()), f"{bad.sum()} of {ok.sum()} differ"
    # the fused CRLB and chi^2 (Fisher matrix from the same row sums) agree as well
    import torch
    d = torch.from_numpy(smp).cuda()
    est.SetGroupSize(Gw)
    ow = est.estimate_device(d, crlb=True)
    est.SetGroupSize(Gs)
    os_ = est.estimate_device(d, crlb=True)
    cw, cs = ow["crlb"].cpu().numpy(), os_["crlb"].cpu().numpy()
    xw, xs = ow["chisq"].cpu().numpy(), os_["chisq"].cpu().numpy()
    m = ok & np.all(np.isfinite(cw), axis=1) & (cw[:, 0] < 0.5)
    assert m.mean() > 0.5
    same = np.all(np.isclose(cs[m], cw[m], rtol=1e-3), axis=1) & np.isclose(xs[m], xw[m], rtol=1e-3, atol=1e-3)
    assert same.mean() > 0.98, f"{(~same).sum()} of {m.sum()} CRLB/chi^2 differ"


def test_refill_batches_do_not_change_results(ctx):
    # a spot's fit does not depend on when its group is refilled
    th, smp = MD.samples("sigma", 3000, 23)
    est, _, _ = MD.make("sigma", ctx)
    outs = []
    for rf in (1, 4, 32):
        est.SetRefill(rf)
        p, _, _ = est.Estimate(smp)
        c, x = est.CRLBAndChiSquare(p, smp)
        outs.append((p, est.last_iterations.copy()))
    est.SetRefill(0)
    for p, it in outs[1:]:
        assert np.array_equal(p, outs[0][0], equal_nan=True) and np.array_equal(it, outs[0][1])


@pytest.mark.parametrize("kind", ["sigma", "astig"])
def test_work_segments_do_not_change_results(ctx, kind):
    # a launch larger than the resident grid hands out spots from one counter or from
    # eight segment counters; which group fits a spot does not change its fit
    import torch
    from photonpy_amd import synthetic
    R = {"sigma": 11, "astig": 16}[kind]
    est = synthetic.make_estimator(ctx, kind, R)
    est.SetLevMarParams(1e-18, 50)
    _, smp = synthetic.make_samples(est, kind, 400_000, seed=5)
    outs = []
    for segs in (1, 8):
        est.SetWorkSegments(segs)
        assert est.GetWorkSegments() == segs
        o = est.estimate_device(smp, crlb=True)
        torch.cuda.synchronize()
        outs.append({k: v.cpu().numpy() for k, v in o.items()})
    with pytest.raises(ValueError):
        est.SetWorkSegments(3)
    for k in ("params", "iterations", "crlb"):
        assert np.array_equal(outs[0][k], outs[1][k], equal_nan=True), k


def test_spline_grid_form_matches_coefficient_form(ctx):
    """The spline's grid form (default for CSpline_Compute tables) and its coefficient form
    agree on expected values, derivatives, CRLB and fits."""
    th, smp = MD.samples("spline", 600, 31)
    est, om, _ = MD.make("spline", ctx)
    assert est.spline_form() == "grid"
    ev_g = est.ExpectedValue(th)
    dv_g, _ = est.Derivatives(th)
    crlb_g, chi_g = est.CRLBAndChiSquare(th, smp)
    est.SetLevMarParams(1e-18, 50)
    p_g, _, _ = est.Estimate(smp)
    it_g = est.last_iterations.copy()
    est.set_spline_mode("coefficients")
    assert est.spline_form() == "coefficients"
    ev_c = est.ExpectedValue(th)
    dv_c, _ = est.Derivatives(th)
    crlb_c, chi_c = est.CRLBAndChiSquare(th, smp)
    p_c, _, _ = est.Estimate(smp)
    it_c = est.last_iterations.copy()
    est.set_spline_mode("auto")
    scale = np.abs(ev_c).max(axis=(1, 2))[:, None, None]
    assert np.max(np.abs(ev_g - ev_c) / scale) < 2e-6
    dscale = np.maximum(np.abs(dv_c).max(axis=(2, 3), keepdims=True), 1e-6)
    assert np.max(np.abs(dv_g - dv_c) / dscale) < 5e-5
    ok = np.all(np.isfinite(crlb_c), axis=1) & (crlb_c[:, 0] < 0.5)
    assert np.allclose(crlb_g[ok], crlb_c[ok], rtol=1e-3)
    assert np.allclose(chi_g, chi_c, rtol=1e-4, atol=1e-3)
    lo, hi = est.ParamLimits()
    last_ok = ~_on_bounds(p_c, lo, hi) & ~_on_bounds(p_g, lo, hi)
    agree = _compare("spline", p_g[last_ok], p_c[last_ok])
    assert agree.mean() > 0.97, agree.mean()
    # the grid form also matches the oracle (coefficient form) through test_estimate_parity


@pytest.mark.parametrize("R,G", [(20, 16), (20, 32), (16, 32), (11, 16), (7, 4), (25, 64)])
def test_spline_grid_group_sizes_agree(ctx, R, G):
    th, smp = MD.samples("spline", 400, 33, R=R)
    est, _, _ = MD.make("spline", ctx, R=R)
    est.SetLevMarParams(1e-18, 50)
    p1, _, _ = est.Estimate(smp)
    c1, x1 = est.CRLBAndChiSquare(p1, smp)
    est.SetGroupSize(G)
    p2, _, _ = est.Estimate(smp)
    # a spot's arithmetic does not depend on G except the order of the group sums
    lo, hi = est.ParamLimits()
    use = ~_on_bounds(p1, lo, hi) & ~_on_bounds(p2, lo, hi)
    assert _compare("spline", p2[use], p1[use]).mean() > 0.97


def test_spline_coefficients_match_oracle(ctx):
    stack, _ = C.spline_zstack(nz=12, ny=15, nx=15)
    c_gpu = pa.CSplineMethods(ctx).ComputeSplineCoeff(stack)
    c_ref = O.spline_compute(stack)
    assert c_gpu.shape == c_ref.shape
    assert np.max(np.abs(c_gpu - c_ref)) <= 1e-6 * max(np.abs(c_ref).max(), 1)
    c_b = pa.CSplineMethods(ctx).ComputeSplineCoeff(stack, border=1)
    assert np.allclose(c_b, O.spline_compute(stack, border=1), atol=1e-6)


# ------------------------------------------------------------- full size
def _full_size_fixed_point(ctx, kind, n, seed):
    import torch
    rng = np.random.default_rng(seed)
    R = MD.ROI[kind]
    th = C.synth_params(kind, R, n, rng).astype(np.float32)
    est, _, _ = MD.make(kind, ctx)
    d_th = torch.from_numpy(th).cuda()
    mu = est.expected_value_device(d_th)
    start = th.copy()
    start[:, :2] += 0.2
    iI = 3 if kind in ("astig", "spline") else 2
    start[:, iI] *= 1.05
    out = est.estimate_device(mu, initial=torch.from_numpy(start).cuda(), crlb=True)
    p = out["params"].cpu().numpy()
    del mu
    err = np.abs(p - th)
    ok = (err[:, 0] < 1e-3) & (err[:, 1] < 1e-3) & (err[:, iI] < 1e-3 * th[:, iI])
    if kind in ("astig", "spline"):
        ok &= err[:, 2] < 1.0
    return ok.mean(), out


@pytest.mark.parametrize("kind,n", [("sigma", 1_000_000), ("astig", 2_000_000), ("fixed", 10_000),
                                    ("spline", 500_000)])
def test_full_size_noise_free_fixed_point(ctx, kind, n):
    frac, out = _full_size_fixed_point(ctx, kind, n, 99)
    assert frac > 0.999, frac
    crlb = out["crlb"].cpu().numpy()
    assert np.mean(np.all(np.isfinite(crlb) & (crlb > 0), axis=1)) > 0.999


def test_full_size_error_statistics_match_crlb(ctx):
    """C2 at full size: normalised errors (fit - truth)/CRLB have unit spread."""
    import torch
    n = 1_000_000
    rng = np.random.default_rng(5)
    th = C.synth_params("sigma", 11, n, rng).astype(np.float32)
    est, _, _ = MD.make("sigma", ctx)
    mu = est.expected_value_device(torch.from_numpy(th).cuda())
    gen = torch.Generator(device="cuda").manual_seed(5)
    smp = torch.poisson(mu, generator=gen)
    out = est.estimate_device(smp, crlb=True)
    p = out["params"].cpu().numpy()
    crlb = out["crlb"].cpu().numpy()
    it = out["iterations"].cpu().numpy()
    good = (it <= 50) & (th[:, 2] > 2000) & np.all(np.isfinite(crlb), axis=1)
    # CRLB of x, y, I (sigma's CRLB carries the reference's 1/sigma scaling)
    for k in (0, 1, 2):
        z = (p[good, k] - th[good, k]) / crlb[good, k]
        assert abs(np.std(z) - 1) < 0.1, (k, np.std(z))
        assert abs(np.mean(z)) < 0.05


@pytest.mark.parametrize("kind", KINDS)
def test_error_against_fp64_fit(ctx, kind):
    """Each model's error against an FP64 CPU fit (the oracle's estimator in FP64) on spots
    that both fits converge on, away from the parameter bounds.  Prints one line per model
    (`FP64-ERR ...`, shown with pytest -s) and asserts the 99th percentile is inside the
    parity tolerance."""
    n = {"fixed": 4000, "sigma": 4000, "astig": 1500, "spline": 600}[kind]
    iters = 30 if kind == "fixed" else 50
    lam = 1e-15 if kind == "fixed" else 1e-18
    th, smp = MD.samples(kind, n, 21)
    est, om, _ = MD.make(kind, ctx)
    est.SetLevMarParams(lam, iters)
    p_gpu, _, _ = est.Estimate(smp)
    it_gpu = est.last_iterations
    p64, it64 = om.estimate(smp, iterations=200, lambda_=lam, dtype=np.float64)
    lo, hi = est.ParamLimits()
    use = (it64 <= 200) & ~_on_bounds(p64, lo, hi) & ~_on_bounds(p_gpu, lo, hi)
    use &= np.all(np.isfinite(p64), axis=1) & np.all(np.isfinite(p_gpu), axis=1)
    if kind in ("fixed", "sigma"):
        use &= it_gpu <= iters
    assert use.mean() > 0.5
    d = np.abs(p_gpu[use] - p64[use])
    iI = 3 if _zcol(kind) else 2
    rel = d.copy()
    rel[:, iI] /= np.maximum(np.abs(p64[use, iI]), 1)
    rel[:, iI + 1] /= np.maximum(np.abs(p64[use, iI + 1]), 1)
    names = est.param_names
    p99 = np.percentile(rel, 99, axis=0)
    print(f"FP64-ERR {kind} spots={use.sum()} " + " ".join(
        f"{nm}:p50={np.median(rel[:, k]):.2e},p99={p99[k]:.2e},max={rel[:, k].max():.2e}"
        for k, nm in enumerate(names)))
    assert p99[0] <= TOL["xy"] and p99[1] <= TOL["xy"]
    assert p99[iI] <= TOL["rel"] and p99[iI + 1] <= TOL["rel"]
    if _zcol(kind):
        assert p99[2] <= TOL["z"]


def test_grid_spline_window_covers_rounded_cells(ctx):
    """A pixel's spline cell is (int)(y - ty) in FP32; for positions just above an
    integer (e.g. y = 10.000001, so -ty = 2.9999990) the sum rounds up and the cell lands
    one past the window origin + y.  The grid form's LDS window must cover that cell:
    its fused CRLB (per-pixel evaluation from the window) agrees with the stand-alone
    CRLB kernel (evaluation from the grid in HBM), and repeated fits are identical."""
    import torch
    est, _, _ = MD.make("spline", ctx)
    assert est.spline_form() == "grid"
    rng = np.random.default_rng(41)
    vals = np.array([k + d for k in (8, 9, 10, 11) for d in (1e-7, 5e-7, 1e-6, 2e-6, 0.3)], np.float32)
    n = 32768
    th = np.zeros((n, 5), np.float32)
    th[:, 0] = rng.choice(vals, n)
    th[:, 1] = rng.choice(vals, n)
    th[:, 2] = rng.uniform(-300, 300, n)
    th[:, 3] = rng.uniform(1000, 5000, n)
    th[:, 4] = rng.choice(np.array([1e-6, 2.0], np.float32), n)
    smp = torch.poisson(est.expected_value_device(torch.from_numpy(th).cuda()))
    d_th = torch.from_numpy(th).cuda()
    est.SetLevMarParams(1e-18, 0)
    fused = est.estimate_device(smp, initial=d_th, crlb=True)
    c, x = est.crlb_device(d_th, smp)
    torch.cuda.synchronize()
    a = torch.cat([fused["crlb"], fused["chisq"][:, None]], 1).cpu().numpy()
    b = torch.cat([c, x[:, None]], 1).cpu().numpy()
    fin = np.all(np.isfinite(b), axis=1)
    assert fin.mean() > 0.9
    # the two kernels sum in different orders, which near-singular Fisher matrices
    # (bg = 1e-6) amplify; a wrong window entry moves a spot's CRLB by orders of magnitude
    close = np.all(np.isclose(a, b, rtol=1e-2, atol=1e-6), axis=1)
    assert close[fin].mean() > 0.999, f"{(~close[fin]).sum()} of {fin.sum()} spots differ"
    est.SetLevMarParams(1e-18, 50)
    p1 = est.estimate_device(smp, crlb=False)["params"].cpu().numpy()
    p2 = est.estimate_device(smp, crlb=False)["params"].cpu().numpy()
    assert np.array_equal(p1, p2, equal_nan=True)


@pytest.mark.parametrize("R,G", [(20, 32), (25, 64), (20, 64), (16, 16), (9, 16), (7, 8), (22, 32), (28, 64),
                                 (32, 64)])
def test_grid_spline_window_at_grid_edges(ctx, R, G):
    """Spots at the corners of the position range and at both ends of the z range.  Their
    LDS windows run past the grid's last sample rows and columns (on the last z-slab, into
    the zero padding behind the table), entries that no pixel reads; the lanes fill the
    window by columns, one lane per column (20/32, 9/16), two with an odd or even window
    width (25/64, 20/64) or two columns per lane (16/16, 7/8).  The fused CRLB (pixels
    evaluated from the window) agrees with the stand-alone CRLB kernel (pixels evaluated
    from the grid in HBM)."""
    import torch
    est, _, _ = MD.make("spline", ctx, R=R)
    assert est.spline_form() == "grid"
    est.SetGroupSize(G)
    lo, hi = est.ParamLimits()
    rng = np.random.default_rng(43)
    n = 8192
    xy = np.array([lo[0], lo[0] + 0.1, lo[0] + 0.6, hi[0] - 0.6, hi[0] - 0.1, hi[0]], np.float32)
    dz = hi[2] - lo[2]
    zs = np.array([lo[2] + 1e-3, lo[2] + 0.02 * dz, 0.5 * (lo[2] + hi[2]), hi[2] - 0.02 * dz, hi[2] - 1e-3],
                  np.float32)
    th = np.zeros((n, 5), np.float32)
    th[:, 0] = rng.choice(xy, n)
    th[:, 1] = rng.choice(xy, n)
    th[:, 2] = rng.choice(zs, n)
    th[:, 3] = rng.uniform(1000, 5000, n)
    th[:, 4] = rng.choice(np.array([1.0, 5.0], np.float32), n)
    d_th = torch.from_numpy(th).cuda()
    smp = torch.poisson(est.expected_value_device(d_th))
    est.SetLevMarParams(1e-18, 0)
    fused = est.estimate_device(smp, initial=d_th, crlb=True)
    c, x = est.crlb_device(d_th, smp)
    torch.cuda.synchronize()
    a = torch.cat([fused["crlb"], fused["chisq"][:, None]], 1).cpu().numpy()
    b = torch.cat([c, x[:, None]], 1).cpu().numpy()
    fin = np.all(np.isfinite(b), axis=1)
    assert fin.mean() > 0.9
    # summation order alone (the two kernels differ in it) moves a well-conditioned CRLB
    # by ~1e-6 relative; a wrong window entry by far more
    close = np.all(np.isclose(a, b, rtol=1e-3, atol=1e-6), axis=1)
    assert close[fin].all(), f"{(~close[fin]).sum()} of {fin.sum()} spots differ, e.g. {th[fin][~close[fin]][:3]}"
