"""Adjoint of the batched Poisson NLL (torch.autograd.Function + HIP kernel).

Bar: gradients match central finite differences of the FP64 oracle NLL (eps = 1e-3
relative to each shared PSF parameter, the spline coefficients included) and the
reference's own analytic derivatives (oracle Derivatives) to 1e-4 relative.
"""
import numpy as np
import pytest

import photonpy_amd as pa
from photonpy_amd.autograd import poisson_nll
from oracle import oracle as O
from tests import _closed as C
from tests import _models as MD

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    with pa.Context() as c:
        yield c


def _nll64(om, th, smp):
    return -om.loglik(th, smp, dtype=np.float64)


def _setup(ctx, kind, n=48, seed=17):
    import torch
    th, smp = MD.samples(kind, n, seed)
    # evaluate away from the optimum so every gradient is well away from zero
    th = th.copy()
    th[:, 0] += 0.15
    th[:, 1] -= 0.1
    est, om, _ = MD.make(kind, ctx)
    d_th = torch.from_numpy(th.astype(np.float32)).cuda().requires_grad_(True)
    d_smp = torch.from_numpy(smp).cuda()
    return th.astype(np.float32).astype(np.float64), smp, est, om, d_th, d_smp


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30)


@pytest.mark.parametrize("kind", ["fixed", "sigma", "astig", "spline"])
def test_nll_and_param_gradients(ctx, kind):
    import torch
    th, smp, est, om, d_th, d_smp = _setup(ctx, kind)
    nll = poisson_nll(est, d_th, d_smp)
    ref = _nll64(om, th, smp)
    assert _rel(nll.detach().cpu().numpy(), ref) < 1e-5
    nll.sum().backward()
    g = d_th.grad.cpu().numpy()
    # analytic: sum_px (1 - x/mu) J with the oracle's FP64 Jacobian (true sigma column)
    dv, ev = om.derivatives(th, np.float64)
    if kind == "sigma":
        dv[:, 4] /= th[:, 4, None, None]
    w = 1 - np.maximum(smp, 1e-8) / np.maximum(ev, 1e-8)
    if kind == "spline":
        unit = th.copy()
        unit[:, 3], unit[:, 4] = 1, 0
        clamped = om.expected_value(unit, np.float64) <= 0
        dv[:, :3] = np.where(clamped[:, None], 0, dv[:, :3])
    ana = np.einsum("npq,nkpq->nk", w, dv)
    assert _rel(g, ana) < 1e-4
    # finite differences of the FP64 NLL
    eps = np.array([1e-4, 1e-4, 1e-3, 1e-3, 1e-4]) if kind != "fixed" else np.array([1e-4, 1e-4, 1e-3, 1e-4])
    if kind in ("astig", "spline"):
        eps = np.array([1e-4, 1e-4, 1e-2, 1e-3, 1e-4])
    fd = np.zeros_like(g, dtype=np.float64)
    for k in range(th.shape[1]):
        tp, tm = th.copy(), th.copy()
        tp[:, k] += eps[k]
        tm[:, k] -= eps[k]
        fd[:, k] = (_nll64(om, tp, smp) - _nll64(om, tm, smp)) / (2 * eps[k])
    for k in range(th.shape[1]):
        assert _rel(g[:, k], fd[:, k]) < 1e-4, (k, _rel(g[:, k], fd[:, k]))


def _oracle_with_shared(kind, R, shared):
    if kind == "fixed":
        return O.gauss2d(R, shared[0], shared[1])
    if kind == "astig":
        return O.astig(R, shared[0:4], shared[5:9], C.ASTIG_ZRANGE, B=(shared[4], shared[9]))
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["fixed", "astig"])
def test_shared_gaussian_gradients(ctx, kind):
    import torch
    th, smp, est, om, d_th, d_smp = _setup(ctx, kind)
    shared = est.shared_parameters()
    shared.requires_grad_(True)
    nll = poisson_nll(est, d_th, d_smp, shared)
    nll.sum().backward()
    g = shared.grad.cpu().numpy().astype(np.float64)
    base = shared.detach().cpu().numpy().astype(np.float64)
    R = MD.ROI[kind]
    def nll_sum_diff(i, h):
        sp, sm = base.copy(), base.copy()
        sp[i] += h
        sm[i] -= h
        # per-spot differences first: the total NLL is large and would swamp them
        return np.sum(_nll64(_oracle_with_shared(kind, R, sp), th, smp) -
                      _nll64(_oracle_with_shared(kind, R, sm), th, smp))

    for i in range(len(base)):
        h = 1e-3 * max(abs(base[i]), 1e-3)
        if kind == "astig" and i in (4, 9):
            h = 1e-9  # B is ~3e-7
        # 4-point central stencil (truncation O(h^4))
        fd = (8 * nll_sum_diff(i, h) - nll_sum_diff(i, 2 * h)) / (12 * h)
        assert abs(g[i] - fd) <= 1e-4 * max(abs(fd), 1e-3 * np.abs(g).max()), (i, g[i], fd)


def test_spline_coefficient_gradients(ctx):
    import torch
    from tests import _spline_np as SN
    th, smp, est, om, d_th, d_smp = _setup(ctx, "spline", n=32)
    coefs, zmin, zmax = MD.spline_coefs()
    R = MD.ROI["spline"]
    shared = torch.from_numpy(coefs.ravel().copy()).cuda().requires_grad_(True)
    wts = np.linspace(0.5, 1.5, len(th))
    nll = poisson_nll(est, d_th, d_smp, shared)
    (nll * torch.from_numpy(wts.astype(np.float32)).cuda()).sum().backward()
    g = shared.grad.cpu().numpy().reshape(-1, 64).astype(np.float64)
    # analytic: d NLL / d c[cell, q] = sum over pixels in the cell of w (1 - x/mu) I basis_q, psf > 0
    cell, B = SN.cells(th, R, coefs.shape[:3], zmin, zmax)
    C2 = coefs.reshape(-1, 64).astype(np.float64)
    f = np.einsum("nyxq,nq->nyx", C2[cell], B)
    ev = om.expected_value(th, np.float64)
    w = 1 - np.maximum(smp, 1e-8) / np.maximum(ev, 1e-8)
    contrib = (wts[:, None, None] * w * th[:, 3, None, None] * (f > 0))[..., None] * B[:, None, None, :]
    ana = np.zeros_like(C2)
    np.add.at(ana, cell.ravel(), contrib.reshape(-1, 64))
    assert np.max(np.abs(g - ana)) <= 1e-4 * np.abs(ana).max()
    # relative check on the components that carry weight (smaller ones are sums of
    # cancelling pixel terms, covered by the absolute check above)
    touched = np.argwhere(np.abs(ana) > 1e-2 * np.abs(ana).max())
    rel = np.abs(g - ana)[tuple(touched.T)] / np.abs(ana)[tuple(touched.T)]
    assert np.max(rel) < 1e-4, np.max(rel)
    assert np.count_nonzero(g) < g.size  # untouched cells get no gradient
    # central finite differences with eps = 1e-3 relative to the coefficient (an absolute
    # 1e-3 is up to a third of a coefficient here, where the NLL's third derivative makes a
    # central difference several percent off), on coefficients of cells whose pixels stay
    # away from the psf >= 0 clamp within the perturbation (mu is piecewise linear in a
    # coefficient)
    kink = np.zeros(C2.shape[0], bool)
    near = np.abs(f) < 1e-3 * np.abs(B).max(axis=1)[:, None, None] * 2
    kink[cell[near]] = True
    rng = np.random.default_rng(0)
    cand = [tuple(t) for t in touched if not kink[t[0]]]
    pick = [cand[i] for i in rng.choice(len(cand), min(24, len(cand)), replace=False)]
    cmax = float(np.abs(coefs).max())
    for (c, q) in pick:
        idx = np.unravel_index(c, coefs.shape[:3]) + (q,)
        c0 = np.float64(coefs[idx])
        h = 1e-3 * max(abs(c0), 1e-3 * cmax)
        cp, cm = coefs.copy(), coefs.copy()
        cp[idx] = np.float32(c0 + h)
        cm[idx] = np.float32(c0 - h)
        h2 = np.float64(cp[idx]) - np.float64(cm[idx])  # the steps as stored in FP32
        d = wts * (_nll64(O.spline(R, cp, zmin, zmax), th, smp) - _nll64(O.spline(R, cm, zmin, zmax), th, smp))
        fd = d.sum() / h2
        assert abs(g[c, q] - fd) <= 1e-4 * abs(fd) + 1e-7 * np.abs(g).max(), ((c, q), g[c, q], fd)


@pytest.mark.parametrize("R", [7, 20, 25, 30])
def test_spline_coefficient_gradients_many_spots(ctx, R):
    """The coefficient gradient of thousands of spots, half of them packed into a few
    windows (same z-slab, x and y within +-0.2 px) so that windows hold several tiles
    and many rounds (R <= 28 run by window rounds, one spot per lane; R = 30 takes the
    one-wavefront-per-spot form and its per-pixel scatter), against an FP64 sparse
    restatement of
    d NLL / d c[cell, q] = sum over the pixels of the cell of g (1 - x/mu) I basis_q (psf > 0).
    The forward NLL of every spot matches the FP64 oracle's, and the per-spot gradients of
    that pass equal those of a pass that forms no coefficient gradient."""
    import scipy.sparse as sp
    import torch
    from tests import _spline_np as SN
    n = 3000
    rng = np.random.default_rng(40 + R)
    th = C.synth_params("spline", R, n, rng)
    c = (R - 1) / 2
    half = n // 2
    th[:half, 0] = c + rng.uniform(-0.2, 0.2, half)
    th[:half, 1] = c + rng.uniform(-0.2, 0.2, half)
    th[:half, 2] = rng.uniform(0.0, 8.0, half)
    est, om, _ = MD.make("spline", ctx, R=R)
    smp = rng.poisson(om.expected_value(th, np.float64)).astype(np.float32)
    th = th.astype(np.float32)
    th[:, 0] += 0.1  # away from the optimum
    coefs, zmin, zmax = MD.spline_coefs()
    shared = torch.from_numpy(coefs.ravel().copy()).cuda().requires_grad_(True)
    d_th = torch.from_numpy(th).cuda().requires_grad_(True)
    d_smp = torch.from_numpy(smp).cuda()
    gout = np.linspace(0.5, 1.5, n).astype(np.float32)
    nll = poisson_nll(est, d_th, d_smp, shared)
    assert _rel(nll.detach().cpu().numpy(), _nll64(om, th.astype(np.float64), smp)) < 1e-5
    (nll * torch.from_numpy(gout).cuda()).sum().backward()
    g = shared.grad.cpu().numpy().reshape(-1, 64).astype(np.float64)
    gp_tiled = d_th.grad.cpu().numpy()
    # the same per-spot gradients from a pass that forms no coefficient gradient
    d_th2 = torch.from_numpy(th).cuda().requires_grad_(True)
    (poisson_nll(est, d_th2, d_smp, shared.detach()) * torch.from_numpy(gout).cuda()).sum().backward()
    gp_direct = d_th2.grad.cpu().numpy()
    assert np.allclose(gp_tiled, gp_direct, rtol=1e-5, atol=1e-5 * np.abs(gp_direct).max())
    th64 = th.astype(np.float64)
    cell, B = SN.cells(th64, R, coefs.shape[:3], zmin, zmax)
    C2 = coefs.reshape(-1, 64).astype(np.float64)
    f = np.einsum("nyxq,nq->nyx", C2[cell], B)
    ev = om.expected_value(th64, np.float64)
    w = 1 - np.maximum(smp, 1e-8) / np.maximum(ev, 1e-8)
    W = gout[:, None, None] * w * th64[:, 3, None, None] * (f > 0)
    A = sp.coo_matrix((W.ravel(), (cell.ravel(), np.repeat(np.arange(n), R * R))), shape=(C2.shape[0], n)).tocsr()
    ana = A @ B
    assert np.max(np.abs(g - ana)) <= 1e-4 * np.abs(ana).max()
    touched = np.abs(ana) > 1e-2 * np.abs(ana).max()
    assert np.max(np.abs(g - ana)[touched] / np.abs(ana)[touched]) < 1e-4
    assert np.all(g[np.asarray(A.sum(axis=1)).ravel() == 0] == 0)  # untouched cells get nothing


def test_spline_gradients_do_not_depend_on_the_window_sort(ctx):
    """A gradient call with upstream weights reuses the window sort of the previous call
    for as many spots on the same stream (an autograd step's forward sorts, its backward
    reuses).  With
    the sort of other parameters (most spots then outside their round's window, evaluated
    from their own cells) the per-spot gradients are bit-identical to those after a fresh
    sort, and the coefficient gradient agrees to FP32 summation order."""
    import torch
    R, n = 20, 2000
    rng = np.random.default_rng(77)
    th_a = C.synth_params("spline", R, n, rng).astype(np.float32)
    th_b = th_a.copy()
    th_b[:, :3] = th_a[rng.permutation(n), :3]  # other windows for most spots
    est, om, _ = MD.make("spline", ctx, R=R)
    smp = rng.poisson(om.expected_value(th_b.astype(np.float64), np.float64)).astype(np.float32)
    coefs, _, _ = MD.spline_coefs()
    shared = torch.from_numpy(coefs.ravel().copy()).cuda()
    d_a, d_b = torch.from_numpy(th_a).cuda(), torch.from_numpy(th_b).cuda()
    d_smp = torch.from_numpy(smp).cuda()
    gout = torch.from_numpy(np.linspace(0.5, 1.5, n).astype(np.float32)).cuda()
    nll = torch.empty(n, device="cuda")

    def grads():
        gp, gs = torch.empty_like(d_b), torch.zeros_like(shared)
        est._nll_grad(d_b, d_smp, shared, gout, None, gp, gs)
        torch.cuda.synchronize()
        return gp.cpu().numpy(), gs.cpu().numpy()

    est._nll_grad(d_a, d_smp, shared, None, nll, None, None)  # sorts the windows of th_a
    gp_stale, gs_stale = grads()                             # reuses that sort for th_b
    est._nll_grad(d_b, d_smp, shared, None, nll, None, None)  # sorts the windows of th_b
    gp_fresh, gs_fresh = grads()
    assert np.array_equal(gp_stale, gp_fresh)
    assert np.max(np.abs(gs_stale - gs_fresh)) <= 1e-5 * np.abs(gs_fresh).max()


@pytest.mark.parametrize("kind", ["spline", "astig"])
def test_autograd_broadcast_weight_matches_direct_kernel(ctx, kind):
    """The forward runs the adjoint once with unit weights (it also yields the NLL) and
    the backward scales its gradients.  For one upstream weight broadcast over the spots
    (.sum() passes a stride-0 gradient) the results equal one direct adjoint pass with
    that weight per spot: the per-spot gradients bit for bit (0.5 is a power of two, so
    scaling before or after the sums gives the same bits), the shared gradient to the
    order of its atomic sums over spots, which differs from run to run.  The NLL equals
    an NLL-only pass."""
    import torch
    R = 20 if kind == "spline" else 16
    n = 1500
    rng = np.random.default_rng(91)
    th = C.synth_params(kind, R, n, rng)
    est, om, _ = MD.make(kind, ctx, R=R)
    smp = rng.poisson(om.expected_value(th, np.float64)).astype(np.float32)
    th = th.astype(np.float32)
    th[:, 0] += 0.1
    shared0 = est.shared_parameters().detach().clone()
    shared = shared0.clone().requires_grad_(True)
    d_th = torch.from_numpy(th).cuda().requires_grad_(True)
    d_smp = torch.from_numpy(smp).cuda()
    nll = poisson_nll(est, d_th, d_smp, shared)
    (nll.sum() * 0.5).backward()
    gp, gs = torch.empty_like(d_th), torch.zeros_like(shared0)
    nll_only = torch.empty(n, device="cuda")
    est._nll_grad(d_th.detach(), d_smp, shared0, None, nll_only, None, None)
    est._nll_grad(d_th.detach(), d_smp, shared0, torch.full((n,), 0.5, device="cuda"), None, gp, gs)
    torch.cuda.synchronize()
    assert np.array_equal(nll.detach().cpu().numpy(), nll_only.cpu().numpy())
    assert np.array_equal(d_th.grad.cpu().numpy(), gp.cpu().numpy())
    a, b = shared.grad.cpu().numpy(), gs.cpu().numpy()
    assert np.max(np.abs(a - b)) <= 1e-5 * np.abs(b).max()


def test_calibration_step_reduces_nll(ctx):
    """A few gradient steps on sigma of the fixed-sigma model move it towards the truth."""
    import torch
    R = 10
    rng = np.random.default_rng(3)
    th = C.synth_params("fixed", R, 4000, rng)
    smp = rng.poisson(C.mu_fixed(R, th, 1.62)).astype(np.float32)
    est = pa.GaussianPSFMethods(ctx).CreatePSF_XYIBg(R, 1.4, True)
    d_th = torch.from_numpy(th.astype(np.float32)).cuda()
    d_smp = torch.from_numpy(smp).cuda()
    sig = torch.tensor([1.4, 1.4], device="cuda", requires_grad=True)
    opt = torch.optim.Adam([sig], lr=0.02)
    first = None
    for _ in range(60):
        opt.zero_grad()
        loss = poisson_nll(est, d_th, d_smp, sig).sum()
        if first is None:
            first = loss.item()
        loss.backward()
        opt.step()
    assert loss.item() < first
    assert np.allclose(sig.detach().cpu().numpy(), 1.62, atol=0.02)
