"""Estimator API, mirroring photonpy/cpp/estimator.py:27-396.

Same method names, argument order, dtypes, output shapes and ValueError checks
as the reference; every computation runs in the HIP library.  The
``*_device`` methods take torch tensors that already live on the GPU and run
asynchronously on the current torch stream (or a given one); they have no
reference counterpart.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from .context import Context
from .lib import EstimatorProperties, LevMarParams


class Estimator:
    def __init__(self, ctx: Context, psfInst, calib=None):
        if not psfInst:
            ctx.smlm.check()
            raise RuntimeError("photonpy_amd: estimator creation failed")
        self.inst = psfInst
        self.ctx = ctx
        self.calib = calib
        self._lib = ctx.smlm.lib
        self._smlm = ctx.smlm

        props = EstimatorProperties()
        self._lib.Estimator_GetProperties(self.inst, props)
        self.numparams = props.numParams
        self.numconst = props.numConst
        self.samplecount = props.sampleCount
        self.indexdims = props.sampleIndexDims
        self.sampleshape = np.zeros(self.indexdims, dtype=np.int32)
        self._lib.Estimator_SampleDims(self.inst, self.sampleshape)
        self.numdiag = props.numDiag
        self.fmt = self.ParamFormat()
        self.param_names = [s.strip() for s in self.fmt.split(',')]
        self.limits = np.zeros((self.numparams, 2), dtype=np.float32)
        self._lib.Estimator_GetParamLimits(self.inst, self.limits)
        self.SetLevMarParams(1e-18, 50)  # estimator.py:151

    def _check(self):
        self._smlm.check()

    # ------------------------------------------------------------ validation
    # One validator per argument kind.  Shapes and the ValueError behaviour follow
    # the reference API (estimator.py:155-187, 317-337): a single parameter row is
    # promoted to [1, K]; constants may be None only when the model has none and
    # otherwise are broadcast to [n, NumConstants]; roi positions default to zeros.
    def _spot_params(self, params):
        p = np.asarray(params, dtype=np.float32)
        if p.ndim == 1:
            p = p[np.newaxis]
        if p.ndim != 2 or p.shape[1] != self.numparams:
            raise ValueError(f"parameters must be [n, {self.numparams}] in the order {self.fmt!r}; "
                             f"got shape {p.shape}")
        return np.ascontiguousarray(p)

    def _spot_constants(self, constants, n):
        nc = self.numconst
        if constants is None:
            if nc:
                raise ValueError(f"this estimator needs {nc} constants per spot: pass an array of shape [{n}, {nc}]")
            return np.zeros((n, 0), np.float32)
        c = np.asarray(constants, dtype=np.float32)
        if c.size == n * nc:
            return np.ascontiguousarray(c.reshape(n, nc))
        try:
            return np.ascontiguousarray(np.broadcast_to(c, (n, nc)))
        except ValueError:
            raise ValueError(f"constants of shape {c.shape} cannot be broadcast to [{n}, {nc}]") from None

    def _spot_roipos(self, roipos, n):
        if roipos is None:
            return np.zeros((n, self.indexdims), np.int32)
        r = np.ascontiguousarray(roipos, dtype=np.int32)
        if r.shape != (n, self.indexdims):
            raise ValueError(f"roi positions must be [{n}, {self.indexdims}]; got shape {r.shape}")
        return r

    def _spot_stack(self, imgs, n=None, what="samples"):
        s = np.ascontiguousarray(imgs, dtype=np.float32)
        want = tuple(int(v) for v in self.sampleshape)
        if s.ndim != 1 + len(want) or tuple(s.shape[1:]) != want or (n is not None and len(s) != n):
            lead = "n" if n is None else str(n)
            raise ValueError(f"{what} must be [{lead}, {', '.join(map(str, want))}]; got shape {s.shape}")
        return s

    def _checkparams(self, params, constants, roipos):
        """(params [n,K], constants [n,NumConstants], roipos [n,indexdims]) as contiguous arrays."""
        p = self._spot_params(params)
        return p, self._spot_constants(constants, len(p)), self._spot_roipos(roipos, len(p))

    @staticmethod
    def _fisher(deriv, ev):
        """Fisher matrices sum_px J_k J_l / mu of [n,K,...] derivatives and [n,...] expected
        values, mu floored at 1e-9 as in the reference's numpy route."""
        n, K = deriv.shape[:2]
        J = deriv.reshape(n, K, -1).astype(np.float64)
        w = 1.0 / np.maximum(ev.reshape(n, -1).astype(np.float64), 1e-9)
        return np.einsum("nkp,nlp,np->nkl", J, J, w)

    @staticmethod
    def _sd_of(cov):
        return np.sqrt(np.abs(np.diagonal(cov, axis1=1, axis2=2)))

    # ------------------------------------------------------------------ model
    def ExpectedValue(self, params, roipos=None, constants=None):
        """Expected value of each pixel of the ROIs: [len(params), *sampleshape]."""
        p, c, r = self._checkparams(params, constants, roipos)
        ev = np.zeros((len(p), *self.sampleshape), dtype=np.float32)
        if len(p):
            self._lib.Estimator_ComputeExpectedValue(self.inst, len(p), p, c, r, ev)
            self._check()
        return ev

    def GenerateSample(self, params, roipos=None, constants=None):
        return np.random.poisson(self.ExpectedValue(params, roipos, constants))

    def Derivatives(self, params, roipos=None, constants=None):
        """(deriv [N,K,*shape], ev [N,*shape])"""
        p, c, r = self._checkparams(params, constants, roipos)
        deriv = np.zeros((len(p), self.numparams, *self.sampleshape), dtype=np.float32)
        ev = np.zeros((len(p), *self.sampleshape), dtype=np.float32)
        if len(p):
            self._lib.Estimator_ComputeDerivatives(self.inst, len(p), p, c, r, deriv, ev)
            self._check()
        return deriv, ev

    def FisherMatrix(self, params, roipos=None, constants=None):
        """[N,K,K] Fisher matrices from the analytic derivatives (float32)."""
        p, c, r = self._checkparams(params, constants, roipos)
        if not len(p):
            return np.zeros((0, self.numparams, self.numparams), np.float32)
        deriv, ev = self.Derivatives(p, r, c)
        return self._fisher(deriv, ev).astype(np.float32)

    def CRLBAndChiSquare(self, params, samples=None, roipos=None, constants=None):
        """(crlb, chisq) when samples are given, otherwise crlb; computed by the library."""
        p, c, r = self._checkparams(params, constants, roipos)
        n = len(p)
        smp = None if samples is None else self._spot_stack(samples, n)
        crlb = np.zeros((n, self.numparams), dtype=np.float32)
        chisq = np.zeros(n, dtype=np.float32)
        if n:
            self._lib.Estimator_ChiSquareAndCRLB(self.inst, p, smp, c, r, n, crlb, chisq)
            self._check()
        return crlb if smp is None else (crlb, chisq)

    def ChiSquare(self, params, samples, roipos=None, constants=None):
        return self.CRLBAndChiSquare(params, samples, roipos, constants)[1]

    def CRLB(self, params, roipos=None, constants=None):
        """CRLB from the pseudo-inverse of FisherMatrix (the reference's numpy route)."""
        return self._sd_of(np.linalg.pinv(self.FisherMatrix(params, roipos, constants)))

    def NumDeriv(self, params, roipos=None, constants=None, eps=1e-6):
        """Central differences of ExpectedValue, (deriv [N,K,*shape], ev [N,*shape]).
        All 2K shifted parameter sets go to the library in one call."""
        p, c, r = self._checkparams(params, constants, roipos)
        n, K = p.shape
        ev = self.ExpectedValue(p, r, c)
        if not n:
            return np.zeros((0, K, *self.sampleshape), np.float32), ev
        shift = np.zeros((2, K, 1, K), np.float32)
        k = np.arange(K)
        shift[0, k, 0, k] = eps
        shift[1, k, 0, k] = -eps
        stacked = (p[np.newaxis, np.newaxis] + shift).reshape(2 * K * n, K)
        evs = self.ExpectedValue(stacked, np.tile(r, (2 * K, 1)), np.tile(c, (2 * K, 1)))
        evs = evs.reshape(2, K, n, *self.sampleshape)
        deriv = np.moveaxis((evs[0] - evs[1]) / (2 * eps), 0, 1)
        return np.ascontiguousarray(deriv, dtype=np.float32), ev

    def NumCRLB(self, params, roipos=None, constants=None, eps=1e-4, useNumDeriv=True):
        """CRLB from the inverse Fisher matrix of numerical (or analytic) derivatives."""
        if useNumDeriv:
            deriv, ev = self.NumDeriv(params, roipos, constants, eps)
        else:
            deriv, ev = self.Derivatives(params, roipos, constants)
        return self._sd_of(np.linalg.inv(self._fisher(deriv, ev)))

    # ---------------------------------------------------------------- fitting
    def GetLevMarParams(self):
        lmp = LevMarParams()
        self._lib.Estimator_GetLMParams(self.inst, lmp)
        self.lmparams = lmp
        return lmp

    def SetLevMarParams(self, lambda_=1e-15, iterations=50, lmFactor=0.0):
        self.lmparams = LevMarParams(iterations=int(iterations), lambda_=lambda_, factor=lmFactor)
        self._lib.Estimator_SetLMParams(self.inst, self.lmparams)

    def Estimate(self, imgs, roipos=None, constants=None, initial=None):
        """Fit ROIs [N,*sampleshape]: (params [N,K], diag [N,numdiag], traces), traces[i]
        being spot i's parameters before each iteration it ran ([iterations_i, K])."""
        smp = self._spot_stack(imgs, what="sample images")
        n, K = len(smp), self.numparams
        c, r = self._spot_constants(constants, n), self._spot_roipos(roipos, n)
        if initial is not None:
            initial = np.ascontiguousarray(initial, dtype=np.float32)
            if initial.shape != (n, K):
                raise ValueError(f"initial values must be [{n}, {K}]; got shape {initial.shape}")
        tl = self.lmparams.iterations
        fitted = np.zeros((n, K), np.float32)
        diag = np.zeros((n, self.numdiag), np.float32)
        its = np.zeros(n, np.int32)
        trace = np.zeros((n, tl, K), np.float32)
        if n:
            self._lib.Estimator_Estimate(self.inst, n, smp, c, r, initial, fitted, diag, its, trace, tl)
            self._check()
        self.last_iterations = its
        return fitted, diag, [trace[i, :its[i]] for i in range(n)]

    # ------------------------------------------------------- device tensors
    def _stream(self, stream):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream()
        return stream.cuda_stream

    def _dev(self, t, dtype, shape):
        import torch
        if not (isinstance(t, torch.Tensor) and t.is_cuda):
            raise ValueError("expected a CUDA (HIP) tensor")
        if t.dtype != dtype:
            raise ValueError(f"expected dtype {dtype}, got {t.dtype}")
        if not t.is_contiguous():
            raise ValueError("expected a contiguous tensor")
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"expected shape {tuple(shape)}, got {tuple(t.shape)}")
        return t.data_ptr()

    def _dev_constants(self, constants, n):
        import torch
        if self.numconst == 0:
            return None
        if constants is None:
            raise ValueError(f"this estimator needs constants [{n}, {self.numconst}]")
        return self._dev(constants, torch.float32, (n, self.numconst))

    def estimate_device(self, samples, initial=None, crlb=True, iterations_out=True, trace_len=0, stream=None,
                        constants=None):
        """Fit device-resident ROIs [N,H,W] (float32).  Returns a dict of tensors
        params [N,K], iterations [N], crlb [N,K] + chisq [N] when crlb=True, and
        diag [N,numdiag] for estimators with diagnostics (a GLRT wrapper).
        constants [N, NumConstants] for the models that have per-spot constants."""
        import torch
        n = samples.shape[0]
        K = self.numparams
        dev = samples.device
        ps = self._dev(samples, torch.float32, (n, *self.sampleshape))
        pc = self._dev_constants(constants, n)
        pi = None if initial is None else self._dev(initial, torch.float32, (n, K))
        out = {"params": torch.empty((n, K), dtype=torch.float32, device=dev)}
        out["iterations"] = torch.empty(n, dtype=torch.int32, device=dev) if iterations_out else None
        if crlb:
            out["crlb"] = torch.empty((n, K), dtype=torch.float32, device=dev)
            out["chisq"] = torch.empty(n, dtype=torch.float32, device=dev)
        tr = torch.zeros((n, trace_len, K), dtype=torch.float32, device=dev) if trace_len > 0 else None
        if tr is not None:
            out["trace"] = tr
        rc = self._lib.Estimator_EstimateDevice(
            self.inst, n, ps, pc, pi, out["params"].data_ptr(),
            out["iterations"].data_ptr() if out["iterations"] is not None else None,
            tr.data_ptr() if tr is not None else None, trace_len,
            out["crlb"].data_ptr() if crlb else None, out["chisq"].data_ptr() if crlb else None,
            self._stream(stream))
        if rc != 0:
            self._check()
            raise RuntimeError(f"Estimator_EstimateDevice failed ({rc})")
        if self.numdiag > 0:
            out["diag"] = torch.zeros((n, self.numdiag), dtype=torch.float32, device=dev)
            rc = self._lib.Estimator_DiagnosticsDevice(self.inst, n, ps, pc, out["params"].data_ptr(),
                                                       out["diag"].data_ptr(), self._stream(stream))
            if rc != 0:
                self._check()
                raise RuntimeError(f"Estimator_DiagnosticsDevice failed ({rc})")
        return out

    def expected_value_device(self, params, stream=None, constants=None):
        import torch
        n = params.shape[0]
        pp = self._dev(params, torch.float32, (n, self.numparams))
        pc = self._dev_constants(constants, n)
        ev = torch.empty((n, *self.sampleshape), dtype=torch.float32, device=params.device)
        rc = self._lib.Estimator_ExpectedValueDevice(self.inst, n, pp, pc, ev.data_ptr(), self._stream(stream))
        if rc != 0:
            self._check()
            raise RuntimeError(f"Estimator_ExpectedValueDevice failed ({rc})")
        return ev

    def derivatives_device(self, params, stream=None, constants=None):
        import torch
        n = params.shape[0]
        pp = self._dev(params, torch.float32, (n, self.numparams))
        pc = self._dev_constants(constants, n)
        ev = torch.empty((n, *self.sampleshape), dtype=torch.float32, device=params.device)
        dv = torch.empty((n, self.numparams, *self.sampleshape), dtype=torch.float32, device=params.device)
        rc = self._lib.Estimator_DerivativesDevice(self.inst, n, pp, pc, dv.data_ptr(), ev.data_ptr(),
                                                   self._stream(stream))
        if rc != 0:
            self._check()
            raise RuntimeError(f"Estimator_DerivativesDevice failed ({rc})")
        return dv, ev

    def crlb_device(self, params, samples=None, stream=None, constants=None):
        import torch
        n = params.shape[0]
        pp = self._dev(params, torch.float32, (n, self.numparams))
        pc = self._dev_constants(constants, n)
        ps = None if samples is None else self._dev(samples, torch.float32, (n, *self.sampleshape))
        crlb = torch.empty((n, self.numparams), dtype=torch.float32, device=params.device)
        chisq = torch.zeros(n, dtype=torch.float32, device=params.device)
        rc = self._lib.Estimator_ChiSquareAndCRLBDevice(self.inst, n, pp, pc, ps, crlb.data_ptr(), chisq.data_ptr(),
                                                        self._stream(stream))
        if rc != 0:
            self._check()
            raise RuntimeError(f"Estimator_ChiSquareAndCRLBDevice failed ({rc})")
        return crlb, chisq

    # ------------------------------------------------ adjoint (autograd.py)
    def num_shared(self):
        """Number of shared PSF parameters (sigma_x/y, astigmatism coefficients or spline table)."""
        return self._lib.Estimator_NumShared(self.inst)

    def shared_parameters(self, device="cuda"):
        """Current shared PSF parameters as a float32 tensor (a copy)."""
        import torch
        k = self.num_shared()
        out = np.zeros(k, np.float32)
        if k:
            self._lib.Estimator_GetShared(self.inst, out.ctypes.data)
        return torch.from_numpy(out).to(device)

    def _nll_grad(self, params, samples, shared, gout, nll, gparams, gshared, stream=None):
        ptr = lambda t: None if t is None else t.data_ptr()
        rc = self._lib.Estimator_NLLGradDevice(self.inst, params.shape[0], ptr(params), ptr(samples), ptr(shared),
                                               ptr(gout), ptr(nll), ptr(gparams), ptr(gshared), self._stream(stream))
        if rc != 0:
            self._check()
            raise RuntimeError(f"Estimator_NLLGradDevice failed ({rc})")

    def SetGroupSize(self, group):
        """Lane-group size of the fit kernel (power of two, 0 = built-in choice)."""
        if self._lib.Estimator_SetGroupSize(self.inst, int(group)) != 0:
            raise ValueError(f"invalid group size {group}")

    def SetRefill(self, refill):
        """Stopped lane groups per batched refill in the fit kernel (0 = default)."""
        if self._lib.Estimator_SetRefill(self.inst, int(refill)) != 0:
            raise ValueError(f"invalid refill batch {refill}")

    def GetGroupSize(self):
        return self._lib.Estimator_GetGroupSize(self.inst)

    def SetWorkSegments(self, segments):
        """Work counters of a fit launch larger than the resident grid: 1, or 8 (the
        spots in contiguous segments with a counter each); results do not depend on it."""
        if self._lib.Estimator_SetWorkSegments(self.inst, int(segments)) != 0:
            raise ValueError(f"invalid work segment count {segments}")

    def GetWorkSegments(self):
        return self._lib.Estimator_GetWorkSegments(self.inst)

    def spline_form(self):
        """'grid' or 'coefficients' for a spline estimator (None otherwise): the form its
        fit and expected-value kernels evaluate the PSF in."""
        f = self._lib.Estimator_SplineForm(self.inst)
        return None if f < 0 else ("grid" if f == 1 else "coefficients")

    def spline_strip_length(self):
        """Grid-form spline fits: the strip length of the pixel layout at the current
        group size (0: flat layout), None for other estimators."""
        L = self._lib.Estimator_SplineStripLength(self.inst)
        return None if L < 0 else L

    def set_spline_mode(self, mode):
        """'auto' (the grid form when the table is a CSpline_Compute interpolant) or
        'coefficients'."""
        m = {"auto": 0, "coefficients": 1}[mode]
        if self._lib.Estimator_SetSplineMode(self.inst, m) != 0:
            raise ValueError("set_spline_mode applies to spline estimators only")

    def set_phase_timing(self, enable=True):
        """Record HIP events around the fit and CRLB kernels of every fit launch."""
        if self._lib.Estimator_SetPhaseTiming(self.inst, 1 if enable else 0) != 0:
            raise RuntimeError("Estimator_SetPhaseTiming failed")

    def phase_times(self):
        """(launches, mean fit-kernel ms, mean CRLB-kernel ms) of the launches recorded
        since timing was enabled or last read; waits for them and releases the events."""
        fit, crlb = ct.c_float(), ct.c_float()
        n = self._lib.Estimator_GetPhaseTimes(self.inst, ct.byref(fit), ct.byref(crlb))
        if n < 0:
            self._check()
            raise RuntimeError("Estimator_GetPhaseTimes failed")
        return n, fit.value, crlb.value

    # --------------------------------------------------------------- misc
    def Destroy(self):
        if self.inst is not None:
            self._lib.Estimator_Delete(self.inst)
            self.inst = None

    def NumConstants(self):
        return self.numconst

    def ParamFormat(self):
        return self._lib.Estimator_ParamFormat(self.inst).decode("utf-8")

    def ParamIndex(self, paramName):
        if isinstance(paramName, (list, tuple)):
            return [self.ParamIndex(p) for p in paramName]
        return self.param_names.index(paramName)

    def NumParams(self):
        return self.numparams

    def SampleCount(self):
        return self.samplecount

    def NumDiag(self):
        return self.numdiag

    def ParamLimits(self):
        """(min[NumParams], max[NumParams])"""
        return self.limits[:, 0], self.limits[:, 1]

    def __enter__(self):
        return self

    def __exit__(self, *args):
        self.Destroy()
