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

    def _checkparams(self, params, constants, roipos):
        """estimator.py:155-187"""
        params = np.array(params)
        if len(params.shape) == 1:
            params = [params]
        params = np.ascontiguousarray(params, dtype=np.float32)
        numspots = len(params)
        if constants is None:
            if self.numconst != 0:
                raise ValueError(f'Estimator is expecting constant array with shape {[len(params), self.numconst]} (None given) ')
            constants = np.zeros((numspots, self.numconst), dtype=np.float32)
        else:
            constants = np.ascontiguousarray(constants, dtype=np.float32)
            if constants.size != numspots * self.numconst:
                try:
                    constants_ = np.zeros([numspots, self.numconst], dtype=np.float32)
                    constants_[:, :] = constants
                    constants = constants_
                except Exception:
                    raise ValueError(f'Provided constants with shape {constants.shape}, expecting {[numspots, self.numconst]}')
        if roipos is None:
            roipos = np.zeros((numspots, self.indexdims), dtype=np.int32)
        else:
            roipos = np.ascontiguousarray(roipos, dtype=np.int32)
            if not np.array_equal(roipos.shape, [numspots, self.indexdims]):
                raise ValueError(f'Provided roi positions have incorrect shape: {roipos.shape} should be {[numspots, self.indexdims]}')
        if params.ndim != 2 or params.shape[1] != self.numparams:
            raise ValueError(f"{self.fmt} expected, {params.shape[-1]} parameters given")
        return params, constants, roipos

    # ------------------------------------------------------------------ model
    def ExpectedValue(self, params, roipos=None, constants=None):
        """Expected value of each pixel of the ROIs: [len(params), *sampleshape]."""
        params, constants, roipos = self._checkparams(params, constants, roipos)
        ev = np.zeros((len(params), *self.sampleshape), dtype=np.float32)
        if len(params) > 0:
            self._lib.Estimator_ComputeExpectedValue(self.inst, len(params), params, constants, roipos, ev)
            self._check()
        return ev

    def GenerateSample(self, params, roipos=None, constants=None):
        ev = self.ExpectedValue(params, roipos, constants)
        return np.random.poisson(ev)

    def Derivatives(self, params, roipos=None, constants=None):
        """(deriv [N,K,*shape], ev [N,*shape])"""
        params, constants, roipos = self._checkparams(params, constants, roipos)
        deriv = np.zeros((len(params), self.numparams, *self.sampleshape), dtype=np.float32)
        ev = np.zeros((len(params), *self.sampleshape), dtype=np.float32)
        if len(params) > 0:
            self._lib.Estimator_ComputeDerivatives(self.inst, len(params), params, constants, roipos, deriv, ev)
            self._check()
        return deriv, ev

    def FisherMatrix(self, params, roipos=None, constants=None):
        """estimator.py:205-220 (ev clamped to >= 1e-9)."""
        params, constants, roipos = self._checkparams(params, constants, roipos)
        fi = np.zeros((len(params), self.numparams, self.numparams), dtype=np.float32)
        if len(params) > 0:
            deriv, ev = self.Derivatives(params, roipos, constants)
            K = self.numparams
            ev = np.reshape(ev, (len(params), self.samplecount))
            deriv = np.reshape(deriv, (len(params), K, self.samplecount))
            ev[ev < 1e-9] = 1e-9
            for i in range(K):
                for j in range(K):
                    fi[:, i, j] = np.sum(1 / ev * (deriv[:, i] * deriv[:, j]), axis=1)
        return fi

    def CRLBAndChiSquare(self, params, samples=None, roipos=None, constants=None):
        """(crlb, chisq) if samples are given, otherwise crlb (estimator.py:222-244)."""
        params, constants, roipos = self._checkparams(params, constants, roipos)
        crlb = np.zeros((len(params), self.numparams), dtype=np.float32)
        chisq = np.zeros(len(params), dtype=np.float32)
        if samples is not None:
            samples = np.ascontiguousarray(samples, dtype=np.float32)
            if not np.array_equal(samples.shape, [len(params), *self.sampleshape]):
                raise ValueError(f"Sample data was expected to have shape {[len(params), *self.sampleshape]}. Given: {samples.shape}")
        if len(params) > 0:
            self._lib.Estimator_ChiSquareAndCRLB(self.inst, params, samples, constants, roipos, len(params), crlb, chisq)
            self._check()
        if samples is not None:
            return crlb, chisq
        return crlb

    def ChiSquare(self, params, samples, roipos=None, constants=None):
        return self.CRLBAndChiSquare(params, samples, roipos, constants)[1]

    def CRLB(self, params, roipos=None, constants=None):
        """Pseudo-inverse route of estimator.py:249-253."""
        fisher = self.FisherMatrix(params, roipos, constants)
        var = np.linalg.pinv(fisher)
        return np.sqrt(np.abs(var[:, np.arange(self.numparams), np.arange(self.numparams)]))

    def NumDeriv(self, params, roipos=None, constants=None, eps=1e-6):
        """Central differences of ExpectedValue (estimator.py:267-284)."""
        params, constants, roipos = self._checkparams(params, constants, roipos)
        ev = self.ExpectedValue(params, roipos, constants)
        deriv = np.zeros((len(params), self.numparams, *self.sampleshape), dtype=np.float32)
        for k in range(self.numparams):
            params_min = params * 1
            params_min[:, k] -= eps
            params_max = params * 1
            params_max[:, k] += eps
            ev_min = self.ExpectedValue(params_min, roipos, constants)
            ev_max = self.ExpectedValue(params_max, roipos, constants)
            deriv[:, k] = (ev_max - ev_min) / (2 * eps)
        return deriv, ev

    def NumCRLB(self, params, roipos=None, constants=None, eps=1e-4, useNumDeriv=True):
        """estimator.py:286-300"""
        if useNumDeriv:
            deriv, ev = self.NumDeriv(params, roipos, constants, eps)
        else:
            deriv, ev = self.Derivatives(params, roipos, constants)
        K = self.numparams
        deriv = np.reshape(deriv, (len(params), K, self.samplecount))
        ev = np.reshape(ev, (len(params), self.samplecount))
        ev[ev < 1e-9] = 1e-9
        fi = np.zeros((len(params), K, K))
        for i in range(K):
            for j in range(K):
                fi[:, i, j] = np.sum(1 / ev * (deriv[:, i] * deriv[:, j]), axis=1)
        var = np.linalg.inv(fi)
        return np.sqrt(np.abs(var[:, np.arange(self.numparams), np.arange(self.numparams)]))

    # ---------------------------------------------------------------- fitting
    def GetLevMarParams(self):
        lmp = LevMarParams()
        self._lib.Estimator_GetLMParams(self.inst, lmp)
        self.lmparams = lmp
        return lmp

    def SetLevMarParams(self, lambda_=1e-15, iterations=50, lmFactor=0.0):
        self.lmparams = LevMarParams()
        self.lmparams.lambda_ = lambda_
        self.lmparams.iterations = iterations
        self.lmparams.factor = lmFactor
        self._lib.Estimator_SetLMParams(self.inst, self.lmparams)

    def Estimate(self, imgs, roipos=None, constants=None, initial=None):
        """(params [N,K], diag [N,numdiag], traces: list of [iterations_k, K]) — estimator.py:317-361."""
        numspots = len(imgs)
        traceBufLen = self.lmparams.iterations
        imgs = np.ascontiguousarray(imgs, dtype=np.float32)
        if not np.array_equal(imgs.shape[1:], self.sampleshape):
            raise ValueError(f"Sample images are expected to have shape {self.sampleshape}. Given: {imgs.shape[1:]}")
        iterations = np.zeros(numspots, np.int32)
        params = np.zeros((numspots, self.numparams), dtype=np.float32)
        trace = np.zeros((numspots, traceBufLen, self.numparams), dtype=np.float32)
        diag = np.zeros((numspots, self.numdiag), dtype=np.float32)
        params, constants, roipos = self._checkparams(params, constants, roipos)
        if initial is not None:
            initial = np.ascontiguousarray(initial, dtype=np.float32)
            if not np.array_equal(initial.shape, [numspots, self.numparams]):
                raise ValueError(f'Initial value array is expected to have shape {[numspots, self.numparams]}. Given: {initial.shape}')
        if numspots > 0:
            self._lib.Estimator_Estimate(self.inst, numspots, imgs, constants, roipos, initial, params, diag,
                                         iterations, trace, traceBufLen)
            self._check()
        self.last_iterations = iterations
        traces = [trace[k, 0:iterations[k], :] for k in range(numspots)]
        return params, diag, traces

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
            out["diag"] = torch.empty((n, self.numdiag), dtype=torch.float32, device=dev)
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

    def spline_form(self):
        """'grid' or 'coefficients' for a spline estimator (None otherwise): the form its
        fit and expected-value kernels evaluate the PSF in."""
        f = self._lib.Estimator_SplineForm(self.inst)
        return None if f < 0 else ("grid" if f == 1 else "coefficients")

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
        if type(paramName) == list:
            return [self.param_names.index(n) for n in paramName]
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
