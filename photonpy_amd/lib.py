"""Loader for the HIP library ``libphotonpy_amd.so`` (C ABI in include/photonpy_amd.h).

Mirrors the role of the reference's ``photonpy/cpp/lib.py`` (``SMLM`` class,
``lib.py:41-280``): one object owning the ``ctypes.CDLL`` with the argument
types of every exported function.  There is no fallback: if the library has
not been built the import fails with instructions.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np
import numpy.ctypeslib as ctl

_HERE = os.path.dirname(os.path.abspath(__file__))
# PHOTONPY_AMD_LIB selects another build of the library (e.g. to time two builds on one machine)
LIB_PATH = os.environ.get("PHOTONPY_AMD_LIB") or os.path.join(_HERE, "_lib", "libphotonpy_amd.so")


def _nullable(dtype):
    base = ctl.ndpointer(dtype, flags="aligned, c_contiguous")

    def from_param(cls, obj):
        if obj is None:
            return obj
        return base.from_param(obj)

    return type(f"Nullable{np.dtype(dtype).name}Array", (base,), {"from_param": classmethod(from_param)})


NullableFloatArrayType = _nullable(np.float32)
NullableIntArrayType = _nullable(np.int32)
FloatArray = ctl.ndpointer(np.float32, flags="aligned, c_contiguous")
IntArray = ctl.ndpointer(np.int32, flags="aligned, c_contiguous")


class EstimatorProperties(ct.Structure):
    _fields_ = [("numParams", ct.c_int), ("sampleCount", ct.c_int), ("numDiag", ct.c_int),
                ("numConst", ct.c_int), ("sampleIndexDims", ct.c_int)]


class LevMarParams(ct.Structure):
    _fields_ = [("iterations", ct.c_int32), ("lambda_", ct.c_float), ("factor", ct.c_float)]


# EstimatorResult (Estimators/EstimationQueue.h:21-25)
EstimResultDType = np.dtype([("id", "<i4"), ("chisq", "<f4"), ("iterations", "<i4")])


class NativeError(RuntimeError):
    """Raised when a call into libphotonpy_amd reports an error."""


_shared = None


class SMLM:
    """Owns the loaded library.  ``SMLM()`` returns a handle to the shared instance."""

    def __init__(self, debugMode=False):
        global _shared
        if _shared is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"photonpy_amd: HIP library not found at {LIB_PATH}; build it with "
                    "`python -c 'import __graft_entry__ as g; g.build()'` (or make -C photonpy_amd/csrc)")
            _shared = _bind(ct.CDLL(LIB_PATH))
        self.lib = _shared
        self.debugMode = debugMode

    def Close(self):
        # the library stays loaded for the life of the process
        self.lib = None

    def check(self):
        msg = self.lib.SMLM_GetLastError()
        if msg:
            self.lib.SMLM_ClearError()
            raise NativeError(msg.decode("utf-8", "replace"))

    def device_count(self):
        return self.lib.SMLM_DeviceCount()

    def GetCudaMemUse(self):
        """(pinned bytes, pitched bytes, device bytes, pitched allocations) held by the library."""
        vals = [ct.c_uint64() for _ in range(4)]
        self.lib.CudaGetMemoryUse(*[ct.byref(v) for v in vals])
        return tuple(v.value for v in vals)

    def GetCudaDevices(self):
        """Names of the visible HIP devices."""
        names = []
        for d in range(self.lib.CudaGetNumDevices()):
            cus, buf = ct.c_int32(), ct.create_string_buffer(256)
            if self.lib.CudaGetDeviceInfo(d, ct.byref(cus), buf, len(buf)):
                names.append(buf.value.decode("utf-8", "replace"))
        return names

    def SetCudaDevice(self, idx):
        return bool(self.lib.CudaSetDevice(int(idx)))


def _bind(lib):
    V, I, F, B = ct.c_void_p, ct.c_int32, ct.c_float, ct.c_bool
    sig = {
        "SMLM_GetLastError": ([], ct.c_char_p),
        "SMLM_ClearError": ([], None),
        "SMLM_DeviceCount": ([], ct.c_int),
        "Context_Create": ([], V),
        "Context_Destroy": ([V], None),
        "Gauss2D_CreatePSF_XYIBg": ([I, F, F, I, V], V),
        "Gauss2D_CreatePSF_XYIBgSigma": ([I, F, I, V], V),
        "Gauss2D_CreatePSF_XYZIBg": ([I, V, I, V], V),
        "Gauss2D_CreatePSF_XYZIBgQ": ([I, V, I, V], V),
        "CSpline_CreatePSF": ([I, V, NullableFloatArrayType, I, B, V], V),
        "CSpline_Compute": ([IntArray, FloatArray, FloatArray, I], None),
        "Gauss2D_CreatePSF_XYIBgConstSigma": ([I, B, V], V),
        "Gauss2D_CreatePSF_XYITiltedBg": ([I, B, V], V),
        "Gauss2D_CreatePSF_XYIBgSigmaXY": ([I, ct.c_float, ct.c_float, B, V], V),
        "Gauss2D_CreatePSF_XYITiltedBgSigmaXY": ([I, ct.c_float, ct.c_float, B, V], V),
        "Gauss2D_EstimateIntensityBg": ([FloatArray, FloatArray, FloatArray, I, FloatArray, IntArray, FloatArray, I, I,
                                         B], None),
        "CSpline_CreatePSF_FromFile": ([I, ct.c_char_p, I, B, V], V),
        "CSpline_Evaluate": ([IntArray, FloatArray, FloatArray, I, FloatArray], None),
        "Estimator_Delete": ([V], None),
        "Estimator_ParamFormat": ([V], ct.c_char_p),
        "Estimator_GetProperties": ([V, ct.POINTER(EstimatorProperties)], None),
        "Estimator_SampleDims": ([V, IntArray], None),
        "Estimator_GetParamLimits": ([V, FloatArray], None),
        "Estimator_SetLMParams": ([V, ct.POINTER(LevMarParams)], None),
        "Estimator_GetLMParams": ([V, ct.POINTER(LevMarParams)], None),
        "Estimator_SetGroupSize": ([V, I], ct.c_int),
        "Estimator_GetGroupSize": ([V], ct.c_int),
        "Estimator_SetRefill": ([V, I], ct.c_int),
        "Estimator_SetWorkSegments": ([V, I], ct.c_int),
        "Estimator_GetWorkSegments": ([V], ct.c_int),
        "Estimator_SetPhaseTiming": ([V, I], ct.c_int),
        "Estimator_SetSplineMode": ([V, I], ct.c_int),
        "Estimator_SplineForm": ([V], ct.c_int),
        "Estimator_SplineStripLength": ([V], ct.c_int),
        "Estimator_GetPhaseTimes": ([V, ct.POINTER(ct.c_float), ct.POINTER(ct.c_float)], ct.c_int),
        "Estimator_ComputeExpectedValue": ([V, I, FloatArray, FloatArray, IntArray, FloatArray], None),
        "Estimator_Estimate": ([V, I, FloatArray, FloatArray, IntArray, NullableFloatArrayType, FloatArray,
                                FloatArray, IntArray, NullableFloatArrayType, I], None),
        "Estimator_ComputeDerivatives": ([V, I, FloatArray, FloatArray, IntArray, FloatArray, FloatArray], None),
        "Estimator_ChiSquareAndCRLB": ([V, FloatArray, NullableFloatArrayType, FloatArray, IntArray, I,
                                        FloatArray, FloatArray], None),
        "Estimator_EstimateDevice": ([V, I, V, V, V, V, V, V, I, V, V, V], ct.c_int),
        "Estimator_ExpectedValueDevice": ([V, I, V, V, V, V], ct.c_int),
        "Estimator_DerivativesDevice": ([V, I, V, V, V, V, V], ct.c_int),
        "Estimator_ChiSquareAndCRLBDevice": ([V, I, V, V, V, V, V, V], ct.c_int),
        "Estimator_NLLGradDevice": ([V, I, V, V, V, V, V, V, V, V], ct.c_int),
        "Estimator_NumShared": ([V], ct.c_int),
        "Estimator_DiagnosticsDevice": ([V, I, V, V, V, V, V], ct.c_int),
        "GLRT_CreatePSF": ([V, V], V),
        "GLRT_SetMeanOfFirstSpot": ([V, I], ct.c_int),
        "EstimQueue_Create": ([V, I, I, ct.c_bool, I, V], V),
        "EstimQueue_Delete": ([V], None),
        "EstimQueue_Schedule": ([V, I, IntArray, FloatArray, NullableFloatArrayType, NullableFloatArrayType,
                                 IntArray], None),
        "EstimQueue_Flush": ([V], None),
        "EstimQueue_IsIdle": ([V], ct.c_bool),
        "EstimQueue_GetResultCount": ([V], ct.c_int),
        "EstimQueue_GetQueueLength": ([V], ct.c_int),
        "EstimQueue_GroupSize": ([V], ct.c_int),
        "EstimQueue_GetResults": ([V, I, FloatArray, NullableFloatArrayType, NullableFloatArrayType, IntArray,
                                   NullableFloatArrayType, ctl.ndpointer(EstimResultDType, flags="aligned, c_contiguous")],
                                  ct.c_int),
        "MixedBatch_Create": ([I, ct.POINTER(V), ct.POINTER(ct.c_int), I], V),
        "MixedBatch_Delete": ([V], None),
        "MixedBatch_NumStreams": ([V], ct.c_int),
        "MixedBatch_Run": ([V] + [ct.POINTER(V)] * 6 + [I, V], ct.c_int),
        "MixedBatch_RunPlain": ([V] + [ct.POINTER(V)] * 6 + [V], ct.c_int),
        "Estimator_GetShared": ([V, V], ct.c_int),
        "Estimator_EstimateDeviceNoAlloc": ([V, I, V, V, V, V, V, V, I, V, V, V], ct.c_int),
        "CudaGetNumDevices": ([], ct.c_int32),
        "CudaSetDevice": ([I], ct.c_bool),
        "CudaGetDeviceInfo": ([I, ct.POINTER(ct.c_int32), ct.c_char_p, I], ct.c_bool),
        "CudaGetMemoryUse": ([ct.POINTER(ct.c_uint64)] * 4, None),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib


def exported_symbols():
    """Names of every function declared in include/photonpy_amd.h (for the ABI tests)."""
    import re
    hdr = os.path.join(_HERE, "..", "include", "photonpy_amd.h")
    text = open(hdr).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", text, flags=re.M)) - {"if"})
