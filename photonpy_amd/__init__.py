"""photonpy_amd — MI355X-native batched Poisson-MLE spot fitting.

Drop-in for photonpy's Estimator API (photonpy/cpp/estimator.py and the
Gaussian / cubic-spline factories), backed by hand-written HIP kernels for
gfx950 in libphotonpy_amd.so.
"""
from .lib import SMLM, NativeError, LIB_PATH
from .context import Context
from .estimator import Estimator
from .gaussian import Gauss3D_Calibration, GaussianPSFMethods, Gaussian
from .cspline import CSplineCalibration, CSplineMethods, CSplinePSF
from .estim_queue import EstimQueue, EstimQueue_Results

__all__ = ["SMLM", "NativeError", "LIB_PATH", "Context", "Estimator", "Gauss3D_Calibration",
           "GaussianPSFMethods", "Gaussian", "CSplineCalibration", "CSplineMethods", "CSplinePSF", "EstimQueue",
           "EstimQueue_Results"]
