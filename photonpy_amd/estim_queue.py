"""Asynchronous estimation queue over host memory (photonpy/cpp/estim_queue.py:76-246).

``EstimQueue(estimator)`` schedules ROIs (with ids and ROI positions) into pinned
batches that a native worker thread fits on a few HIP streams; ``GetResults``
returns the finished ones as an ``EstimQueue_Results``.  Same names, arguments,
defaults and result fields as the reference module; the native side is
``EstimQueue_*`` in include/photonpy_amd.h.
"""
from __future__ import annotations

import copy
import time

import numpy as np

from .estimator import Estimator
from .lib import EstimResultDType


class EstimQueue_Results:
    """Results of a queue, one row per spot (estim_queue.py:14-70)."""

    def __init__(self, param_names, sampleshape, estim, diag, crlb, iterations, chisq, roipos, samples, ids):
        self.estim = estim
        self.sampleshape = sampleshape
        self.diagnostics = diag
        self.crlb = crlb
        self.roipos = roipos
        self.iterations = iterations
        self.chisq = chisq
        self.ids = ids
        self.samples = samples
        self.param_names = param_names

    def CRLB(self):
        return self.crlb

    def SortByID(self, isUnique=False):
        """Reorders the rows by id; with isUnique the ids must be a permutation of 0..n-1."""
        if isUnique:
            order = np.empty(len(self.ids), dtype=np.int64)
            order[self.ids] = np.arange(len(self.ids))
        else:
            order = np.argsort(self.ids, kind="stable")
        self.Filter(order)
        return order

    def Filter(self, indices):
        indices = np.asarray(indices)
        if indices.dtype == bool:
            indices = np.nonzero(indices)[0]
        for name in ("estim", "diagnostics", "crlb", "roipos", "iterations", "chisq", "ids"):
            setattr(self, name, getattr(self, name)[indices])
        if self.samples is not None:
            self.samples = self.samples[indices]
        return indices

    def FilterXY(self, minX, minY, maxX, maxY):
        x, y = self.estim[:, 0], self.estim[:, 1]
        return self.Filter(np.nonzero((x > minX) & (y > minY) & (x < maxX) & (y < maxY))[0])

    def Clone(self):
        return copy.deepcopy(self)

    def ColIdx(self, *names):
        return np.squeeze(np.array([self.param_names.index(n) for n in names], dtype=np.int64))


class EstimQueue:
    """estim_queue.py:76-246.  maxQueueLenInBatches is passed to the native queue as its
    queue-length limit, as the reference does."""

    def __init__(self, estim: Estimator, batchSize=256, maxQueueLenInBatches=5, numStreams=-1, keepSamples=False,
                 ctx=None):
        self.ctx = estim.ctx if ctx is None else ctx
        self.estim = estim
        self.batchSize = batchSize
        self.keepSamples = keepSamples
        self._lib = estim._lib
        self.param_names = estim.param_names
        cinst = self.ctx.inst if self.ctx is not None else None
        self.inst = self._lib.EstimQueue_Create(estim.inst, int(batchSize), int(maxQueueLenInBatches),
                                                bool(keepSamples), int(numStreams), cinst)
        if not self.inst:
            estim._smlm.check()
            raise RuntimeError("Unable to create PSF MLE Queue with given PSF")

    def __enter__(self):
        return self

    def __exit__(self, *args):
        self.Destroy()

    def Destroy(self):
        if self.inst:
            self._lib.EstimQueue_Delete(self.inst)
            self.inst = None

    def Flush(self):
        self._lib.EstimQueue_Flush(self.inst)

    def WaitUntilDone(self):
        self.Flush()
        while not self.IsIdle():
            time.sleep(0.005)

    def IsIdle(self):
        return bool(self._lib.EstimQueue_IsIdle(self.inst))

    def Schedule(self, samples, roipos=None, ids=None, initial=None, constants=None):
        samples = np.ascontiguousarray(samples, dtype=np.float32)
        numspots = len(samples)
        e = self.estim
        if roipos is None:
            roipos = np.zeros((numspots, e.indexdims), dtype=np.int32)
        if constants is not None:
            constants = np.ascontiguousarray(constants, dtype=np.float32)
            if constants.size != numspots * e.numconst:
                raise ValueError(f"Estimator is expecting constants array with shape {(numspots, e.numconst)}. "
                                 f"Given: {constants.shape}")
        elif e.numconst != 0:
            raise ValueError("this estimator needs per-spot constants")
        if initial is not None:
            initial = np.ascontiguousarray(initial, dtype=np.float32)
            if tuple(initial.shape) != (numspots, e.numparams):
                raise ValueError(f"initial must have shape {(numspots, e.numparams)}, got {initial.shape}")
        roipos = np.ascontiguousarray(roipos, dtype=np.int32)
        if tuple(roipos.shape) != (numspots, e.indexdims):
            raise ValueError(f"Incorrect shape for ROI positions: {roipos.shape} given, expecting "
                             f"{[numspots, e.indexdims]}")
        if e.samplecount * numspots != samples.size:
            raise ValueError(f"samples must hold {numspots} ROIs of {e.sampleshape}")
        if ids is None:
            ids = np.zeros(numspots, dtype=np.int32)
        else:
            if len(ids) != numspots:
                raise ValueError("one id per spot is required")
            ids = np.ascontiguousarray(ids, dtype=np.int32)
        self._lib.EstimQueue_Schedule(self.inst, numspots, ids, samples, initial, constants, roipos)
        e._check()

    def GetQueueLength(self):
        return self._lib.EstimQueue_GetQueueLength(self.inst)

    def GroupSize(self):
        """Lanes per spot of the queue's batch fits (no reference counterpart)."""
        return self._lib.EstimQueue_GroupSize(self.inst)

    def GetResultCount(self):
        return self._lib.EstimQueue_GetResultCount(self.inst)

    def GetResults(self, maxResults=None, getSampleData=False) -> EstimQueue_Results:
        count = self.GetResultCount()
        if maxResults is not None and count > maxResults:
            count = maxResults
        e = self.estim
        K = e.NumParams()
        estim = np.zeros((count, K), dtype=np.float32)
        diag = np.zeros((count, e.NumDiag()), dtype=np.float32)
        crlb = np.zeros((count, K), dtype=np.float32)
        roipos = np.zeros((count, e.indexdims), dtype=np.int32)
        results = np.zeros(count, dtype=EstimResultDType)
        samples = np.zeros((count, *e.sampleshape), dtype=np.float32) if getSampleData else None
        copied = self._lib.EstimQueue_GetResults(self.inst, count, estim, diag if diag.size else None, crlb, roipos,
                                                 samples, results)
        e._check()
        if copied < count:  # whole batches only: keep what was copied
            estim, diag, crlb, roipos, results = estim[:copied], diag[:copied], crlb[:copied], roipos[:copied], \
                results[:copied]
            if samples is not None:
                samples = samples[:copied]
        return EstimQueue_Results(self.param_names, e.sampleshape, estim, diag, crlb, results["iterations"],
                                  results["chisq"], roipos, samples, results["id"])
