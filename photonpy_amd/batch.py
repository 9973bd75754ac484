"""Mixed-model, mixed-size batch (no reference counterpart).

A ``MixedBatch`` fits several bins of spots, one bin per estimator (model and
ROI size).  The bins run concurrently on a few HIP streams (longest first), and
with ``use_graph`` the whole fork / launches / join sequence is one HIP graph,
replayed with a single launch.  ``run_plain`` fits the same bins one launch
after another on one stream, which is the comparison form.  Each estimator's
lane-group size and refill batch come from ``photonpy_amd.autotune`` when a
tuning table entry exists for its (model, ROI size).
"""
from __future__ import annotations

import ctypes as ct


class MixedBatch:
    def __init__(self, estimators, counts, streams=4, tune=True):
        if len(estimators) != len(counts) or not estimators:
            raise ValueError("one spot count per estimator is required")
        self.estimators = list(estimators)
        self.counts = [int(c) for c in counts]
        self._lib = estimators[0]._lib
        self._smlm = estimators[0]._smlm
        if tune:
            from .autotune import apply
            for e in self.estimators:
                apply(e)
        nb = len(estimators)
        arr = (ct.c_void_p * nb)(*[e.inst for e in estimators])
        cnt = (ct.c_int * nb)(*self.counts)
        self.inst = self._lib.MixedBatch_Create(nb, arr, cnt, int(streams))
        if not self.inst:
            self._smlm.check()
            raise RuntimeError("MixedBatch_Create failed")

    def num_streams(self):
        return self._lib.MixedBatch_NumStreams(self.inst)

    def allocate_outputs(self, device="cuda", crlb=True):
        """Per-bin output tensors (params, iterations, crlb, chisq)."""
        import torch
        outs = []
        for e, n in zip(self.estimators, self.counts):
            o = {"params": torch.empty((n, e.numparams), dtype=torch.float32, device=device),
                 "iterations": torch.empty(n, dtype=torch.int32, device=device)}
            if crlb:
                o["crlb"] = torch.empty((n, e.numparams), dtype=torch.float32, device=device)
                o["chisq"] = torch.empty(n, dtype=torch.float32, device=device)
            outs.append(o)
        return outs

    def _ptrs(self, samples, outs, initial=None):
        nb = len(self.estimators)
        for e, n, s in zip(self.estimators, self.counts, samples):
            if tuple(s.shape) != (n, *e.sampleshape) or not s.is_contiguous():
                raise ValueError(f"bin samples must be contiguous [{n}, {e.sampleshape[0]}, {e.sampleshape[1]}]")
        if initial is not None:
            for e, n, t in zip(self.estimators, self.counts, initial):
                if tuple(t.shape) != (n, e.numparams) or not t.is_contiguous() or t.dtype != s.dtype:
                    raise ValueError(f"bin initial values must be contiguous float32 [{n}, {e.numparams}]")

        def arr(key, seq):
            vals = [None if (o is None or o.get(key) is None) else o[key].data_ptr() for o in seq]
            if all(v is None for v in vals):
                return None
            return (ct.c_void_p * nb)(*vals)

        sp = (ct.c_void_p * nb)(*[s.data_ptr() for s in samples])
        ini = None if initial is None else (ct.c_void_p * nb)(*[t.data_ptr() for t in initial])
        return sp, ini, arr("params", outs), arr("iterations", outs), arr("crlb", outs), arr("chisq", outs)

    def run(self, samples, outs, use_graph=True, stream=None, initial=None):
        """Fit every bin: samples[b] [n_b,R_b,R_b] and, optionally, initial[b] [n_b,K_b]
        (device tensors); results go to the tensors of outs (allocate_outputs)."""
        import torch
        stream = stream or torch.cuda.current_stream()
        sp, ini, pp, ip, cp, xp = self._ptrs(samples, outs, initial)
        rc = self._lib.MixedBatch_Run(self.inst, sp, ini, pp, ip, cp, xp, int(use_graph), stream.cuda_stream)
        if rc != 0:
            self._smlm.check()
            raise RuntimeError(f"MixedBatch_Run failed ({rc})")

    def run_plain(self, samples, outs, stream=None, initial=None):
        import torch
        stream = stream or torch.cuda.current_stream()
        sp, ini, pp, ip, cp, xp = self._ptrs(samples, outs, initial)
        rc = self._lib.MixedBatch_RunPlain(self.inst, sp, ini, pp, ip, cp, xp, stream.cuda_stream)
        if rc != 0:
            self._smlm.check()
            raise RuntimeError(f"MixedBatch_RunPlain failed ({rc})")

    def destroy(self):
        if self.inst:
            self._lib.MixedBatch_Delete(self.inst)
            self.inst = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.destroy()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
