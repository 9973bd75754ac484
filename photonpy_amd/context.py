"""Native object lifetime, mirroring photonpy/cpp/context.py:6-48.

Estimators created with a context are destroyed together with it
(Context.cpp:41-55)."""
from __future__ import annotations

from .lib import SMLM


class Context:
    def __init__(self, smlm: SMLM = None, debugMode=False):
        self.libOwner = smlm is None
        if self.libOwner:
            smlm = SMLM(debugMode)
        self.smlm = smlm
        self.lib = smlm.lib
        self.inst = self.lib.Context_Create()

    def Create(self):
        return Context(self.smlm)

    def Destroy(self):
        if self.inst:
            self.lib.Context_Destroy(self.inst)
            self.inst = None
            if self.libOwner:
                self.smlm.Close()
                self.smlm = None
                self.lib = None

    def __enter__(self):
        return self

    def __exit__(self, *args):
        self.Destroy()
