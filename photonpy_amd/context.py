"""Native object lifetime (the role of photonpy/cpp/context.py:6-48).

A ``Context`` is a handle to a native ``Context`` object (Context_Create,
Context.cpp:65-73).  Estimators and queues created with it are registered on
the native side and freed by ``Destroy`` (Context_Destroy, Context.cpp:41-55),
so a script can drop a whole set of objects at once.  The library itself is
loaded once per process (``SMLM``) and stays loaded; a context only borrows it.
"""
from __future__ import annotations

from .lib import SMLM


class Context:
    """``Context(smlm=None, debugMode=False)``; usable as a context manager.

    Attributes ``smlm`` (the library handle), ``lib`` (its ctypes CDLL) and
    ``inst`` (the native pointer, None once destroyed)."""

    def __init__(self, smlm: SMLM = None, debugMode=False):
        # a context made without a library handle opens its own and lets go of it
        # when destroyed; one given a handle leaves it to the caller
        self._borrowed = smlm is not None
        self.smlm = smlm if self._borrowed else SMLM(debugMode)
        self.lib = self.smlm.lib
        self.inst = self.lib.Context_Create()

    def Create(self):
        """A further context on the same library handle."""
        return Context(self.smlm)

    @property
    def alive(self):
        return self.inst is not None

    def Destroy(self):
        handle, self.inst = self.inst, None
        if handle is None:
            return
        self.lib.Context_Destroy(handle)
        if not self._borrowed:
            own, self.smlm, self.lib = self.smlm, None, None
            own.Close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.Destroy()
