"""The reference's Python loader binds these symbols when it opens the library.

photonpy/cpp/lib.py:71-153 (SMLM.__init__) and gaussian.py:76-84
(GaussianPSFMethods.__init__) look up CudaGetNumDevices, CudaSetDevice,
CudaGetDeviceInfo, CudaGetMemoryUse, SetDebugPrintCallback,
SetDebugImageCallback, ShowDebugImage, FFT, FFT2, DrawROIs and Gauss2D_Draw
with the argument types below; the lookups must succeed on libphotonpy_amd.so.
Errors caught by C-ABI calls reach the debug-print callback, as the
reference's DebugPrintf delivers them (StringUtils.cpp:26-49).  No GPU is used.
"""
import ctypes as ct

import numpy as np
import numpy.ctypeslib as ctl
import pytest

from photonpy_amd import lib as L

PRINT_CB = ct.CFUNCTYPE(ct.c_int32, ct.c_char_p)
IMAGE_CB = ct.CFUNCTYPE(None, ct.c_int32, ct.c_int32, ct.c_int32, ct.POINTER(ct.c_float), ct.c_char_p)
F32 = ctl.ndpointer(np.float32, flags="aligned, c_contiguous")
F64 = ctl.ndpointer(np.float64, flags="aligned, c_contiguous")
C64 = ctl.ndpointer(np.complex64, flags="aligned, c_contiguous")
I32 = ctl.ndpointer(np.int32, flags="aligned, c_contiguous")

# (argtypes, restype) as the reference's loader declares them
LOADER = {
    "CudaGetNumDevices": ([], ct.c_int32),
    "CudaSetDevice": ([ct.c_int32], ct.c_bool),
    "CudaGetDeviceInfo": ([ct.c_int32, ct.POINTER(ct.c_int32), ct.c_char_p, ct.c_int32], ct.c_bool),
    "CudaGetMemoryUse": ([ct.POINTER(ct.c_uint64)] * 4, None),
    "SetDebugPrintCallback": ([PRINT_CB], None),
    "SetDebugImageCallback": ([IMAGE_CB], None),
    "ShowDebugImage": ([ct.c_int32, ct.c_int32, ct.c_int32, F32, ct.c_char_p], None),
    "FFT": ([C64, C64, ct.c_int32, ct.c_int32, ct.c_int32], None),
    "FFT2": ([C64, C64, ct.c_int32, ct.c_int32, ct.c_int32, ct.c_int32], None),
    "DrawROIs": ([F32, ct.c_int32, ct.c_int32, F32, ct.c_int32, ct.c_int32, I32], None),
    "Gauss2D_Draw": ([F64, ct.c_int32, ct.c_int32, F32, ct.c_int32, ct.c_float], None),
    "Gauss2D_CreatePSF_XYIBg": ([ct.c_int32, ct.c_float, ct.c_float, ct.c_int32, ct.c_void_p], ct.c_void_p),
}


@pytest.fixture
def so():
    lib = ct.CDLL(L.LIB_PATH)  # a handle of its own, bound as the reference binds it
    for name, (args, res) in LOADER.items():
        f = getattr(lib, name)
        f.argtypes, f.restype = args, res
    yield lib
    lib.SetDebugPrintCallback(PRINT_CB())  # NULL: back to the default
    lib.SetDebugImageCallback(IMAGE_CB())


@pytest.fixture
def messages(so):
    got = []

    def cb(msg):
        got.append(msg.decode())
        return 0  # handled: no echo to stderr

    keep = PRINT_CB(cb)  # the callback object must outlive its registration
    so.SetDebugPrintCallback(keep)
    yield got
    so.SetDebugPrintCallback(PRINT_CB())
    del keep


def test_every_loader_symbol_binds(so):
    for name in LOADER:
        assert getattr(so, name) is not None


def test_device_queries_without_a_gpu(so):
    n = so.CudaGetNumDevices()
    assert n >= 0
    cus, name = ct.c_int32(-1), ct.create_string_buffer(100)
    # an index past the device count is reported as a failure, not a crash
    assert so.CudaGetDeviceInfo(n, ct.byref(cus), name, 100) is False
    vals = [ct.c_uint64(123) for _ in range(4)]
    so.CudaGetMemoryUse(*[ct.byref(v) for v in vals])
    assert vals[1].value == 0 and vals[3].value == 0  # no pitched allocations exist


def test_errors_reach_the_print_callback(so, messages):
    # roisize 0 is refused by the creator (NULL) before any device work
    assert not so.Gauss2D_CreatePSF_XYIBg(0, 1.5, 1.5, 1, None)
    assert any("roisize" in m for m in messages), messages
    assert all(m.endswith("\n") for m in messages)


def test_out_of_scope_functions_report_themselves(so, messages):
    src = np.zeros((2, 8), np.complex64)
    dst = np.zeros_like(src)
    so.FFT(src, dst, 2, 8, 1)
    so.FFT2(np.zeros((1, 4, 4), np.complex64), np.zeros((1, 4, 4), np.complex64), 1, 4, 4, 1)
    img = np.zeros((8, 8), np.float32)
    so.DrawROIs(img, 8, 8, np.ones((1, 3, 3), np.float32), 1, 3, np.zeros((1, 2), np.int32))
    so.Gauss2D_Draw(np.zeros((8, 8)), 8, 8, np.ones((1, 5), np.float32), 1, 0.0)
    for name in ("FFT", "FFT2", "DrawROIs", "Gauss2D_Draw"):
        assert any(name + " is not provided" in m for m in messages), (name, messages)
    assert not img.any() and not dst.any()  # nothing was written


def test_debug_image_callback_roundtrip(so):
    seen = []

    def cb(w, h, n, data, title):
        seen.append((w, h, n, np.ctypeslib.as_array(data, shape=(n, h, w)).copy(), title.decode()))

    keep = IMAGE_CB(cb)
    so.SetDebugImageCallback(keep)
    img = np.arange(2 * 3 * 4, dtype=np.float32).reshape(2, 3, 4)
    so.ShowDebugImage(4, 3, 2, img, b"probe")
    assert len(seen) == 1
    w, h, n, data, title = seen[0]
    assert (w, h, n, title) == (4, 3, 2, "probe") and np.array_equal(data, img)


def test_memory_use_tracks_library_allocations(so):
    # the Python loader's helper reads the same tallies
    smlm = L.SMLM()
    pinned, pitched, dev, nalloc = smlm.GetCudaMemUse()
    assert pitched == 0 and nalloc == 0 and pinned >= 0 and dev >= 0
