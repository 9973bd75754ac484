// Process-level entry points that the reference's Python loader binds when it opens
// the library (photonpy/cpp/lib.py:71-153, gaussian.py:76-84): device count,
// selection and description, memory-use tallies, and the debug-print and
// debug-image callbacks.  Every error that a C-ABI call catches is delivered through
// the debug-print callback (the reference reports its caught exceptions with
// DebugPrintf, StringUtils.cpp:26-49) and is also kept for SMLM_GetLastError.
//
// The loader also binds a few functions that belong to subsystems this library does
// not provide (FFT, rendering).  They exist so that the loader binds unchanged; each
// reports "not provided" through the same channel and does nothing else.
#include "capi_internal.h"

#include <atomic>
#include <unordered_map>

namespace {

std::mutex g_printMutex;
int (*g_printCallback)(const char*) = nullptr;

std::mutex g_imageMutex;
void (*g_imageCallback)(int, int, int, const float*, const char*) = nullptr;

// Bytes held by the library's own allocations, by kind (the reference's
// CudaGetMemoryUse tallies, CudaUtils.cu:38-46).  Pitched allocations do not exist
// here; those tallies stay zero.
std::mutex g_allocMutex;
std::unordered_map<void*, std::pair<int, size_t>> g_allocs;  // pointer -> (kind, bytes)
uint64_t g_bytes[2] = {0, 0};                                  // 0 device, 1 pinned host

void remember(void* p, int kind, size_t bytes) {
    std::lock_guard<std::mutex> l(g_allocMutex);
    g_allocs[p] = {kind, bytes};
    g_bytes[kind] += bytes;
}

void forget(void* p) {
    std::lock_guard<std::mutex> l(g_allocMutex);
    auto it = g_allocs.find(p);
    if (it == g_allocs.end()) return;
    g_bytes[it->second.first] -= it->second.second;
    g_allocs.erase(it);
}

void not_provided(const char* name) {
    set_error(std::string("photonpy_amd: ") + name +
              " is not provided by this library (only the spot-fitting path is built)");
}

}  // namespace

void deliver_message(const std::string& msg) {
    std::lock_guard<std::mutex> l(g_printMutex);
    std::string line = msg;
    if (line.empty() || line.back() != '\n') line += '\n';
    // a callback that returns 0 has handled the message; otherwise (or without one)
    // it also goes to stderr
    const bool echo = g_printCallback ? g_printCallback(line.c_str()) != 0 : true;
    if (echo) std::fputs(line.c_str(), stderr);
}

hipError_t dev_alloc(void** p, size_t bytes) {
    *p = nullptr;
    const hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) remember(*p, 0, bytes);
    return e;
}

hipError_t dev_release(void* p) {
    if (!p) return hipSuccess;
    forget(p);
    return hipFree(p);
}

hipError_t pinned_alloc(void** p, size_t bytes) {
    *p = nullptr;
    const hipError_t e = hipHostMalloc(p, bytes, hipHostMallocDefault);
    if (e == hipSuccess) remember(*p, 1, bytes);
    return e;
}

hipError_t pinned_release(void* p) {
    if (!p) return hipSuccess;
    forget(p);
    return hipHostFree(p);
}

extern "C" {

int CudaGetNumDevices(void) {
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess ? n : 0;
}

bool CudaSetDevice(int index) { return hipSetDevice(index) == hipSuccess; }

bool CudaGetDeviceInfo(int index, int* numMultiprocessors, char* name, int namelen) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, index) != hipSuccess) return false;
    if (numMultiprocessors) *numMultiprocessors = prop.multiProcessorCount;
    if (name && namelen > 0) {
        std::strncpy(name, prop.name, (size_t)namelen - 1);
        name[namelen - 1] = '\0';
    }
    return true;
}

void CudaGetMemoryUse(uint64_t* pinnedBytes, uint64_t* devicePitchedBytes, uint64_t* deviceBytes,
                      uint64_t* pitchedNumAllocs) {
    std::lock_guard<std::mutex> l(g_allocMutex);
    if (pinnedBytes) *pinnedBytes = g_bytes[1];
    if (devicePitchedBytes) *devicePitchedBytes = 0;
    if (deviceBytes) *deviceBytes = g_bytes[0];
    if (pitchedNumAllocs) *pitchedNumAllocs = 0;
}

void SetDebugPrintCallback(int (*cb)(const char* msg)) {
    std::lock_guard<std::mutex> l(g_printMutex);
    g_printCallback = cb;
}

void SetDebugImageCallback(void (*cb)(int width, int height, int numImages, const float* data, const char* title)) {
    std::lock_guard<std::mutex> l(g_imageMutex);
    g_imageCallback = cb;
}

void ShowDebugImage(int w, int h, int numImages, const float* data, const char* title) {
    std::lock_guard<std::mutex> l(g_imageMutex);
    if (g_imageCallback) g_imageCallback(w, h, numImages, data, title);
}

void FFT(const void*, void*, int, int, int) { not_provided("FFT"); }
void FFT2(const void*, void*, int, int, int, int) { not_provided("FFT2"); }
void DrawROIs(float*, int, int, const float*, int, int, const int*) { not_provided("DrawROIs"); }
void Gauss2D_Draw(double*, int, int, const float*, int, float) { not_provided("Gauss2D_Draw"); }

}  // extern "C"
