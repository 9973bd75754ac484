// Objects and helpers shared by the C-ABI sources (capi.hip, batch.hip, queue.hip):
// the last-error string, HIP error checks, Context and Estimator, and the fused
// fit launch every entry point goes through.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/photonpy_amd.h"
#include "dispatch.h"

// Context and Estimator are the opaque types of include/photonpy_amd.h, so they and
// these helpers live in the global namespace.
using namespace ppa;

// ---------------------------------------------------------------- errors
inline thread_local std::string g_lastError;

// Passes a message to the debug-print callback (runtime.hip).
void deliver_message(const std::string& msg);

// Records the calling thread's last error and delivers it through the debug-print
// callback, as the reference reports the exceptions its C-ABI calls catch.
inline void set_error(const std::string& msg) {
    g_lastError = msg;
    deliver_message(msg);
}

// Device and pinned host allocations of the library, tallied for CudaGetMemoryUse
// (runtime.hip).  The release functions accept NULL.
hipError_t dev_alloc(void** p, size_t bytes);
hipError_t dev_release(void* p);
hipError_t pinned_alloc(void** p, size_t bytes);
hipError_t pinned_release(void* p);
template <class T>
inline hipError_t dev_alloc(T** p, size_t bytes) {
    return dev_alloc(reinterpret_cast<void**>(p), bytes);
}
template <class T>
inline hipError_t pinned_alloc(T** p, size_t bytes) {
    return pinned_alloc(reinterpret_cast<void**>(p), bytes);
}

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------- objects
struct EstimationQueue;
void queue_destroy(EstimationQueue* q);  // queue.hip
// memcpy over the process's copy threads (queue.hip); host memory only
void parallel_copy(void* dst, const void* src, size_t bytes);

struct Context {
    std::mutex mtx;
    std::unordered_set<Estimator*> objects;
    std::unordered_set<EstimationQueue*> queues;  // deleted before the estimators they use
};

struct Estimator {
    Context* ctx = nullptr;
    std::mutex mtx;
    int roisize = 0;
    ModelDesc desc{};
    std::string fmt;
    std::vector<ParamLimit> limits;
    LMParams lm{30, 1e-15f, 0.0f};  // Estimator.cpp:24
    // A wrapper (GLRT_CreatePSF) computes with `inner`, whose LM parameters it uses
    // (GLRT_PSF::Estimate calls the wrapped estimator, GLRT_PSF.cu:91-94); it adds
    // numDiag diagnostics per spot.
    Estimator* inner = nullptr;
    int numDiag = 0;
    bool hostFlavour = false;  // created with cuda = false (see set_flavour)
    int glrtMeanOfSpot0 = 1;  // the reference's GLRT mean (first spot of the call)
    int group = 0;                  // 0: default_group(roisize)
    int refill = 0;                 // stopped groups per batched refill (0: kernel default)
    int segShift = 0;               // log2 of the fit's work segments (0: one counter, 3: eight)
    // spline coefficients: host copy (owned) and lazily uploaded device copy
    std::vector<float> hostCoefs;
    float* devCoefs = nullptr;
    // grid form of the spline (the samples the table interpolates), when the table
    // is a CSpline_Compute interpolant; used unless the caller asks for coefficients
    std::vector<float> hostGrid;
    float* devGrid = nullptr;
    bool gridForm = false;     // table verified to be a grid interpolant
    int splineMode = 0;        // 0: grid form when available, 1: always coefficients
    // Pipelined upload of a large host call's samples (Estimator_Estimate): two
    // pinned staging buffers, a copy stream and a fit stream, made on first use.
    struct Upload {
        float* staging[2] = {nullptr, nullptr};
        size_t floats = 0;  // per staging buffer
        hipStream_t copy = nullptr, fit = nullptr;
        hipEvent_t copied[2] = {nullptr, nullptr};  // staging buffer's H2D done
    } up;

    void ensure_upload(size_t floats) {
        if (!up.copy) {
            hip_check(hipStreamCreateWithFlags(&up.copy, hipStreamNonBlocking), "hipStreamCreate(copy)");
            hip_check(hipStreamCreateWithFlags(&up.fit, hipStreamNonBlocking), "hipStreamCreate(fit)");
            for (auto& ev : up.copied) hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
        }
        if (floats > up.floats) {
            for (auto& b : up.staging) {
                if (b) hip_check(pinned_release(b), "hipHostFree(staging)");
                b = nullptr;
            }
            up.floats = 0;
            for (auto& b : up.staging)
                hip_check(pinned_alloc(&b, floats * sizeof(float)), "hipHostMalloc(staging)");
            up.floats = floats;
        }
    }

    void release_upload() {
        if (up.copy) (void)hipStreamSynchronize(up.copy);
        if (up.fit) (void)hipStreamSynchronize(up.fit);
        for (auto& b : up.staging)
            if (b) (void)pinned_release(b);
        for (auto& ev : up.copied)
            if (ev) (void)hipEventDestroy(ev);
        if (up.copy) (void)hipStreamDestroy(up.copy);
        if (up.fit) (void)hipStreamDestroy(up.fit);
        up = Upload{};
    }

    // grow-only device workspace for the host-pointer API
    void* ws = nullptr;
    size_t wsBytes = 0;
    // grow-only device scratch of the spline adjoint's window sort
    void* adjWs = nullptr;
    size_t adjWsBytes = 0;
    // the window sort held in adjWs: its spot count (-1: none) and the stream it was made
    // on; a gradient call with upstream weights for as many spots on that stream reuses
    // it (the forward of an autograd step sorts, the backward of the same parameters
    // reuses)
    int adjSortN = -1;
    void* adjSortStream = nullptr;

    void* adjoint_workspace(size_t bytes) {
        if (bytes > adjWsBytes) {
            adjSortN = -1;
            if (adjWs) hip_check(dev_release(adjWs), "hipFree(adjoint scratch)");
            adjWs = nullptr;
            adjWsBytes = 0;
            hip_check(dev_alloc(&adjWs, bytes), "hipMalloc(adjoint scratch)");
            adjWsBytes = bytes;
        }
        return adjWs;
    }

    float* devShared = nullptr;  // device copy of the shared PSF parameters (Gaussian models)
    // Work counters of the fit kernel: one slot of kWorkSegs counters (256 bytes
    // apart) per launch, cycling, so launches of this estimator on different streams
    // do not share counters unless 16 are in flight.
    static constexpr int kCounterSlots = 16;
    int* devCounters = nullptr;
    unsigned counterSlot = 0;

    // Optional kernel timing: three events per fit launch (before the fit kernel,
    // after it, after the CRLB kernel), read and released by Estimator_GetPhaseTimes.
    bool phaseTiming = false;
    std::vector<hipEvent_t> phaseEvents;

    void release_phase_events() {
        for (hipEvent_t ev : phaseEvents) (void)hipEventDestroy(ev);
        phaseEvents.clear();
    }

    void ensure_counters() {
        if (!devCounters)
            hip_check(dev_alloc(&devCounters, kCounterSlots * kCounterInts * sizeof(int)), "hipMalloc(counters)");
    }

    int* next_counter() {
        ensure_counters();
        return devCounters + (size_t)(counterSlot++ % kCounterSlots) * kCounterInts;
    }

    int K() const { return num_params(desc.kind); }
    int NC() const { return num_consts(desc.kind, roisize); }

    // Shared PSF parameters in the adjoint's layout
    std::vector<float> shared_values() const {
        if (desc.kind == kKindFixed) return {desc.gc.sx, desc.gc.sy};
        if (desc.kind == kKindAstig) {
            std::vector<float> v(desc.gc.ax, desc.gc.ax + 5);
            v.insert(v.end(), desc.gc.ay, desc.gc.ay + 5);
            return v;
        }
        if (desc.kind == kKindSpline) return hostCoefs;
        return {};
    }
    size_t num_shared() const {
        if (desc.kind == kKindSpline) return hostCoefs.size();
        return shared_values().size();
    }
    const float* device_shared() {
        ensure_device();
        if (desc.kind == kKindSpline) return devCoefs;
        std::vector<float> v = shared_values();
        if (v.empty()) return nullptr;
        if (!devShared) {
            hip_check(dev_alloc(&devShared, v.size() * sizeof(float)), "hipMalloc(shared)");
            hip_check(hipMemcpy(devShared, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice), "upload shared");
        }
        return devShared;
    }
    int S() const { return roisize * roisize; }
    int G() const { return group > 0 ? group : default_group_for(desc, roisize); }

    ~Estimator() {
        release_upload();
        if (devCoefs) (void)dev_release(devCoefs);
        if (devGrid) (void)dev_release(devGrid);
        if (devShared) (void)dev_release(devShared);
        if (devCounters) (void)dev_release(devCounters);
        if (ws) (void)dev_release(ws);
        if (adjWs) (void)dev_release(adjWs);
        release_phase_events();
    }

    void ensure_device() {
        if (is_spline(desc.kind) && !devCoefs) {
            const size_t bytes = hostCoefs.size() * sizeof(float);
            hip_check(dev_alloc(&devCoefs, bytes), "hipMalloc(spline coefficients)");
            hip_check(hipMemcpy(devCoefs, hostCoefs.data(), bytes, hipMemcpyHostToDevice), "upload spline coefficients");
            desc.sm.coefs = devCoefs;
        }
        if (is_spline(desc.kind) && desc.useGrid && !devGrid) {
            // the table, then the zeros the fit kernels' window loads may reach past it
            const size_t bytes = hostGrid.size() * sizeof(float);
            const size_t pad = SplineGridModel::pad_floats(desc.sg.Nx) * sizeof(float);
            hip_check(dev_alloc(&devGrid, bytes + pad), "hipMalloc(spline grid)");
            hip_check(hipMemcpy(devGrid, hostGrid.data(), bytes, hipMemcpyHostToDevice), "upload spline grid");
            hip_check(hipMemset(reinterpret_cast<char*>(devGrid) + bytes, 0, pad), "clear spline grid padding");
            desc.sg.grid = devGrid;
        }
    }

    // Whether the device tables a fit reads are resident (nothing to upload).
    bool device_ready() const {
        if (!is_spline(desc.kind)) return true;
        return devCoefs && (!desc.useGrid || devGrid);
    }

    void* workspace(size_t bytes) {
        if (bytes > wsBytes) {
            if (ws) hip_check(dev_release(ws), "hipFree(workspace)");
            ws = nullptr;
            wsBytes = 0;
            hip_check(dev_alloc(&ws, bytes), "hipMalloc(workspace)");
            wsBytes = bytes;
        }
        return ws;
    }
};

// The estimator that does the computing: the innermost one a wrapper refers to.
inline Estimator* model_of(Estimator* e) {
    while (e && e->inner) e = e->inner;
    return e;
}

inline void attach(Estimator* e, Context* ctx) {
    if (!ctx) return;
    std::lock_guard<std::mutex> l(ctx->mtx);
    ctx->objects.insert(e);
    e->ctx = ctx;
}

inline void detach(Estimator* e) {
    if (!e->ctx) return;
    std::lock_guard<std::mutex> l(e->ctx->mtx);
    e->ctx->objects.erase(e);
    e->ctx = nullptr;
}

template <class F>
inline auto guarded(F fn, decltype(fn()) fallback) -> decltype(fn()) {
    try {
        return fn();
    } catch (const std::exception& ex) {
        set_error(ex.what());
        return fallback;
    }
}

// cuda = false asks the reference for its CPU estimator (PSFImpl, EstimatorImpl.h:108-185).
// Such an estimator computes on the GPU like the others; it keeps the CPU estimator's
// interface: NumDiag = 1 (EstimatorImpl.h:120), a diagnostics column its Estimate
// never writes (EstimatorImpl.h:166-182).
inline void set_flavour(Estimator* e, bool cuda) {
    e->hostFlavour = !cuda;
    e->numDiag = cuda ? 0 : 1;
}

// ---------------------------------------------------------------- launch helpers
inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Fused fit (+ CRLB / chi^2 when d_crlb / d_chisq are given) of n spots on st with the
// estimator's model, LM settings, lane-group size, refill batch and next work counter;
// timing = false keeps phase-timing events out (graph capture).
// Lanes per spot of a fit: the largest compiled group size from groupHint down to
// the estimator's own (its own when groupHint <= its own).
int fit_group(Estimator* e, int groupHint);
// groupHint > 0: try that many lanes per spot first (halving down to the estimator's
// own group size while no kernel is compiled for it) -- small batches fill more of
// the GPU with more lanes per spot.
// counter != null: use those kWorkSegs work counters and allocate nothing (the
// estimator's device tables must already be resident; no timing events are recorded).
void run_fit(Estimator* e, int n, const float* d_sample, const float* d_constants, const float* d_initial,
             float* d_params, int* d_iter, float* d_trace, int traceLen, float* d_crlb, float* d_chisq, hipStream_t st,
             bool timing = true, int groupHint = 0, int* counter = nullptr);
