// Host-side interface between the C-ABI layer and the kernel instantiations.
#pragma once

#include <hip/hip_runtime.h>

#include "fit_kernel.h"
#include "eval_kernels.h"

namespace ppa {

enum ModelKind { kKindFixed = 0, kKindSigma = 1, kKindAstig = 2, kKindSpline = 3 };

struct ModelDesc {
    int kind;
    GaussCalib gc;
    SplineModel sm;  // coefs must point to device memory when launched
};

inline int num_params(int kind) { return kind == kKindFixed ? 4 : 5; }

// Lane-group size used for an R x R ROI unless the tuning table says otherwise:
// the smallest power of two that keeps <= ~64 samples per lane in registers.
inline int default_group(int R) {
    if (R <= 4) return 1;
    if (R <= 11) return 2;
    if (R <= 16) return 4;
    if (R <= 24) return 8;
    return 16;
}

constexpr int kMinFitR = 3, kMaxFitR = 32;

// Per-model launchers, defined in kernels_<model>.hip.  Return hipErrorInvalidValue
// when no kernel is compiled for (R, G).
hipError_t launch_fit_fixed(const ModelDesc& d, int R, int G, const FitArgs& a, int maxBlocks, hipStream_t st);
hipError_t launch_fit_sigma(const ModelDesc& d, int R, int G, const FitArgs& a, int maxBlocks, hipStream_t st);
hipError_t launch_fit_astig(const ModelDesc& d, int R, int G, const FitArgs& a, int maxBlocks, hipStream_t st);
hipError_t launch_fit_spline(const ModelDesc& d, int R, int G, const FitArgs& a, int maxBlocks, hipStream_t st);

hipError_t launch_ev_any(const ModelDesc& d, int R, int n, const float* params, float* ev, float* deriv, hipStream_t st);
hipError_t launch_crlb_any(const ModelDesc& d, int R, int n, const float* params, const float* sample, float* crlb,
                           float* chisq, hipStream_t st);

// Poisson NLL and its adjoint (adjoint.hip)
struct AdjArgs {
    int n, R;
    const float* params;   // [n,K]
    const float* sample;   // [n,R,R]
    const float* gout;     // [n] upstream gradient or null (= 1)
    const float* shared;   // device copy of the shared PSF parameters
    float* nll;            // [n] or null
    float* gparams;        // [n,K] or null
    float* gshared;        // accumulated (+=) or null
};
hipError_t launch_adjoint_any(const ModelDesc& d, const AdjArgs& a, hipStream_t st);
hipError_t launch_spline_coefs(const int* shape, const float* d_data, float* d_out, int border, hipStream_t st);
// Sets *counter = value on the stream (a one-thread kernel, so it can also be captured in a graph).
hipError_t launch_set_counter(int* counter, int value, hipStream_t st);

inline hipError_t launch_fit_any(const ModelDesc& d, int R, int G, const FitArgs& a, int maxBlocks, hipStream_t st) {
    switch (d.kind) {
    case kKindFixed: return launch_fit_fixed(d, R, G, a, maxBlocks, st);
    case kKindSigma: return launch_fit_sigma(d, R, G, a, maxBlocks, st);
    case kKindAstig: return launch_fit_astig(d, R, G, a, maxBlocks, st);
    case kKindSpline: return launch_fit_spline(d, R, G, a, maxBlocks, st);
    default: return hipErrorInvalidValue;
    }
}

// Persistent grid for a fit kernel: enough blocks to cover every group once,
// capped at the number of blocks the device keeps resident.
template <class M, int R, int G>
hipError_t launch_fit_t(const M& m, const FitArgs& a, int maxBlocks, hipStream_t st) {
    constexpr int kBlock = 256;
    static int resident = 0;  // blocks resident on the whole device
    if (resident == 0) {
        int dev = 0, perCU = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) != hipSuccess) return hipErrorNoDevice;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return hipErrorNoDevice;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fit_kernel<M, R, G>, kBlock, 0) != hipSuccess || perCU < 1)
            perCU = 1;
        resident = perCU * prop.multiProcessorCount;
    }
    const long groupsPerBlock = kBlock / G;
    long blocks = (a.n + groupsPerBlock - 1) / groupsPerBlock;
    long cap = maxBlocks > 0 ? maxBlocks : resident;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    if (a.counter) {
        // spots [0, groups) are the groups' first spots; the counter hands out the rest
        const hipError_t e = launch_set_counter(a.counter, (int)(blocks * groupsPerBlock), st);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((fit_kernel<M, R, G>), dim3((unsigned)blocks), dim3(kBlock), 0, st, m, a, 0);
    if (a.crlb || a.chisq)
        hipLaunchKernelGGL((fit_kernel<M, R, G>), dim3((unsigned)blocks), dim3(kBlock), 0, st, m, a, 1);
    return hipGetLastError();
}

}  // namespace ppa
