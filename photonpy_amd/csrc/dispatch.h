// Host-side interface between the C-ABI layer and the kernel instantiations.
#pragma once

#include <hip/hip_runtime.h>

#include "fit_kernel.h"
#include "eval_kernels.h"

namespace ppa {

enum ModelKind {
    kKindFixed = 0,
    kKindSigma = 1,
    kKindAstig = 2,
    kKindSpline = 3,
    kKindConstSigma = 4,
    kKindTiltedBg = 5,
    kKindSigmaXY = 6,
    kKindTiltedBgSigmaXY = 7,
    kKindIBg = 8,
    kKindSplineTilted = 9,  // CSpline_Model_XYZITiltedBg (CSpline_CreatePSF mode 1)
    kKindSplineImage = 10,  // CSpline_Model_XYZIBgImage (mode 2)
};

inline bool is_spline(int kind) { return kind == kKindSpline || kind == kKindSplineTilted || kind == kKindSplineImage; }

struct ModelDesc {
    int kind;
    GaussCalib gc;
    SplineModel sm;      // coefs must point to device memory when launched
    SplineGridModel sg;  // grid form of the same spline (when grid != null)
    bool useGrid;        // spline: fit / evaluate from sg instead of sm
};

inline int num_params(int kind) {
    switch (kind) {
    case kKindFixed: case kKindConstSigma: return GaussModel<kFixed>::K;
    case kKindTiltedBg: return GaussModel<kTiltedBg>::K;
    case kKindSigmaXY: return GaussModel<kSigmaXY>::K;
    case kKindTiltedBgSigmaXY: return GaussModel<kTiltedBgSigmaXY>::K;
    case kKindIBg: return IBgModel::K;
    case kKindSplineTilted: return SplineModelT<kSplineTilted>::K;
    default: return 5;
    }
}

// Per-spot constants of a model on R x R ROIs (the reference's NumConstants):
// the background-image spline takes an R x R image per spot.
inline int num_consts(int kind, int R) {
    switch (kind) {
    case kKindConstSigma: return GaussModel<kConstSigma>::NC;
    case kKindTiltedBg: return GaussModel<kTiltedBg>::NC;
    case kKindIBg: return IBgModel::NC;
    case kKindSplineImage: return R * R;
    default: return 0;
    }
}

// Lane-group size used for an R x R ROI unless the tuning table says otherwise:
// the smallest power of two that keeps <= ~64 samples per lane in registers.
inline int default_group(int R) {
    if (R <= 4) return 1;
    if (R <= 11) return 2;
    if (R <= 16) return 4;
    if (R <= 24) return 8;
    return 16;
}

// Lane-group size of the grid-form spline fit (FlatLayout: any G divides the
// pixels evenly enough; larger groups keep the per-group LDS window per lane small).
inline int default_group_grid(int R) {
    if (R <= 6) return 4;
    if (R <= 12) return 8;
    if (R <= 17) return 16;
    if (R <= 22) return 32;
    return 64;
}

inline int default_group_for(const ModelDesc& d, int R) {
    return (is_spline(d.kind) && d.useGrid) ? default_group_grid(R) : default_group(R);
}

constexpr int kMinFitR = 3, kMaxFitR = 32;

// Whether a fit kernel is compiled for (R, G) for this model and form (the size
// lists of fit_sizes.h).
bool fit_compiled(const ModelDesc& d, int R, int G);

// Per-model launchers, defined in fit_<model>_<part>.hip.  Return hipErrorInvalidValue
// when no kernel is compiled for (R, G).  `events` is null or three events that are
// recorded on the stream before the fit kernel, after it and after the CRLB kernel.
hipError_t launch_fit_fixed(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st);
hipError_t launch_fit_sigma(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st);
hipError_t launch_fit_astig(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st);
hipError_t launch_fit_spline(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st);
hipError_t launch_fit_splinegrid(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events,
                                 hipStream_t st);
hipError_t launch_fit_constsigma(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events,
                                 hipStream_t st);
hipError_t launch_fit_tiltedbg(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st);
hipError_t launch_fit_sigmaxy(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st);
hipError_t launch_fit_tiltedbgsigmaxy(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events,
                                      hipStream_t st);
hipError_t launch_fit_ibg(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st);
hipError_t launch_fit_splinetilted(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events,
                                  hipStream_t st);
hipError_t launch_fit_splinetiltedgrid(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events,
                                      hipStream_t st);
hipError_t launch_fit_splineimage(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events,
                                  hipStream_t st);
hipError_t launch_fit_splineimagegrid(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events,
                                      hipStream_t st);

// Fit (+ CRLB / chi^2) of a roisize with no compiled fit kernel (one wavefront per spot).
hipError_t launch_fit_large_any(const ModelDesc& d, int R, const FitArgs& a, hipEvent_t* events, hipStream_t st);
hipError_t launch_ev_any(const ModelDesc& d, int R, int n, const float* params, const float* constants, float* ev,
                         float* deriv, hipStream_t st);
hipError_t launch_crlb_any(const ModelDesc& d, int R, int n, const float* params, const float* constants,
                           const float* sample, float* crlb, float* chisq, hipStream_t st);
// GLRT diagnostics {ll_on, ll_off, mean} per spot into diag [n, 3].
hipError_t launch_glrt_any(const ModelDesc& d, int R, int n, const float* params, const float* constants,
                           const float* sample, int meanOfSpot0, float* diag, hipStream_t st);

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
// Sort scratch of the spline coefficient gradient by window tiles (adjoint.hip).
struct SplineAdjTiles {
    int* key;     // [n] window key (k, j0, i0) of each spot
    int* order;   // [n] spots sorted by key
    int* count;   // [nkeys] spots per key, then the scatter cursor
    int* offset;  // [nkeys] first sorted position of each key
    int2* tiles;  // [maxTiles] (first sorted position, spots)
    int* ntiles;  // [2]: tiles, then the sorted positions handed out (allocation counters)
    int nkeys, maxTiles;
};
// Whether the spline NLL and adjoint of roisize R run by window rounds (spots sorted by
// window), and the bytes of scratch that needs for n spots and nkeys = Nz Ny Nx window keys.
bool spline_adjoint_tiled(int R);
size_t spline_adjoint_workspace(int n, int nkeys);
// ws: spline_adjoint_workspace bytes of device scratch, or null (direct atomics).
// sorted: ws already holds a window sort of a.n spots (made by an earlier call on the same
// stream): it is used as it is.  Any order of the spots gives the same per-spot results (a
// spot outside its round's window evaluates its own cells with the same arithmetic); a
// sort of other parameters only costs speed.
hipError_t launch_adjoint_any(const ModelDesc& d, const AdjArgs& a, void* ws, bool sorted, hipStream_t st);
hipError_t launch_spline_coefs(const int* shape, const float* d_data, float* d_out, int border, hipStream_t st);
// Sets the 2^shift work counters of a fit launch of `blocks` blocks on the stream
// (a one-wave kernel, so it can also be captured in a graph).
hipError_t launch_set_counters(int* counters, int blocks, int groupsPerBlock, int shift, hipStream_t st);

inline hipError_t launch_fit_any(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st) {
    switch (d.kind) {
    case kKindFixed: return launch_fit_fixed(d, R, G, a, events, st);
    case kKindSigma: return launch_fit_sigma(d, R, G, a, events, st);
    case kKindAstig: return launch_fit_astig(d, R, G, a, events, st);
    case kKindSpline:
        return d.useGrid ? launch_fit_splinegrid(d, R, G, a, events, st) : launch_fit_spline(d, R, G, a, events, st);
    case kKindConstSigma: return launch_fit_constsigma(d, R, G, a, events, st);
    case kKindTiltedBg: return launch_fit_tiltedbg(d, R, G, a, events, st);
    case kKindSigmaXY: return launch_fit_sigmaxy(d, R, G, a, events, st);
    case kKindTiltedBgSigmaXY: return launch_fit_tiltedbgsigmaxy(d, R, G, a, events, st);
    case kKindIBg: return launch_fit_ibg(d, R, G, a, events, st);
    case kKindSplineTilted:
        return d.useGrid ? launch_fit_splinetiltedgrid(d, R, G, a, events, st)
                         : launch_fit_splinetilted(d, R, G, a, events, st);
    case kKindSplineImage:
        return d.useGrid ? launch_fit_splineimagegrid(d, R, G, a, events, st)
                         : launch_fit_splineimage(d, R, G, a, events, st);
    default: return hipErrorInvalidValue;
    }
}

// Blocks of a kernel that the whole device keeps resident (block threads per
// block, ldsBytes of dynamic LDS each).
inline int resident_blocks(const void* kernel, int block, size_t ldsBytes) {
    int dev = 0, perCU = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kernel, block, ldsBytes) != hipSuccess || perCU < 1)
        perCU = 1;
    return perCU * prop.multiProcessorCount;
}

// Threads per block of a fit kernel: 256, or fewer for models with per-group LDS
// tables so that one block's tables stay within 64 KiB.
constexpr int fit_block(int G, int ldsFloatsPerGroup) {
    return (ldsFloatsPerGroup == 0 || (256 / G) * ldsFloatsPerGroup * 4 <= 65536)   ? 256
           : ((128 / G) * ldsFloatsPerGroup * 4 <= 65536 && 128 >= G)               ? 128
                                                                                     : 64;
}

// Persistent grids: enough blocks to cover every group once, capped at what the
// device keeps resident.  With work counters the capped fit kernel starts block b's
// groups on segment b % kWorkSegs and hands out the rest from the segments' counters;
// without them (or uncapped), on spot b * groupsPerBlock + group with the grid as
// stride.
template <class M, int R, int G>
hipError_t launch_fit_t(const M& m, const FitArgs& a, hipEvent_t* events, hipStream_t st) {
    constexpr int kBlock = fit_block(G, M::lds_floats(R, G));
    constexpr size_t kLds = (size_t)(kBlock / G) * M::lds_floats(R, G) * sizeof(float);
    static_assert(kLds <= 65536, "per-group LDS table too large for one lane group per 64 lanes");
    static int residentFit = 0, residentCrlb = 0;  // per process; one device model (gfx950)
    if (residentFit == 0) {
        residentFit = resident_blocks(reinterpret_cast<const void*>(fit_kernel<M, R, G>), kBlock, kLds);
        residentCrlb = resident_blocks(reinterpret_cast<const void*>(crlb_phase_kernel<M, R, G>), kBlock, kLds);
        if (residentFit == 0 || residentCrlb == 0) return hipErrorNoDevice;
    }
    const long groupsPerBlock = kBlock / G;
    const long need = (a.n + groupsPerBlock - 1) / groupsPerBlock;
    const long fitBlocks = need < 1 ? 1 : (need > residentFit ? residentFit : need);
    const long crlbBlocks = need < 1 ? 1 : (need > residentCrlb ? residentCrlb : need);
    // A grid that covers every spot once hands out nothing: its groups take spots
    // b * groupsPerBlock + group and stop (no counters, no atomics).  A capped
    // (persistent) grid takes the rest from the work counters.
    FitArgs fa = a;
    if (need <= residentFit) fa.counter = nullptr;
    if (fa.counter) {
        const hipError_t e = launch_set_counters(fa.counter, (int)fitBlocks, (int)groupsPerBlock, fa.segShift, st);
        if (e != hipSuccess) return e;
    }
    if (events) (void)hipEventRecord(events[0], st);
    hipLaunchKernelGGL((fit_kernel<M, R, G>), dim3((unsigned)fitBlocks), dim3(kBlock), kLds, st, m, fa);
    if (events) (void)hipEventRecord(events[1], st);
    if (a.crlb || a.chisq)
        hipLaunchKernelGGL((crlb_phase_kernel<M, R, G>), dim3((unsigned)crlbBlocks), dim3(kBlock), kLds, st, m, a);
    if (events) (void)hipEventRecord(events[2], st);
    return hipGetLastError();
}

}  // namespace ppa
