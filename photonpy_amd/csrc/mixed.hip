// Persistent mixed-model, mixed-size fit (no reference counterpart).
//
// One launch fits every bin of a batch: bins are (model, ROI size) groups of
// spots, each bin cut into tiles of consecutive spots.  The persistent grid's
// wavefronts claim tiles from one atomic counter (a tile is claimed by a whole
// wavefront, so the model/size switch is wave-uniform) and fit the tile's spots
// with the same lane-group code as the per-model kernel (fit_spots).  Tiles are
// ordered longest-first (estimated cost), so the expensive bins start early and
// the tail of the launch is made of short tiles.  The launch and the counter
// reset are captured once into a HIP graph; each run only refreshes the bin
// table (buffer pointers) before replaying it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "dispatch.h"
#include "mixed.h"

namespace ppa {

// (id, model, R, G) of the variants compiled into the persistent kernel: the
// batch workload's ROI sizes {7,9,11,13,16,20,25} for all four models.
#define PPA_MIXED_VARIANTS(X)                                                                              \
    X(0, kFixed, 7, 2) X(1, kFixed, 9, 2) X(2, kFixed, 11, 2) X(3, kFixed, 13, 4) X(4, kFixed, 16, 4)     \
    X(5, kFixed, 20, 8) X(6, kFixed, 25, 16)                                                               \
    X(7, kSigma, 7, 2) X(8, kSigma, 9, 2) X(9, kSigma, 11, 2) X(10, kSigma, 13, 4) X(11, kSigma, 16, 4)   \
    X(12, kSigma, 20, 8) X(13, kSigma, 25, 16)                                                             \
    X(14, kAstig, 7, 2) X(15, kAstig, 9, 2) X(16, kAstig, 11, 2) X(17, kAstig, 13, 4) X(18, kAstig, 16, 4) \
    X(19, kAstig, 20, 8) X(20, kAstig, 25, 16)                                                             \
    X(21, kSpline, 7, 2) X(22, kSpline, 9, 2) X(23, kSpline, 11, 2) X(24, kSpline, 13, 4)                  \
    X(25, kSpline, 16, 4) X(26, kSpline, 20, 8) X(27, kSpline, 25, 16)

enum { kSpline = 3 };  // model tag for the variant list (Gaussian tags are GaussVariant)

int mixed_variant(int kind, int R, int G) {
#define PPA_MV_FIND(id, m, r, g) \
    if (kind == (int)(m) && R == r && G == g) return id;
    PPA_MIXED_VARIANTS(PPA_MV_FIND)
#undef PPA_MV_FIND
    return -1;
}

int mixed_default_group(int kind, int R) {
#define PPA_MV_G(id, m, r, g) \
    if (kind == (int)(m) && R == r) return g;
    PPA_MIXED_VARIANTS(PPA_MV_G)
#undef PPA_MV_G
    return -1;
}

template <int V>
struct ModelFor {
    using type = GaussModel<V>;
    static PP_DEV type make(const MixedBin& b) { return type{b.gc}; }
};
template <>
struct ModelFor<kSpline> {
    using type = SplineModel;
    static PP_DEV type make(const MixedBin& b) { return b.sm; }
};

__global__ void mixed_reset_kernel(int* counter) { *counter = 0; }

// One non-inlined function per variant: each keeps the register allocation of
// its own per-model kernel instead of the union of all variants.
#define PPA_MV_FUNC(id, m, r, g)                                                                         \
    __device__ __noinline__ void mixed_variant_##id(const MixedBin& b, int begin, int end) {            \
        FitArgs a = b.a;                                                                                 \
        a.n = end;                                                                                       \
        const auto model = ModelFor<m>::make(b);                                                         \
        fit_spots<typename ModelFor<m>::type, r, g>(model, a, begin + (int)(threadIdx.x & (kWave - 1)) / g, \
                                                    kWave / g);                                          \
    }
PPA_MIXED_VARIANTS(PPA_MV_FUNC)
#undef PPA_MV_FUNC

__global__ void __launch_bounds__(256) mixed_kernel(const MixedBin* __restrict__ bins,
                                                    const MixedTile* __restrict__ tiles, int ntiles, int* counter) {
    const int lane = threadIdx.x & (kWave - 1);
    for (;;) {
        int t = 0;
        if (lane == 0) t = atomicAdd(counter, 1);
        t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, kWave));
        if (t >= ntiles) break;
        const MixedTile tile = tiles[t];
        const MixedBin& b = bins[tile.bin];
        switch (b.variant) {
#define PPA_MV_CASE(id, m, r, g) \
    case id: mixed_variant_##id(b, tile.begin, tile.end); break;
            PPA_MIXED_VARIANTS(PPA_MV_CASE)
#undef PPA_MV_CASE
        default: break;
        }
    }
}

hipError_t launch_mixed(const MixedBin* d_bins, const MixedTile* d_tiles, int ntiles, int* d_counter, int blocks,
                        hipStream_t st) {
    hipLaunchKernelGGL(mixed_reset_kernel, dim3(1), dim3(1), 0, st, d_counter);
    hipLaunchKernelGGL(mixed_kernel, dim3(blocks), dim3(256), 0, st, d_bins, d_tiles, ntiles, d_counter);
    return hipGetLastError();
}

int mixed_resident_blocks() {
    static int resident = 0;
    if (resident == 0) {
        int dev = 0, perCU = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, mixed_kernel, 256, 0) != hipSuccess || perCU < 1)
            perCU = 1;
        resident = perCU * prop.multiProcessorCount;
    }
    return resident;
}

}  // namespace ppa
