// Bin and tile descriptors of the persistent mixed-batch fit (mixed.hip).
#pragma once

#include "fit_kernel.h"

namespace ppa {

struct MixedBin {
    int variant;   // compiled (model, R, G) variant
    int kind;
    GaussCalib gc;
    SplineModel sm;
    FitArgs a;     // buffers of this bin; a.n = spots in the bin
};

struct MixedTile {
    int bin;
    int begin, end;  // spot range inside the bin
};

int mixed_variant(int kind, int R, int G);
int mixed_default_group(int kind, int R);
int mixed_resident_blocks();
hipError_t launch_mixed(const MixedBin* d_bins, const MixedTile* d_tiles, int ntiles, int* d_counter, int blocks,
                        hipStream_t st);

}  // namespace ppa
