// Fit-kernel instantiations: splinetiltedgrid model, size list part 2 (written by gen_fit_sources.py).
#include "dispatch.h"
#include "fit_sizes.h"

namespace ppa {


hipError_t launch_fit_splinetiltedgrid_2(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st) {
    using Model = SplineGridModelT<kSplineTilted>;
    Model m = spline_as<kSplineTilted>(d.sg);
    PPA_GRID_VAR_SIZES_2(PPA_FIT_CASE)
    return hipErrorInvalidValue;
}

}  // namespace ppa
