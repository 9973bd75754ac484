// Fit-kernel instantiations: splineimagegrid model, size list part 3 (written by gen_fit_sources.py).
#include "dispatch.h"
#include "fit_sizes.h"

namespace ppa {


hipError_t launch_fit_splineimagegrid_3(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st) {
    using Model = SplineGridModelT<kSplineImage>;
    Model m = spline_as<kSplineImage>(d.sg);
    PPA_GRID_VAR_SIZES_3(PPA_FIT_CASE)
    return hipErrorInvalidValue;
}

}  // namespace ppa
