// Fit-kernel instantiations: splineimage model, size list part 1 (written by gen_fit_sources.py).
#include "dispatch.h"
#include "fit_sizes.h"

namespace ppa {


hipError_t launch_fit_splineimage_1(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st) {
    using Model = SplineModelT<kSplineImage>;
    Model m = spline_as<kSplineImage>(d.sm);
    PPA_VAR_SIZES_1(PPA_FIT_CASE)
    return hipErrorInvalidValue;
}

}  // namespace ppa
