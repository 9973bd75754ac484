// Fit-kernel instantiations: splineimagegrid model, size list part 0 (written by gen_fit_sources.py).
#include "dispatch.h"
#include "fit_sizes.h"

namespace ppa {

hipError_t launch_fit_splineimagegrid_1(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st);
hipError_t launch_fit_splineimagegrid_2(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st);
hipError_t launch_fit_splineimagegrid_3(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st);

hipError_t launch_fit_splineimagegrid(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st) {
    using Model = SplineGridModelT<kSplineImage>;
    Model m = spline_as<kSplineImage>(d.sg);
    PPA_GRID_VAR_SIZES_0(PPA_FIT_CASE)
    {
        hipError_t e = launch_fit_splineimagegrid_1(d, R, G, a, events, st);
        if (e != hipErrorInvalidValue) return e;
    }
    {
        hipError_t e = launch_fit_splineimagegrid_2(d, R, G, a, events, st);
        if (e != hipErrorInvalidValue) return e;
    }
    {
        hipError_t e = launch_fit_splineimagegrid_3(d, R, G, a, events, st);
        if (e != hipErrorInvalidValue) return e;
    }
    return hipErrorInvalidValue;
}

}  // namespace ppa
