// Fit-kernel instantiations: splinetiltedgrid model, size list part 0 (written by gen_fit_sources.py).
#include "dispatch.h"
#include "fit_sizes.h"

namespace ppa {

hipError_t launch_fit_splinetiltedgrid_1(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st);
hipError_t launch_fit_splinetiltedgrid_2(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st);
hipError_t launch_fit_splinetiltedgrid_3(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st);

hipError_t launch_fit_splinetiltedgrid(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st) {
    using Model = SplineGridModelT<kSplineTilted>;
    Model m = spline_as<kSplineTilted>(d.sg);
    PPA_GRID_VAR_SIZES_0(PPA_FIT_CASE)
    {
        hipError_t e = launch_fit_splinetiltedgrid_1(d, R, G, a, events, st);
        if (e != hipErrorInvalidValue) return e;
    }
    {
        hipError_t e = launch_fit_splinetiltedgrid_2(d, R, G, a, events, st);
        if (e != hipErrorInvalidValue) return e;
    }
    {
        hipError_t e = launch_fit_splinetiltedgrid_3(d, R, G, a, events, st);
        if (e != hipErrorInvalidValue) return e;
    }
    return hipErrorInvalidValue;
}

}  // namespace ppa
