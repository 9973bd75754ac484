"""Writes the fit-kernel instantiation sources fit_<model>_<part>.hip.

Run from this directory after changing the model list; the outputs are
committed.  Each file instantiates one part of PPA_FIT_SIZES for one model so
that `make -j` compiles them in parallel.
"""
MODELS = [("fixed", "GaussModel<kFixed>", "Model m{d.gc};", "PPA_FIT_SIZES"),
          ("sigma", "GaussModel<kSigma>", "Model m{d.gc};", "PPA_FIT_SIZES"),
          ("astig", "GaussModel<kAstig>", "Model m{d.gc};", "PPA_FIT_SIZES"),
          ("spline", "SplineModel", "Model m = d.sm;", "PPA_FIT_SIZES"),
          ("splinegrid", "SplineGridModel", "Model m = d.sg;", "PPA_GRID_SIZES"),
          ("constsigma", "GaussModel<kConstSigma>", "Model m{d.gc};", "PPA_VAR_SIZES"),
          ("tiltedbg", "GaussModel<kTiltedBg>", "Model m{d.gc};", "PPA_VAR_SIZES"),
          ("sigmaxy", "GaussModel<kSigmaXY>", "Model m{d.gc};", "PPA_VAR_SIZES"),
          ("tiltedbgsigmaxy", "GaussModel<kTiltedBgSigmaXY>", "Model m{d.gc};", "PPA_VAR_SIZES"),
          ("ibg", "IBgModel", "Model m{};", "PPA_VAR_SIZES"),
          ("splinetilted", "SplineModelT<kSplineTilted>", "Model m = spline_as<kSplineTilted>(d.sm);",
           "PPA_VAR_SIZES"),
          ("splinetiltedgrid", "SplineGridModelT<kSplineTilted>", "Model m = spline_as<kSplineTilted>(d.sg);",
           "PPA_GRID_VAR_SIZES"),
          ("splineimage", "SplineModelT<kSplineImage>", "Model m = spline_as<kSplineImage>(d.sm);", "PPA_VAR_SIZES"),
          ("splineimagegrid", "SplineGridModelT<kSplineImage>", "Model m = spline_as<kSplineImage>(d.sg);",
           "PPA_GRID_VAR_SIZES")]
PARTS = 4

for name, ty, ctor, sizes in MODELS:
    for part in range(PARTS):
        decls = ""
        tail = "    return hipErrorInvalidValue;\n"
        if part == 0:
            decls = "".join(
                f"hipError_t launch_fit_{name}_{p}(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events,"
                f" hipStream_t st);\n" for p in range(1, PARTS))
            tail = "".join(
                f"    {{\n        hipError_t e = launch_fit_{name}_{p}(d, R, G, a, events, st);\n"
                f"        if (e != hipErrorInvalidValue) return e;\n    }}\n" for p in range(1, PARTS)) + tail
            fname = f"launch_fit_{name}"
        else:
            fname = f"launch_fit_{name}_{part}"
        src = f"""// Fit-kernel instantiations: {name} model, size list part {part} (written by gen_fit_sources.py).
#include "dispatch.h"
#include "fit_sizes.h"

namespace ppa {{

{decls}
hipError_t {fname}(const ModelDesc& d, int R, int G, const FitArgs& a, hipEvent_t* events, hipStream_t st) {{
    using Model = {ty};
    {ctor}
    {sizes}_{part}(PPA_FIT_CASE)
{tail}}}

}}  // namespace ppa
"""
        with open(f"fit_{name}_{part}.hip", "w") as f:
            f.write(src)
