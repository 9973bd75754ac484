// PSF models of the MI355X spot fitter (device side).
//
// Four models, with the reference's parameter order, clamps, stop limits and
// initial estimates:
//   GaussModel<kFixed>  x,y,I,bg        Gauss2D_Model_XYIBg      GaussianPSFModels.h:35-152
//   GaussModel<kSigma>  x,y,I,bg,sigma  Gauss2D_Model_XYIBgSigma GaussianPSFModels.h:308-387
//   GaussModel<kAstig>  x,y,z,I,bg      Gauss2D_Model_XYZIBg     GaussianPSFModels.h:636-751
//                       (sigma(z) gains a quartic term B*t^4/d^2 under the square root;
//                        B = 0 is the reference model)
//   GaussModel<kConstSigma>      x,y,I,bg + sigma x,y per spot    GaussianPSFModels.h:155-229
//   GaussModel<kTiltedBg>        x,y,I,bg,bgx,bgy + sigma x,y     GaussianPSFModels.h:232-304
//   GaussModel<kSigmaXY>         x,y,I,bg,sx,sy                   GaussianPSFModels.h:388-467
//   GaussModel<kTiltedBgSigmaXY> x,y,I,bg,bgx,bgy,sx,sy           GaussianPSFModels.h:471-553
//   IBgModel            I,bg at a given x,y,sigma   IntensityBgModel GaussianPSFModels.h:557-623
// Per-spot constants (NC of them) follow the K parameters in the parameter array
// the kernels pass around; the Levenberg-Marquardt step never touches them.
//   SplineModelT<kSplineFlat>    x,y,z,I,bg          CSpline_Model_XYZIBg        CubicSplinePSFModels.h:14-154
//   SplineModelT<kSplineTilted>  x,y,z,I,bg,bgx,bgy  CSpline_Model_XYZITiltedBg  CubicSplinePSFModels.h:157-298
//   SplineModelT<kSplineImage>   x,y,z,I,bg + a background image per spot (R*R constants)
//                                                    CSpline_Model_XYZIBgImage   CubicSplinePSFModels.h:300-443
//   (SplineGridModelT<...>: the same models evaluated from the spline's sample grid)
// Models with kPixelBg take their background per pixel: mu = (model mu) + bg * b[p]
// and d mu / d bg = b[p], with b the spot's background image, which the kernels
// read like the samples and apply to what the model returns.
//
// The Gaussians are pixel-integrated: for pixel edge e the argument is
// a_e = (e - t - 1/2) / (sqrt(2) sigma), and each pixel's terms are differences
// of erf(a_e), exp(-a_e^2) and a_e exp(-a_e^2) at its two edges.  A fit kernel
// evaluates the R+1 edges of each axis once per iteration (the reference
// re-evaluates both edges for every pixel); eval_pixel() evaluates the same
// expressions for a single pixel so both paths give bit-identical mu and J.
#pragma once

#include <type_traits>

#include "device_common.h"

namespace ppa {

// Pixel ownership inside a lane group of G lanes fitting one R x R spot.  Lane g
// keeps NS samples ("slots"); slot s of lane g is pixel index(g, s) = y*R + x.
// RowLayout: lane g owns the rows [g*RL, g*RL+RL) (the Gaussian and coefficient
// spline models, whose per-row and per-column terms are shared along rows).
// FlatLayout: lane g owns pixels g, g+G, g+2G, ... (models that read their
// per-pixel terms from LDS, where any pixel order costs the same).
//
// The row layouts also split each slot's coordinates as x = xbase(g, region(s)) +
// xoff(s) (and y likewise), region and offsets known at compile time, so a sum over
// the slots such as the centre of mass needs only a few per-lane values.
template <int R, int G>
struct RowLayout {
    static constexpr int RL = (R + G - 1) / G;
    static constexpr int NS = RL * R;
    static constexpr bool kStrided = false;
    static constexpr int NREG = 1;
    PP_DEV static int index(int g, int s) { return g * RL * R + s; }
    PP_DEV static int x(int, int s) { return s % R; }
    PP_DEV static int y(int g, int s) { return g * RL + s / R; }
    static constexpr int region(int) { return 0; }
    static constexpr int xoff(int s) { return s % R; }
    static constexpr int yoff(int s) { return s / R; }
    PP_DEV static int xbase(int, int) { return 0; }
    PP_DEV static int ybase(int g, int) { return g * RL; }
};

// SplitRowLayout (row-factored models where the rows do not divide evenly): the
// lanes work in pairs, pair b owning the band of B = 2R/G rows starting at row b*B,
// with B odd.  RowLayout would give each lane (B+1)/2 rows and leave slots empty;
// here the pair's lane l owns H = (B-1)/2 whole rows (lane 0 band rows [0, H), lane 1
// [H+1, B)) in slots [0, H*R), and half of the band's middle row H in slots H*R + j:
// lane 0 columns j in [0, M), lane 1 columns M + j, M = (R+1)/2.  Lane 1's slot past
// column R - 1 (odd R) is padding: index() puts it past the ROI.
template <int R, int G>
struct SplitRowLayout {
    static constexpr int B = 2 * R / G;  // rows per lane pair
    static_assert(G >= 2 && 2 * R % G == 0 && B % 2 == 1, "an odd band of rows per lane pair");
    static constexpr int H = (B - 1) / 2, M = (R + 1) / 2;
    static constexpr int NS = H * R + M;
    static constexpr bool kStrided = false;
    PP_DEV static int y(int g, int s) {
        const int y0 = (g >> 1) * B;
        return s < H * R ? y0 + (g & 1) * (H + 1) + s / R : y0 + H;
    }
    PP_DEV static int x(int g, int s) { return s < H * R ? s % R : (g & 1) * M + (s - H * R); }
    PP_DEV static int index(int g, int s) { return x(g, s) < R ? y(g, s) * R + x(g, s) : R * R; }
    // region 0: the whole rows; region 1: the half of the middle row
    static constexpr int NREG = 2;
    static constexpr int region(int s) { return s < H * R ? 0 : 1; }
    static constexpr int xoff(int s) { return s < H * R ? s % R : s - H * R; }
    static constexpr int yoff(int s) { return s < H * R ? s / R : 0; }
    PP_DEV static int xbase(int g, int r) { return r == 0 ? 0 : (g & 1) * M; }
    PP_DEV static int ybase(int g, int r) { return (g >> 1) * B + (r == 0 ? (g & 1) * (H + 1) : H); }
};

template <int R, int G>
struct FlatLayout {
    static constexpr int NS = (R * R + G - 1) / G;
    static constexpr bool kStrided = true;
    PP_DEV static int index(int g, int s) { return g + G * s; }
    PP_DEV static int x(int g, int s) { return index(g, s) % R; }
    PP_DEV static int y(int g, int s) { return index(g, s) / R; }
};

// StripLayout (grid-form spline): lane g owns M vertical strips of up to LS pixels.
// Each of the R columns is cut into P = ceil(R / LS) pieces of LS rows (the last one
// shorter); piece p = g + G j (j < M) of the R P pieces is column p % R, rows
// [(p / R) LS, (p / R) LS + LS).  Slot s = j LS + i is row i of lane g's strip j;
// slots past the ROI's last row or past the last piece are padding (index() puts
// them past the ROI).
template <int R, int G, int LS>
struct StripLayout {
    static constexpr int P = (R + LS - 1) / LS;  // pieces per column
    static constexpr int NP = R * P;             // pieces
    static constexpr int M = (NP + G - 1) / G;   // strips per lane
    static constexpr int NS = M * LS;
    static constexpr bool kStrided = true;
    PP_DEV static int piece(int g, int s) { return g + G * (s / LS); }
    PP_DEV static int x(int g, int s) { return piece(g, s) % R; }
    PP_DEV static int y(int g, int s) { return (piece(g, s) / R) * LS + s % LS; }
    PP_DEV static int index(int g, int s) {
        const int p = piece(g, s);
        const int yy = (p / R) * LS + s % LS;
        return (p < NP && yy < R) ? yy * R + p % R : R * R;
    }
};

enum GaussVariant { kFixed = 0, kSigma = 1, kAstig = 2, kConstSigma = 3, kTiltedBg = 4, kSigmaXY = 5, kTiltedBgSigmaXY = 6 };

constexpr float kPiF = 3.141592653589793f;

struct Edge {
    float E, G, H;  // erf(a), exp(-a^2), a*exp(-a^2)
};

// exp(-a^2) as 2^(-a^2 log2 e) on the hardware exp2: relative error about
// a^2 * 2^-24 from rounding the argument (the value itself is below 2^-24 once
// a^2 > 16, so the absolute error stays under one FP32 ulp of 1).  Results
// below FLT_MIN flush to zero.
PP_DEV float exp_neg_sq(float a) { return __builtin_amdgcn_exp2f(-(a * a) * 1.44269504088896341f); }

// erf(a) given g = exp(-a^2), without branches:
//   |a| <  1:  a P(a^2), P of degree 6;
//   |a| >= 1:  1 - g t Q(t), t = 1 / (1 + 0.71 |a|), Q of degree 3
// (the second is the form of Abramowitz & Stegun 7.1.26, erfc = e^{-x^2} t Q(t)).  P and Q
// were fitted for this library to erf(x)/x on [0, 1] and to erfc(x) e^{x^2} / t on [1, 12]
// by weighted least squares with Lawson's minimax reweighting: approximation error below
// 6e-9; the FP32 evaluation is within 1.3e-7 (2.3 ulp) of erf.  It reuses the edge's
// exp(-a^2), where a library erff evaluates its own exponential behind a branch.
PP_DEV float erf_with_exp(float a, float g) {
    const float x = fabsf(a), u = a * a;
    float p = 7.647113349610659e-05f;
    p = fmaf(p, u, -0.0007942355074335833f);
    p = fmaf(p, u, 0.005179749416557941f);
    p = fmaf(p, u, -0.02684859124920865f);
    p = fmaf(p, u, 0.11283430117086246f);
    p = fmaf(p, u, -0.3761260598505243f);
    p = fmaf(p, u, 1.128379158410482f);
    const float lo = x * p;
    const float t = rcpf(fmaf(0.71f, x, 1.0f));
    float q = -0.2933092747783785f;
    q = fmaf(q, t, 0.5490077620704806f);
    q = fmaf(q, t, 0.33261061953102816f);
    q = fmaf(q, t, 0.40756540382343115f);
    const float hi = 1.0f - (g * t) * q;
    return copysignf(x < 1.0f ? lo : hi, a);
}

PP_DEV Edge gauss_edge(float e, float t, float c2) {
    float a = (e - t - 0.5f) * c2;
    float g = exp_neg_sq(a);
    return {erf_with_exp(a, g), g, a * g};
}

// Per-axis terms of one pixel from its left (e0) and right (e1) edges.  dE and S
// are scaled by the intensity I (scale_by) before pixel() uses them, once per
// column and row instead of once per pixel.
struct AxisTerms {
    float D, dE, S;  // integrated Gaussian, d/dt, shape-derivative term

    PP_DEV void scale_by(float I) {
        dE *= I;
        S *= I;
    }
};

PP_DEV AxisTerms axis_terms(const Edge& e0, const Edge& e1, float cp, float cs) {
    AxisTerms t;
    t.D = 0.5f * e1.E - 0.5f * e0.E;
    t.dE = cp * (e0.G - e1.G);
    t.S = (e0.H - e1.H) * cs;
    return t;
}

// Compile-time loop: calls f(std::integral_constant<int, I>) for I in [B, E).
template <int B, int E, class F>
PP_DEV void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Edge E of a group-shared edge set: held by lane E % G in slot E / G.
template <int G, int E, int PER>
PP_DEV Edge fetch_edge(const float (&Ev)[PER], const float (&Gv)[PER], const float (&Hv)[PER]) {
    return {group_bcast<G, E % G>(Ev[E / G]), group_bcast<G, E % G>(Gv[E / G]), group_bcast<G, E % G>(Hv[E / G])};
}

// Calibration of the Gaussian models.  ax/ay = {s0, gamma, d, A, B} per axis.
struct GaussCalib {
    float sx, sy;          // kFixed
    float initialSigma;    // kSigma; x for kSigmaXY / kTiltedBgSigmaXY
    float ax[5], ay[5];    // kAstig
    float minz, maxz;      // kAstig
    float initialSigmaY;   // kSigmaXY / kTiltedBgSigmaXY
};

// Per-spot, per-iteration shape constants of one axis.
struct AxisShape {
    float cp;   // 1/(sqrt(2 pi) sigma)
    float c2;   // 1/(sqrt(2) sigma)
    float cs;   // multiplier of (H0 - H1) in the shape column
};

template <int V>
struct GaussModel {
    static constexpr bool kTilt = (V == kTiltedBg || V == kTiltedBgSigmaXY);
    static constexpr bool kSigXY = (V == kSigmaXY || V == kTiltedBgSigmaXY);
    static constexpr int K = (V == kFixed || V == kConstSigma) ? 4
                             : (V == kTiltedBg || V == kSigmaXY) ? 6
                             : (V == kTiltedBgSigmaXY) ? 8 : 5;
    static constexpr int NC = (V == kConstSigma || V == kTiltedBg) ? 2 : 0;  // sigma x, y per spot
    static constexpr bool kPixelBg = false;
    static constexpr int KT = K + NC;
    static constexpr int iX = 0, iY = 1;
    static constexpr int iZ = 2;                               // kAstig only
    static constexpr int iI = (V == kAstig) ? 3 : 2;
    static constexpr int iBg = (V == kAstig) ? 4 : 3;
    static constexpr int iS = (V == kSigma) ? 4 : (V == kAstig ? 2 : -1);  // column fed by S terms
    static constexpr int iBgX = 4, iBgY = 5;                   // tilted background
    static constexpr int iSx = (V == kSigmaXY) ? 4 : 6;        // sigma x, y columns
    static constexpr int iSy = iSx + 1;
    static constexpr bool kCrlbAbs = false;
    // Models whose fit accumulates the normal equations per row (accumulate_rows).
    static constexpr bool kRowFactored = (V == kFixed || V == kConstSigma || V == kSigma || V == kAstig);

    GaussCalib c;

    PP_DEV static float stop_limit(int k) {
        if (V == kFixed || V == kConstSigma) return (k == 2) ? 1e-1f : 1e-5f;
        if (V == kTiltedBg) return (k == 2) ? 1e-1f : 1e-5f;
        if (V == kSigma || V == kSigmaXY) return (k == 2) ? 1e-1f : (k == 3 ? 1e-4f : 1e-5f);
        if (V == kTiltedBgSigmaXY) return (k == 2) ? 1e-1f : ((k >= 3 && k <= 5) ? 1e-4f : 1e-5f);
        return (k == 2) ? 1e-1f : 1e-4f;
    }

    PP_DEV void check_limits(float (&t)[KT], int R) const {
        const float hi = (float)R - 1.5f - 1.0f;
        t[0] = clampf(t[0], 1.5f, hi);
        t[1] = clampf(t[1], 1.5f, hi);
        if (V == kFixed || V == kConstSigma || V == kTiltedBg) {
            if (t[2] < 25.0f) t[2] = 25.0f;
            if (t[3] < 0.0f) t[3] = 0.0f;
        } else if (V == kSigma) {
            if (t[2] < 10.0f) t[2] = 10.0f;
            if (t[3] < 0.0f) t[3] = 0.0f;
            if (t[4] < 1.0f) t[4] = 1.0f;
            if (t[4] > (float)R * 0.5f) t[4] = (float)R * 0.5f;
        } else if (kSigXY) {
            if (t[2] < 10.0f) t[2] = 10.0f;
            if (t[3] < 0.0f) t[3] = 0.0f;
            t[iSx] = clampf(t[iSx], 1.0f, (float)R * 0.5f);
            t[iSy] = clampf(t[iSy], 1.0f, (float)R * 0.5f);
        } else {
            t[2] = clampf(t[2], c.minz, c.maxz);
            t[3] = fmaxf(t[3], 10.f);
            t[4] = fmaxf(t[4], 0.0f);
        }
    }

    PP_DEV void initial(float cx, float cy, float sum, float (&t)[KT]) const {
        t[0] = cx;
        t[1] = cy;
        if (V == kFixed || V == kConstSigma) { t[2] = sum * 0.7f; t[3] = 0.0f; }
        else if (V == kTiltedBg) { t[2] = sum * 0.7f; t[3] = 0.0f; t[4] = 0.0f; t[5] = 0.0f; }
        else if (V == kSigma) { t[2] = sum; t[3] = 0.0f; t[4] = c.initialSigma; }
        else if (kSigXY) {
            t[2] = sum;
            t[3] = 0.0f;
            if (kTilt) { t[4] = 0.0f; t[5] = 0.0f; }
            t[iSx] = c.initialSigma;
            t[iSy] = c.initialSigmaY;
        }
        else { t[2] = 0.0f; t[3] = sum; t[4] = 0.01f; }
    }

    // sigma_x(z) and its z-derivative for the astigmatic model.
    PP_DEV static void astig_sigma(const float* a, float z, float& sigma, float& dsdz) {
        const float s0 = a[0], gam = a[1], d = a[2], A = a[3], B = a[4];
        const float d2 = d * d;
        const float t = z - gam, t2 = t * t, t3 = t2 * t, t4 = t3 * t;
        const float q = 1.0f + t2 / d2 + A * t3 / d2 + B * t4 / d2;
        const float sq = sqrtf(q);
        sigma = s0 * sq;
        dsdz = s0 * (2 * t / d2 + A * 3 * t2 / d2 + B * 4 * t3 / d2) / (2 * sq);
    }

    PP_DEV static void fixed_shape(float sigma, float r2pi, float r2, AxisShape& A) {
        A.cp = 1.0f / (r2pi * sigma);
        A.c2 = 1.0f / (r2 * sigma);
        A.cs = 0.0f;
    }

    PP_DEV void shapes(const float (&t)[KT], AxisShape& X, AxisShape& Y) const {
        const float r2pi = sqrtf(2 * kPiF), r2 = sqrtf(2.0f), ispi = 1.0f / sqrtf(kPiF);
        if (V == kFixed) {
            fixed_shape(c.sx, r2pi, r2, X);
            fixed_shape(c.sy, r2pi, r2, Y);
        } else if (NC == 2) {
            fixed_shape(t[K], r2pi, r2, X);
            fixed_shape(t[K + 1], r2pi, r2, Y);
        } else if (V == kSigma) {
            const float s = t[4];
            X.cp = Y.cp = 1.0f / (r2pi * s);
            X.c2 = Y.c2 = 1.0f / (r2 * s);
            // The reference's d(DeltaX)/d(sigma) omits the 1/sigma factor
            // (GaussianPSFModels.h:359,368); kept for parity of Estimate,
            // Derivatives and CRLB.
            X.cs = Y.cs = ispi;
        } else if (kSigXY) {
            // same 1/sigma omission per axis (GaussianPSFModels.h:434-447)
            fixed_shape(t[iSx], r2pi, r2, X);
            fixed_shape(t[iSy], r2pi, r2, Y);
            X.cs = Y.cs = ispi;
        } else {
            float sgx, dsx, sgy, dsy;
            astig_sigma(c.ax, t[iZ], sgx, dsx);
            astig_sigma(c.ay, t[iZ], sgy, dsy);
            X.cp = 1.0f / (r2pi * sgx); X.c2 = 1.0f / (r2 * sgx); X.cs = ispi / sgx * dsx;
            Y.cp = 1.0f / (r2pi * sgy); Y.c2 = 1.0f / (r2 * sgy); Y.cs = ispi / sgy * dsy;
        }
    }

    // mu and Jacobian of pixel (x, y) from its column (X) and row (Y) terms, whose dE
    // and S are already multiplied by I; half = R / 2 (the tilt's origin).  The
    // kTiltedBgSigmaXY reference leaves the tilt out of mu while its Jacobian has
    // the tilt columns (GaussianPSFModels.h:537-540); kept for parity.
    PP_DEV void pixel(const AxisTerms& X, const AxisTerms& Y, float x, float y, float half, const float (&t)[KT],
                      float& mu, float (&J)[K]) const {
        const float I = t[iI], bg = t[iBg];
        const float p = X.D * Y.D;
        mu = bg + I * p;
        if (V == kTiltedBg) mu = bg + t[iBgX] * (x - half) + t[iBgY] * (y - half) + I * p;
        J[0] = X.dE * Y.D;
        J[1] = Y.dE * X.D;
        J[iI] = p;
        J[iBg] = 1.0f;
        if (V == kSigma || V == kAstig) J[iS] = X.S * Y.D + Y.S * X.D;
        if (kSigXY) {
            J[iSx] = X.S * Y.D;
            J[iSy] = Y.S * X.D;
        }
        if (kTilt) {
            J[iBgX] = x - half;
            J[iBgY] = y - half;
        }
    }

    PP_DEV static AxisTerms scaled(AxisTerms a, float I) {
        a.scale_by(I);
        return a;
    }

    // Single-pixel evaluation (expected value / derivatives kernels).
    PP_DEV void eval_pixel(const float (&t)[KT], int x, int y, int R, float& mu, float (&J)[K]) const {
        AxisShape SX, SY;
        shapes(t, SX, SY);
        const Edge x0 = gauss_edge((float)x, t[0], SX.c2), x1 = gauss_edge((float)(x + 1), t[0], SX.c2);
        const Edge y0 = gauss_edge((float)y, t[1], SY.c2), y1 = gauss_edge((float)(y + 1), t[1], SY.c2);
        pixel(scaled(axis_terms(x0, x1, SX.cp, SX.cs), t[iI]), scaled(axis_terms(y0, y1, SY.cp, SY.cs), t[iI]),
              (float)x, (float)y, (float)R * 0.5f, t, mu, J);
    }

    // The row-factored models split the middle row of each lane pair's band when
    // the band has an odd number of rows (2R/G odd: odd R at 2 lanes, R = 10, 14, ...
    // at 4, R = 12, 20, 28 at 8, R = 8, 24 at 16).
    template <int R, int G>
    static constexpr bool kSplit = kRowFactored && G >= 2 && 2 * R % G == 0 && (2 * R / G) % 2 == 1;
    template <int R, int G>
    using Layout = std::conditional_t<kSplit<R, G>, SplitRowLayout<R, G>, RowLayout<R, G>>;
    // Register budgets that measured faster than the compiler's own choice, a few
    // spilled registers included: three wavefronts per SIMD (<= 168 VGPRs) for the
    // two-lane split fits (11x11 free sigma: fit kernel 4.27 -> 4.20 ms) and the
    // 16x16 astigmatic fit at 4 lanes (11.4 -> 11.2 ms), four (<= 128) for it at 8
    // lanes (13.1 -> 12.6 ms).  One more wavefront than that spills several times as
    // much and runs 27-50% slower.
    static constexpr int min_waves(int R, int G) {
        if (kRowFactored && G == 2 && R % 2 == 1) return 3;
        if (V == kAstig && R == 16 && G == 8) return 4;
        if (V == kAstig && R == 16 && G == 4) return 3;
        return 1;
    }
    // The row-factored fits of groups of >= 2 lanes keep the R column terms of an
    // iteration in the group's LDS region, (dE, D, S, pad) per column (sigma / z
    // models) or (dE, D) (fixed sigma): each lane of the group evaluates its share of
    // the columns once and every lane reads them per pixel, instead of every lane
    // holding all of them in VGPRs (which spilled under the three-wave budgets).  A
    // one-lane group shares nothing and keeps them in VGPRs.
    static constexpr int kColFloats = (V == kSigma || V == kAstig) ? 4 : 2;
    static constexpr bool cols_in_lds(int G) { return kRowFactored && G >= 2; }
    // The groups of a half-wave read the same column at once from their own tables;
    // the table stride keeps those reads on disjoint banks (64 banks of 4 bytes, each
    // read spanning kColFloats banks): the smallest stride >= kColFloats R whose
    // multiples by 1 .. groups - 1 stay kColFloats away from 0 mod 64.
    static constexpr int col_stride(int R, int G) {
        const int groups = G >= 32 ? 1 : 32 / G;
        for (int per = kColFloats * R;; per += kColFloats) {
            bool ok = true;
            for (int k = 1; k < groups; k++) {
                const int d = (k * per) % 64;
                if (d < kColFloats || 64 - d < kColFloats) ok = false;
            }
            if (ok) return per;
        }
    }
    static constexpr int lds_floats(int R, int G) { return cols_in_lds(G) ? col_stride(R, G) : 0; }

    // Fit-kernel evaluation of lane g's pixels: fn(slot, mu, J) for every slot of
    // RowLayout<R, G> that lies inside the ROI.
    template <int R, int G, class Fn>
    PP_DEV void for_pixels(const float (&t)[KT], int g, float*, Fn&& fn) const {
        static_assert(!kSplit<R, G>, "the split layout is accumulated by accumulate_rows");
        constexpr int RL = RowLayout<R, G>::RL;
        for_rows<R, RL, G>(t, g * RL, [&](int r, int x, float mu, const float (&J)[K]) { fn(r * R + x, mu, J); });
    }

    // The R column terms (I folded into dE and S), the group sharing the R+1 column
    // edges: lane g evaluates edges g, g+G, ... and the others receive them by lane
    // broadcast.  Called by every lane of an aligned group of G lanes.
    template <int R, int G>
    PP_DEV void columns(const float (&t)[KT], const AxisShape& SX, AxisTerms (&col)[R]) const {
        if constexpr (G == 1) {
            Edge prev = gauss_edge(0.0f, t[0], SX.c2);
#pragma unroll
            for (int x = 0; x < R; x++) {
                Edge nx = gauss_edge((float)(x + 1), t[0], SX.c2);
                col[x] = scaled(axis_terms(prev, nx, SX.cp, SX.cs), t[iI]);
                prev = nx;
            }
        } else {
            constexpr int PER = (R + 1 + G - 1) / G;  // edges evaluated per lane
            const int g = threadIdx.x & (G - 1);
            float Ev[PER], Gv[PER], Hv[PER];
#pragma unroll
            for (int j = 0; j < PER; j++) {
                const Edge ed = gauss_edge((float)(g + j * G), t[0], SX.c2);
                Ev[j] = ed.E;
                Gv[j] = ed.G;
                Hv[j] = ed.H;
            }
            Edge prev = fetch_edge<G, 0>(Ev, Gv, Hv);
            static_for<1, R + 1>([&](auto e) {
                constexpr int E = decltype(e)::value;
                const Edge nx = fetch_edge<G, E>(Ev, Gv, Hv);
                col[E - 1] = scaled(axis_terms(prev, nx, SX.cp, SX.cs), t[iI]);
                prev = nx;
            });
        }
    }

    // The R column terms (I folded into dE and S) into the group's LDS table: lane g
    // evaluates columns [g C, g C + C), C = ceil(R / G), from its C + 1 edges.  Called
    // by every lane of an aligned group of G lanes; the group reads the table after it.
    template <int R, int G>
    PP_DEV void columns_lds(const float (&t)[KT], const AxisShape& SX, int g, float* lds) const {
        constexpr int C = (R + G - 1) / G;
        const int x0 = g * C;
        // the previous iteration's reads of the table complete before it is overwritten
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        Edge prev = gauss_edge((float)x0, t[0], SX.c2);
#pragma unroll
        for (int j = 0; j < C; j++) {
            const int x = x0 + j;
            const Edge nx = gauss_edge((float)(x + 1), t[0], SX.c2);
            const AxisTerms a = scaled(axis_terms(prev, nx, SX.cp, SX.cs), t[iI]);
            prev = nx;
            if (C * G == R || x < R) {
                if constexpr (kColFloats == 4)
                    reinterpret_cast<float4*>(lds)[x] = make_float4(a.dE, a.D, a.S, 0.0f);
                else
                    reinterpret_cast<float2*>(lds)[x] = make_float2(a.dE, a.D);
            }
        }
        // the group's writes complete before any of its lanes reads the table (one
        // wavefront: workgroup-scope fences order its LDS operations, no workgroup barrier)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }

    // Column x's terms from the LDS table for the j-th pixel of a row: a = dE, b = D,
    // c = S (0 for fixed sigma).  The reads are issued kColAhead pixels at a time (two
    // (dE, D) pairs make one ds_read2_b64): the compiler barrier (it emits no
    // instruction) keeps them from being hoisted over the whole row and held in VGPRs.
    static constexpr int kColAhead = 2;
    PP_DEV void column_from_lds(const float* lds, int x, int j, float& a, float& b, float& c) const {
        if (j % kColAhead == 0) asm volatile("" ::: "memory");
        if constexpr (kColFloats == 4) {
            const float4 v = reinterpret_cast<const float4*>(lds)[x];
            a = v.x; b = v.y; c = v.z;
        } else {
            const float2 v = reinterpret_cast<const float2*>(lds)[x];
            a = v.x; b = v.y; c = 0.0f;
        }
    }

    // Row sums of accumulate_rows: sum w p q over p, q in {a, b, c, 1} and sum v p
    // (FISHER: T1 holds chi^2 and the other T sums are unused).
    struct RowSums {
        float Saa, Sab, Sa1, Sbb, Sb1, S11, Ta, Tb, T1;
        float Sac = 0.0f, Sbc = 0.0f, Scc = 0.0f, Sc1 = 0.0f, Tc = 0.0f;
    };

    // Sums over the N pixels j of one row (or row part) with row term D; pix(j, a, b,
    // c, x, ok) gives pixel j's column terms, sample and whether it is a pixel at all.
    // mu = bg + (I D) b: one FMA per pixel.
    template <bool FISHER, int N, class Pix>
    PP_DEV static void row_sums(float I, float bg, float D, Pix&& pix, RowSums& z) {
        const float ID = I * D;
        constexpr bool SH = (V == kSigma || V == kAstig);  // has the c column
#pragma unroll
        for (int j = 0; j < N; j++) {
            float a, b, c, xs;
            bool ok;
            pix(j, a, b, c, xs, ok);
            const float mu = bg + ID * b;
            float w, v = 0.0f, chi = 0.0f;
            if constexpr (FISHER) {
                const float m = fmaxf(1e-8f, mu);
                const float inv = rcpf(m);
                const float e = xs - m;
                w = inv;
                chi = e * e * inv;
            } else {
                const float inv = rcpf(fmaxf(mu, 1e-6f));
                const float tv = xs * inv;  // samples clamped at load
                w = tv * inv;
                v = tv - 1.0f;
            }
            if (!ok) w = v = chi = 0.0f;  // padding slot (folds away where ok is constant)
            const float wa = w * a, wb = w * b;
            if (j == 0) {
                z.Saa = wa * a; z.Sab = wa * b; z.Sa1 = wa;
                z.Sbb = wb * b; z.Sb1 = wb; z.S11 = w;
                if constexpr (FISHER) {
                    z.T1 = chi;
                } else {
                    z.Ta = v * a; z.Tb = v * b; z.T1 = v;
                }
            } else {
                z.Saa += wa * a; z.Sab += wa * b; z.Sa1 += wa;
                z.Sbb += wb * b; z.Sb1 += wb; z.S11 += w;
                if constexpr (FISHER) {
                    z.T1 += chi;
                } else {
                    z.Ta += v * a; z.Tb += v * b; z.T1 += v;
                }
            }
            if constexpr (SH) {
                const float wc = w * c;
                if (j == 0) {
                    z.Sac = wa * c; z.Sbc = wb * c; z.Scc = wc * c; z.Sc1 = wc;
                    if constexpr (!FISHER) z.Tc = v * c;
                } else {
                    z.Sac += wa * c; z.Sbc += wb * c; z.Scc += wc * c; z.Sc1 += wc;
                    if constexpr (!FISHER) z.Tc += v * c;
                }
            }
        }
    }

    // alpha_ij += sum_pq Ri_p Rj_q (sum w p q), beta_i += sum_p Ri_p (sum v p) for one
    // row (FISHER: the Fisher matrix and chi^2); the first row sets acc.
    template <bool FISHER, int NACC>
    PP_DEV static void fold_row(const AxisTerms& row, const RowSums& z, bool first, float (&acc)[NACC]) {
        constexpr bool SH = (V == kSigma || V == kAstig);
        constexpr int NA = K * (K + 1) / 2;
        // packed upper-triangle index of alpha(i, j), i <= j
        auto q = [](int i, int j) { return i * K - i * (i - 1) / 2 + (j - i); };
        auto put = [&](int i, float val) {
            if (first)
                acc[i] = val;
            else
                acc[i] += val;
        };
        const float D = row.D, E = row.dE;
        const float DD = D * D, DE = D * E;
        constexpr int X = 0, Y = 1;
        put(q(X, X), DD * z.Saa);
        put(q(X, Y), DE * z.Sab);
        put(q(X, iI), DD * z.Sab);
        put(q(X, iBg), D * z.Sa1);
        put(q(Y, Y), (E * E) * z.Sbb);
        put(q(Y, iI), DE * z.Sbb);
        put(q(Y, iBg), E * z.Sb1);
        put(q(iI, iI), DD * z.Sbb);
        put(q(iI, iBg), D * z.Sb1);
        put(q(iBg, iBg), z.S11);
        if constexpr (FISHER) {
            put(NA, z.T1);
        } else {
            put(NA + X, D * z.Ta);
            put(NA + Y, E * z.Tb);
            put(NA + iI, D * z.Tb);
            put(NA + iBg, z.T1);
        }
        if constexpr (SH) {
            // J_s = D c + Sy b
            const float Sy = row.S;
            const float DS = D * Sy;
            constexpr int s0 = iS;
            put(q(X < s0 ? X : s0, X < s0 ? s0 : X), DS * z.Sab + DD * z.Sac);
            put(q(Y < s0 ? Y : s0, Y < s0 ? s0 : Y), (E * Sy) * z.Sbb + DE * z.Sbc);
            put(q(iI < s0 ? iI : s0, iI < s0 ? s0 : iI), DS * z.Sbb + DD * z.Sbc);
            put(q(iBg < s0 ? iBg : s0, iBg < s0 ? s0 : iBg), Sy * z.Sb1 + D * z.Sc1);
            put(q(s0, s0), (Sy * Sy) * z.Sbb + (2.0f * DS) * z.Sbc + DD * z.Scc);
            if constexpr (!FISHER) put(NA + s0, Sy * z.Tb + D * z.Tc);
        }
    }

    // Fit-kernel normal equations of lane g's rows (Layout<R, G>), for the models
    // whose Jacobian columns are sums of (row term) x (column term):
    //   J_x = Dy a, J_y = Ey b, J_I = Dy b, J_bg = 1, J_s = Dy c + Sy b
    // with column terms a = I dEx, b = Dx, c = I Sx and row terms Dy, Ey = I dEy,
    // Sy = I Sy (J_s is the sigma / z column; absent for fixed sigma).  Per pixel
    // it forms a weight w and accumulates the row sums sum w p q over p, q in
    // {a, b, c, 1}; per row it adds alpha_ij += sum_pq Ri_p Rj_q (sum w p q) with Ri
    // the row coefficients of J_i.  The sums are those of the per-pixel form,
    // sum w J_i J_j, regrouped.  acc is overwritten: the first row sets it.
    //   FISHER = false (LM step): w = x/mu^2 with mu >= 1e-6, and also
    //     beta_i = sum (x/mu - 1) J_i; acc = packed alpha (i <= j), then beta.
    //   FISHER = true (CRLB): w = 1/mu with mu >= 1e-8 (the Fisher matrix), and
    //     chi^2 = sum (x - mu)^2 / mu; acc = packed Fisher matrix, then chi^2.
    template <int R, int G, bool FISHER = false, int NS, int NACC>
    PP_DEV void accumulate_rows(const float (&t)[KT], int g, const float (&s)[NS], float (&acc)[NACC],
                                float* lds) const {
        constexpr int NA = K * (K + 1) / 2;
        static_assert(NACC == NA + (FISHER ? 1 : K), "accumulator layout");
        static_assert(NS == Layout<R, G>::NS, "sample slots of the layout");
        AxisShape SX, SY;
        shapes(t, SX, SY);
        constexpr bool kLds = cols_in_lds(G);
        AxisTerms col[kLds ? 1 : R];
        if constexpr (kLds)
            columns_lds<R, G>(t, SX, g, lds);
        else
            columns<R, G>(t, SX, col);
        // column x's terms for the j-th pixel of a row
        auto colv = [&](int x, int j, float& a, float& b, float& c) {
            if constexpr (kLds) {
                column_from_lds(lds, x, j, a, b, c);
            } else {
                a = col[x].dE; b = col[x].D; c = col[x].S;
            }
        };
        const float I = t[iI], bg = t[iBg];
        if constexpr (kSplit<R, G>) {
            using SL = SplitRowLayout<R, G>;
            constexpr int H = SL::H, M = SL::M;
            const int l = g & 1;  // lane in the pair; the pair's band starts at row yb
            const int yb = (g >> 1) * SL::B;
            const int y0 = yb + l * (H + 1);
            const Edge first = gauss_edge((float)y0, t[1], SY.c2);
            Edge prev = first;
#pragma unroll
            for (int r = 0; r < H; r++) {
                const Edge nx = gauss_edge((float)(y0 + r + 1), t[1], SY.c2);
                const AxisTerms row = scaled(axis_terms(prev, nx, SY.cp, SY.cs), I);
                prev = nx;
                RowSums z;
                row_sums<FISHER, R>(I, bg, row.D, [&](int x, float& a, float& b, float& c, float& xs, bool& ok) {
                    colv(x, x, a, b, c);
                    xs = s[r * R + x];
                    ok = true;
                }, z);
                fold_row<FISHER>(row, z, r == 0, acc);
            }
            // the band's middle row yb + H: lane 0 holds its upper edge (its last) and lane
            // 1 its lower edge (its first); each evaluates the other one
            const Edge en = gauss_edge((float)(yb + (l == 0 ? H + 1 : H)), t[1], SY.c2);
            const Edge lo = l == 0 ? prev : en, hi = l == 0 ? en : first;
            const AxisTerms row = scaled(axis_terms(lo, hi, SY.cp, SY.cs), I);
            RowSums z;
            row_sums<FISHER, M>(I, bg, row.D, [&](int j, float& a, float& b, float& c, float& xs, bool& ok) {
                if (M + j < R) {  // lane 0 column j, lane 1 column M + j
                    colv(j + (l != 0 ? M : 0), j, a, b, c);
                    ok = true;
                } else {  // odd R: column M - 1 for lane 0, padding for lane 1
                    colv(j, j, a, b, c);
                    ok = l == 0;
                }
                xs = s[H * R + j];
            }, z);
            fold_row<FISHER>(row, z, H == 0, acc);
        } else {
            constexpr int RL = RowLayout<R, G>::RL;
            const int y0 = g * RL;
            Edge prev = gauss_edge((float)y0, t[1], SY.c2);
#pragma unroll
            for (int r = 0; r < RL; r++) {
                const int y = y0 + r;
                const Edge nx = gauss_edge((float)(y + 1), t[1], SY.c2);
                const AxisTerms row = scaled(axis_terms(prev, nx, SY.cp, SY.cs), I);
                prev = nx;
                if (RL * G == R || y < R) {
                    RowSums z;
                    row_sums<FISHER, R>(I, bg, row.D, [&](int x, float& a, float& b, float& c, float& xs, bool& ok) {
                        colv(x, x, a, b, c);
                        xs = s[r * R + x];
                        ok = true;
                    }, z);
                    fold_row<FISHER>(row, z, r == 0, acc);
                } else if (r == 0) {
#pragma unroll
                    for (int i = 0; i < NACC; i++) acc[i] = 0.0f;
                }
            }
        }
    }

    // Fit-kernel evaluation over a block of RL rows starting at y0 of an RxR ROI,
    // called by every lane of an aligned group of G lanes.  f(r, x, mu, J) is called
    // for every pixel of rows y0+r < R.
    template <int R, int RL, int G, class F>
    PP_DEV void for_rows(const float (&t)[KT], int y0, F&& f) const {
        AxisShape SX, SY;
        shapes(t, SX, SY);
        AxisTerms col[R];
        columns<R, G>(t, SX, col);
        Edge prev = gauss_edge((float)y0, t[1], SY.c2);
#pragma unroll
        for (int r = 0; r < RL; r++) {
            const int y = y0 + r;
            Edge nx = gauss_edge((float)(y + 1), t[1], SY.c2);
            const AxisTerms row = scaled(axis_terms(prev, nx, SY.cp, SY.cs), t[iI]);
            prev = nx;
            if (y < R) {
#pragma unroll
                for (int x = 0; x < R; x++) {
                    float mu, J[K];
                    pixel(col[x], row, (float)x, (float)y, (float)R * 0.5f, t, mu, J);
                    f(r, x, mu, J);
                }
            }
        }
    }
};

// ---------------------------------------------------------------------------
// Cubic-spline PSF: coefs[Nz][Ny][Nx][64], coefficient q = 16*m + 4*n + o
// multiplies z^m y^n x^o of the sub-voxel position.
// Grid, clamps, stop limits and initial values shared by both spline forms.
enum SplineBg { kSplineFlat = 0, kSplineTilted = 1, kSplineImage = 2 };

template <int BG>
struct SplineCommon {
    static constexpr int K = BG == kSplineTilted ? 7 : 5;
    static constexpr int NC = 0, KT = K;  // the background image is read per pixel, not as th
    static constexpr bool kCrlbAbs = false;
    static constexpr bool kPixelBg = BG == kSplineImage;
    static constexpr int iI = 3, iBg = 4;
    int Nz, Ny, Nx;  // cells per axis
    float zmin, zmax;

    // stop limits {1e-4, 1e-4, 1e-7, 1e-4, ...} (CubicSplinePSFModels.h:34-38, 175-179, 321-325)
    PP_DEV static float stop_limit(int k) { return (k == 2) ? 1e-7f : 1e-4f; }

    PP_DEV void check_limits(float (&t)[K], int R) const {
        const float hi = (float)R - 1.0f - 1.5f;
        t[0] = clampf(t[0], 1.5f, hi);
        t[1] = clampf(t[1], 1.5f, hi);
        t[2] = clampf(t[2], zmin + 1e-3f, zmax - 1e-3f);
        t[3] = fmaxf(t[3], 10.f);
        t[4] = fmaxf(t[4], 0.0f);
    }

    PP_DEV void initial(float cx, float cy, float sum, float (&t)[K]) const {
        t[0] = cx; t[1] = cy; t[2] = 0.0f; t[3] = sum; t[4] = 0.0f;
#pragma unroll
        for (int k = 5; k < K; k++) t[k] = 0.0f;
    }

    // Spot coordinate relative to the spline grid of n cells: t - n/2, or, for the
    // background-image model, ((t - R/2) + (R - n + 1)/2) + 1/2 (one pixel further,
    // CubicSplinePSFModels.h:79-80 vs 374-377), in the reference's operation order.
    PP_DEV static float rel(float t, int n, int R) {
        if constexpr (BG == kSplineImage)
            return ((t - (float)R / 2.0f) + ((float)R - (float)n + 1.0f) / 2.0f) + 0.5f;
        else
            return t - (float)n * 0.5f;
    }

    // mu and the background columns of J at pixel (x, y) from I * psf: flat bg,
    // bg + bgx (x - R/2) + bgy (y - R/2), or nothing (the image term is added by
    // the kernel, which owns the image).
    PP_DEV static void background(const float (&t)[K], float I, float psf, int x, int y, int R, float& mu,
                                  float (&J)[K]) {
        if constexpr (BG == kSplineTilted) {
            const float half = (float)R * 0.5f;
            const float ox = (float)x - half, oy = (float)y - half;
            J[4] = 1.0f;
            J[5] = ox;
            J[6] = oy;
            mu = I * psf + t[4] + t[5] * ox + t[6] * oy;
        } else if constexpr (BG == kSplineImage) {
            J[4] = 0.0f;
            mu = I * psf;
        } else {
            J[4] = 1.0f;
            mu = I * psf + t[4];
        }
    }
};

template <int BG>
struct SplineModelT : SplineCommon<BG> {
    using Base = SplineCommon<BG>;
    using Base::K;
    using Base::Nz;
    using Base::Ny;
    using Base::Nx;
    using Base::zmin;
    using Base::zmax;
    const float* coefs;

    struct Frame {
        float tx, ty;              // position relative to the spline grid centre
        float du;
        float x1, x2, x3;          // x, x^2, x^3 of the sub-pixel x
        float dx2, dx3;            // 2x, 3x^2
        float yp[4], dyp[4];       // y^n, n y^(n-1)
        float zp[4], dzp[4];       // z^m, m z^(m-1)
        const float* slab;         // coefficients of z-slab k
    };

    PP_DEV void frame(const float (&t)[K], int R, Frame& F) const {
        F.du = (zmax - zmin) / (float)Nz;
        F.tx = Base::rel(t[0], Nx, R);
        F.ty = Base::rel(t[1], Ny, R);
        float spx = -F.tx; spx -= (float)(int)spx;   // truncation toward zero (reference quirk)
        float spy = -F.ty; spy -= (float)(int)spy;
        float spz = (t[2] - zmin) / F.du;
        const int k = clampi((int)spz, 0, Nz - 1);
        spz = clampf(spz - (float)k, 0.0f, 1.0f);
        F.x1 = spx; F.x2 = spx * spx; F.x3 = F.x2 * spx;
        F.dx2 = 2.0f * spx; F.dx3 = 3.0f * F.x2;
        F.yp[0] = 1.0f; F.yp[1] = spy; F.yp[2] = spy * spy; F.yp[3] = F.yp[2] * spy;
        F.dyp[0] = 0.0f; F.dyp[1] = 1.0f; F.dyp[2] = 2.0f * spy; F.dyp[3] = 3.0f * F.yp[2];
        F.zp[0] = 1.0f; F.zp[1] = spz; F.zp[2] = spz * spz; F.zp[3] = F.zp[2] * spz;
        F.dzp[0] = 0.0f; F.dzp[1] = 1.0f; F.dzp[2] = 2.0f * spz; F.dzp[3] = 3.0f * F.zp[2];
        F.slab = coefs + (size_t)64 * ((size_t)k * Nx * Ny);
    }

    // value and x/y/z derivatives of the spline cell c at the frame's sub-voxel position
    // One z-power of a cell: its 16 coefficients q (4 rows of 4 x-powers) contracted with
    // the x and y powers into the value a and the x / y derivatives ax, ay.
    PP_DEV static void cell_m(const float4 (&q)[4], const Frame& F, float& a, float& ax, float& ay) {
        a = 0.0f, ax = 0.0f, ay = 0.0f;
#pragma unroll
        for (int n = 0; n < 4; n++) {
            const float px = q[n].x + q[n].y * F.x1 + q[n].z * F.x2 + q[n].w * F.x3;
            const float dpx = q[n].y + q[n].z * F.dx2 + q[n].w * F.dx3;
            a += px * F.yp[n];
            ax += dpx * F.yp[n];
            ay += px * F.dyp[n];
        }
    }

    // The four z-powers' parts combined: f = fm0 + fm1 z + fm2 z^2 + fm3 z^3, its x / y
    // derivatives alike, fz = fm1 + fm2 2z + fm3 3z^2.
    PP_DEV static void cell_combine(const float (&fm)[4], const float (&fxm)[4], const float (&fym)[4],
                                    const Frame& F, float& f, float& fx, float& fy, float& fz) {
        f = fm[0] + fm[1] * F.zp[1] + fm[2] * F.zp[2] + fm[3] * F.zp[3];
        fx = fxm[0] + fxm[1] * F.zp[1] + fxm[2] * F.zp[2] + fxm[3] * F.zp[3];
        fy = fym[0] + fym[1] * F.zp[1] + fym[2] * F.zp[2] + fym[3] * F.zp[3];
        fz = fm[1] + fm[2] * F.dzp[2] + fm[3] * F.dzp[3];
    }

    PP_DEV static void cell(const float* c, const Frame& F, float& f, float& fx, float& fy, float& fz) {
        const float4* c4 = reinterpret_cast<const float4*>(c);
        float fm[4], fxm[4], fym[4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const float4 q[4] = {c4[m * 4], c4[m * 4 + 1], c4[m * 4 + 2], c4[m * 4 + 3]};
            cell_m(q, F, fm[m], fxm[m], fym[m]);
        }
        cell_combine(fm, fxm, fym, F, f, fx, fy, fz);
    }

    PP_DEV const float* row_ptr(const Frame& F, int y) const {
        const int j = clampi((int)((float)y - F.ty), 0, Ny - 1);
        return F.slab + (size_t)64 * ((size_t)j * Nx);
    }

    PP_DEV int col_index(const Frame& F, int x) const { return clampi((int)((float)x - F.tx), 0, Nx - 1); }

    // mu and J at pixel (x, y) from the raw spline value and derivatives.
    PP_DEV static void finish(const Frame& F, const float (&t)[K], float f, float fx, float fy, float fz, int x,
                              int y, int R, float& mu, float (&J)[K]) {
        const float I = t[3];
        J[0] = fx * -I;
        J[1] = fy * -I;
        J[2] = fz * (I / F.du);
        const float psf = f < 0.0f ? 0.0f : f;
        J[3] = psf;
        Base::background(t, I, psf, x, y, R, mu, J);
    }

    PP_DEV void eval_pixel(const float (&t)[K], int x, int y, int R, float& mu, float (&J)[K]) const {
        Frame F;
        frame(t, R, F);
        const float* c = row_ptr(F, y) + 64 * col_index(F, x);
        float f, fx, fy, fz;
        cell(c, F, f, fx, fy, fz);
        finish(F, t, f, fx, fy, fz, x, y, R, mu, J);
    }

    template <int R, int G>
    using Layout = RowLayout<R, G>;
    static constexpr int lds_floats(int, int) { return 0; }

    template <int R, int G, class Fn>
    PP_DEV void for_pixels(const float (&t)[K], int g, float*, Fn&& fn) const {
        constexpr int RL = RowLayout<R, G>::RL;
        for_rows<R, RL, G>(t, g * RL, [&](int r, int x, float mu, const float (&J)[K]) { fn(r * R + x, mu, J); });
    }

    template <int R, int RL, int G, class Fn>
    PP_DEV void for_rows(const float (&t)[K], int y0, Fn&& fn) const {
        Frame F;
        frame(t, R, F);
#pragma unroll
        for (int r = 0; r < RL; r++) {
            const int y = y0 + r;
            if (y < R) {
                const float* rp = row_ptr(F, y);
#pragma unroll
                for (int x = 0; x < R; x++) {
                    const float* c = rp + 64 * col_index(F, x);
                    float f, fx, fy, fz, mu, J[K];
                    cell(c, F, f, fx, fy, fz);
                    finish(F, t, f, fx, fy, fz, x, y, R, mu, J);
                    fn(r, x, mu, J);
                }
            }
        }
    }
};

// ---------------------------------------------------------------------------
// Cubic-spline PSF evaluated from the samples it interpolates.  A table made by
// CSpline_Compute (CubicSplinePSF.cu:151-204) holds, per cell, the tricubic
// polynomial through a 4x4x4 block of samples: along each axis, cell c takes the
// samples grid[c..c+3] at the local nodes t = -1, 0, 1, 2.  With the Lagrange
// basis L_m of those nodes the cell's value at sub-voxel (w, v, u) is
//   f = sum_{m,n,q} grid[k+m][j+n][i+q] L_m(w) L_n(v) L_q(u),
// the same polynomial as the coefficient form sum_p c_p w^pz v^py u^px.  The
// grid is [Nz+3][Ny+3][Nx+3] floats (215 KB for a 61x26x26-cell table, which
// stays in L2) against 64 floats per cell (10.5 MB).  Per spot and iteration the
// lane group contracts the grid's four z-slices over the spot's window into an
// LDS table of (value, z-derivative) weight pairs (I / du folded in); each pixel
// then reads its 4x4 pairs.  Cell indices, the truncating sub-pixel offset
// and the psf >= 0 clamp follow CubicSplinePSFModels.h:64-153 as in SplineModelT.
template <int BG>
struct SplineGridModelT : SplineCommon<BG> {
    using Base = SplineCommon<BG>;
    using Base::K;
    using Base::Nz;
    using Base::Ny;
    using Base::Nx;
    using Base::zmin;
    using Base::zmax;
    using Base::iI;
    const float* grid;

    // Strip length of the strip layout for R x R spots at G lanes, or 0 for the flat
    // layout.  A pixel's value and derivatives contract a 4 x 4 block of the window
    // table: 64 FMAs.  Its x-contraction of one window row (12 FMAs for value,
    // x-derivative and z-derivative) is the same for the four pixels of a column that
    // use that row, so a lane walking down a strip of L pixels contracts L + 3 rows and
    // then 16 FMAs per pixel.  Model cost per lane (VALU instructions): strips
    // M (12 (L + 3) + 8) + NS (16 + 35), flat NS (64 + 35), 35 for a pixel's
    // normal-equation terms; strips are taken where they cost under 0.95 of flat and
    // the last piece's rows stay inside the R + 4 window (P L <= R + 1).
    static constexpr int strip_len(int R, int G) {
        const int flat = (R * R + G - 1) / G * (64 + 35);
        int best = 0, cost = flat * 95 / 100;
        for (int L = 3; L <= R; L++) {
            const int P = (R + L - 1) / L;
            if (P * L > R + 1) continue;
            const int M = (R * P + G - 1) / G;
            const int c = M * (12 * (L + 3) + 8) + M * L * (16 + 35);
            if (c < cost) { cost = c; best = L; }
        }
        return best;
    }
    template <int R, int G>
    using Layout = std::conditional_t<strip_len(R, G) != 0, StripLayout<R, G, strip_len(R, G) ? strip_len(R, G) : 1>,
                                      FlatLayout<R, G>>;
    // Four or five wavefronts per SIMD hide the L2 latency of the window loads: the
    // flat-background form with groups of >= 8 lanes and <= 14 pixels per lane gets the
    // register budget of as many wavefronts as its LDS tables allow (at most five, 96
    // VGPRs, for the flat layout; four, 128 VGPRs, for the strip layout, whose four rows
    // in flight spill under five, and where the tables of five would not fit in the
    // CU's 160 KiB).  Other sizes and the tilted / image backgrounds keep the
    // compiler's own choice.
    static constexpr int slots(int R, int G) {
        const int L = strip_len(R, G);
        return L ? (R * ((R + L - 1) / L) + G - 1) / G * L : (R * R + G - 1) / G;
    }
    static constexpr int lds_waves(int R, int G) { return 40960 / (4 * lds_floats(R, G) * (64 / G)); }
    // Zero floats behind the device sample table that for_pixels' window loads may
    // touch: a window of R + 4 samples at the last cell reaches R (Nx + 4) floats past
    // the table's end, and a lane that shares its column with NPL - 1 others loads up
    // to NPL - 1 <= 8 rows more (R <= kMaxWindowR).
    static constexpr int kMaxWindowR = 32;
    __host__ __device__ static size_t pad_floats(int Nx) { return (size_t)(kMaxWindowR + 9) * (size_t)(Nx + 4); }
    static constexpr int min_waves(int R, int G) {
        const int cap = strip_len(R, G) ? 4 : 5;
        const int w = lds_waves(R, G) < cap ? lds_waves(R, G) : cap;
        return (BG == kSplineFlat && G >= 8 && slots(R, G) <= 14 && w >= 4) ? w : 1;
    }
    // window of R + 4 grid samples per axis (see for_pixels) as (value, z-derivative)
    // pairs, and the per-column and per-row cell offsets (2R ints); an even count
    // keeps every group's pairs 8-byte aligned
    //
    // LDS banks: a ds_read_b64 is served 32 lanes at a time from 64 4-byte banks, so 32
    // lanes reading pixels p .. p + 31 of a window stored with row stride S are
    // conflict-free when the (value, z-derivative) entries y S + x are distinct modulo
    // 32, i.e. when S = R (mod 32).  The plain stride W = R + 4 breaks this at every row
    // end (measured: 2 conflict cycles per ds_read_b64, 4-8 per ds_read2_b64, for R = 20).
    // 32-lane groups therefore pair up: the two groups of a wavefront share one table
    // whose rows (stride S = R mod 32, S >= 2W) hold both windows side by side, when that
    // costs no more LDS than four wavefronts per SIMD allow (1280 floats per group);
    // their cell offsets are kept as 16-bit values.  Smaller groups keep the plain
    // layout.
    static constexpr int pair_stride(int R) {
        int S = 2 * (R + 4);
        while (S % 32 != R % 32) S++;
        return S;
    }
    static constexpr bool paired(int R, int G) { return G == 32 && (R + 4) * pair_stride(R) + R <= 1280; }
    static constexpr int lds_floats(int R, int G) {
        return paired(R, G) ? (R + 4) * pair_stride(R) + R : 2 * (R + 4) * (R + 4) + 2 * R;
    }

    // Lagrange weights on the nodes -1, 0, 1, 2 at t, and their derivatives.
    PP_DEV static void lagrange(float t, float (&w)[4], float (&dw)[4]) {
        const float t2 = t * t, t3 = t2 * t;
        w[0] = (-1.0f / 3.0f) * t + 0.5f * t2 - (1.0f / 6.0f) * t3;
        w[1] = 1.0f - 0.5f * t - t2 + 0.5f * t3;
        w[2] = t + 0.5f * t2 - 0.5f * t3;
        w[3] = (-1.0f / 6.0f) * t + (1.0f / 6.0f) * t3;
        dw[0] = (-1.0f / 3.0f) + t - 0.5f * t2;
        dw[1] = -0.5f - 2.0f * t + 1.5f * t2;
        dw[2] = 1.0f + t - 1.5f * t2;
        dw[3] = (-1.0f / 6.0f) + 0.5f * t2;
    }

    struct Frame {
        float tx, ty;          // position relative to the grid centre
        int k;                 // z-slab
        float wx[4], wdx[4];   // x weights; derivative weights times -I
        float wy[4], wdy[4];   // y weights; derivative weights times -I
        float wz[4], wdz[4];   // z weights; derivative weights times I / du
    };

    PP_DEV void frame(const float (&t)[K], int R, Frame& F) const {
        const float du = (zmax - zmin) / (float)Nz;
        F.tx = Base::rel(t[0], Nx, R);
        F.ty = Base::rel(t[1], Ny, R);
        float spx = -F.tx; spx -= (float)(int)spx;   // truncation toward zero (reference quirk)
        float spy = -F.ty; spy -= (float)(int)spy;
        float spz = (t[2] - zmin) / du;
        F.k = clampi((int)spz, 0, Nz - 1);
        spz = clampf(spz - (float)F.k, 0.0f, 1.0f);
        const float I = t[iI], mI = -I, Idu = I / du;
        lagrange(spx, F.wx, F.wdx);
        lagrange(spy, F.wy, F.wdy);
        lagrange(spz, F.wz, F.wdz);
#pragma unroll
        for (int m = 0; m < 4; m++) {
            F.wdx[m] *= mI;
            F.wdy[m] *= mI;
            F.wdz[m] *= Idu;
        }
    }

    PP_DEV int cell_x(const Frame& F, int x) const { return clampi((int)((float)x - F.tx), 0, Nx - 1); }
    PP_DEV int cell_y(const Frame& F, int y) const { return clampi((int)((float)y - F.ty), 0, Ny - 1); }

    PP_DEV static void finish(const float (&t)[K], float f, float fx, float fy, float fz, int x, int y, int R,
                              float& mu, float (&J)[K]) {
        J[0] = fx;  // -I folded into the x/y derivative weights, I / du into the z ones
        J[1] = fy;
        J[2] = fz;
        // max(f, 0) with NaN passed through, as f < 0 ? 0 : f, in one instruction
        // (v_maximum_f32); it gives +0 for f = -0, which no sum below can tell apart
        const float psf = __builtin_elementwise_maximum(f, 0.0f);
        J[3] = psf;
        Base::background(t, t[iI], psf, x, y, R, mu, J);
    }

    // Single-pixel evaluation straight from the grid (expected value / derivatives).
    PP_DEV void eval_pixel(const float (&t)[K], int x, int y, int R, float& mu, float (&J)[K]) const {
        Frame F;
        frame(t, R, F);
        const int i = cell_x(F, x), j = cell_y(F, y);
        const int Dy = Ny + 3, Dx = Nx + 3;
        float f = 0.0f, fx = 0.0f, fy = 0.0f, fz = 0.0f;
#pragma unroll
        for (int n = 0; n < 4; n++) {
            float a = 0.0f, ad = 0.0f, az = 0.0f;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                float e = 0.0f, ez = 0.0f;
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    const float v = grid[((size_t)(F.k + m) * Dy + (j + n)) * Dx + (i + q)];
                    e += v * F.wz[m];
                    ez += v * F.wdz[m];
                }
                a += F.wx[q] * e;
                ad += F.wdx[q] * e;
                az += F.wx[q] * ez;
            }
            f += F.wy[n] * a;
            fx += F.wy[n] * ad;
            fy += F.wdy[n] * a;
            fz += F.wy[n] * az;
        }
        finish(t, f, fx, fy, fz, x, y, R, mu, J);
    }

    // Lane g's pixels of the strip layout (StripLayout<R, G, LS>) from the group's
    // window table T (row stride S) on a spot whose pixel rows take consecutive window
    // rows: per strip, the x-contractions (value, x-derivative, z-derivative weights)
    // of its first three rows, then per pixel one more row and the 16-FMA y-contraction.
    // The four rows in flight rotate through fixed registers (the loops unroll).
    template <int R, int G, int LS, int S, class Off, class Fn>
    PP_DEV void strip_pixels(const float (&t)[K], const Frame& F, const float2* T, const Off* OX, Fn& fn,
                             int g) const {
        using SL = StripLayout<R, G, LS>;
#pragma unroll
        for (int j = 0; j < SL::M; j++) {
            int p = g + G * j;
            const bool live = SL::NP % G == 0 || SL::M * G == SL::NP || p < SL::NP;
            if (!live) p = SL::NP - 1;  // padding lanes walk a real strip and use nothing
            const int c = p % R, y0 = (p / R) * LS;
            // rows r .. r + 3 from one base (ds_read2_b64 reaches 255 x 8 bytes past its
            // address; four rows of stride S <= 52 stay within it), advanced every four
            // rows through an opaque copy, which the compiler would otherwise fold into
            // per-row address additions
            int ob = y0 * S + (int)OX[c];
            float a[4], ad[4], az[4];
            // row r's four (value, z-derivative) entries, read one pixel ahead of their
            // x-contraction so that the LDS latency overlaps the pixel before
            float2 e0, e1, e2, e3;
            auto load = [&](int r) {
                if (r > 0 && (r & 3) == 0) {
                    ob += 4 * S;
                    asm volatile("" : "+v"(ob));
                }
                const float2* e = T + ob + (r & 3) * S;
                e0 = e[0], e1 = e[1], e2 = e[2], e3 = e[3];
            };
            auto contract = [&](int r) {
                const int k = r & 3;
                a[k] = F.wx[0] * e0.x + F.wx[1] * e1.x + F.wx[2] * e2.x + F.wx[3] * e3.x;
                ad[k] = F.wdx[0] * e0.x + F.wdx[1] * e1.x + F.wdx[2] * e2.x + F.wdx[3] * e3.x;
                az[k] = F.wx[0] * e0.y + F.wx[1] * e1.y + F.wx[2] * e2.y + F.wx[3] * e3.y;
            };
#pragma unroll
            for (int r = 0; r < 3; r++) {
                load(r);
                contract(r);
            }
            load(3);
#pragma unroll
            for (int i = 0; i < LS; i++) {
                contract(i + 3);
                if (i + 1 < LS) load(i + 4);
                const int y = y0 + i;
                if (live && (SL::P * LS == R || y < R)) {
                    float f = 0.0f, fx = 0.0f, fy = 0.0f, fz = 0.0f;
#pragma unroll
                    for (int n = 0; n < 4; n++) {
                        const int k = (i + n) & 3;
                        f += F.wy[n] * a[k];
                        fx += F.wy[n] * ad[k];
                        fy += F.wdy[n] * a[k];
                        fz += F.wy[n] * az[k];
                    }
                    float mu, J[K];
                    finish(t, f, fx, fy, fz, c, y, R, mu, J);
                    fn(j * LS + i, mu, J);
                }
                // one pixel at a time (the scheduler would otherwise hoist later rows'
                // LDS reads and run out of VGPRs)
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }

    // Fit-kernel evaluation of lane g's pixels (Layout<R, G>), called by every
    // lane of the group with the group's LDS region (lds_floats(R, G) floats).
    template <int R, int G, class Fn>
    PP_DEV void for_pixels(const float (&t)[K], int g, float* lds, Fn&& fn) const {
        // Window width in grid samples.  A pixel's cell is (int)(x - tx) in FP32, which
        // rounding can push one past (int)(-tx) + x (e.g. -tx = 2.9999990: x = 19 gives
        // 19 + 2.9999990 = 22.0), so cells run from i0 to i0 + R and the 4x4 stencil to
        // i0 + R + 3: R + 4 samples.
        constexpr int W = R + 4;
        constexpr int NE = W * W;  // window entries
        static_assert(R <= kMaxWindowR, "window loads past the table's padding");
        using L = Layout<R, G>;
        constexpr int LS = strip_len(R, G);
        // Pixel coordinates depend only on g and the slot; without this opaque copy
        // the compiler hoists all of them out of the iteration loop and keeps them
        // in VGPRs for the whole fit.  (An empty asm: it emits no instruction.)
        asm volatile("" : "+v"(g));
        Frame F;
        frame(t, R, F);
        const int i0 = cell_x(F, 0), j0 = cell_y(F, 0);
        const int Dy = Ny + 3, Dx = Nx + 3;
        const size_t zs = (size_t)Dy * Dx;
        // (value, z-derivative) weight pairs: one 8-byte LDS read per stencil entry,
        // two entries per ds_read2_b64 at immediate offsets from one address; row
        // stride S (the paired layout: this group's window starts W entries into the
        // shared rows when it is the wavefront's second group)
        constexpr bool kPaired = paired(R, G);
        constexpr int S = kPaired ? pair_stride(R) : W;
        constexpr int PER = lds_floats(R, G);
        const int half = kPaired ? (int)((threadIdx.x / G) & 1) : 0;
        float* base = lds - half * PER;  // the pair's region (the group's own when unpaired)
        float2* T = reinterpret_cast<float2*>(base) + half * W;
        // cell offsets of the pixel columns and rows from the window origin, rows times S
        using Off = std::conditional_t<kPaired, unsigned short, int>;
        Off* OX = reinterpret_cast<Off*>(base + (kPaired ? 2 * (R + 4) * S : 2 * NE)) + half * 2 * R;
        Off* OY = OX + R;
#pragma unroll
        for (int u0 = 0; u0 < 2 * R; u0 += G) {
            const int u = u0 + g;
            if (u < R)
                OX[u] = (Off)(cell_x(F, u) - i0);
            else if (u < 2 * R)
                OY[u - R] = (Off)((cell_y(F, u - R) - j0) * S);
        }
        // z-contraction of the window: rows j0 .., columns i0 .. of slab k.  A lane takes
        // one window column (a group narrower than the window: a second one; a group
        // at least twice as wide: every NPL-th row of one) and walks its rows, so a
        // row's and a z-plane's grid offset is the same for every lane: buffer loads take
        // it as their scalar offset beside the lane's 32-bit byte offset, and an entry
        // costs its 8 FMAs (with per-entry addresses it cost about 25 VALU instructions:
        // 64-bit address arithmetic per load, the division by W, the clamps).  Nothing is
        // clamped: an entry past the grid's last sample column or row holds whatever
        // follows in memory (the next row, or the zero padding of pad_floats behind the
        // table) and is never read, since a pixel's cell is clamped to the grid and its
        // 4x4 stencil stays inside it; every read offset (cell - origin) * S + stencil
        // stays inside the group's W x W window.
        {
            constexpr int NPL = G >= W ? G / W : 1;             // lanes per column
            constexpr int NCOL = G >= W ? 1 : (W + G - 1) / G;  // columns per lane
            constexpr int NROW = (W + NPL - 1) / NPL;           // rows per lane
            constexpr int QB = 3;                               // rows per batch of loads
            constexpr int NB = (NROW + QB - 1) / QB;
            const int col = G >= W ? g % W : g, ph = G >= W ? g / W : 0;
            const bool on = G < W || g < NPL * W;
            const unsigned off = (unsigned)(((F.k * Dy + j0 + ph) * Dx + i0 + col) * 4);
            // the table and its padding (under 2^31 bytes, checked where it is made)
            const int nbytes = (int)((zs * (size_t)(Nz + 3) + pad_floats(Nx)) * 4);
            const __amdgpu_buffer_rsrc_t rs =
                __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(grid), (short)0, nbytes, 0x00020000);
            // opaque per call: the row and plane offsets are then formed on the scalar unit
            // beside their loads, not hoisted out of the fit loop into spilled SGPRs
            unsigned zb = (unsigned)zs * 4, rb = (unsigned)(NPL * Dx * 4);
            asm volatile("" : "+s"(zb), "+s"(rb));
            float2* Tl = T + ph * S + col;
#pragma unroll
            for (int j = 0; j < NCOL; j++) {
                if (on && (NCOL == 1 || col + G * j < W)) {
                    // rows in batches of QB, the next batch's loads in flight while one is
                    // contracted (a lane with NPL > 1 loads up to NPL - 1 rows past the
                    // window, inside the padding, and stores only the window's rows)
                    float v[2][QB][4];
                    auto load = [&](int b) {
#pragma unroll
                        for (int u = 0; u < QB; u++) {
                            const int q = b * QB + u;
                            if (q < NROW) {
#pragma unroll
                                for (int m = 0; m < 4; m++)
                                    v[b & 1][u][m] = __builtin_bit_cast(
                                        float, __builtin_amdgcn_raw_buffer_load_b32(
                                                   rs, off, (unsigned)(m * zb + q * rb + G * j * 4), 0));
                            }
                        }
                    };
                    load(0);
#pragma unroll
                    for (int b = 0; b < NB; b++) {
                        if (b + 1 < NB) load(b + 1);
#pragma unroll
                        for (int u = 0; u < QB; u++) {
                            const int q = b * QB + u;
                            if (q < NROW && (W % NPL == 0 || ph + q * NPL < W)) {
                                const float* e = v[b & 1][u];
                                Tl[q * NPL * S + G * j] =
                                    make_float2(e[0] * F.wz[0] + e[1] * F.wz[1] + e[2] * F.wz[2] + e[3] * F.wz[3],
                                                e[0] * F.wdz[0] + e[1] * F.wdz[1] + e[2] * F.wdz[2] + e[3] * F.wdz[3]);
                            }
                        }
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
            }
        }
        // Strip layout: the strips walk the window one row per pixel, which holds when
        // the pixel rows' cells are consecutive (cell_y(y) = j0 + y for every row y; not
        // where the grid edge clamps them or FP32 rounding skips a cell).  A group with
        // any other row pattern takes the per-pixel path below.
        bool strips = false;
        if constexpr (LS != 0) {
            bool ok = true;
#pragma unroll
            for (int u0 = 0; u0 < R; u0 += G) {
                const int u = u0 + g;
                if (u < R && cell_y(F, u) - j0 != u) ok = false;
            }
            const unsigned long long bad = __ballot(!ok);
            const int first = (int)(threadIdx.x & (kWave - 1)) & ~(G - 1);
            const unsigned long long mine = G == 64 ? ~0ull : (((1ull << G) - 1ull) << first);
            strips = (bad & mine) == 0;
        }
        // the group is inside one wavefront: its writes must have completed before any
        // lane reads the tables (workgroup-scope fences wait for the wave's LDS
        // operations; no workgroup barrier is needed)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if constexpr (LS != 0) {
            if (strips) {
                strip_pixels<R, G, LS, S>(t, F, T, OX, fn, g);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                return;
            }
        }
#pragma unroll
        for (int s = 0; s < L::NS; s++) {
            const int p = L::index(g, s);
            if (L::NS * G == R * R || p < R * R) {
                // p / R as a 20-bit multiply-shift, exact for p < 2048 and R <= 32
                constexpr int kDivR = (1 << 20) / R + 1;
                const int y = (p * kDivR) >> 20, x = p - y * R;
                const int o = (int)OY[y] + (int)OX[x];
                float f = 0.0f, fx = 0.0f, fy = 0.0f, fz = 0.0f;
#pragma unroll
                for (int n = 0; n < 4; n++) {
                    float a = 0.0f, ad = 0.0f, az = 0.0f;
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const float2 v = T[o + n * S + q];
                        const float e = v.x, ez = v.y;
                        a += F.wx[q] * e;
                        ad += F.wdx[q] * e;
                        az += F.wx[q] * ez;
                    }
                    f += F.wy[n] * a;
                    fx += F.wy[n] * ad;
                    fy += F.wdy[n] * a;
                    fz += F.wy[n] * az;
                }
                float mu, J[K];
                finish(t, f, fx, fy, fz, x, y, R, mu, J);
                // the pixel's value and derivatives are formed here: without this (an
                // empty asm, no instruction) the compiler sinks every pixel's arithmetic
                // below the LDS reads of all the lane's pixels, which at 28 or more slots
                // held 16 x 28 pairs at once (20x20 at 16 lanes: 512 registers and 1.4 KB
                // of scratch per lane; 191 registers with it)
                asm volatile("" : "+v"(f), "+v"(fx), "+v"(fy), "+v"(fz) :: "memory");
                fn(s, mu, J);
            }
            // one pixel at a time: without this the scheduler hoists the LDS reads of
            // many pixels and runs out of VGPRs
            __builtin_amdgcn_sched_barrier(0);
        }
        // the next call overwrites the tables: every lane's reads complete first
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
};

using SplineModel = SplineModelT<kSplineFlat>;
using SplineGridModel = SplineGridModelT<kSplineFlat>;

// The same spline table under another background mode (host side).
template <int BG>
inline SplineModelT<BG> spline_as(const SplineModel& s) {
    SplineModelT<BG> d{};
    d.Nz = s.Nz; d.Ny = s.Ny; d.Nx = s.Nx;
    d.zmin = s.zmin; d.zmax = s.zmax;
    d.coefs = s.coefs;
    return d;
}

template <int BG>
inline SplineGridModelT<BG> spline_as(const SplineGridModel& s) {
    SplineGridModelT<BG> d{};
    d.Nz = s.Nz; d.Ny = s.Ny; d.Nx = s.Nx;
    d.zmin = s.zmin; d.zmax = s.zmax;
    d.grid = s.grid;
    return d;
}

// ---------------------------------------------------------------------------
// Intensity and background of a spot at a known position and width
// (IntensityBgModel, GaussianPSFModels.h:557-623, fitted by
// Gauss2D_EstimateIntensityBg, GaussianPSF.cu:216-240): mu = I * DX * DY + bg with
// the pixel-integrated Gaussian of width sigma at (x, y), the three per-spot
// constants.  Start {I, bg} = {1, 0}; stop limits {1e-2, 1e-4}; I >= 1e-5,
// bg >= 1e-8; the CRLB takes |diag(FI^-1)| (ComputeCRLB, Estimation.h:318-333).
struct IBgModel {
    static constexpr int K = 2, NC = 3, KT = 5;
    static constexpr bool kPixelBg = false;
    static constexpr int iI = 0, iBg = 1;
    static constexpr bool kCrlbAbs = true;

    template <int R, int G>
    using Layout = RowLayout<R, G>;
    static constexpr int lds_floats(int, int) { return 0; }

    PP_DEV static float stop_limit(int k) { return k == 0 ? 1e-2f : 1e-4f; }

    PP_DEV void check_limits(float (&t)[KT], int) const {
        if (t[0] < 1e-5f) t[0] = 1e-5f;
        if (t[1] < 1e-8f) t[1] = 1e-8f;
    }

    PP_DEV void initial(float, float, float, float (&t)[KT]) const {
        t[0] = 1.0f;
        t[1] = 0.0f;
    }

    PP_DEV static float delta(float px, float pos, float c2) {
        return 0.5f * erff(((px - pos) + 0.5f) * c2) - 0.5f * erff(((px - pos) - 0.5f) * c2);
    }

    PP_DEV void eval_pixel(const float (&t)[KT], int x, int y, int, float& mu, float (&J)[K]) const {
        const float c2 = 1.0f / (sqrtf(2.0f) * t[4]);
        const float p = delta((float)x, t[2], c2) * delta((float)y, t[3], c2);
        mu = t[0] * p + t[1];
        J[0] = p;
        J[1] = 1.0f;
    }

    template <int R, int G, class Fn>
    PP_DEV void for_pixels(const float (&t)[KT], int g, float*, Fn&& fn) const {
        constexpr int RL = RowLayout<R, G>::RL;
        const float c2 = 1.0f / (sqrtf(2.0f) * t[4]);
        float dx[R];
#pragma unroll
        for (int x = 0; x < R; x++) dx[x] = delta((float)x, t[2], c2);
#pragma unroll
        for (int r = 0; r < RL; r++) {
            const int y = g * RL + r;
            if (y < R) {
                const float dy = delta((float)y, t[3], c2);
#pragma unroll
                for (int x = 0; x < R; x++) {
                    const float p = dx[x] * dy;
                    const float J[K] = {p, 1.0f};
                    fn(r * R + x, t[0] * p + t[1], J);
                }
            }
        }
    }
};

}  // namespace ppa
