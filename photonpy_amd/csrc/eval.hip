// Launchers of the any-size kernels: expected value / derivatives, stand-alone
// CRLB + chi^2, and the spline coefficient solve.
#include <algorithm>

#include "dispatch.h"
#include "fit_sizes.h"

namespace ppa {

template <class M>
static hipError_t ev_t(const M& m, int R, int n, const float* params, const float* constants, float* ev, float* deriv,
                       hipStream_t st) {
    const size_t total = (size_t)n * R * R;
    const unsigned blocks = (unsigned)((total + 255) / 256);
    hipLaunchKernelGGL(ev_kernel<M>, dim3(blocks), dim3(256), 0, st, m, n, R, params, constants, ev, deriv);
    return hipGetLastError();
}

template <class M>
static hipError_t crlb_t(const M& m, int R, int n, const float* params, const float* constants, const float* sample,
                         float* crlb, float* chisq, hipStream_t st) {
    const unsigned blocks = (unsigned)(((size_t)n * kWave + 255) / 256);
    hipLaunchKernelGGL(crlb_kernel<M>, dim3(blocks), dim3(256), 0, st, m, n, R, params, constants, sample, crlb,
                       chisq);
    return hipGetLastError();
}

template <class M>
static hipError_t glrt_t(const M& m, int R, int n, const float* params, const float* constants, const float* sample,
                         int meanOfSpot0, float* diag, hipStream_t st) {
    const unsigned blocks = (unsigned)(((size_t)n * kWave + 255) / 256);
    hipLaunchKernelGGL(glrt_kernel<M>, dim3(blocks), dim3(256), 0, st, m, n, R, params, constants, sample, meanOfSpot0,
                       diag);
    return hipGetLastError();
}

// Calls f(model) with the model object of a descriptor.
template <class F>
static hipError_t with_model(const ModelDesc& d, F&& f) {
    switch (d.kind) {
    case kKindFixed: return f(GaussModel<kFixed>{d.gc});
    case kKindSigma: return f(GaussModel<kSigma>{d.gc});
    case kKindAstig: return f(GaussModel<kAstig>{d.gc});
    case kKindSpline: return d.useGrid ? f(d.sg) : f(d.sm);
    case kKindConstSigma: return f(GaussModel<kConstSigma>{d.gc});
    case kKindTiltedBg: return f(GaussModel<kTiltedBg>{d.gc});
    case kKindSigmaXY: return f(GaussModel<kSigmaXY>{d.gc});
    case kKindTiltedBgSigmaXY: return f(GaussModel<kTiltedBgSigmaXY>{d.gc});
    case kKindIBg: return f(IBgModel{});
    case kKindSplineTilted:
        return d.useGrid ? f(spline_as<kSplineTilted>(d.sg)) : f(spline_as<kSplineTilted>(d.sm));
    case kKindSplineImage:
        return d.useGrid ? f(spline_as<kSplineImage>(d.sg)) : f(spline_as<kSplineImage>(d.sm));
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ev_any(const ModelDesc& d, int R, int n, const float* params, const float* constants, float* ev,
                         float* deriv, hipStream_t st) {
    return with_model(d, [&](const auto& m) { return ev_t(m, R, n, params, constants, ev, deriv, st); });
}

hipError_t launch_crlb_any(const ModelDesc& d, int R, int n, const float* params, const float* constants,
                           const float* sample, float* crlb, float* chisq, hipStream_t st) {
    return with_model(d, [&](const auto& m) { return crlb_t(m, R, n, params, constants, sample, crlb, chisq, st); });
}

template <class M>
static hipError_t fit_large_t(const M& m, int R, const FitArgs& a, hipEvent_t* events, hipStream_t st) {
    static int resident = 0;  // per model type; one device model (gfx950)
    if (resident == 0) resident = resident_blocks(reinterpret_cast<const void*>(fit_large_kernel<M>), 256, 0);
    if (resident <= 0) return hipErrorNoDevice;
    const long need = ((long)a.n * kWave + 255) / 256;
    const unsigned blocks = (unsigned)std::max(1L, std::min(need, (long)resident));
    if (events) (void)hipEventRecord(events[0], st);
    hipLaunchKernelGGL(fit_large_kernel<M>, dim3(blocks), dim3(256), 0, st, m, a, R);
    if (events) (void)hipEventRecord(events[1], st);
    if (a.crlb || a.chisq) {
        const unsigned cb = (unsigned)(((size_t)a.n * kWave + 255) / 256);
        hipLaunchKernelGGL(crlb_kernel<M>, dim3(cb), dim3(256), 0, st, m, a.n, R, a.params, a.constants, a.sample,
                           a.crlb, a.chisq);
    }
    if (events) (void)hipEventRecord(events[2], st);
    return hipGetLastError();
}

hipError_t launch_fit_large_any(const ModelDesc& d, int R, const FitArgs& a, hipEvent_t* events, hipStream_t st) {
    return with_model(d, [&](const auto& m) { return fit_large_t(m, R, a, events, st); });
}

bool fit_compiled(const ModelDesc& d, int R, int G) {
#define PPA_HAS(R_, G_) \
    if (R == R_ && G == G_) return true;
    const bool variant = !(d.kind == kKindFixed || d.kind == kKindSigma || d.kind == kKindAstig || d.kind == kKindSpline);
    if (is_spline(d.kind) && d.useGrid) {
        if (variant) {
            PPA_GRID_VAR_SIZES_0(PPA_HAS) PPA_GRID_VAR_SIZES_1(PPA_HAS) PPA_GRID_VAR_SIZES_2(PPA_HAS)
            PPA_GRID_VAR_SIZES_3(PPA_HAS)
        } else {
            PPA_GRID_SIZES_0(PPA_HAS) PPA_GRID_SIZES_1(PPA_HAS) PPA_GRID_SIZES_2(PPA_HAS) PPA_GRID_SIZES_3(PPA_HAS)
        }
    } else if (variant) {
        PPA_VAR_SIZES_0(PPA_HAS) PPA_VAR_SIZES_1(PPA_HAS) PPA_VAR_SIZES_2(PPA_HAS) PPA_VAR_SIZES_3(PPA_HAS)
    } else {
        PPA_FIT_SIZES_0(PPA_HAS) PPA_FIT_SIZES_1(PPA_HAS) PPA_FIT_SIZES_2(PPA_HAS) PPA_FIT_SIZES_3(PPA_HAS)
    }
#undef PPA_HAS
    return false;
}

hipError_t launch_glrt_any(const ModelDesc& d, int R, int n, const float* params, const float* constants,
                           const float* sample, int meanOfSpot0, float* diag, hipStream_t st) {
    return with_model(d, [&](const auto& m) {
        return glrt_t(m, R, n, params, constants, sample, meanOfSpot0, diag, st);
    });
}

// ---------------------------------------------------------------------------
// Spline coefficients (CSpline_Compute, CubicSplinePSF.cu:151-204).  Cell
// (i,j,k) interpolates the 4x4x4 data neighbourhood starting at it, at offsets
// t = -1..2 per axis.  The reference solves the 64x64 tensor-product
// Vandermonde system A = V (x) V (x) V with V[t][p] = (t-1)^p by LU; here the
// exact inverse Vi of the 4x4 factor is applied along each axis in FP64.
__constant__ double kVinv[16] = {
    0.0,        1.0,  0.0,        0.0,         // c0 = f(0)
    -1.0 / 3.0, -0.5, 1.0,        -1.0 / 6.0,  // c1
    0.5,        -1.0, 0.5,        0.0,         // c2
    -1.0 / 6.0, 0.5,  -0.5,       1.0 / 6.0};  // c3

__global__ void __launch_bounds__(256) spline_coef_kernel(int s0, int s1, int s2, const float* __restrict__ data,
                                                          float* __restrict__ out, int o0, int o1, int o2, int border) {
    const size_t cell = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= (size_t)o0 * o1 * o2) return;
    const int k = (int)(cell % o2);
    const int j = (int)((cell / o2) % o1);
    const int i = (int)(cell / ((size_t)o1 * o2));
    double b[64];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int z = clampi(i + m - border, 0, s0 - 1);
                const int y = clampi(j + n - border, 0, s1 - 1);
                const int x = clampi(k + q - border, 0, s2 - 1);
                b[m * 16 + n * 4 + q] = (double)data[((size_t)z * s1 + y) * s2 + x];
            }
    double t1[64], t2[64];
    // along x (last index)
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int p = 0; p < 4; p++) {
                double s = 0.0;
#pragma unroll
                for (int q = 0; q < 4; q++) s += kVinv[p * 4 + q] * b[m * 16 + n * 4 + q];
                t1[m * 16 + n * 4 + p] = s;
            }
    // along y
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                double s = 0.0;
#pragma unroll
                for (int n = 0; n < 4; n++) s += kVinv[p * 4 + n] * t1[m * 16 + n * 4 + q];
                t2[m * 16 + p * 4 + q] = s;
            }
    // along z
    float* dst = out + cell * 64;
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                double s = 0.0;
#pragma unroll
                for (int m = 0; m < 4; m++) s += kVinv[p * 4 + m] * t2[m * 16 + n * 4 + q];
                dst[p * 16 + n * 4 + q] = (float)s;
            }
}

// CSpline_Evaluate (CubicSplinePSF.cu:208-271): value and x/y/z derivatives of the
// spline at every grid voxel (z, y, x) of [border, shape - border), shifted by
// `shift` (z, y, x).  The sub-voxel offset d = -shift - trunc(-shift) is the same
// for every voxel; voxel (z, y, x) uses cell (clamp(trunc(z - shift_z)), ...).
__global__ void __launch_bounds__(256) spline_evaluate_kernel(int s0, int s1, int s2, const float* __restrict__ coefs,
                                                              float h0, float h1, float h2, int border,
                                                              float* __restrict__ out) {
    const int o0 = s0 - 2 * border, o1 = s1 - 2 * border, o2 = s2 - 2 * border;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)o0 * o1 * o2) return;
    const int x = border + (int)(idx % o2);
    const int y = border + (int)((idx / o2) % o1);
    const int z = border + (int)(idx / ((size_t)o1 * o2));
    float dz = -h0; dz -= (float)(int)dz;
    float dy = -h1; dy -= (float)(int)dy;
    float dx = -h2; dx -= (float)(int)dx;
    const int k = clampi((int)((float)z - h0), 0, s0 - 1);
    const int j = clampi((int)((float)y - h1), 0, s1 - 1);
    const int i = clampi((int)((float)x - h2), 0, s2 - 1);
    const float* c = coefs + 64 * (((size_t)k * s1 + j) * s2 + i);
    const float zp[4] = {1.0f, dz, dz * dz, dz * dz * dz}, yp[4] = {1.0f, dy, dy * dy, dy * dy * dy};
    const float xp[4] = {1.0f, dx, dx * dx, dx * dx * dx};
    float v = 0.0f, vx = 0.0f, vy = 0.0f, vz = 0.0f;
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int o = 0; o < 4; o++) {
                const float q = c[m * 16 + n * 4 + o];
                v += zp[m] * yp[n] * xp[o] * q;
                if (o > 0) vx += (float)o * zp[m] * yp[n] * xp[o - 1] * q;
                if (n > 0) vy += (float)n * zp[m] * yp[n - 1] * xp[o] * q;
                if (m > 0) vz += (float)m * zp[m - 1] * yp[n] * xp[o] * q;
            }
    float* dst = out + 4 * idx;
    dst[0] = v;
    dst[1] = vx;
    dst[2] = vy;
    dst[3] = vz;
}

hipError_t launch_spline_evaluate(const int* shape, const float* d_coefs, const float* shift, int border,
                                  float* d_out, hipStream_t st) {
    const int o0 = shape[0] - 2 * border, o1 = shape[1] - 2 * border, o2 = shape[2] - 2 * border;
    if (o0 <= 0 || o1 <= 0 || o2 <= 0) return hipErrorInvalidValue;
    const size_t n = (size_t)o0 * o1 * o2;
    hipLaunchKernelGGL(spline_evaluate_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, shape[0],
                       shape[1], shape[2], d_coefs, shift[0], shift[1], shift[2], border, d_out);
    return hipGetLastError();
}

// counters[s] = the spots of segment s (of 2^shift) that the blocks starting there
// take first: blocks b < blocks with b % 2^shift == s, groupsPerBlock spots each
__global__ void set_counters_kernel(int* counters, int blocks, int groupsPerBlock, int shift) {
    const int s = (int)threadIdx.x;
    if (s < (1 << shift))
        counters[s * kCounterStride] = (s < blocks ? ((blocks - 1 - s) >> shift) + 1 : 0) * groupsPerBlock;
}

hipError_t launch_set_counters(int* counters, int blocks, int groupsPerBlock, int shift, hipStream_t st) {
    hipLaunchKernelGGL(set_counters_kernel, dim3(1), dim3(kWave), 0, st, counters, blocks, groupsPerBlock, shift);
    return hipGetLastError();
}

hipError_t launch_spline_coefs(const int* shape, const float* d_data, float* d_out, int border, hipStream_t st) {
    const int o0 = shape[0] - 3 + 2 * border, o1 = shape[1] - 3 + 2 * border, o2 = shape[2] - 3 + 2 * border;
    if (o0 <= 0 || o1 <= 0 || o2 <= 0) return hipErrorInvalidValue;
    const size_t cells = (size_t)o0 * o1 * o2;
    hipLaunchKernelGGL(spline_coef_kernel, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, st, shape[0], shape[1],
                       shape[2], d_data, d_out, o0, o1, o2, border);
    return hipGetLastError();
}

}  // namespace ppa
