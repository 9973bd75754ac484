// Poisson negative log-likelihood and its adjoint (no reference counterpart).
//
// NLL of a spot: sum over pixels of mu' - x' log mu', with mu' = max(mu, 1e-8)
// and x' = max(x, 1e-8) (the convention of ComputeLogLikelihood,
// Estimation.h:138-156, negated).  The backward pass recomputes mu and the
// Jacobian from the parameters (nothing is stored by the forward pass) and
// returns TRUE derivatives:
//   d NLL / d theta_k  = sum_px g * (1 - x'/mu') * dmu/dtheta_k
//   d NLL / d shared   = sum_spots sum_px g * (1 - x'/mu') * dmu/dshared
// where g is the upstream gradient of the spot's NLL.  Shared parameters:
//   fixed sigma   {sigma_x, sigma_y}
//   astigmatic    {s0, gamma, d, A, B} for x then y (sigma(z) = s0 sqrt(1 + (t^2 + A t^3 + B t^4)/d^2), t = z - gamma)
//   spline        the coefficient table [Nz,Ny,Nx,64]
// Unlike the reference's Derivatives output, the free-sigma column is the true
// derivative and clamped spline pixels (negative spline value) contribute no
// position derivative.
//
// The Gaussian adjoint evaluates the model in FP64 (erf/exp of double); the
// spline adjoint in FP32 (its value has no erf-difference cancellation).
//
// Mapping: one wavefront per spot, lanes stride over pixels.  Scalar shared
// gradients are reduced per block in LDS and added to HBM once per block; the
// spline coefficient gradient is scattered pixel by pixel with the 64 lanes
// covering the 64 coefficients of the pixel's cell, so every atomic
// instruction updates one contiguous 256-byte cell.
#include "dispatch.h"

namespace ppa {

constexpr int kMaxAdjSlots = 16;  // pixels per lane: R <= 32

PP_DEV float nll_weight(float mu, float x, float& nll) {
    const float m = fmaxf(mu, 1e-8f), xv = fmaxf(x, 1e-8f);
    nll += m - xv * logf(m);
    return mu > 1e-8f ? 1.0f - xv / m : 0.0f;
}

// Gaussian pixel with true derivatives, evaluated in FP64: the gradient is a
// sum of terms of both signs, and FP32 erf differences in the PSF tails would
// otherwise limit its relative accuracy.  dS[0], dS[1]: dmu/dsigma_x, dmu/dsigma_y.
struct EdgeD {
    double E, G, H;
};

PP_DEV EdgeD gauss_edge_d(double e, double t, double c2) {
    const double a = (e - t - 0.5) * c2;
    const double g = exp(-a * a);
    return {erf(a), g, a * g};
}

template <int V>
PP_DEV void gauss_true(const GaussModel<V>& M, const float (&tf)[GaussModel<V>::K], int x, int y, double& mu,
                       double (&J)[GaussModel<V>::K], double (&dS)[2]) {
    using Mod = GaussModel<V>;
    constexpr int K = Mod::K;
    double t[K];
#pragma unroll
    for (int k = 0; k < K; k++) t[k] = (double)tf[k];
    double sgx, sgy, dsx = 0.0, dsy = 0.0;
    if (V == kFixed) {
        sgx = M.c.sx;
        sgy = M.c.sy;
    } else if (V == kSigma) {
        sgx = sgy = t[4];
    } else {
        auto sig = [&](const float* a, double& s, double& ds) {
            const double s0 = a[0], gam = a[1], d = a[2], A = a[3], B = a[4], d2 = d * d;
            const double u = t[2] - gam, u2 = u * u, u3 = u2 * u, u4 = u3 * u;
            const double sq = sqrt(1.0 + u2 / d2 + A * u3 / d2 + B * u4 / d2);
            s = s0 * sq;
            ds = s0 * (2 * u / d2 + A * 3 * u2 / d2 + B * 4 * u3 / d2) / (2 * sq);
        };
        sig(M.c.ax, sgx, dsx);
        sig(M.c.ay, sgy, dsy);
    }
    const double r2 = 1.4142135623730951, r2pi = 2.5066282746310002, ispi = 0.5641895835477563;
    const double c2x = 1.0 / (r2 * sgx), c2y = 1.0 / (r2 * sgy);
    const EdgeD x0 = gauss_edge_d(x, t[0], c2x), x1 = gauss_edge_d(x + 1, t[0], c2x);
    const EdgeD y0 = gauss_edge_d(y, t[1], c2y), y1 = gauss_edge_d(y + 1, t[1], c2y);
    const double DX = 0.5 * (x1.E - x0.E), DY = 0.5 * (y1.E - y0.E);
    const double dEX = (x0.G - x1.G) / (r2pi * sgx), dEY = (y0.G - y1.G) / (r2pi * sgy);
    const double SX = (x0.H - x1.H) * ispi / sgx, SY = (y0.H - y1.H) * ispi / sgy;  // d(Delta)/d(sigma)
    const double I = t[Mod::iI], bg = t[Mod::iBg];
    const double p = DX * DY;
    mu = bg + I * p;
    J[0] = I * dEX * DY;
    J[1] = I * dEY * DX;
    J[Mod::iI] = p;
    J[Mod::iBg] = 1.0;
    dS[0] = I * SX * DY;
    dS[1] = I * SY * DX;
    if (V == kSigma) J[4] = dS[0] + dS[1];
    if (V == kAstig) J[2] = dS[0] * dsx + dS[1] * dsy;
}

PP_DEV double nll_weight_d(double mu, float x, double& nll) {
    const double m = fmax(mu, 1e-8), xv = fmax((double)x, 1e-8);
    nll += m - xv * log(m);
    return mu > 1e-8 ? 1.0 - xv / m : 0.0;
}

// d sigma / d {s0, gamma, d, A, B} of one astigmatic axis
PP_DEV void astig_sigma_grads(const float* a, double z, double (&g)[5]) {
    const double s0 = a[0], gam = a[1], d = a[2], A = a[3], B = a[4];
    const double d2 = d * d;
    const double t = z - gam, t2 = t * t, t3 = t2 * t, t4 = t3 * t;
    const double poly = t2 + A * t3 + B * t4;
    const double sq = sqrt(1.0 + poly / d2);
    g[0] = sq;
    g[1] = -s0 * (2 * t + 3 * A * t2 + 4 * B * t3) / (2 * d2 * sq);
    g[2] = -s0 * poly / (d2 * d * sq);
    g[3] = s0 * t3 / (2 * d2 * sq);
    g[4] = s0 * t4 / (2 * d2 * sq);
}

template <int V>
__global__ void __launch_bounds__(256) gauss_adjoint_kernel(GaussModel<V> M, AdjArgs a) {
    using Mod = GaussModel<V>;
    constexpr int K = Mod::K;
    constexpr int NS = (V == kFixed) ? 2 : (V == kAstig ? 10 : 0);
    __shared__ float blockShared[NS > 0 ? NS : 1];
    if (NS > 0 && threadIdx.x < NS) blockShared[threadIdx.x] = 0.0f;
    if (V == kFixed) { M.c.sx = a.shared[0]; M.c.sy = a.shared[1]; }
    if (V == kAstig) {
#pragma unroll
        for (int i = 0; i < 5; i++) { M.c.ax[i] = a.shared[i]; M.c.ay[i] = a.shared[5 + i]; }
    }
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int spot = (int)(((size_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    if (spot < a.n) {
        const int R = a.R, S = R * R;
        float t[K];
#pragma unroll
        for (int k = 0; k < K; k++) t[k] = a.params[(size_t)spot * K + k];
        double acc[K + 3];
#pragma unroll
        for (int i = 0; i < K + 3; i++) acc[i] = 0.0;
        for (int p = lane; p < S; p += kWave) {
            const int y = p / R, x = p - (p / R) * R;
            double mu, J[K], dS[2];
            gauss_true<V>(M, t, x, y, mu, J, dS);
            const double w = nll_weight_d(mu, a.sample[(size_t)spot * S + p], acc[K]);
#pragma unroll
            for (int k = 0; k < K; k++) acc[k] += w * J[k];
            acc[K + 1] += w * dS[0];
            acc[K + 2] += w * dS[1];
        }
#pragma unroll
        for (int i = 0; i < K + 3; i++) {
#pragma unroll
            for (int m = 1; m < kWave; m <<= 1) acc[i] += __shfl_xor(acc[i], m, kWave);
        }
        if (lane == 0) {
            const double go = a.gout ? a.gout[spot] : 1.0f;
            if (a.nll) a.nll[spot] = (float)acc[K];
            if (a.gparams) {
#pragma unroll
                for (int k = 0; k < K; k++) a.gparams[(size_t)spot * K + k] = (float)(go * acc[k]);
            }
            if (a.gshared && NS > 0) {
                if (V == kFixed) {
                    atomicAdd(&blockShared[0], (float)(go * acc[K + 1]));
                    atomicAdd(&blockShared[1], (float)(go * acc[K + 2]));
                } else if (V == kAstig) {
                    double gx[5], gy[5];
                    astig_sigma_grads(M.c.ax, t[2], gx);
                    astig_sigma_grads(M.c.ay, t[2], gy);
#pragma unroll
                    for (int i = 0; i < 5; i++) {
                        atomicAdd(&blockShared[i], (float)(go * acc[K + 1] * gx[i]));
                        atomicAdd(&blockShared[5 + i], (float)(go * acc[K + 2] * gy[i]));
                    }
                }
            }
        }
    }
    if (NS > 0 && a.gshared) {
        __syncthreads();
        if (threadIdx.x < NS) atomicAdd(&a.gshared[threadIdx.x], blockShared[threadIdx.x]);
    }
}

__global__ void __launch_bounds__(256) spline_adjoint_kernel(SplineModel M, AdjArgs a) {
    constexpr int K = 5;
    const int lane = threadIdx.x & (kWave - 1);
    const int spot = (int)(((size_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    if (spot >= a.n) return;
    M.coefs = a.shared;
    const int R = a.R, S = R * R;
    float t[K];
#pragma unroll
    for (int k = 0; k < K; k++) t[k] = a.params[(size_t)spot * K + k];
    SplineModel::Frame F;
    M.frame(t, R, F);
    const float go = a.gout ? a.gout[spot] : 1.0f;
    float acc[K + 1];
#pragma unroll
    for (int i = 0; i < K + 1; i++) acc[i] = 0.0f;
    int cellOf[kMaxAdjSlots];
    float wI[kMaxAdjSlots];
#pragma unroll
    for (int s = 0; s < kMaxAdjSlots; s++) {
        const int p = lane + s * kWave;
        cellOf[s] = 0;
        wI[s] = 0.0f;
        if (p < S) {
            const int y = p / R, x = p - (p / R) * R;
            const float* rp = M.row_ptr(F, y);
            const int i = M.col_index(F, x);
            const float* c = rp + 64 * i;
            float f, fx, fy, fz, mu, J[K];
            SplineModel::cell(c, F, f, fx, fy, fz);
            SplineModel::finish(F, t, f, fx, fy, fz, x, y, R, mu, J);
            if (f < 0.0f) J[0] = J[1] = J[2] = 0.0f;  // mu does not depend on position where clamped
            const float w = nll_weight(mu, a.sample[(size_t)spot * S + p], acc[K]);
#pragma unroll
            for (int k = 0; k < K; k++) acc[k] += w * J[k];
            cellOf[s] = (int)((c - M.coefs) >> 6);
            wI[s] = (f > 0.0f) ? go * w * t[3] : 0.0f;
        }
    }
    group_sum_n<kWave>(acc);
    if (lane == 0) {
        if (a.nll) a.nll[spot] = acc[K];
        if (a.gparams) {
#pragma unroll
            for (int k = 0; k < K; k++) a.gparams[(size_t)spot * K + k] = go * acc[k];
        }
    }
    if (!a.gshared) return;
    // lane q owns coefficient q = 16m + 4n + o of every cell: basis z^m y^n x^o
    const int m = lane >> 4, nn = (lane >> 2) & 3, o = lane & 3;
    const float xo[4] = {1.0f, F.x1, F.x2, F.x3};
    const float basis = F.zp[m] * F.yp[nn] * xo[o];
#pragma unroll
    for (int s = 0; s < kMaxAdjSlots; s++) {
        if (s * kWave >= S) break;
        for (int src = 0; src < kWave; src++) {
            if (s * kWave + src >= S) break;
            const float g = __shfl(wI[s], src, kWave);
            if (g != 0.0f) {
                const int cell = __shfl(cellOf[s], src, kWave);
                atomicAdd(&a.gshared[(size_t)cell * 64 + lane], g * basis);
            }
        }
    }
}

hipError_t launch_adjoint_any(const ModelDesc& d, const AdjArgs& a, hipStream_t st) {
    if (a.R > 32) return hipErrorInvalidValue;
    const unsigned blocks = (unsigned)(((size_t)a.n * kWave + 255) / 256);
    switch (d.kind) {
    case kKindFixed: hipLaunchKernelGGL(gauss_adjoint_kernel<kFixed>, dim3(blocks), dim3(256), 0, st, GaussModel<kFixed>{d.gc}, a); break;
    case kKindSigma: hipLaunchKernelGGL(gauss_adjoint_kernel<kSigma>, dim3(blocks), dim3(256), 0, st, GaussModel<kSigma>{d.gc}, a); break;
    case kKindAstig: hipLaunchKernelGGL(gauss_adjoint_kernel<kAstig>, dim3(blocks), dim3(256), 0, st, GaussModel<kAstig>{d.gc}, a); break;
    case kKindSpline: hipLaunchKernelGGL(spline_adjoint_kernel, dim3(blocks), dim3(256), 0, st, d.sm, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace ppa
