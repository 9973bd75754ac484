// Expected value / derivatives, stand-alone CRLB + chi^2, and the fit of ROI sizes
// outside the compiled fit kernels (any ROI size).
//
// ev_kernel: one lane per (spot, pixel); writes ev[n,R,R] and deriv[n,K,R,R]
//   (layout of ComputeDerivatives, Estimation.h:169-181).
// crlb_kernel: one wavefront per spot; Fisher matrix FI_ij = sum J_i J_j / max(mu,1e-8)
//   and chi^2 = sum (x-mu)^2/mu (Estimation.h:63-102), FP64 inversion.
// Models with a background image (kPixelBg) read it from constants [n, R*R].
#pragma once

#include "fit_kernel.h"

namespace ppa {

template <class M>
__global__ void __launch_bounds__(256) ev_kernel(M model, int n, int R, const float* __restrict__ params,
                                                 const float* __restrict__ constants, float* __restrict__ ev,
                                                 float* __restrict__ deriv) {
    constexpr int K = M::K, KT = M::KT;
    const int S = R * R;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)n * S) return;
    const int spot = (int)(idx / S);
    const int p = (int)(idx - (size_t)spot * S);
    const int y = p / R, x = p - (p / R) * R;
    float th[KT];
#pragma unroll
    for (int k = 0; k < K; k++) th[k] = params[(size_t)spot * K + k];
#pragma unroll
    for (int c = 0; c < M::NC; c++) th[K + c] = constants[(size_t)spot * M::NC + c];
    float mu, J[K];
    model.eval_pixel(th, x, y, R, mu, J);
    if constexpr (M::kPixelBg) {  // background image: mu + bg * b, d mu / d bg = b
        const float b = constants[idx];
        mu += th[M::iBg] * b;
        J[M::iBg] = b;
    }
    if (ev) ev[idx] = mu;
    if (deriv) {
#pragma unroll
        for (int k = 0; k < K; k++) deriv[((size_t)spot * K + k) * S + p] = J[k];
    }
}

// Levenberg-Marquardt fit of ROIs whose size has no compiled fit kernel (roisize < 3 or
// > 32; the reference's models take any roisize, GaussianPSFModels.h:76-97).  One
// wavefront per spot; each iteration re-reads the spot's samples (they stay in L2)
// and evaluates every pixel on its own (eval_pixel), so it trades the per-size
// kernels' register-resident samples and shared edge terms for generality.  The
// iteration, clamp, stop and trace semantics are those of fit_spots.
template <class M>
__global__ void __launch_bounds__(256) fit_large_kernel(M model, FitArgs a, int R) {
    constexpr int K = M::K, KT = M::KT, NA = K * (K + 1) / 2;
    const int lane = threadIdx.x & (kWave - 1);
    const int S = R * R;
    const size_t waves = (size_t)gridDim.x * blockDim.x / kWave;
    for (size_t sp = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; sp < (size_t)a.n; sp += waves) {
        const float* smp = a.sample + sp * S;
        const float* img = M::kPixelBg ? a.constants + sp * S : nullptr;
        float th[KT];
#pragma unroll
        for (int c = 0; c < M::NC; c++) th[K + c] = a.constants[sp * M::NC + c];
        if (a.initial) {
#pragma unroll
            for (int k = 0; k < K; k++) th[k] = a.initial[sp * K + k];
        } else {
            float sx = 0.0f, sy = 0.0f, sum = 0.0f;
            for (int p = lane; p < S; p += kWave) {
                const float v = smp[p];
                const int y = p / R, x = p - y * R;
                sx += (float)x * v;
                sy += (float)y * v;
                sum += v;
            }
            sx = group_sum<kWave>(sx);
            sy = group_sum<kWave>(sy);
            sum = group_sum<kWave>(sum);
            model.initial(sx / sum, sy / sum, sum, th);
        }
        int iter = 0;
        for (;;) {
            if (iter >= a.maxIter) {
                iter = a.maxIter + 1;  // loop exhausted without a stop: the reference reports max + 1
                break;
            }
            if (a.trace && iter < a.traceLen && lane == 0) {
#pragma unroll
                for (int k = 0; k < K; k++) a.trace[(sp * a.traceLen + iter) * K + k] = th[k];
            }
            float acc[NA + K];
#pragma unroll
            for (int i = 0; i < NA + K; i++) acc[i] = 0.0f;
            for (int p = lane; p < S; p += kWave) {
                const int y = p / R, x = p - y * R;
                float mu, J[K];
                model.eval_pixel(th, x, y, R, mu, J);
                if constexpr (M::kPixelBg) {
                    mu += th[M::iBg] * img[p];
                    J[M::iBg] = img[p];
                }
                const float v = smp[p];
                const float xv = v < 1e-6f ? 1e-6f : v;  // keeps a NaN (Estimation.h:114)
                const float inv = rcpf(fmaxf(mu, 1e-6f));
                const float t = xv * inv;
                const float xf2 = t * inv;
                const float bf = t - 1.0f;
                int q = 0;
#pragma unroll
                for (int i = 0; i < K; i++) {
                    const float w = J[i] * xf2;
#pragma unroll
                    for (int j = i; j < K; j++) acc[q++] += w * J[j];
                }
#pragma unroll
                for (int i = 0; i < K; i++) acc[NA + i] += bf * J[i];
            }
            group_sum_n<kWave>(acc);
            float step[K];
            bool stop;
            if (!lm_step<K>(acc, acc + NA, a.lambda, step)) {
                stop = true;  // zero pivot: the reference breaks before updating
            } else {
                float last[K];
#pragma unroll
                for (int k = 0; k < K; k++) {
                    last[k] = th[k];
                    th[k] += step[k];
                }
                model.check_limits(th, R);
                bool small = true;
#pragma unroll
                for (int k = 0; k < K; k++)
                    if (fabsf(last[k] - th[k]) > M::stop_limit(k)) small = false;
                stop = (iter > 0 && small);
            }
            iter++;
            if (stop) break;
        }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < K; k++) a.params[sp * K + k] = th[k];
            if (a.iterations) a.iterations[sp] = iter;
        }
    }
}

template <class M>
__global__ void __launch_bounds__(256) crlb_kernel(M model, int n, int R, const float* __restrict__ params,
                                                   const float* __restrict__ constants, const float* __restrict__ sample,
                                                   float* __restrict__ crlb, float* __restrict__ chisq) {
    constexpr int K = M::K, KT = M::KT, NA = K * (K + 1) / 2;
    const int lane = threadIdx.x & (kWave - 1);
    const int spot = (int)(((size_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    if (spot >= n) return;
    const int S = R * R;
    float th[KT];
#pragma unroll
    for (int k = 0; k < K; k++) th[k] = params[(size_t)spot * K + k];
#pragma unroll
    for (int c = 0; c < M::NC; c++) th[K + c] = constants[(size_t)spot * M::NC + c];
    float fi[NA + 1];
#pragma unroll
    for (int i = 0; i < NA + 1; i++) fi[i] = 0.0f;
    for (int p = lane; p < S; p += kWave) {
        const int y = p / R, x = p - (p / R) * R;
        float mu, J[K];
        model.eval_pixel(th, x, y, R, mu, J);
        if constexpr (M::kPixelBg) {
            const float b = constants[(size_t)spot * S + p];
            mu += th[M::iBg] * b;
            J[M::iBg] = b;
        }
        const float m = fmaxf(1e-8f, mu);
        const float inv = rcpf(m);
        if (sample) {
            const float e = sample[(size_t)spot * S + p] - m;
            fi[NA] += e * e * inv;
        }
        int q = 0;
#pragma unroll
        for (int i = 0; i < K; i++) {
            const float w = J[i] * inv;
#pragma unroll
            for (int j = i; j < K; j++) fi[q++] += w * J[j];
        }
    }
    group_sum_n<kWave>(fi);
    if (lane == 0) {
        float c[K];
        crlb_from_fisher<K, M::kCrlbAbs>(fi, c);
        if (crlb) {
#pragma unroll
            for (int k = 0; k < K; k++) crlb[(size_t)spot * K + k] = c[k];
        }
        if (chisq) chisq[spot] = fi[NA];
    }
}

// Generalized likelihood-ratio test diagnostics (GLRT_PSF::Estimate,
// GLRT_PSF.cu:103-129): per spot {ll_on, ll_off, mean} with
//   ll_on  = sum_p d log(mu) - mu,        mu = max(model mu, 1e-8), d = max(x, 1e-8)
//   ll_off = sum_p d log(mean) - mean,
// the log-likelihoods of the fitted model and of a flat background at `mean`.
// The reference's mean sums the first spot's samples for every spot (its loop
// indexes the whole batch's sample array with the pixel counter); meanOfSpot0 = 1
// keeps that, 0 takes each spot's own mean.  One wavefront per spot.
template <class M>
__global__ void __launch_bounds__(256) glrt_kernel(M model, int n, int R, const float* __restrict__ params,
                                                   const float* __restrict__ constants,
                                                   const float* __restrict__ sample, int meanOfSpot0,
                                                   float* __restrict__ diag) {
    constexpr int K = M::K, KT = M::KT;
    const int lane = threadIdx.x & (kWave - 1);
    const int spot = (int)(((size_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    if (spot >= n) return;
    const int S = R * R;
    const float* own = sample + (size_t)spot * S;
    const float* ms = meanOfSpot0 ? sample : own;
    float th[KT];
#pragma unroll
    for (int k = 0; k < K; k++) th[k] = params[(size_t)spot * K + k];
#pragma unroll
    for (int c = 0; c < M::NC; c++) th[K + c] = constants[(size_t)spot * M::NC + c];
    float sum = 0.0f;
    for (int p = lane; p < S; p += kWave) sum += ms[p];
    sum = group_sum<kWave>(sum);
    const float mean = sum / (float)S;
    const float logMean = logf(mean);
    float ll[2] = {0.0f, 0.0f};
    for (int p = lane; p < S; p += kWave) {
        const int y = p / R, x = p - (p / R) * R;
        float mu, J[K];
        model.eval_pixel(th, x, y, R, mu, J);
        if constexpr (M::kPixelBg) mu += th[M::iBg] * constants[(size_t)spot * S + p];
        mu = fmaxf(mu, 1e-8f);
        const float d = fmaxf(own[p], 1e-8f);
        ll[0] += d * logf(mu) - mu;
        ll[1] += d * logMean - mean;
    }
    group_sum_n<kWave>(ll);
    if (lane == 0) {
        float* o = diag + (size_t)spot * 3;
        o[0] = ll[0];
        o[1] = ll[1];
        o[2] = mean;
    }
}

}  // namespace ppa
