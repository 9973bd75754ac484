// Batched Levenberg-Marquardt MLE fit with fused CRLB / chi-square.
//
// Mapping (gfx950, wave64): a spot is fitted by an aligned group of G lanes of
// one wavefront (G = 1..16, a power of two chosen per ROI size).  Lane g of a
// group owns the rows [g*RL, g*RL+RL) of the R x R ROI and keeps those samples
// in VGPRs for the whole fit, so HBM is read once per spot.  Per iteration
// each lane evaluates its pixels, accumulates the packed normal equations
// alpha (K(K+1)/2) and beta (K) in FP32, the group all-reduces them with lane
// shuffles, and every lane solves the FP64 damped Cholesky system redundantly
// (identical inputs give identical steps, so the group stays convergent).
// When a spot stops, its group writes params / iterations (and CRLB + chi^2
// from the same registers) and refills with the next spot of a static stride,
// so a wavefront keeps all its lanes busy until the batch is exhausted.
//
// Semantics follow LevMarOptimize (Estimation.h:194-251): trace before the
// step, damping, Cholesky break on a zero pivot, clamp, stop when i > 0 and
// every |delta| <= StopLimit, iterations = i + 1.
#pragma once

#include "models.h"

namespace ppa {

struct FitArgs {
    int n;                  // number of spots
    const float* sample;    // [n, R*R]
    const float* initial;   // [n, K] or null (centre-of-mass start)
    float* params;          // [n, K]
    int* iterations;        // [n] or null
    float* trace;           // [n, traceLen, K] or null
    int traceLen;
    float* crlb;            // [n, K] or null: skip the CRLB pass
    float* chisq;           // [n] or null
    int maxIter;
    float lambda;
};

template <class M, int R, int G>
struct FitTraits {
    static constexpr int K = M::K;
    static constexpr int S = R * R;
    static constexpr int RL = (R + G - 1) / G;   // rows per lane
    static constexpr int NA = K * (K + 1) / 2;   // packed alpha
};

// Fits spots first, first+stride, ... < a.n with the lane groups of the calling
// wavefront; g_first is this lane group's first spot.  Every lane of the
// wavefront must call it (the loop runs until all groups of the wave are done).
template <class M, int R, int G>
__device__ __forceinline__ void fit_spots(const M& model, const FitArgs& a, int g_first, int stride) {
    using T = FitTraits<M, R, G>;
    constexpr int K = T::K, RL = T::RL, NA = T::NA;
    static_assert(G >= 1 && G <= 64 && (G & (G - 1)) == 0, "G must be a power of two <= 64");

    const int lane = threadIdx.x & (kWave - 1);
    const int g = lane & (G - 1);
    const int y0 = g * RL;
    int spot = g_first;

    float s[RL][R];
    float th[K];
    int iter = 0;

    auto load_and_init = [&](int sp) {
        const float* src = a.sample + (size_t)sp * T::S;
        float sx = 0.0f, sy = 0.0f, sum = 0.0f;
#pragma unroll
        for (int r = 0; r < RL; r++) {
            const int y = y0 + r;
#pragma unroll
            for (int x = 0; x < R; x++) {
                float v = (y < R) ? __builtin_nontemporal_load(src + y * R + x) : 0.0f;
                s[r][x] = v;
                sx += (float)x * v;
                sy += (float)y * v;
                sum += v;
            }
        }
        if (a.initial) {
#pragma unroll
            for (int k = 0; k < K; k++) th[k] = a.initial[(size_t)sp * K + k];
        } else {
            sx = group_sum<G>(sx);
            sy = group_sum<G>(sy);
            sum = group_sum<G>(sum);
            model.initial(sx / sum, sy / sum, sum, th);
        }
        iter = 0;
    };

    bool active = spot < a.n;
    if (active) load_and_init(spot);

    while (__any(active)) {
        if (active) {
            bool done = false;
            if (iter >= a.maxIter) {
                // loop exhausted without a stop: the reference reports i + 1 = max + 1
                done = true;
                iter = a.maxIter + 1;
            } else {
                if (a.trace && iter < a.traceLen && g == 0) {
                    float* tp = a.trace + ((size_t)spot * a.traceLen + iter) * K;
#pragma unroll
                    for (int k = 0; k < K; k++) tp[k] = th[k];
                }
                float acc[NA + K];
#pragma unroll
                for (int i = 0; i < NA + K; i++) acc[i] = 0.0f;
                model.template for_rows<R, RL>(th, y0, [&](int r, int x, float mu, const float (&J)[K]) {
                    const float xv = fmaxf(s[r][x], 1e-6f);
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
                });
                group_sum_n<G>(acc);

                double step[K];
                if (!lm_step_f64<K>(acc, acc + NA, a.lambda, step)) {
                    done = true;  // zero pivot: the reference breaks before updating
                } else {
                    float last[K];
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        last[k] = th[k];
                        th[k] += (float)step[k];
                    }
                    model.check_limits(th, R);
                    bool small = true;
#pragma unroll
                    for (int k = 0; k < K; k++)
                        if (fabsf(last[k] - th[k]) > M::stop_limit(k)) small = false;
                    done = (iter > 0 && small);
                }
                iter++;
            }
            if (done) {
                float crlb[K], chi = 0.0f;
                if (a.crlb || a.chisq) {
                    float fi[NA + 1];
#pragma unroll
                    for (int i = 0; i < NA + 1; i++) fi[i] = 0.0f;
                    model.template for_rows<R, RL>(th, y0, [&](int r, int x, float mu, const float (&J)[K]) {
                        const float m = fmaxf(1e-8f, mu);
                        const float inv = rcpf(m);
                        const float e = s[r][x] - m;
                        fi[NA] += e * e * inv;
                        int q = 0;
#pragma unroll
                        for (int i = 0; i < K; i++) {
                            const float w = J[i] * inv;
#pragma unroll
                            for (int j = i; j < K; j++) fi[q++] += w * J[j];
                        }
                    });
                    group_sum_n<G>(fi);
                    chi = fi[NA];
                    if (a.crlb) crlb_from_fisher<K>(fi, crlb);
                }
                if (g == 0) {
                    float* pp = a.params + (size_t)spot * K;
#pragma unroll
                    for (int k = 0; k < K; k++) pp[k] = th[k];
                    if (a.iterations) a.iterations[spot] = iter;
                    if (a.crlb) {
                        float* cp = a.crlb + (size_t)spot * K;
#pragma unroll
                        for (int k = 0; k < K; k++) cp[k] = crlb[k];
                    }
                    if (a.chisq) a.chisq[spot] = chi;
                }
                spot += stride;
                active = spot < a.n;
                if (active) load_and_init(spot);
            }
        }
    }
}

template <class M, int R, int G>
__global__ void __launch_bounds__(256) fit_kernel(M model, FitArgs a) {
    const int groupsPerBlock = blockDim.x / G;
    fit_spots<M, R, G>(model, a, blockIdx.x * groupsPerBlock + threadIdx.x / G, gridDim.x * groupsPerBlock);
}

}  // namespace ppa
