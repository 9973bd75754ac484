// Batched Levenberg-Marquardt MLE fit, followed by CRLB / chi-square.
//
// Mapping (gfx950, wave64): a spot is fitted by an aligned group of G lanes of
// one wavefront (G = 1..32, a power of two chosen per ROI size).  Each lane
// keeps its share of the R x R samples in VGPRs for the whole fit (the model's
// pixel layout says which: whole rows, or every G-th pixel), so HBM is read once
// per spot.  Per iteration each lane evaluates its pixels, accumulates the packed
// normal equations alpha (K(K+1)/2) and beta (K) in FP32, the group all-reduces
// them with lane shuffles, and every lane solves the FP32 damped Cholesky system
// redundantly (identical inputs give identical steps, so the group stays
// convergent).  Models that stage per-spot tables in LDS get a region of
// M::lds_floats(R, G) floats per group.
//
// Spots stop after different numbers of iterations.  A group whose spot has
// stopped waits until `refill` groups of its wavefront (or every still-running
// one) have stopped; those groups then write their results and load their next
// spots together, so the wavefront pays the partly-masked refill once per batch
// of groups instead of once per spot.  Next spots come from work counters in
// HBM (one atomic per batch), so every group stays busy until the batch is
// exhausted and the launch ends within about one spot's fit of each other.
// The spots can be split into 2^segShift (1 or kWorkSegs) contiguous segments
// with a counter each, block b starting on segment b % 2^segShift: the hardware
// deals a launch's blocks to the 8 XCDs in turn, so neighbouring spots are fitted
// under one L2, and the claims of a fit that refills often are spread over eight
// counters (one counter measured 6% slower on 16x16 astigmatic fits, eight 3%
// slower on 11x11 free-sigma fits; the tuner picks per model and size).  A wave
// whose segment runs out moves on to the next one with spots left.  The
// CRLB and chi^2 are computed by a second kernel that walks all spots in
// lockstep (every group does exactly one spot per pass), reading the samples and
// the fitted parameters back from HBM.
//
// Semantics follow LevMarOptimize (Estimation.h:194-251): trace before the
// step, damping, Cholesky break on a zero pivot, clamp, stop when i > 0 and
// every |delta| <= StopLimit, iterations = i + 1 (max + 1 when it never stops).
#pragma once

#include "models.h"

namespace ppa {

struct FitArgs {
    int n;                  // spots [0, n) are fitted (the persistent batch passes a tile's end)
    const float* sample;    // [*, R*R]
    const float* initial;   // [*, K] or null (centre-of-mass start)
    float* params;          // [*, K]
    int* iterations;        // [*] or null
    float* trace;           // [*, traceLen, K] or null
    int traceLen;
    float* crlb;            // [*, K] or null
    float* chisq;           // [*] or null
    int maxIter;
    float lambda;
    int refill;             // stopped groups per batched refill (<= 0: one eighth of the wave's groups)
    int* counter;           // next unclaimed spot of each segment (relative to the segment) at
                            // counter[s * kCounterStride] (set by launch_set_counters), or null:
                            // static stride
    int segShift;           // log2 of the work segments (0 or 3)
    const float* constants; // [*, NC] per-spot model constants (models with NC > 0)
};

// Work segments of a fit launch (at most one per XCD of gfx950) and segment s's first
// spot of 2^shift segments.
constexpr int kWorkSegs = 8;
// ints from one segment counter to the next (each on a 256-byte line of its own) and
// ints of one launch's counter set
constexpr int kCounterStride = 64;
constexpr int kCounterInts = kWorkSegs * kCounterStride;
PP_DEV int seg_begin(int n, int s, int shift) { return (int)(((long long)n * s) >> shift); }

template <class M, int R, int G>
struct FitTraits {
    using L = typename M::template Layout<R, G>;
    static constexpr int K = M::K;
    static constexpr int KT = M::KT;             // parameters + per-spot constants
    static constexpr int S = R * R;
    static constexpr int NS = L::NS;             // samples per lane
    static constexpr int NA = K * (K + 1) / 2;   // packed alpha
    static constexpr int LDS = M::lds_floats(R, G); // LDS floats per group
    static constexpr int NB = M::kPixelBg ? NS : 1;  // background-image values per lane
};

// Models that accumulate the fit's normal equations themselves (M::kRowFactored).
template <class M, class = void>
struct RowFactored : std::false_type {};
template <class M>
struct RowFactored<M, std::void_t<decltype(M::kRowFactored)>> : std::bool_constant<M::kRowFactored> {};
template <class M>
constexpr bool row_factored() { return RowFactored<M>::value; }

// Lane g's share (layout L) of a per-pixel array of the spot, e.g. its background image.
template <int R, class L, int NS>
PP_DEV void load_pixels(const float* src, int g, float (&v)[NS]) {
#pragma unroll
    for (int k = 0; k < NS; k++) {
        const int idx = L::index(g, k);
        v[k] = idx < R * R ? src[idx] : 0.0f;
    }
}

// The model's callback with the background image applied for kPixelBg models:
// mu + bg * b[k] and d mu / d bg = b[k] (CubicSplinePSFModels.h:433-436).
template <class M, int NB, class Fn>
PP_DEV auto pixel_bg(const float (&th)[M::KT], const float (&b)[NB], Fn& fn) {
    return [&](int k, float mu, const float (&J)[M::K]) {
        if constexpr (M::kPixelBg) {
            float Jb[M::K];
#pragma unroll
            for (int i = 0; i < M::K; i++) Jb[i] = J[i];
            Jb[M::iBg] = b[k];
            fn(k, mu + th[M::iBg] * b[k], Jb);
        } else {
            fn(k, mu, J);
        }
    };
}

// The Levenberg-Marquardt terms' sample clamp `if (x < 1e-6) x = 1e-6` (a NaN stays
// NaN, Estimation.h:114) as one NaN-propagating maximum (v_maximum3_f32).  Written
// as a compare and select, the compiler evaluated all of a lane's compares before
// the selects and parked the masks in VGPR lanes (122 SGPRs for 61 samples).
PP_DEV float clamp_sample(float v) { return __builtin_elementwise_maximum(v, 1e-6f); }

// Loads lane g's samples of a spot (layout L) and returns the lane's
// centre-of-mass partial sums.  With CLAMP the stored samples are max(x, 1e-6)
// as the Levenberg-Marquardt terms use them (`if (x < 1e-6) x = 1e-6`, which
// keeps a NaN, Estimation.h:114); the sums use the raw values.  The fit kernel
// loads with the non-temporal hint (a few groups refill at a time and each ROI
// is read once); the CRLB pass, where every group loads at once, uses ordinary
// loads so that the lines its lanes share stay in L2 between loads.
template <int R, class L, int NS, bool CLAMP, bool NT, bool FACTORED = false>
PP_DEV void load_samples(const float* src, int g, float (&s)[NS], float& sx, float& sy, float& sum) {
    // The per-slot offsets and coordinates depend only on g; for a strided layout
    // the opaque copy keeps the compiler from hoisting all of them out of the fit
    // loop into VGPRs (an empty asm: it emits no instruction).
    if constexpr (L::kStrided) asm volatile("" : "+v"(g));
    sx = sy = sum = 0.0f;
    if constexpr (L::kStrided || !FACTORED) {
#pragma unroll
        for (int k = 0; k < NS; k++) {
            const int idx = L::index(g, k);
            float v = 0.0f;
            if (idx < R * R) v = NT ? __builtin_nontemporal_load(src + idx) : src[idx];
            s[k] = CLAMP ? clamp_sample(v) : v;
            sx += (float)L::x(g, k) * v;
            sy += (float)L::y(g, k) * v;
            sum += v;
        }
    } else {
        // row layouts under a register budget: sum x v = sum over regions xbase * (sum v)
        // + sum xoff v (and y likewise), with compile-time offsets, so that no per-slot
        // coordinates are held (hoisted out of the fit loop they were spilled)
        float part[L::NREG];
#pragma unroll
        for (int r = 0; r < L::NREG; r++) part[r] = 0.0f;
        float sxo = 0.0f, syo = 0.0f;
#pragma unroll
        for (int k = 0; k < NS; k++) {
            const int idx = L::index(g, k);
            float v = 0.0f;
            if (idx < R * R) v = NT ? __builtin_nontemporal_load(src + idx) : src[idx];
            s[k] = CLAMP ? clamp_sample(v) : v;
            part[L::region(k)] += v;
            if (L::xoff(k) != 0) sxo += (float)L::xoff(k) * v;
            if (L::yoff(k) != 0) syo += (float)L::yoff(k) * v;
        }
        sx = sxo;
        sy = syo;
#pragma unroll
        for (int r = 0; r < L::NREG; r++) {
            sx += (float)L::xbase(g, r) * part[r];
            sy += (float)L::ybase(g, r) * part[r];
            sum += part[r];
        }
    }
}

// Fisher matrix (packed) and chi^2 of one spot, written by lane g == 0 of its group.
// Whether a group of G >= 32 lanes sums N values through its LDS region (group_sum_lds):
// the model's region, free between its per-iteration tables, holds what the sum uses.
template <class M, int R, int G, int N>
constexpr bool group_sum_in_lds() {
    if constexpr (G >= 32)
        return M::lds_floats(R, G) >= group_sum_lds_floats<G, N>();
    else
        return false;
}

template <class M, int R, int G>
PP_DEV void crlb_spot(const M& model, const FitArgs& a, const float (&s)[FitTraits<M, R, G>::NS],
                      const float (&pb)[FitTraits<M, R, G>::NB], const float (&th)[M::KT], int g, float* lds,
                      int spot) {
    using T = FitTraits<M, R, G>;
    constexpr int K = T::K, NA = T::NA;
    float fi[NA + 1];
#pragma unroll
    for (int i = 0; i < NA + 1; i++) fi[i] = 0.0f;
    auto body = [&](int k, float mu, const float (&J)[K]) {
        const float m = fmaxf(1e-8f, mu);
        const float inv = rcpf(m);
        const float e = s[k] - m;
        fi[NA] += e * e * inv;
        int q = 0;
#pragma unroll
        for (int i = 0; i < K; i++) {
            const float w = J[i] * inv;
#pragma unroll
            for (int j = i; j < K; j++) fi[q++] += w * J[j];
        }
    };
    if constexpr (row_factored<M>())
        model.template accumulate_rows<R, G, true>(th, g, s, fi, lds);
    else
        model.template for_pixels<R, G>(th, g, lds, pixel_bg<M>(th, pb, body));
    // (the LDS form measured 5% slower here: one sum per spot and pass, on a kernel that
    // waits on its loads, pays the LDS round trips in full)
    group_sum_n<G>(fi);
    if (g == 0) {
        if (a.crlb) {
            float c[K];
            crlb_from_fisher<K, M::kCrlbAbs>(fi, c);
            float* cp = a.crlb + (size_t)spot * K;
#pragma unroll
            for (int k = 0; k < K; k++) cp[k] = c[k];
        }
        if (a.chisq) a.chisq[spot] = fi[NA];
    }
}

// Minimum wavefronts per SIMD the compiler must fit a model's fit kernel into
// (its register budget): M::min_waves(R, G) where the model names one, else 1
// (no bound).
template <class M, int R, int G, class = void>
struct MinWaves {
    static constexpr int value = 1;
};
template <class M, int R, int G>
struct MinWaves<M, R, G, std::void_t<decltype(M::min_waves(R, G))>> {
    static constexpr int value = M::min_waves(R, G);
};

// Fits spots first, first+stride, ... < a.n with the lane groups of the calling
// wavefront, or, with work counters, its group's first spot g_first of segment seg
// and then the spots the counters hand out; lds is the group's LDS region.  Every
// lane of the wavefront must call it (the loops run until all groups of the wave
// are done) with the same seg.
template <class M, int R, int G>
__device__ __forceinline__ void fit_spots(const M& model, const FitArgs& a, int g_first, int stride, float* lds,
                                          int seg = 0) {
    using T = FitTraits<M, R, G>;
    using L = typename T::L;
    constexpr int K = T::K, NS = T::NS, NA = T::NA;
    constexpr int P = kWave / G;  // groups per wavefront
    static_assert(G >= 1 && G <= 64 && (G & (G - 1)) == 0, "G must be a power of two <= 64");

    const int lane = threadIdx.x & (kWave - 1);
    const int g = lane & (G - 1);
    const bool lead = (g == 0);
    const int refill = a.refill > 0 ? a.refill : (P >= 8 ? P / 8 : 1);

    constexpr int KT = T::KT;
    float s[NS];
    float pb[T::NB];
    float th[KT];
    int iter = 0;
    int spot = g_first;
    const int nseg = 1 << a.segShift;
    if (a.counter) {
        const int lo = seg_begin(a.n, seg, a.segShift);
        spot = g_first < seg_begin(a.n, seg + 1, a.segShift) - lo ? lo + g_first : a.n;
    }

    auto start_spot = [&](int sp) {
        float sx, sy, sum;
        load_samples<R, L, NS, true, true, (MinWaves<M, R, G>::value > 1)>(a.sample + (size_t)sp * T::S, g, s, sx,
                                                                          sy, sum);
        if constexpr (M::kPixelBg) load_pixels<R, L, NS>(a.constants + (size_t)sp * T::S, g, pb);
#pragma unroll
        for (int c = 0; c < M::NC; c++) th[K + c] = a.constants[(size_t)sp * M::NC + c];
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

    // state 0 = iterating, 1 = stopped (results not yet written; spot < 0: waiting for a
    // first spot), 2 = no more spots.  With work counters a group whose segment has no
    // first spot for it claims one from the counters like a stopped group.
    int state = 2;
    if (spot < a.n) {
        start_spot(spot);
        state = 0;
    } else if (a.counter) {
        spot = -1;
        state = 1;
    }
    for (;;) {
        const unsigned long long stopped = __ballot(lead && state == 1);
        const unsigned long long running = __ballot(lead && state == 0);
        if (!stopped && !running) break;
        if (stopped && (!running || __popcll(stopped) >= (unsigned)refill)) {
            // next spots: claimed together from the segment's work counter (one atomic
            // per batch, handed out in lane order; the rest from the next segment when
            // this one runs out), or the static stride
            // (the claim bookkeeping is wave-uniform: kept in scalar registers)
            int next = a.n;
            if (a.counter) {
                const int need = (int)__popcll(stopped);
                const int leadLane = lane & ~(G - 1);
                const int rank = (int)__popcll(stopped & ((1ull << leadLane) - 1ull));
                int done = 0;
                // every try serves the batch or uses a segment up, so nseg tries visit
                // every segment that still has spots
                for (int tries = 0; done < need && tries < nseg; tries++) {
                    int base = 0;
                    if (lane == 0) base = atomicAdd(a.counter + seg * kCounterStride, need - done);
                    base = __builtin_amdgcn_readfirstlane(base);
                    const int lo = seg_begin(a.n, seg, a.segShift);
                    const int avail = max(0, min(need - done, seg_begin(a.n, seg + 1, a.segShift) - lo - base));
                    if (rank >= done && rank < done + avail) next = lo + base + (rank - done);
                    done += avail;
                    if (done < need) {
                        // this segment is used up: read the other counters at once (lane k
                        // the k-th next segment's) and go on with the first that has spots
                        // left; counters only grow, so one seen used up stays so
                        bool left = false;
                        if (lane < nseg - 1) {
                            const int s = (seg + 1 + lane) & (nseg - 1);
                            const int c = __hip_atomic_load(a.counter + s * kCounterStride, __ATOMIC_RELAXED,
                                                            __HIP_MEMORY_SCOPE_AGENT);
                            left = c < seg_begin(a.n, s + 1, a.segShift) - seg_begin(a.n, s, a.segShift);
                        }
                        const unsigned long long has = __ballot(left);
                        if (!has) break;
                        seg = (seg + __ffsll((unsigned long long)has)) & (nseg - 1);
                        seg = __builtin_amdgcn_readfirstlane(seg);
                    }
                }
            }
            if (state == 1) {
                if (lead && spot >= 0) {
                    float* pp = a.params + (size_t)spot * K;
#pragma unroll
                    for (int k = 0; k < K; k++) pp[k] = th[k];
                    if (a.iterations) a.iterations[spot] = iter;
                }
                if (a.counter) {
                    spot = next;
                } else {
                    spot += stride;
                }
                if (spot < a.n) {
                    start_spot(spot);
                    state = 0;
                } else {
                    state = 2;
                }
            }
            continue;
        }
        if (state != 0) continue;
        if (iter >= a.maxIter) {
            // loop exhausted without a stop: the reference reports i + 1 = max + 1
            iter = a.maxIter + 1;
            state = 1;
            continue;
        }
        if (a.trace && iter < a.traceLen && lead) {
            float* tp = a.trace + ((size_t)spot * a.traceLen + iter) * K;
#pragma unroll
            for (int k = 0; k < K; k++) tp[k] = th[k];
        }
        float acc[NA + K];
#pragma unroll
        for (int i = 0; i < NA + K; i++) acc[i] = 0.0f;
        auto body = [&](int k, float mu, const float (&J)[K]) {
            const float xv = s[k];  // clamped at load
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
        };
        // The lane's pixel coordinates, edge positions and table addresses depend only
        // on g, and lambda is fixed.  Under a register budget, opaque copies made per
        // iteration (empty asm, no instruction) keep the compiler from hoisting them out
        // of the loop, where they were spilled to scratch and reloaded every iteration;
        // recomputing them takes a few VALU instructions (11x11 free sigma: fit kernel
        // 3.69 -> 3.56 ms, 120 -> 0 bytes of scratch).  Without a budget nothing spills
        // and the hoisted values are cheaper (10x10 fixed sigma: 48.2 vs 49.8 us).
        int gi = g;
        float lam = a.lambda;
        if constexpr (MinWaves<M, R, G>::value > 1) {
            asm volatile("" : "+v"(gi));
            asm volatile("" : "+s"(lam));
        }
        if constexpr (row_factored<M>()) {
            model.template accumulate_rows<R, G>(th, gi, s, acc, lds);
        } else {
            model.template for_pixels<R, G>(th, gi, lds, pixel_bg<M>(th, pb, body));
        }
        if constexpr (group_sum_in_lds<M, R, G, NA + K>())
            group_sum_lds<G>(acc, lds);
        else
            group_sum_n<G>(acc);

        float step[K];
        bool stop;
        if (!lm_step<K>(acc, acc + NA, lam, step)) {
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
        if (stop) state = 1;
    }
}

// CRLB and chi^2 of spots first, first+stride, ... < a.n from the samples and the
// fitted parameters in HBM, one spot per lane group per pass (no divergence).
template <class M, int R, int G>
__device__ __forceinline__ void crlb_spots(const M& model, const FitArgs& a, int g_first, int stride, float* lds) {
    using T = FitTraits<M, R, G>;
    constexpr int K = T::K, NS = T::NS, KT = T::KT;
    const int g = threadIdx.x & (G - 1);
    float s[NS];
    float pb[T::NB];
    float th[KT];
    for (int sp = g_first; __any(sp < a.n); sp += stride) {
        if (sp < a.n) {
            float sx, sy, sum;
            load_samples<R, typename T::L, NS, false, false>(a.sample + (size_t)sp * T::S, g, s, sx, sy, sum);
            if constexpr (M::kPixelBg) load_pixels<R, typename T::L, NS>(a.constants + (size_t)sp * T::S, g, pb);
#pragma unroll
            for (int k = 0; k < K; k++) th[k] = a.params[(size_t)sp * K + k];
#pragma unroll
            for (int c = 0; c < M::NC; c++) th[K + c] = a.constants[(size_t)sp * M::NC + c];
            crlb_spot<M, R, G>(model, a, s, pb, th, g, lds, sp);
        }
    }
}

// LDS region of this thread's lane group (models with per-group tables).
template <class M, int R, int G>
PP_DEV float* group_lds() {
    extern __shared__ float ppa_lds[];
    constexpr int per = M::lds_floats(R, G);
    return per > 0 ? ppa_lds + (size_t)(threadIdx.x / G) * per : nullptr;
}

// Phase 1: fit.  Phase 2 (CRLB and chi^2 of the fitted spots) is a kernel of its
// own, launched after it on the same stream, so that it sees every wavefront's
// results and runs at the occupancy of its own, smaller register set.
template <class M, int R, int G>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MinWaves<M, R, G>::value)))
fit_kernel(M model, FitArgs a) {
    const int groupsPerBlock = blockDim.x / G;
    // with work counters block b starts as the (b >> segShift)-th block of segment
    // b & (2^segShift - 1) (one call site: the fit loop is inlined once)
    const bool segs = a.counter != nullptr;
    const int seg = segs ? (int)(blockIdx.x & ((1u << a.segShift) - 1)) : 0;
    const int blk = segs ? (int)(blockIdx.x >> a.segShift) : (int)blockIdx.x;
    fit_spots<M, R, G>(model, a, blk * groupsPerBlock + threadIdx.x / G, gridDim.x * groupsPerBlock,
                       group_lds<M, R, G>(), seg);
}

template <class M, int R, int G>
__global__ void __launch_bounds__(256) crlb_phase_kernel(M model, FitArgs a) {
    const int groupsPerBlock = blockDim.x / G;
    crlb_spots<M, R, G>(model, a, blockIdx.x * groupsPerBlock + threadIdx.x / G, gridDim.x * groupsPerBlock,
                        group_lds<M, R, G>());
}

}  // namespace ppa
