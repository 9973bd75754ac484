// Device-side building blocks shared by every kernel of the MI355X spot fitter:
// lane-group reductions, the FP64 normal-equation solve and the FP64 Fisher
// inversion.  Written for gfx950 (wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PP_DEV __device__ __forceinline__

namespace ppa {

constexpr int kWave = 64;

// IEEE minNum/maxNum clamp: a NaN value maps to the upper bound, which is what
// fmaxf(a, fminf(v, b)) does (reference CudaMath.h:1152-1155).
PP_DEV float clampf(float v, float a, float b) { return fmaxf(a, fminf(v, b)); }
PP_DEV int clampi(int v, int a, int b) { return max(a, min(v, b)); }

PP_DEV float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

// All-reduce a value over the G lanes of an aligned lane group (G power of 2,
// G <= 64).  Partners combine partial sums commutatively, so every lane of the
// group ends with bit-identical totals.  Steps inside a row of 16 lanes use DPP
// moves (quad_perm xor 1 / xor 2, row_half_mirror, row_mirror: each pairs every
// lane with one in the other half of its block); the steps across rows use the
// gfx950 row swaps (v_permlane16_swap: rows 0-1 and 2-3, v_permlane32_swap: the
// two halves of the wavefront), which leave both partners' values in every lane.
// None of them goes through LDS.  The moved value is written with old = 0 and
// bound_ctrl set (every lane of a group is active, so no lane reads a disabled
// source); that form lets the compiler fold the move into the add as its DPP
// source operand.
template <int CTRL>
PP_DEV float dpp_add(float v) {
    return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}

// {value of the even row, value of the odd row} of this lane's pair of 16-lane rows.
PP_DEV void row_pair(float v, float& even, float& odd) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    even = __uint_as_float(r[0]);
    odd = __uint_as_float(r[1]);
}

// {value of the lower half, value of the upper half} of the wavefront at this lane.
PP_DEV void half_pair(float v, float& lo, float& hi) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    lo = __uint_as_float(r[0]);
    hi = __uint_as_float(r[1]);
}

template <int G>
PP_DEV float group_sum(float v) {
    if constexpr (G >= 2) v = dpp_add<0xB1>(v);   // quad_perm [1,0,3,2]
    if constexpr (G >= 4) v = dpp_add<0x4E>(v);   // quad_perm [2,3,0,1]
    if constexpr (G >= 8) v = dpp_add<0x141>(v);  // row_half_mirror
    if constexpr (G >= 16) v = dpp_add<0x140>(v); // row_mirror
    if constexpr (G >= 32) {
        float a, b;
        row_pair(v, a, b);
        v = a + b;
    }
    if constexpr (G >= 64) {
        float a, b;
        half_pair(v, a, b);
        v = a + b;
    }
    return v;
}

template <int G, int N>
PP_DEV void group_sum_n(float (&v)[N]) {
    if constexpr (G > 1) {
#pragma unroll
        for (int i = 0; i < N; i++) v[i] = group_sum<G>(v[i]);
    }
}

// Row stride, in floats, of the partial sums of group_sum_lds: N rounded up to 2 mod 4.
// A lane's row stays 8-byte aligned (its writes pair into 64-bit stores), and the rows of
// 16 consecutive lanes start on 16 different even banks, so each 64-bit store of a half
// wavefront covers every bank twice, the least it can.  With a stride of 20 (20 sums),
// lanes g and g + 8 started on the same bank: four-way conflicts.
template <int N>
constexpr int group_sum_stride() {
    return N + ((2 - N) % 4 + 4) % 4;
}

// LDS floats group_sum_lds<G, N> uses: the G rows of partial sums, then the N totals.
template <int G, int N>
constexpr int group_sum_lds_floats() {
    return G * group_sum_stride<N>() + N;
}

// The same all-reduce for groups of 32 or 64 lanes through the group's LDS region
// (group_sum_lds_floats<G, N>() floats, free while the group holds no table in it): every
// lane writes its N partial sums, lane g adds the G partials of values g, g + G, ... (< N)
// in lane order and writes their totals, and every lane reads the N totals back, so all
// lanes end with bit-identical totals.  About 2G/64 + N adds per wavefront where the
// DPP/row-swap butterfly takes (log2 G + 1) N instructions (20 sums at 32 lanes: about 35
// VALU instructions instead of 120).
template <int G, int N>
PP_DEV void group_sum_lds(float (&v)[N], float* lds) {
    static_assert(G >= 32, "the butterfly is cheaper for small groups");
    constexpr int S = group_sum_stride<N>();
    const int g = threadIdx.x & (G - 1);
#pragma unroll
    for (int i = 0; i < N; i++) lds[g * S + i] = v[i];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
#pragma unroll
    for (int r = 0; r < (N + G - 1) / G; r++) {
        const int i = g + r * G;
        if (N % G == 0 || i < N) {
            float t = lds[i];
#pragma unroll
            for (int k = 1; k < G; k++) t += lds[k * S + i];
            lds[G * S + i] = t;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
#pragma unroll
    for (int i = 0; i < N; i++) v[i] = lds[G * S + i];
    // every lane's reads complete before the region is written again
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Value of v held by lane SRC of this lane's aligned group of G lanes.
// G <= 4 uses a quad_perm DPP move (a full-rate VALU op); larger groups use
// ds_bpermute, whose lane offset is an immediate, so the broadcasts of a kernel
// that is bound by VALU issue take no VALU slot.
template <int G, int SRC>
PP_DEV float group_bcast(float v) {
    static_assert(SRC >= 0 && SRC < G, "source lane outside the group");
    if constexpr (G == 1) {
        return v;
    } else if constexpr (G == 2) {
        constexpr int ctrl = SRC | (SRC << 2) | ((2 + SRC) << 4) | ((2 + SRC) << 6);
        return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctrl, 0xF, 0xF, false));
    } else if constexpr (G == 4) {
        constexpr int ctrl = SRC * 0x55;
        return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctrl, 0xF, 0xF, false));
    } else {
        const int lane = threadIdx.x & (kWave - 1);
        return __shfl(v, (lane & ~(G - 1)) + SRC, kWave);
    }
}

// Packed index of (a,b), a<=b, in the upper triangle of a KxK matrix.
template <int K>
constexpr int tri(int a, int b) { return a * K - a * (a - 1) / 2 + (b - a); }

// ---------------------------------------------------------------------------
// Levenberg-Marquardt step in FP32, the reference's precision (Estimation.h:213-228
// and SolveMatrix.h:107-163).  alpha is the packed upper triangle, beta the gradient
// term.  Returns false where the reference breaks out of the loop (an exactly zero
// pivot).  The factorisation works in place on the packed lower triangle; each pivot
// takes 1/sqrt from v_rsq_f32 (1 ulp) and the reference's divisions by the pivot
// become products with it.  A negative pivot gives NaN, which the parameter clamps
// turn into a bound, as the reference's fmax/fmin do.  (An FP64 form with two Newton
// steps per pivot measured 6% slower on the 20x20 spline fit: its 170 FP64 and
// conversion instructions against about 100.)
template <int K>
constexpr int ltri(int i, int j) { return i * (i + 1) / 2 + j; }  // j <= i

template <int K>
PP_DEV bool lm_step(const float* alphaTri, const float* beta, float lambda, float* step) {
    constexpr int NP = K * (K + 1) / 2;
    float P[NP];
#pragma unroll
    for (int i = 0; i < K; i++)
#pragma unroll
        for (int j = 0; j <= i; j++) P[ltri<K>(i, j)] = alphaTri[tri<K>(j, i)];
    // scale-invariant (lambda >= 0) or additive (lambda < 0) damping; every column
    // sum uses the undamped matrix, as the reference's loop does
    float dg[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        if (lambda < 0.0f) {
            dg[k] = -lambda;
        } else {
            float s = 0.0f;
#pragma unroll
            for (int j = 0; j < K; j++) {
                const float v = P[j >= k ? ltri<K>(j, k) : ltri<K>(k, j)];
                s += v * v;
            }
            dg[k] = s * lambda;
        }
    }
#pragma unroll
    for (int k = 0; k < K; k++) P[ltri<K>(k, k)] += dg[k];
    // in-place Cholesky: P <- L
    float invd[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        float v = P[ltri<K>(k, k)];
#pragma unroll
        for (int p = 0; p < k; p++) v -= P[ltri<K>(k, p)] * P[ltri<K>(k, p)];
        if (v == 0.0f) return false;
        invd[k] = __builtin_amdgcn_rsqf(v);
#pragma unroll
        for (int i = k + 1; i < K; i++) {
            float w = P[ltri<K>(i, k)];
#pragma unroll
            for (int p = 0; p < k; p++) w -= P[ltri<K>(i, p)] * P[ltri<K>(k, p)];
            P[ltri<K>(i, k)] = w * invd[k];
        }
    }
    // L y = beta, then L^T x = y (x overwrites y)
#pragma unroll
    for (int i = 0; i < K; i++) {
        float v = beta[i];
#pragma unroll
        for (int k = 0; k < i; k++) v -= P[ltri<K>(i, k)] * step[k];
        step[i] = v * invd[i];
    }
#pragma unroll
    for (int i = K - 1; i >= 0; i--) {
        float v = step[i];
#pragma unroll
        for (int k = i + 1; k < K; k++) v -= P[ltri<K>(k, i)] * step[k];
        step[i] = v * invd[i];
    }
    return true;
}

// CRLB from the packed Fisher matrix: FP64 LU with partial pivoting and the
// reference's failure rule (largest pivot < 1e-9 -> Inf), sqrt of the
// diagonal of the inverse, without abs (SolveMatrix.h:219-263, 344-365,
// 441-452; Estimation.h:101) or with it (ABS: ComputeCRLB, Estimation.h:318-333).
template <int K, bool ABS = false>
PP_DEV void crlb_from_fisher(const float* fiTri, float* crlb) {
    double A[K][K];
#pragma unroll
    for (int a = 0; a < K; a++)
#pragma unroll
        for (int b = a; b < K; b++) {
            double v = (double)fiTri[tri<K>(a, b)];
            A[a][b] = v;
            A[b][a] = v;
        }
    int P[K];
#pragma unroll
    for (int i = 0; i < K; i++) P[i] = i;
    bool ok = true;
#pragma unroll
    for (int i = 0; i < K; i++) {
        double maxA = 0.0;
        int imax = i;
#pragma unroll
        for (int k = i; k < K; k++) {
            double v = fabs(A[k][i]);
            if (v > maxA) { maxA = v; imax = k; }
        }
        if (maxA < 1e-9) ok = false;
        // row swap i <-> imax (select-based so the arrays stay in registers)
#pragma unroll
        for (int k = i + 1; k < K; k++) {
            if (k == imax) {
#pragma unroll
                for (int j = 0; j < K; j++) { double t = A[i][j]; A[i][j] = A[k][j]; A[k][j] = t; }
                int t = P[i]; P[i] = P[k]; P[k] = t;
            }
        }
        double inv = 1.0 / A[i][i];
#pragma unroll
        for (int j = i + 1; j < K; j++) {
            A[j][i] *= inv;
#pragma unroll
            for (int k = i + 1; k < K; k++) A[j][k] -= A[j][i] * A[i][k];
        }
    }
    if (!ok) {
#pragma unroll
        for (int k = 0; k < K; k++) crlb[k] = __builtin_inff();
        return;
    }
    // Only the diagonal of the inverse is needed: column j of inv(A) solves
    // A x = e_j, i.e. L U x = P e_j.
#pragma unroll
    for (int j = 0; j < K; j++) {
        double x[K];
#pragma unroll
        for (int i = 0; i < K; i++) {
            double v = (P[i] == j) ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < i; k++) v -= A[i][k] * x[k];
            x[i] = v;
        }
#pragma unroll
        for (int i = K - 1; i >= 0; i--) {
            double v = x[i];
#pragma unroll
            for (int k = i + 1; k < K; k++) v -= A[i][k] * x[k];
            x[i] = v / A[i][i];
        }
        crlb[j] = (float)sqrt(ABS ? fabs(x[j]) : x[j]);
    }
}

}  // namespace ppa
