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
// group ends with bit-identical totals.
template <int G>
PP_DEV float group_sum(float v) {
#pragma unroll
    for (int m = 1; m < G; m <<= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

template <int G, int N>
PP_DEV void group_sum_n(float (&v)[N]) {
    if constexpr (G > 1) {
#pragma unroll
        for (int i = 0; i < N; i++) v[i] = group_sum<G>(v[i]);
    }
}

// Packed index of (a,b), a<=b, in the upper triangle of a KxK matrix.
template <int K>
constexpr int tri(int a, int b) { return a * K - a * (a - 1) / 2 + (b - a); }

// ---------------------------------------------------------------------------
// Levenberg-Marquardt step in FP64 (reference: Estimation.h:213-228 and
// SolveMatrix.h:107-163).  alpha is the packed upper triangle, beta the
// gradient term.  Returns false where the reference breaks out of the loop
// (an exactly zero pivot).
template <int K>
PP_DEV bool lm_step_f64(const float* alphaTri, const float* beta, float lambda, double* step) {
    double A[K][K];
#pragma unroll
    for (int a = 0; a < K; a++)
#pragma unroll
        for (int b = a; b < K; b++) {
            double v = (double)alphaTri[tri<K>(a, b)];
            A[a][b] = v;
            A[b][a] = v;
        }
    // scale-invariant (lambda >= 0) or additive (lambda < 0) damping
    double dg[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        if (lambda < 0.0f) {
            dg[k] = A[k][k] - (double)lambda;
        } else {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < K; j++) s += A[j][k] * A[j][k];
            dg[k] = A[k][k] + s * (double)lambda;
        }
    }
#pragma unroll
    for (int k = 0; k < K; k++) A[k][k] = dg[k];
    // Cholesky A = L L^T
    double L[K][K];
    double invd[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = 0.0;
#pragma unroll
        for (int p = 0; p < k; p++) s += L[k][p] * L[k][p];
        double v = A[k][k] - s;
        if (v == 0.0) return false;
        double d = sqrt(v);
        L[k][k] = d;
        invd[k] = 1.0 / d;
#pragma unroll
        for (int i = k + 1; i < K; i++) {
            double s2 = 0.0;
#pragma unroll
            for (int p = 0; p < k; p++) s2 += L[i][p] * L[k][p];
            L[i][k] = (A[i][k] - s2) * invd[k];
        }
    }
    double y[K];
#pragma unroll
    for (int i = 0; i < K; i++) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < i; k++) s += L[i][k] * y[k];
        y[i] = ((double)beta[i] - s) * invd[i];
    }
#pragma unroll
    for (int i = K - 1; i >= 0; i--) {
        double s = 0.0;
#pragma unroll
        for (int k = i + 1; k < K; k++) s += L[k][i] * step[k];
        step[i] = (y[i] - s) * invd[i];
    }
    return true;
}

// CRLB from the packed Fisher matrix: FP64 LU with partial pivoting and the
// reference's failure rule (largest pivot < 1e-9 -> Inf), sqrt of the
// diagonal of the inverse without abs (SolveMatrix.h:219-263, 344-365,
// 441-452; Estimation.h:101).
template <int K>
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
        crlb[j] = (float)sqrt(x[j]);
    }
}

}  // namespace ppa
