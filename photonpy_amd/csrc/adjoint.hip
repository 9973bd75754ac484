// Poisson negative log-likelihood and its adjoint (no reference counterpart).
//
// NLL of a spot: sum over pixels of mu' - x' log mu', with mu' = max(mu, 1e-8)
// and x' = max(x, 1e-8) (the convention of ComputeLogLikelihood,
// Estimation.h:138-156, negated).  The adjoint recomputes mu and the Jacobian from
// the parameters (nothing of an NLL pass is stored) and returns TRUE derivatives:
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
// Mapping: one wavefront per spot, lanes stride over pixels (any roisize).  Scalar
// shared gradients are reduced per block in LDS and added to HBM once per block.
// The spline NLL and adjoint of roisize <= 28 run by window rounds, one spot per lane
// (spline_adjoint_round_kernel, below); larger ROIs take the one-wavefront-per-spot
// form and scatter the coefficient gradient pixel by pixel.
#include "dispatch.h"

namespace ppa {


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

// Value and x/y/z derivatives of spline cell c at the frame's sub-voxel position, one
// z-power (16 coefficients) after another, for a lane's own cell in the round kernel
// (rarely taken; with all 64 coefficients in registers it set the kernel's register
// peak).  The arithmetic and its order are SplineModel::cell's
//   f = fm0 + fm1 z + fm2 z^2 + fm3 z^3,  fz = fm1 + fm2 2z + fm3 3z^2
// (z-powers formed as the frame forms them), so it gives SplineModel::cell's bits.
template <int UNROLL_M>
PP_DEV void spline_cell_by_z(const float* c, const SplineModel::Frame& F, float& f, float& fx, float& fy,
                             float& fz) {
    const float4* c4 = reinterpret_cast<const float4*>(c);
    // the frame's z-powers, formed from z by the frame's own operations (z z, (z z) z,
    // 2 z, 3 (z z)) and selected per m in registers: indexing the frame's arrays with m
    // put the frame in scratch memory
    const float z = F.zp[1], z2 = z * z, z3 = z2 * z, dz2 = 2.0f * z, dz3 = 3.0f * z2;
    f = fx = fy = fz = 0.0f;
#pragma unroll UNROLL_M
    for (int m = 0; m < 4; m++) {
        float a, ax, ay;
        const float4 q[4] = {c4[m * 4], c4[m * 4 + 1], c4[m * 4 + 2], c4[m * 4 + 3]};
        SplineModel::cell_m(q, F, a, ax, ay);
        const float zm = m == 1 ? z : m == 2 ? z2 : z3;
        const float dzm = m == 2 ? dz2 : dz3;
        if (m == 0) {
            f = a;
            fx = ax;
            fy = ay;
        } else {
            f += a * zm;
            fx += ax * zm;
            fy += ay * zm;
            fz = m == 1 ? a : fz + a * dzm;
        }
    }
}

// One wavefront per spot (any roisize): the NLL, the per-spot gradient and, unless the
// tiled form below takes it, the coefficient gradient, scattered pixel by pixel with
// the 64 lanes covering the 64 coefficients of the pixel's cell (each atomic
// instruction updates one contiguous 256-byte cell).
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
    // lane q owns coefficient q = 16m + 4n + o of every cell: basis z^m y^n x^o
    const int m = lane >> 4, nn = (lane >> 2) & 3, o = lane & 3;
    const float zb = m == 0 ? 1.0f : m == 1 ? F.zp[1] : m == 2 ? F.zp[2] : F.zp[3];
    const float yb = nn == 0 ? 1.0f : nn == 1 ? F.yp[1] : nn == 2 ? F.yp[2] : F.yp[3];
    const float xb = o == 0 ? 1.0f : o == 1 ? F.x1 : o == 2 ? F.x2 : F.x3;
    const float basis = zb * yb * xb;
    float acc[K + 1];
#pragma unroll
    for (int i = 0; i < K + 1; i++) acc[i] = 0.0f;
    // every lane runs every trip (the scatter is a wavefront operation)
    for (int p0 = 0; p0 < S; p0 += kWave) {
        const int p = p0 + lane;
        float W = 0.0f;
        int cell = 0;
        if (p < S) {
            const int y = p / R, x = p - y * R;
            const float* c = M.row_ptr(F, y) + 64 * M.col_index(F, x);
            float f, fx, fy, fz, mu, J[K];
            SplineModel::cell(c, F, f, fx, fy, fz);  // all 64 coefficient loads in flight
            SplineModel::finish(F, t, f, fx, fy, fz, x, y, R, mu, J);
            if (f < 0.0f) J[0] = J[1] = J[2] = 0.0f;  // mu does not depend on position where clamped
            const float w = nll_weight(mu, a.sample[(size_t)spot * S + p], acc[K]);
#pragma unroll
            for (int k = 0; k < K; k++) acc[k] += w * J[k];
            cell = (int)((c - M.coefs) >> 6);
            W = (f > 0.0f) ? go * w * t[3] : 0.0f;
        }
        if (a.gshared) {
            unsigned long long any = __ballot(W != 0.0f);
            while (any) {
                const int src = __ffsll((long long)any) - 1;
                any &= any - 1;
                const float g = __shfl(W, src, kWave);
                const int cs = __shfl(cell, src, kWave);
                atomicAdd(&a.gshared[(size_t)cs * 64 + lane], g * basis);
            }
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
}

// ---------------------------------------------------------------------------
// Spline NLL and adjoint by window rounds (roisize <= kMaxTiledR).
//
// A spot's pixels all share its sub-voxel position (x_s, y_s, z_s), and pixel (y, x)
// takes cell (k, j0 + y, i0 + x) of its z-slab k, (j0, i0) the cell of pixel (0, 0):
// the spots of one window (k, j0, i0) read the same cell at the same pixel.  Spots
// are sorted by window and cut into tiles of up to kTileSpots spots of one window; a
// workgroup of kRoundWaves wavefronts takes up to 64 spots of a tile, one spot per
// lane, each wavefront a band of pixel rows.  At each pixel the cell's address is the
// same for all lanes, so its 64 coefficients are loaded once for 64 spots (the
// per-spot form gathered them once per spot), and each lane evaluates its own spot.
// A lane whose own cell differs (FP32 rounding of the cell index, or the grid edge
// clamping it) evaluates from its own cell instead.  The spot's NLL and parameter
// gradient stay in the lane.
//
// Coefficient gradient: spot s adds W_s(p) z_s^m y_s^n x_s^o to coefficient (m, n, o)
// of the cell of its pixel p, W_s(p) = g (1 - x/mu) I (zero where the spline value is
// clamped).  Per pixel row the lanes store their W_s in LDS; then lane q forms, for
// each cell of the row, sum_s W_s B_s[q] over the 64 spots (B_s[q] = the spot's basis
// value, held in the lane's registers for the round) and adds it with one atomic per
// cell (a wavefront atomic over the cell's 256 contiguous bytes).  A lane whose pixel
// fell in another cell adds its 64 terms to that cell directly.
constexpr int kMaxTiledR = 28;
constexpr int kTileSpots = 256;              // spots per tile
constexpr int kRoundWaves = 4;               // wavefronts per round (pixel-row bands)
constexpr int kRoundBlock = kRoundWaves * kWave;

PP_DEV int spline_window_key(const SplineModel& M, const float (&t)[5], int R, SplineModel::Frame& F, int& i0,
                             int& j0, int& k) {
    M.frame(t, R, F);
    i0 = M.col_index(F, 0);
    j0 = clampi((int)(0.0f - F.ty), 0, M.Ny - 1);
    k = (int)((F.slab - M.coefs) / ((size_t)64 * M.Nx * M.Ny));
    return (k * M.Ny + j0) * M.Nx + i0;
}

__global__ void adj_fill_kernel(int* p, int n, int v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// window key of every spot and the number of spots per key
__global__ void __launch_bounds__(256) adj_key_kernel(SplineModel M, AdjArgs a, SplineAdjTiles w) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.n) return;
    M.coefs = a.shared;
    float t[5];
#pragma unroll
    for (int k = 0; k < 5; k++) t[k] = a.params[(size_t)s * 5 + k];
    SplineModel::Frame F;
    int i0, j0, k;
    const int key = spline_window_key(M, t, a.R, F, i0, j0, k);
    w.key[s] = key;
    atomicAdd(&w.count[key], 1);
}

// Each window with spots gets a contiguous range of the sorted order and its tiles,
// both taken from counters (cursor[0]: tiles, cursor[1]: sorted positions), so that
// only the windows with spots do any work; the count array is reset to zero for use
// as the scatter cursor.  The order of the windows is arbitrary.
__global__ void __launch_bounds__(256) adj_alloc_kernel(SplineAdjTiles w) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < w.nkeys; k += gridDim.x * blockDim.x) {
        const int n = w.count[k];
        if (n == 0) continue;
        const int pos = atomicAdd(&w.ntiles[1], n);
        const int nt = (n + kTileSpots - 1) / kTileSpots;
        int tile = atomicAdd(&w.ntiles[0], nt);
        w.offset[k] = pos;
        for (int f = 0; f < n; f += kTileSpots) w.tiles[tile++] = make_int2(pos + f, min(kTileSpots, n - f));
        w.count[k] = 0;
    }
}

// spots in key order (the order inside a key is arbitrary)
__global__ void __launch_bounds__(256) adj_scatter_kernel(AdjArgs a, SplineAdjTiles w) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.n) return;
    const int key = w.key[s];
    w.order[w.offset[key] + atomicAdd(&w.count[key], 1)] = s;
}

// Value of spline cell c at the frame's sub-voxel position (the NLL needs no derivatives),
// UNROLL_M z-powers of 16 coefficients at a time (4: all 64 loads in flight; 1: 16
// registers of coefficients, for the rarely taken per-lane cell of the round kernels,
// which would otherwise set their register peak).
// One z-power's part of the value: its 16 coefficients contracted with the x and y powers.
PP_DEV float spline_value_m(const float4 (&q)[4], const SplineModel::Frame& F) {
    float a = 0.0f;
#pragma unroll
    for (int n = 0; n < 4; n++) a += (q[n].x + q[n].y * F.x1 + q[n].z * F.x2 + q[n].w * F.x3) * F.yp[n];
    return a;
}

template <int UNROLL_M = 4>
PP_DEV float spline_cell_value(const float* c, const SplineModel::Frame& F) {
    const float4* c4 = reinterpret_cast<const float4*>(c);
    float f = 0.0f, zm = 1.0f;
#pragma unroll UNROLL_M
    for (int m = 0; m < 4; m++) {
        const float4 q[4] = {c4[m * 4], c4[m * 4 + 1], c4[m * 4 + 2], c4[m * 4 + 3]};
        f += spline_value_m(q, F) * zm;
        zm *= F.zp[1];
    }
    return f;
}

// LDS floats of a round's workgroup: the 64 spots' powers x^o, y^n, z^m (12 each), the
// per-wave partial sums (K + 1 per spot), the round's spot indices, each wave's samples
// of one pixel row ([R][kSmStride]) and, with the coefficient gradient, each wave's
// weights of one pixel row ([R][64]).
// Only the NLL-only form stages the samples: the gradient form's arithmetic per pixel
// hides its per-lane sample loads (staged, it measured 15% slower: 0.72 vs 0.63 ms on
// 65,536 20x20 ROIs), the NLL form's does not (0.40 vs 0.56 ms).
constexpr int kSmStride = kWave + 1;  // the staging writes of consecutive pixels spread over the banks
constexpr size_t round_lds_floats(int R, bool grad) {
    return 12 * kWave + (size_t)kRoundWaves * 6 * kWave + kWave +
           (grad ? (size_t)kRoundWaves * R * kWave : (size_t)kRoundWaves * R * kSmStride);
}

// cf: the coefficient table (a.shared) as a restrict argument, through which the
// shared cell's wave-uniform loads become scalar loads: 256 bytes into SGPRs per
// pixel and wavefront, where a vector load delivers them to all 64 lanes' VGPRs.
template <bool GRAD>
__global__ void __launch_bounds__(kRoundBlock) spline_adjoint_round_kernel(SplineModel M, AdjArgs a,
                                                                           SplineAdjTiles w,
                                                                           const float* __restrict__ cf) {
    extern __shared__ float lds[];
    constexpr int K = 5;
    M.coefs = a.shared;
    const int R = a.R, S = R * R;
    // v is wave-uniform; readfirstlane tells the compiler, which keeps the row loop and
    // the shared cell's address in scalar registers
    const int lane = threadIdx.x & (kWave - 1), v = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    float* Pw = lds;                          // [12][64]: x^o (4), y^n (4), z^m (4) of each spot
    float* Ps = lds + 12 * kWave;             // [waves][K + 1][64] partial sums
    int* Sp = reinterpret_cast<int*>(Ps + kRoundWaves * 6 * kWave);  // [64] the round's spots
    // NLL form: this wave's samples of one pixel row, [R][65]; gradient form: its row
    // weights, [R][64]
    float* Sm = reinterpret_cast<float*>(Sp + kWave) + (size_t)v * R * kSmStride;
    float* Wt = reinterpret_cast<float*>(Sp + kWave) + (size_t)v * R * kWave;
    // row staging: element i = s R + x of a row's 64 R samples (spot s, pixel x), i = lane + 64 k
    const int sq = kWave / R, sr = kWave - sq * R;
    const int s0 = lane / R, x0 = lane - s0 * R;
    const int rows = (R + kRoundWaves - 1) / kRoundWaves;
    const int y0 = v * rows, y1 = min(R, y0 + rows);
    // lane q's basis index q = 16m + 4n + o (coefficient gradient)
    const int m = lane >> 4, nn = (lane >> 2) & 3, o = lane & 3;
    const int items = w.ntiles[0] * (kTileSpots / kWave);
    for (int item = blockIdx.x; item < items; item += gridDim.x) {
        const int2 td = w.tiles[item / (kTileSpots / kWave)];
        const int base = (item % (kTileSpots / kWave)) * kWave;
        if (base >= td.y) continue;  // the whole workgroup skips together
        const int cnt = min(kWave, td.y - base);
        const bool live = lane < cnt;
        const int spot = w.order[td.x + base + (live ? lane : 0)];
        const int key0 = __builtin_amdgcn_readfirstlane(w.key[w.order[td.x]]);
        const int i0 = key0 % M.Nx, j0 = (key0 / M.Nx) % M.Ny, k0 = key0 / (M.Nx * M.Ny);
        float t[K];
#pragma unroll
        for (int k = 0; k < K; k++) t[k] = a.params[(size_t)spot * K + k];
        SplineModel::Frame F;
        M.frame(t, R, F);
        const float go = a.gout ? a.gout[spot] : 1.0f;
        // the lane's own z-slab: the shared cells are the window's only for spots of the
        // window's slab (always, after the sort of this call; a reused order may be stale)
        const int ks = (int)((F.slab - M.coefs) / ((size_t)64 * M.Nx * M.Ny));
        if (v == 0) Sp[lane] = spot;
        if constexpr (GRAD) {
            if (v == 0) {
                Pw[0 * kWave + lane] = 1.0f;
                Pw[1 * kWave + lane] = F.x1;
                Pw[2 * kWave + lane] = F.x2;
                Pw[3 * kWave + lane] = F.x3;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    Pw[(4 + i) * kWave + lane] = F.yp[i];
                    Pw[(8 + i) * kWave + lane] = F.zp[i];
                }
            }
        }
        __syncthreads();
        float sums[K + 1];
#pragma unroll
        for (int i = 0; i < K + 1; i++) sums[i] = 0.0f;
        for (int y = y0; y < y1; y++) {
            if constexpr (GRAD) {
                // the frame again per row from an opaque copy of the parameters (an empty asm,
                // no instruction), so that its powers are not held through the row's outer
                // product, whose 64 basis values set the kernel's register peak
#pragma unroll
                for (int k = 0; k < K; k++) asm volatile("" : "+v"(t[k]));
                M.frame(t, R, F);
            }
            // the 64 spots' samples of row y into LDS, transposed to [x][spot]: each load
            // instruction reads runs of R consecutive floats of about 64 / R spots (reading
            // its own spot's sample per pixel, a lane would touch a line of its own per load)
            // (in chunks of 8 elements per lane, so that a chunk's index reads, loads and
            // stores are each in flight together)
            if constexpr (!GRAD) {
                int st = s0, xt = x0;
                for (int k0 = 0; k0 < R; k0 += 8) {
                    int si[8], xi[8], sp[8];
                    float sv[8];
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        si[j] = st;
                        xi[j] = xt;
                        xt += sr;
                        st += sq;
                        if (xt >= R) {
                            xt -= R;
                            st++;
                        }
                    }
#pragma unroll
                    for (int j = 0; j < 8; j++) sp[j] = k0 + j < R ? Sp[si[j]] : 0;
#pragma unroll
                    for (int j = 0; j < 8; j++)
                        sv[j] = k0 + j < R ? a.sample[(size_t)sp[j] * S + y * R + xi[j]] : 0.0f;
#pragma unroll
                    for (int j = 0; j < 8; j++)
                        if (k0 + j < R) Sm[xi[j] * kSmStride + si[j]] = sv[j];
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            }
            const int cyu = min(j0 + y, M.Ny - 1);
            const float* rowu = cf + 64 * ((size_t)(k0 * M.Ny + cyu) * M.Nx);
            const int cy = clampi((int)((float)y - F.ty), 0, M.Ny - 1);
            const float* rowo = cf + 64 * ((size_t)(ks * M.Ny + cy) * M.Nx);
            // the shared cell's coefficients arrive by scalar loads one z-power (16 floats)
            // at a time, and the next pixel's first two z-powers are issued as soon as this
            // pixel's last two are, so that their latency overlaps the rest of this pixel
            // instead of stalling the next one (the last pixel of a row loads a clamped,
            // unused cell); the arithmetic is SplineModel::cell's or spline_cell_value's
            float4 qa[4], qb[4];
            auto ld4 = [&](const float* c, int m, float4 (&q)[4]) {
                const float4* c4 = reinterpret_cast<const float4*>(c) + 4 * m;
                q[0] = c4[0], q[1] = c4[1], q[2] = c4[2], q[3] = c4[3];
            };
            ld4(rowu + 64 * min(i0, M.Nx - 1), 0, qa);
            ld4(rowu + 64 * min(i0, M.Nx - 1), 1, qb);
            for (int x = 0; x < R; x++) {
                const int cxu = min(i0 + x, M.Nx - 1);
                const int cx = M.col_index(F, x);
                const bool own = ks == k0 && cy == cyu && cx == cxu;  // the lane's cell is the shared one
                const float xs = GRAD ? a.sample[(size_t)spot * S + y * R + x] : Sm[x * kSmStride + lane];
                float W = 0.0f;
                if constexpr (GRAD) {
                    float f, fx, fy, fz, mu, J[K];
                    // the shared cell by scalar loads, a lane's own one (rare) 16 coefficients
                    // at a time: the same arithmetic in the same order, so a spot's results do
                    // not depend on which round it falls in
                    {
                        const float* cu = rowu + 64 * cxu;
                        const float* cn = rowu + 64 * min(i0 + x + 1, M.Nx - 1);
                        float fm[4], fxm[4], fym[4];
                        float4 qc[4], qd[4];
                        ld4(cu, 2, qc);
                        SplineModel::cell_m(qa, F, fm[0], fxm[0], fym[0]);
                        ld4(cu, 3, qd);
                        SplineModel::cell_m(qb, F, fm[1], fxm[1], fym[1]);
                        ld4(cn, 0, qa);
                        SplineModel::cell_m(qc, F, fm[2], fxm[2], fym[2]);
                        ld4(cn, 1, qb);
                        SplineModel::cell_m(qd, F, fm[3], fxm[3], fym[3]);
                        SplineModel::cell_combine(fm, fxm, fym, F, f, fx, fy, fz);
                    }
                    if (!own) spline_cell_by_z<1>(rowo + 64 * cx, F, f, fx, fy, fz);
                    SplineModel::finish(F, t, f, fx, fy, fz, x, y, R, mu, J);
                    if (f < 0.0f) J[0] = J[1] = J[2] = 0.0f;  // mu does not depend on position where clamped
                    const float wv = nll_weight(mu, xs, sums[K]);
#pragma unroll
                    for (int k = 0; k < K; k++) sums[k] += wv * J[k];
                    W = (live && f > 0.0f) ? go * wv * t[3] : 0.0f;
                    Wt[x * kWave + lane] = own ? W : 0.0f;
                    // a pixel in another cell: its 64 terms straight to that cell
                    unsigned long long any = a.gshared ? __ballot(!own && W != 0.0f) : 0ull;
                    while (any) {
                        const int src = __ffsll((long long)any) - 1;
                        any &= any - 1;
                        const float g = __shfl(W, src, kWave);
                        const int cs = __shfl((ks * M.Ny + cy) * M.Nx + cx, src, kWave);
                        const float b = Pw[(8 + m) * kWave + src] * Pw[(4 + nn) * kWave + src] * Pw[o * kWave + src];
                        atomicAdd(&a.gshared[(size_t)cs * 64 + lane], g * b);
                    }
                } else {
                    float f;
                    {
                        const float* cu = rowu + 64 * cxu;
                        const float* cn = rowu + 64 * min(i0 + x + 1, M.Nx - 1);
                        float4 qc[4], qd[4];
                        ld4(cu, 2, qc);
                        float zm = 1.0f;
                        f = 0.0f;
                        f += spline_value_m(qa, F) * zm;
                        zm *= F.zp[1];
                        ld4(cu, 3, qd);
                        f += spline_value_m(qb, F) * zm;
                        zm *= F.zp[1];
                        ld4(cn, 0, qa);
                        f += spline_value_m(qc, F) * zm;
                        zm *= F.zp[1];
                        ld4(cn, 1, qb);
                        f += spline_value_m(qd, F) * zm;
                    }
                    if (!own) f = spline_cell_value<1>(rowo + 64 * cx, F);
                    const float psf = f < 0.0f ? 0.0f : f;
                    nll_weight(t[3] * psf + t[4], xs, sums[K]);
                }
            }
            if constexpr (GRAD) {
                if (a.gshared) {
                    // the row's weights are in LDS before any lane reads them (one wavefront:
                    // workgroup-scope fences order its LDS operations)
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                    // lane q: B_s[q] = z_s^m y_s^n x_s^o of every spot s of the round, formed
                    // here per row (the fences keep it from being held over the pixel loop)
                    float B[kWave];
#pragma unroll
                    for (int s4 = 0; s4 < kWave; s4 += 4) {
                        const float4 zq = *reinterpret_cast<const float4*>(Pw + (8 + m) * kWave + s4);
                        const float4 yq = *reinterpret_cast<const float4*>(Pw + (4 + nn) * kWave + s4);
                        const float4 xq = *reinterpret_cast<const float4*>(Pw + o * kWave + s4);
                        B[s4] = zq.x * yq.x * xq.x;
                        B[s4 + 1] = zq.y * yq.y * xq.y;
                        B[s4 + 2] = zq.z * yq.z * xq.z;
                        B[s4 + 3] = zq.w * yq.w * xq.w;
                    }
                    for (int x = 0; x < R; x++) {
                        const float4* wc = reinterpret_cast<const float4*>(Wt + x * kWave);
                        float g = 0.0f;
#pragma unroll
                        for (int s4 = 0; s4 < kWave / 4; s4++) {
                            const float4 wv = wc[s4];
                            g += wv.x * B[4 * s4] + wv.y * B[4 * s4 + 1] + wv.z * B[4 * s4 + 2] + wv.w * B[4 * s4 + 3];
                        }
                        if (g != 0.0f) {
                            const int cxu = min(i0 + x, M.Nx - 1);
                            atomicAdd(&a.gshared[((size_t)(k0 * M.Ny + cyu) * M.Nx + cxu) * 64 + lane], g);
                        }
                    }
                }
            }
            // every lane's reads complete before the next row overwrites the samples and weights
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
        // the spot's sums over the waves' row bands, in a fixed order
#pragma unroll
        for (int i = 0; i < K + 1; i++) Ps[(v * (K + 1) + i) * kWave + lane] = sums[i];
        __syncthreads();
        if (v == 0 && live) {
            float tot[K + 1];
#pragma unroll
            for (int i = 0; i < K + 1; i++) {
                tot[i] = Ps[i * kWave + lane];
                for (int u = 1; u < kRoundWaves; u++) tot[i] += Ps[(u * (K + 1) + i) * kWave + lane];
            }
            if (a.nll) a.nll[spot] = tot[K];
            if (GRAD && a.gparams) {
#pragma unroll
                for (int k = 0; k < K; k++) a.gparams[(size_t)spot * K + k] = go * tot[k];
            }
        }
        __syncthreads();  // Pw and Ps are reused by the next round
    }
}


size_t spline_adjoint_workspace(int n, int nkeys) {
    const size_t maxTiles = (size_t)(n + kTileSpots - 1) / kTileSpots + (size_t)std::min(n, nkeys);
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    return al((size_t)n * 4) * 2 + al((size_t)nkeys * 4) * 2 + al(maxTiles * 8) + al(8);
}

static SplineAdjTiles carve_tiles(void* ws, int n, int nkeys) {
    const size_t maxTiles = (size_t)(n + kTileSpots - 1) / kTileSpots + (size_t)std::min(n, nkeys);
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    char* p = static_cast<char*>(ws);
    SplineAdjTiles w{};
    w.key = reinterpret_cast<int*>(p); p += al((size_t)n * 4);
    w.order = reinterpret_cast<int*>(p); p += al((size_t)n * 4);
    w.count = reinterpret_cast<int*>(p); p += al((size_t)nkeys * 4);
    w.offset = reinterpret_cast<int*>(p); p += al((size_t)nkeys * 4);
    w.tiles = reinterpret_cast<int2*>(p); p += al(maxTiles * 8);
    w.ntiles = reinterpret_cast<int*>(p);
    w.nkeys = nkeys;
    w.maxTiles = (int)maxTiles;
    return w;
}

bool spline_adjoint_tiled(int R) { return R >= 1 && R <= kMaxTiledR; }

static hipError_t launch_spline_tiled(const SplineModel& M, const AdjArgs& a, void* ws, bool sorted,
                                      hipStream_t st) {
    const int nkeys = M.Nz * M.Ny * M.Nx;
    SplineAdjTiles w = carve_tiles(ws, a.n, nkeys);
    const unsigned spotBlocks = (unsigned)((a.n + 255) / 256);
    const unsigned keyBlocks = (unsigned)std::min(1024, (nkeys + 255) / 256);
    if (!sorted) {
        // counts and the two allocation counters start at zero (fill kernels, no memset:
        // capturable)
        hipLaunchKernelGGL(adj_fill_kernel, dim3(keyBlocks), dim3(256), 0, st, w.count, nkeys, 0);
        hipLaunchKernelGGL(adj_fill_kernel, dim3(1), dim3(64), 0, st, w.ntiles, 2, 0);
        hipLaunchKernelGGL(adj_key_kernel, dim3(spotBlocks), dim3(256), 0, st, M, a, w);
        hipLaunchKernelGGL(adj_alloc_kernel, dim3(keyBlocks), dim3(256), 0, st, w);
        hipLaunchKernelGGL(adj_scatter_kernel, dim3(spotBlocks), dim3(256), 0, st, a, w);
    }
    const bool grad = a.gshared || a.gparams;
    const size_t lds = round_lds_floats(a.R, grad) * sizeof(float);
    auto go = [&](auto kfn, int* cache) -> hipError_t {
        if (*cache == 0) *cache = resident_blocks(reinterpret_cast<const void*>(kfn), kRoundBlock, lds);
        if (*cache <= 0) return hipErrorNoDevice;
        // one work item per 64 spots of a tile: at most maxTiles x 4
        const int blocks = std::max(1, std::min(*cache, w.maxTiles * (kTileSpots / kWave)));
        hipLaunchKernelGGL(kfn, dim3(blocks), dim3(kRoundBlock), lds, st, M, a, w, a.shared);
        return hipGetLastError();
    };
    // resident workgroups per roisize and form (one device model)
    static int residentGrad[kMaxTiledR + 1] = {}, residentNll[kMaxTiledR + 1] = {};
    return grad ? go(spline_adjoint_round_kernel<true>, &residentGrad[a.R])
                : go(spline_adjoint_round_kernel<false>, &residentNll[a.R]);
}

hipError_t launch_adjoint_any(const ModelDesc& d, const AdjArgs& a, void* ws, bool sorted, hipStream_t st) {
    if (d.kind == kKindSpline && ws && spline_adjoint_tiled(a.R)) {
        SplineModel M = d.sm;
        return launch_spline_tiled(M, a, ws, sorted, st);
    }
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
