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
// Mapping: one wavefront per spot, lanes stride over pixels (any roisize).  Scalar
// shared gradients are reduced per block in LDS and added to HBM once per block.
// The spline coefficient gradient of roisize <= 28 is formed by window rounds
// (spline_adjoint_tiled_kernel, below); larger ROIs scatter it pixel by pixel.
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
// pair of z-powers at a time (32 coefficients in registers instead of 64): the tiled
// kernel's 1024-thread blocks leave it 128 VGPRs.  The same polynomial as SplineModel::cell,
// summed over z by Horner-free running powers.
PP_DEV void spline_cell_lean(const float* c, const SplineModel::Frame& F, float& f, float& fx, float& fy,
                             float& fz) {
    const float4* c4 = reinterpret_cast<const float4*>(c);
    const float z = F.zp[1];
    float zm = 1.0f, dzm = 0.0f;  // z^m and m z^(m-1)
    f = fx = fy = fz = 0.0f;
#pragma unroll 2
    for (int m = 0; m < 4; m++) {
        float a = 0.0f, ax = 0.0f, ay = 0.0f;
#pragma unroll
        for (int n = 0; n < 4; n++) {
            const float4 q = c4[m * 4 + n];
            const float px = q.x + q.y * F.x1 + q.z * F.x2 + q.w * F.x3;
            const float dpx = q.y + q.z * F.dx2 + q.w * F.dx3;
            a += px * F.yp[n];
            ax += dpx * F.yp[n];
            ay += px * F.dyp[n];
        }
        f += a * zm;
        fx += ax * zm;
        fy += ay * zm;
        fz += a * dzm;
        dzm = dzm * z + zm;
        zm *= z;
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
// Coefficient gradient by window rounds (roisize <= kMaxTiledR).
//
// Spot s adds  W_p z_s^m y_s^n x_s^o  to coefficient (m, n, o) of the cell of each of
// its pixels p, with W_p = g (1 - x/mu) I (zero where the spline value is clamped) and
// (x_s, y_s, z_s) the spot's sub-voxel position.  All pixels of a spot share that
// position, and their cells lie in the (R+1) x (R+1) window of z-slab k starting at the
// cell of pixel (0, 0).  So the gradient of a window cell c is a sum over the spots
// with that window of W_s[c] B_s[q] (B_s = the spot's 64 basis values): an outer
// product.  Spots are sorted by window (k, j0, i0) and cut into tiles of up to
// kTileSpots spots of one window.  A workgroup of 16 wavefronts takes a tile in rounds
// of 16 x SPW spots (each wavefront evaluates SPW spots, its lanes over the pixels),
// scatters their W_s[c] into LDS, forms the round's (R+1)^2 x 64 gradient entries
// from LDS (thread = one coefficient of cells v, v + 16, ...) and adds each to HBM
// with one atomic: (R+1)^2 x 64 atomics per round of up to 16 SPW spots instead of
// R^2 x 64 per spot.  Pixels whose cell falls outside the window (none, by the bound
// above; kept as a guard) add straight to HBM as the direct kernel does.
constexpr int kMaxTiledR = 28;
constexpr int kWaves = 16;                   // wavefronts per block
constexpr int kTileSpots = 256;              // spots per tile (a multiple of every round)
constexpr int kTiledBlock = kWaves * kWave;

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

// One block of 1024 threads: exclusive scans of the spot counts (-> first sorted
// position of each key) and of the tile counts, the tile table, and the count array
// reset to zero for use as the scatter cursor.  Thread i owns keys [i*per, (i+1)*per).
__global__ void __launch_bounds__(1024) adj_scan_kernel(SplineAdjTiles w) {
    __shared__ int sc[1024], st[1024];
    const int tid = threadIdx.x;
    const int per = (w.nkeys + 1023) / 1024;
    const int k0 = min(w.nkeys, tid * per), k1 = min(w.nkeys, k0 + per);
    int c = 0, tl = 0;
    for (int k = k0; k < k1; k++) {
        const int n = w.count[k];
        c += n;
        tl += (n + kTileSpots - 1) / kTileSpots;
    }
    sc[tid] = c;
    st[tid] = tl;
    __syncthreads();
    // inclusive Hillis-Steele scan of both sums
    for (int d = 1; d < 1024; d <<= 1) {
        const int vc = tid >= d ? sc[tid - d] : 0, vt = tid >= d ? st[tid - d] : 0;
        __syncthreads();
        sc[tid] += vc;
        st[tid] += vt;
        __syncthreads();
    }
    int pos = sc[tid] - c, tile = st[tid] - tl;
    for (int k = k0; k < k1; k++) {
        const int n = w.count[k];
        w.offset[k] = pos;
        for (int f = 0; f < n; f += kTileSpots) w.tiles[tile++] = make_int2(pos + f, min(kTileSpots, n - f));
        pos += n;
        w.count[k] = 0;
    }
    if (tid == 1023) w.ntiles[0] = st[1023];
}

// spots in key order (the order inside a key is arbitrary)
__global__ void __launch_bounds__(256) adj_scatter_kernel(AdjArgs a, SplineAdjTiles w) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.n) return;
    const int key = w.key[s];
    w.order[w.offset[key] + atomicAdd(&w.count[key], 1)] = s;
}

// Per round of RS = 16 SPW spots, LDS holds per window cell c the round's RS weights
// W_s[c] (at position s ^ (c & 15) inside the cell's RS floats, so that a wavefront's
// scatter over consecutive cells spreads over the banks) and the round's basis vectors
// B[s][q].  Thread (wave v, lane q) then forms the entries (cell v + 16 i, coefficient q).
template <int SPW>
__global__ void __launch_bounds__(kTiledBlock) spline_adjoint_tiled_kernel(SplineModel M, AdjArgs a,
                                                                           SplineAdjTiles w) {
    extern __shared__ float lds[];
    constexpr int K = 5, RS = kWaves * SPW;
    M.coefs = a.shared;
    const int R = a.R, S = R * R, WR = R + 1, ncells = WR * WR;
    float* Wt = lds;                         // [ncells][RS]
    float* Bs = lds + (size_t)ncells * RS;   // [RS][64]
    const int lane = threadIdx.x & (kWave - 1), v = threadIdx.x / kWave;
    const int m = lane >> 4, nn = (lane >> 2) & 3, o = lane & 3;
    const int ntiles = w.ntiles[0];
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int2 td = w.tiles[tile];
        const int first = td.x, cnt = td.y;
        const int key0 = w.key[w.order[first]];
        const int i0 = key0 % M.Nx, j0 = (key0 / M.Nx) % M.Ny, k0 = key0 / (M.Nx * M.Ny);
        for (int c0 = 0; c0 < cnt; c0 += RS) {
            for (int i = threadIdx.x; i < ncells * RS; i += kTiledBlock) Wt[i] = 0.0f;
            __syncthreads();
#pragma unroll 1
            for (int u = 0; u < SPW; u++) {
                const int sl = v + kWaves * u;  // the spot's column in this round
                const int slot = c0 + sl;
                float basis = 0.0f;
                if (slot < cnt) {
                    const int spot = w.order[first + slot];
                    float t[K];
#pragma unroll
                    for (int k = 0; k < K; k++) t[k] = a.params[(size_t)spot * K + k];
                    SplineModel::Frame F;
                    M.frame(t, R, F);
                    const float go = a.gout ? a.gout[spot] : 1.0f;
                    {
                        const float zb = m == 0 ? 1.0f : m == 1 ? F.zp[1] : m == 2 ? F.zp[2] : F.zp[3];
                        const float yb = nn == 0 ? 1.0f : nn == 1 ? F.yp[1] : nn == 2 ? F.yp[2] : F.yp[3];
                        const float xb = o == 0 ? 1.0f : o == 1 ? F.x1 : o == 2 ? F.x2 : F.x3;
                        basis = zb * yb * xb;
                    }
                    float sums[K + 1];
#pragma unroll
                    for (int i = 0; i < K + 1; i++) sums[i] = 0.0f;
                    const int kSpot = (int)((F.slab - M.coefs) / ((size_t)64 * M.Nx * M.Ny));
                    // every lane runs every trip (the guard path below is a wavefront operation)
                    for (int p0 = 0; p0 < S; p0 += kWave) {
                        const int p = p0 + lane;
                        const bool valid = p < S;
                        const int y = valid ? p / R : 0, x = valid ? p - y * R : 0;
                        const int ci = M.col_index(F, x);
                        float W = 0.0f;
                        bool outside = false;
                        if (valid) {
                            float f, fx, fy, fz, mu, J[K];
                            spline_cell_lean(M.row_ptr(F, y) + 64 * ci, F, f, fx, fy, fz);
                            SplineModel::finish(F, t, f, fx, fy, fz, x, y, R, mu, J);
                            if (f < 0.0f) J[0] = J[1] = J[2] = 0.0f;  // mu does not depend on position where clamped
                            const float wv = nll_weight(mu, a.sample[(size_t)spot * S + p], sums[K]);
#pragma unroll
                            for (int k = 0; k < K; k++) sums[k] += wv * J[k];
                            W = (f > 0.0f) ? go * wv * t[3] : 0.0f;
                        }
                        if (W != 0.0f) {
                            const int cj = clampi((int)((float)y - F.ty), 0, M.Ny - 1) - j0;
                            const int cx = ci - i0;
                            if (kSpot == k0 && cx >= 0 && cx < WR && cj >= 0 && cj < WR) {
                                const int c = cj * WR + cx;
                                atomicAdd(&Wt[c * RS + (sl ^ (c & 15))], W);
                            } else {
                                outside = true;
                            }
                        }
                        // guard path: the wavefront adds a flagged pixel's 64 terms straight to HBM
                        unsigned long long any = __ballot(outside);
                        while (any) {
                            const int src = __ffsll((long long)any) - 1;
                            any &= any - 1;
                            const float g = __shfl(W, src, kWave);
                            const int ys = __shfl(y, src, kWave), xs = __shfl(x, src, kWave);
                            const int cell = (int)((M.row_ptr(F, ys) + 64 * M.col_index(F, xs) - M.coefs) >> 6);
                            atomicAdd(&a.gshared[(size_t)cell * 64 + lane], g * basis);
                        }
                    }
                    group_sum_n<kWave>(sums);
                    if (lane == 0) {
                        if (a.nll) a.nll[spot] = sums[K];
                        if (a.gparams) {
#pragma unroll
                            for (int k = 0; k < K; k++) a.gparams[(size_t)spot * K + k] = go * sums[k];
                        }
                    }
                }
                Bs[sl * kWave + lane] = basis;
            }
            __syncthreads();
            // this thread's cells are v + 16 i, whose swizzle is v: spot s sits at s ^ v
            float Bq[RS];
#pragma unroll
            for (int s2 = 0; s2 < RS; s2++) Bq[s2] = Bs[(s2 ^ v) * kWave + lane];
            for (int c = v; c < ncells; c += kWaves) {
                const float4* wc = reinterpret_cast<const float4*>(Wt + (size_t)c * RS);
                float sum = 0.0f;
#pragma unroll
                for (int q4 = 0; q4 < RS / 4; q4++) {
                    const float4 wv = wc[q4];
                    sum += wv.x * Bq[4 * q4] + wv.y * Bq[4 * q4 + 1] + wv.z * Bq[4 * q4 + 2] + wv.w * Bq[4 * q4 + 3];
                }
                const int gi = i0 + c % WR, gj = j0 + c / WR;
                if (sum != 0.0f && gi < M.Nx && gj < M.Ny)
                    atomicAdd(&a.gshared[(((size_t)k0 * M.Ny + gj) * M.Nx + gi) * 64 + lane], sum);
            }
            __syncthreads();
        }
    }
}

// Spots per wavefront and round: 2 where the round's weights fit in 64 KiB of LDS.
static int tiled_spw(int R) {
    const size_t cells = (size_t)(R + 1) * (R + 1);
    return (cells * 2 * kWaves + 2 * kWaves * kWave) * 4 <= 65536 ? 2 : 1;
}


size_t spline_adjoint_workspace(int n, int nkeys) {
    const size_t maxTiles = (size_t)(n + kTileSpots - 1) / kTileSpots + (size_t)std::min(n, nkeys);
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    return al((size_t)n * 4) * 2 + al((size_t)nkeys * 4) * 2 + al(maxTiles * 8) + al(4);
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

static hipError_t launch_spline_tiled(const SplineModel& M, const AdjArgs& a, void* ws, hipStream_t st) {
    const int nkeys = M.Nz * M.Ny * M.Nx;
    SplineAdjTiles w = carve_tiles(ws, a.n, nkeys);
    const unsigned spotBlocks = (unsigned)((a.n + 255) / 256);
    // count and the tile counter start at zero (a fill kernel, no memset: capturable)
    hipLaunchKernelGGL(adj_fill_kernel, dim3(std::min(1024, (nkeys + 255) / 256)), dim3(256), 0, st, w.count, nkeys, 0);
    hipLaunchKernelGGL(adj_key_kernel, dim3(spotBlocks), dim3(256), 0, st, M, a, w);
    hipLaunchKernelGGL(adj_scan_kernel, dim3(1), dim3(1024), 0, st, w);
    hipLaunchKernelGGL(adj_scatter_kernel, dim3(spotBlocks), dim3(256), 0, st, a, w);
    const int WR = a.R + 1, spw = tiled_spw(a.R);
    const size_t lds = ((size_t)WR * WR * kWaves * spw + (size_t)kWaves * spw * kWave) * sizeof(float);
    auto go = [&](auto kfn) -> hipError_t {
        static int resident[kMaxTiledR + 1] = {};  // persistent grid per roisize (one device model)
        if (resident[a.R] == 0) resident[a.R] = resident_blocks(reinterpret_cast<const void*>(kfn), kTiledBlock, lds);
        if (resident[a.R] <= 0) return hipErrorNoDevice;
        const int blocks = std::max(1, std::min(resident[a.R], w.maxTiles));
        hipLaunchKernelGGL(kfn, dim3(blocks), dim3(kTiledBlock), lds, st, M, a, w);
        return hipGetLastError();
    };
    return spw == 2 ? go(spline_adjoint_tiled_kernel<2>) : go(spline_adjoint_tiled_kernel<1>);
}

hipError_t launch_adjoint_any(const ModelDesc& d, const AdjArgs& a, void* ws, hipStream_t st) {
    if (d.kind == kKindSpline && a.gshared && ws && spline_adjoint_tiled(a.R)) {
        SplineModel M = d.sm;
        return launch_spline_tiled(M, a, ws, st);
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
