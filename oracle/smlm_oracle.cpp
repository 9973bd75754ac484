// ORACLE — test infrastructure only.
//
// CPU restatement of photonpy's batched Poisson-MLE spot fitter, used as the
// checker for the HIP implementation in photonpy_amd/.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
// library; nothing in the product path links or calls it.
//
// Every routine follows the reference's operation order for T=float so that
// the float instantiation reproduces the reference CPU estimator
// (PSFImpl, cpp/SMLMLib/Estimators/EstimatorImpl.h:108-185).  The double
// instantiation evaluates the same mathematics in FP64 and is used to state
// each model's error against an FP64 fit.
//
// Followed reference code (paths relative to the reference root):
//   ComputeCOM                 cpp/SMLMLib/Estimators/Estimation.h:336-352
//   ComputeLevMarAlphaBeta     cpp/SMLMLib/Estimators/Estimation.h:104-134
//   LevMarOptimize             cpp/SMLMLib/Estimators/Estimation.h:194-251
//   ComputeChiSquareAndCRLB    cpp/SMLMLib/Estimators/Estimation.h:63-102
//   ComputeLogLikelihood       cpp/SMLMLib/Estimators/Estimation.h:138-156
//   Cholesky / SolveCholesky   cpp/SMLMLib/SolveMatrix.h:107-125, 145-163
//   LUPDecompose / LUPInvert   cpp/SMLMLib/SolveMatrix.h:219-263, 344-365
//   InvertMatrix (tol 1e-9)    cpp/SMLMLib/SolveMatrix.h:411-452
//   clamp (fmaxf(a,fminf(v,b))) cpp/SMLMLib/CudaMath.h:1152-1155 (semantics only)
//   Gauss2D_Model_XYIBg        cpp/SMLMLib/Estimators/Gaussian/GaussianPSFModels.h:35-152
//   Gauss2D_Model_XYIBgSigma   cpp/SMLMLib/Estimators/Gaussian/GaussianPSFModels.h:308-387
//   Gauss2D_Model_XYZIBg       cpp/SMLMLib/Estimators/Gaussian/GaussianPSFModels.h:636-751
//     (+ quartic term B*t^4/d^2 inside the square root; B=0 is the reference)
//   Gauss2D_Model_XYIBgConstSigma   GaussianPSFModels.h:155-229
//   Gauss2D_Model_XYITiltedBg       GaussianPSFModels.h:232-304
//   Gauss2D_Model_XYIBgSigmaXY      GaussianPSFModels.h:388-467
//   Gauss2D_Model_XYITiltedBgSigmaXY GaussianPSFModels.h:471-553 (mu without the tilt, as there)
//   IntensityBgModel + Gauss2D_EstimateIntensityBg  GaussianPSFModels.h:557-623, GaussianPSF.cu:216-240
//   GLRT_PSF diagnostics {ll_on, ll_off, mean}      GLRT_PSF.cu:103-129
//   ComputeFisherMatrix / ComputeCRLB (|diag|)      Estimation.h:22-60, 318-333
//   CSpline_Model_XYZIBg       cpp/SMLMLib/Estimators/CSpline/CubicSplinePSFModels.h:14-154
//   CSpline_Model_XYZITiltedBg CubicSplinePSFModels.h:157-298 (x,y,z,I,bg,bgx,bgy)
//   CSpline_Model_XYZIBgImage  CubicSplinePSFModels.h:300-443 (background image per spot, R*R constants)
//   CSpline_Compute            cpp/SMLMLib/Estimators/CSpline/CubicSplinePSF.cu:151-204
//
// The reference's own per-spot std::async scheduler (ThreadUtils.h:22-32) is
// replaced by a fixed pool with dynamic chunks so the CPU baseline scales.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------- helpers
template <typename T> inline T clampf_(T v, T a, T b) { return std::fmax(a, std::fmin(v, b)); }
inline int clampi_(int v, int a, int b) { return std::max(a, std::min(v, b)); }

template <typename F>
void pool_for(int n, int nthreads, F fn) {
    if (nthreads <= 1 || n < 64) {
        for (int i = 0; i < n; i++) fn(i);
        return;
    }
    // chunks small enough that every thread gets several (per-spot cost varies a lot)
    const int chunk = std::max(1, std::min(256, n / (nthreads * 8)));
    std::atomic<int> next{0};
    auto worker = [&]() {
        for (;;) {
            int s = next.fetch_add(chunk);
            if (s >= n) break;
            int e = std::min(n, s + chunk);
            for (int i = s; i < e; i++) fn(i);
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads - 1; t++) th.emplace_back(worker);
    worker();
    for (auto& t : th) t.join();
}

// MathUtils.h:95 defines pi as a float literal; the FP64 instantiation uses the double value.
template <typename T> constexpr T kPiT = T(3.141592653589793);
template <> constexpr float kPiT<float> = 3.141592653589793f;

// ------------------------------------------------------------- model desc
// kind: 0 = 2D Gaussian fixed sigma (x,y,I,bg)
//       1 = 2D Gaussian free sigma (x,y,I,bg,sigma)
//       2 = astigmatic 3D Gaussian (x,y,z,I,bg)
//       3 = cubic spline 3D (x,y,z,I,bg)
//       4 = 2D Gaussian, sigma x,y per-spot constants (x,y,I,bg)
//       5 = tilted background, sigma x,y per-spot constants (x,y,I,bg,bgx,bgy)
//       6 = free sigma x and y (x,y,I,bg,sx,sy)
//       7 = tilted background + free sigma x and y (x,y,I,bg,bgx,bgy,sx,sy)
//       8 = intensity / background at per-spot constants x, y, sigma (I,bg)
//       9 = cubic spline 3D, tilted background (x,y,z,I,bg,bgx,bgy)
//      10 = cubic spline 3D, background image per spot (x,y,z,I,bg; R*R constants)
struct OrModel {
    int kind;
    int roisize;
    double p[16]; // 0: sx,sy | 1: initialSigma | 2: x[4],y[4],minz,maxz,Bx,By | 3,9,10: zmin,zmax | 6,7: isx,isy
    int shape[3]; // spline (Nz,Ny,Nx)
    const float* coefs;
};

template <typename T>
struct Gauss2DFixed {
    enum { K = 4, NC = 0, CrlbAbs = 0 };
    const T* cst = nullptr;
    int R;
    T sx, sy;
    static T StopLimit(int k) { const T s[4] = {T(1e-5f), T(1e-5f), T(1e-1f), T(1e-5f)}; return s[k]; }
    void CheckLimits(T* t) const {
        t[0] = clampf_<T>(t[0], T(1.5f), T(R) - T(1.5f) - T(1));
        t[1] = clampf_<T>(t[1], T(1.5f), T(R) - T(1.5f) - T(1));
        if (t[2] < T(25.0f)) t[2] = T(25.0f);
        if (t[3] < T(0.0f)) t[3] = T(0.0f);
    }
    void Initial(const T com[3], T* t) const { t[0] = com[0]; t[1] = com[1]; t[2] = com[2] * T(0.7f); t[3] = 0; }
    template <typename CB> void Eval(const T* th, CB cb) const {
        const T cpx = T(1.0f) / (std::sqrt(T(2) * kPiT<T>) * sx);
        const T c2x = T(1.0f) / (std::sqrt(T(2)) * sx);
        const T cpy = T(1.0f) / (std::sqrt(T(2) * kPiT<T>) * sy);
        const T c2y = T(1.0f) / (std::sqrt(T(2)) * sy);
        const T tx = th[0], ty = th[1], tI = th[2], tbg = th[3];
        for (int y = 0; y < R; y++) {
            T a0 = (T(y) - ty + T(.5f)) * c2y;
            T a1 = (T(y) - ty - T(.5f)) * c2y;
            T DY = T(0.5f) * std::erf(a0) - T(0.5f) * std::erf(a1);
            T dEy = cpy * (std::exp(-a1 * a1) - std::exp(-a0 * a0));
            for (int x = 0; x < R; x++) {
                T b0 = (T(x) - tx + T(.5f)) * c2x;
                T b1 = (T(x) - tx - T(.5f)) * c2x;
                T DX = T(0.5f) * std::erf(b0) - T(0.5f) * std::erf(b1);
                T dEx = cpx * (std::exp(-b1 * b1) - std::exp(-b0 * b0));
                T mu = tbg + tI * DX * DY;
                const T J[4] = {tI * dEx * DY, tI * dEy * DX, DX * DY, T(1)};
                cb(y * R + x, mu, J);
            }
        }
    }
};

template <typename T>
struct Gauss2DSigma {
    enum { K = 5, NC = 0, CrlbAbs = 0 };
    const T* cst = nullptr;
    int R;
    T initialSigma;
    static T StopLimit(int k) { const T s[5] = {T(1e-5f), T(1e-5f), T(1e-1f), T(1e-4f), T(1e-5f)}; return s[k]; }
    void CheckLimits(T* t) const {
        t[0] = clampf_<T>(t[0], T(1.5f), T(R) - T(1.5f) - T(1));
        t[1] = clampf_<T>(t[1], T(1.5f), T(R) - T(1.5f) - T(1));
        if (t[2] < T(10.0f)) t[2] = T(10.0f);
        if (t[3] < T(0.0f)) t[3] = T(0.0f);
        if (t[4] < T(1.0f)) t[4] = T(1.0f);
        if (t[4] > T(R) * T(0.5f)) t[4] = T(R) * T(0.5f);
    }
    void Initial(const T com[3], T* t) const { t[0] = com[0]; t[1] = com[1]; t[2] = com[2]; t[3] = 0; t[4] = initialSigma; }
    template <typename CB> void Eval(const T* th, CB cb) const {
        const T tx = th[0], ty = th[1], tI = th[2], tbg = th[3], ts = th[4];
        const T isp = T(1.0f) / std::sqrt(kPiT<T>);
        const T cp = T(1.0f) / (std::sqrt(T(2) * kPiT<T>) * ts);
        const T c2 = T(1.0f) / (std::sqrt(T(2)) * ts);
        for (int y = 0; y < R; y++) {
            T a0 = (T(y) - ty + T(.5f)) * c2;
            T a1 = (T(y) - ty - T(.5f)) * c2;
            T DY = T(0.5f) * std::erf(a0) - T(0.5f) * std::erf(a1);
            T e1 = std::exp(-a1 * a1), e0 = std::exp(-a0 * a0);
            T dEy = cp * (e1 - e0);
            T dDYs = (e1 * a1 - e0 * a0) * isp;
            for (int x = 0; x < R; x++) {
                T b0 = (T(x) - tx + T(.5f)) * c2;
                T b1 = (T(x) - tx - T(.5f)) * c2;
                T DX = T(0.5f) * std::erf(b0) - T(0.5f) * std::erf(b1);
                T f1 = std::exp(-b1 * b1), f0 = std::exp(-b0 * b0);
                T dEx = cp * (f1 - f0);
                T dDXs = (f1 * b1 - f0 * b0) * isp;
                T mu = tbg + tI * DX * DY;
                const T J[5] = {tI * dEx * DY, tI * dEy * DX, DX * DY, T(1), tI * (dDXs * DY + dDYs * DX)};
                cb(y * R + x, mu, J);
            }
        }
    }
};

template <typename T>
struct Astig3D {
    enum { K = 5, NC = 0, CrlbAbs = 0 };
    const T* cst = nullptr;
    int R;
    T cx[4], cy[4], minz, maxz, Bx, By;
    static T StopLimit(int k) { const T s[5] = {T(1e-4f), T(1e-4f), T(1e-1f), T(1e-4f), T(1e-4f)}; return s[k]; }
    void CheckLimits(T* t) const {
        t[0] = clampf_<T>(t[0], T(1.5f), T(R) - T(1.5f) - T(1));
        t[1] = clampf_<T>(t[1], T(1.5f), T(R) - T(1.5f) - T(1));
        t[2] = clampf_<T>(t[2], minz, maxz);
        t[3] = std::fmax(t[3], T(10.f));
        t[4] = std::fmax(t[4], T(0.0f));
    }
    void Initial(const T com[3], T* t) const { t[0] = com[0]; t[1] = com[1]; t[2] = 0; t[3] = com[2]; t[4] = T(0.01f); }
    template <typename CB> void Eval(const T* th, CB cb) const {
        const T tx = th[0], ty = th[1], tz = th[2], tI = th[3], tbg = th[4];
        const T s0x = cx[0], gx = cx[1], dx = cx[2], Ax = cx[3];
        const T s0y = cy[0], gy = cy[1], dy = cy[2], Ay = cy[3];
        const T dy2 = dy * dy, dx2 = dx * dx;
        const T zx = tz - gx, zx2 = zx * zx, zx3 = zx2 * zx, zx4 = zx3 * zx;
        const T zy = tz - gy, zy2 = zy * zy, zy3 = zy2 * zy, zy4 = zy3 * zy;
        const T qx = T(1.0f) + zx2 / dx2 + Ax * zx3 / dx2 + Bx * zx4 / dx2;
        const T qy = T(1.0f) + zy2 / dy2 + Ay * zy3 / dy2 + By * zy4 / dy2;
        const T sgx = s0x * std::sqrt(qx);
        const T sgy = s0y * std::sqrt(qy);
        const T cpx = T(1.0f) / (std::sqrt(T(2) * kPiT<T>) * sgx);
        const T c2x = T(1.0f) / (std::sqrt(T(2)) * sgx);
        const T cpy = T(1.0f) / (std::sqrt(T(2) * kPiT<T>) * sgy);
        const T c2y = T(1.0f) / (std::sqrt(T(2)) * sgy);
        for (int y = 0; y < R; y++) {
            T a0 = (T(y) - ty + T(.5f)) * c2y;
            T a1 = (T(y) - ty - T(.5f)) * c2y;
            T DY = T(0.5f) * std::erf(a0) - T(0.5f) * std::erf(a1);
            T dEy = cpy * (std::exp(-a1 * a1) - std::exp(-a0 * a0));
            T ym = T(y) - ty - T(0.5f), yp = T(y) - ty + T(0.5f);
            T G21y = T(1) / (std::sqrt(T(2.0f) * kPiT<T>) * sgy * sgy) *
                     (ym * std::exp(-ym * ym / (T(2.0f) * sgy * sgy)) - yp * std::exp(-yp * yp / (T(2.0f) * sgy * sgy)));
            for (int x = 0; x < R; x++) {
                T b0 = (T(x) - tx + T(.5f)) * c2x;
                T b1 = (T(x) - tx - T(.5f)) * c2x;
                T DX = T(0.5f) * std::erf(b0) - T(0.5f) * std::erf(b1);
                T dEx = cpx * (std::exp(-b1 * b1) - std::exp(-b0 * b0));
                T mu = tbg + tI * DX * DY;
                T dmdx = tI * dEx * DY;
                T dmdy = tI * dEy * DX;
                T xm = T(x) - tx - T(0.5f), xp = T(x) - tx + T(0.5f);
                T G21x = T(1) / (std::sqrt(T(2.0f) * kPiT<T>) * sgx * sgx) *
                         (xm * std::exp(-xm * xm / (T(2.0f) * sgx * sgx)) - xp * std::exp(-xp * xp / (T(2.0f) * sgx * sgx)));
                T dMsx = tI * DY * G21x;
                T dMsy = tI * DX * G21y;
                T dSxz = s0x * (T(2) * zx / dx2 + Ax * T(3) * zx2 / dx2 + Bx * T(4) * zx3 / dx2) / (T(2) * std::sqrt(qx));
                T dSyz = s0y * (T(2) * zy / dy2 + Ay * T(3) * zy2 / dy2 + By * T(4) * zy3 / dy2) / (T(2) * std::sqrt(qy));
                T dmdz = dMsx * dSxz + dMsy * dSyz;
                const T J[5] = {dmdx, dmdy, dmdz, DX * DY, T(1)};
                cb(y * R + x, mu, J);
            }
        }
    }
};

// BG: 0 flat background, 1 tilted, 2 background image (per-spot constants).
template <typename T, int BG = 0>
struct Spline3D {
    enum { K = BG == 1 ? 7 : 5, NC = 0, CrlbAbs = 0 };
    const T* cst = nullptr;
    int R;
    int Nz, Ny, Nx;
    T zmin, zmax;
    const float* coefs;
    static T StopLimit(int k) { return k == 2 ? T(1e-7f) : T(1e-4f); }
    void CheckLimits(T* t) const {
        t[0] = clampf_<T>(t[0], T(1.5f), T(R) - T(1.0f) - T(1.5f));
        t[1] = clampf_<T>(t[1], T(1.5f), T(R) - T(1.0f) - T(1.5f));
        t[2] = clampf_<T>(t[2], zmin + T(1e-3f), zmax - T(1e-3f));
        t[3] = std::fmax(t[3], T(10.f));
        t[4] = std::fmax(t[4], T(0.0f));
    }
    void Initial(const T com[3], T* t) const {
        t[0] = com[0]; t[1] = com[1]; t[2] = 0; t[3] = com[2]; t[4] = 0;
        for (int k = 5; k < K; k++) t[k] = 0;
    }
    template <typename CB> void Eval(const T* th, CB cb) const {
        const T du = (zmax - zmin) / T(Nz);
        T tx = th[0], ty = th[1], tz = th[2], tI = th[3], tbg = th[4];
        if (BG == 2) {  // CubicSplinePSFModels.h:374-377
            tx = tx - T(R) / T(2.0f) + (T(R) - T(Nx) + T(1)) / T(2.0f);
            ty = ty - T(R) / T(2.0f) + (T(R) - T(Ny) + T(1)) / T(2.0f);
            tx += T(0.5f);
            ty += T(0.5f);
        } else {
            tx -= T(Nx) * T(0.5f);
            ty -= T(Ny) * T(0.5f);
        }
        const T half = T(R) * T(0.5f);
        T bf[64] = {0}, bx[64] = {0}, by[64] = {0}, bz[64] = {0};
        T spx = -tx; spx -= T(int(spx));
        T spy = -ty; spy -= T(int(spy));
        T spz = (tz - zmin) / du;
        int k = clampi_(int(spz), 0, Nz - 1);
        spz = clampf_<T>(spz - T(k), T(0.0f), T(1.0f));
        T zzz = 1;
        for (int m = 0; m < 4; m++) {
            T yyy = 1;
            for (int n = 0; n < 4; n++) {
                T xxx = 1;
                for (int o = 0; o < 4; o++) {
                    const T v = xxx * yyy * zzz;
                    bf[m * 16 + n * 4 + o] = v;
                    if (o < 3) bx[m * 16 + n * 4 + o + 1] = T(o + 1) * v;
                    if (n < 3) by[m * 16 + (n + 1) * 4 + o] = T(n + 1) * v;
                    if (m < 3) bz[(m + 1) * 16 + n * 4 + o] = T(m + 1) * v;
                    xxx *= spx;
                }
                yyy *= spy;
            }
            zzz *= spz;
        }
        for (int y = 0; y < R; y++) {
            T yy = T(y) - ty;
            int j = clampi_(int(yy), 0, Ny - 1);
            for (int x = 0; x < R; x++) {
                T xx = T(x) - tx;
                int i = clampi_(int(xx), 0, Nx - 1);
                T psf = 0, gx = 0, gy = 0, gz = 0;
                const float* c = coefs + 64 * ((size_t)k * Nx * Ny + (size_t)j * Nx + i);
                for (int q = 0; q < 64; q++) {
                    T cf = T(c[q]);
                    psf += bf[q] * cf;
                    gx += bx[q] * cf;
                    gy += by[q] * cf;
                    gz += bz[q] * cf;
                }
                gx *= -tI;
                gy *= -tI;
                gz *= tI / du;
                if (psf < T(0.0f)) psf = T(0.0f);
                if (BG == 1) {
                    T mu = tI * psf + tbg + th[5] * (T(x) - half) + th[6] * (T(y) - half);
                    const T J[7] = {gx, gy, gz, psf, T(1), T(x) - half, T(y) - half};
                    cb(y * R + x, mu, J);
                } else if (BG == 2) {
                    const T b = cst[y * R + x];
                    T mu = tI * psf + tbg * b;
                    const T J[5] = {gx, gy, gz, psf, b};
                    cb(y * R + x, mu, J);
                } else {
                    T mu = tI * psf + tbg;
                    const T J[5] = {gx, gy, gz, psf, T(1)};
                    cb(y * R + x, mu, J);
                }
            }
        }
    }
};

// ---------------------------------------------------- small dense solvers
// One axis of the pixel-integrated Gaussian (GaussianPSFModels.h:76-94 pattern).
template <typename T>
struct GAxis {
    T D, dE, dS;  // DeltaX, dEx (without I), (e1 a1 - e0 a0)/sqrt(pi) (the sigma term without I)
};

template <typename T>
GAxis<T> gaxis(int px, T t, T sigma) {
    const T cp = T(1.0f) / (std::sqrt(T(2) * kPiT<T>) * sigma);
    const T c2 = T(1.0f) / (std::sqrt(T(2)) * sigma);
    const T isp = T(1.0f) / std::sqrt(kPiT<T>);
    T a0 = (T(px) - t + T(.5f)) * c2;
    T a1 = (T(px) - t - T(.5f)) * c2;
    GAxis<T> g;
    g.D = T(0.5f) * std::erf(a0) - T(0.5f) * std::erf(a1);
    T e1 = std::exp(-a1 * a1), e0 = std::exp(-a0 * a0);
    g.dE = cp * (e1 - e0);
    g.dS = (e1 * a1 - e0 * a0) * isp;
    return g;
}

template <typename T>
void xy_clamp(T* t, int R) {
    t[0] = clampf_<T>(t[0], T(1.5f), T(R) - T(1.5f) - T(1));
    t[1] = clampf_<T>(t[1], T(1.5f), T(R) - T(1.5f) - T(1));
}

// XYIBgConstSigma (GaussianPSFModels.h:155-229): sigma x, y per spot in cst.
template <typename T>
struct Gauss2DConstSigma {
    enum { K = 4, NC = 2, CrlbAbs = 0 };
    const T* cst = nullptr;
    int R;
    static T StopLimit(int k) { const T s[4] = {T(1e-5f), T(1e-5f), T(1e-1f), T(1e-5f)}; return s[k]; }
    void CheckLimits(T* t) const {
        xy_clamp(t, R);
        if (t[2] < T(25.0f)) t[2] = T(25.0f);
        if (t[3] < T(0.0f)) t[3] = T(0.0f);
    }
    void Initial(const T com[3], T* t) const { t[0] = com[0]; t[1] = com[1]; t[2] = com[2] * T(0.7f); t[3] = 0; }
    template <typename CB> void Eval(const T* th, CB cb) const {
        const T tx = th[0], ty = th[1], tI = th[2], tbg = th[3];
        for (int y = 0; y < R; y++) {
            GAxis<T> gy = gaxis<T>(y, ty, cst[1]);
            for (int x = 0; x < R; x++) {
                GAxis<T> gx = gaxis<T>(x, tx, cst[0]);
                T mu = tbg + tI * gx.D * gy.D;
                const T J[4] = {tI * gx.dE * gy.D, tI * gy.dE * gx.D, gx.D * gy.D, T(1)};
                cb(y * R + x, mu, J);
            }
        }
    }
};

// XYITiltedBg (GaussianPSFModels.h:232-304): mu = bg + bgx (x - R/2) + bgy (y - R/2) + I G.
template <typename T>
struct Gauss2DTiltedBg {
    enum { K = 6, NC = 2, CrlbAbs = 0 };
    const T* cst = nullptr;
    int R;
    static T StopLimit(int k) { return k == 2 ? T(1e-1f) : T(1e-5f); }
    void CheckLimits(T* t) const {
        xy_clamp(t, R);
        if (t[2] < T(25.0f)) t[2] = T(25.0f);
        if (t[3] < T(0.0f)) t[3] = T(0.0f);
    }
    void Initial(const T com[3], T* t) const {
        t[0] = com[0]; t[1] = com[1]; t[2] = com[2] * T(0.7f); t[3] = 0; t[4] = 0; t[5] = 0;
    }
    template <typename CB> void Eval(const T* th, CB cb) const {
        const T tx = th[0], ty = th[1], tI = th[2], tbg = th[3], bgx = th[4], bgy = th[5];
        const float half = R * 0.5f;
        for (int y = 0; y < R; y++) {
            GAxis<T> gy = gaxis<T>(y, ty, cst[1]);
            for (int x = 0; x < R; x++) {
                GAxis<T> gx = gaxis<T>(x, tx, cst[0]);
                T mu = tbg + bgx * (T(x) - T(half)) + bgy * (T(y) - T(half)) + tI * gx.D * gy.D;
                const T J[6] = {tI * gx.dE * gy.D, tI * gy.dE * gx.D, gx.D * gy.D, T(1), T(x) - T(half),
                                T(y) - T(half)};
                cb(y * R + x, mu, J);
            }
        }
    }
};

// XYIBgSigmaXY (GaussianPSFModels.h:388-467) and, with TILT, XYITiltedBgSigmaXY
// (:471-553), whose mu leaves the tilt out while its Jacobian has the tilt columns.
template <typename T, bool TILT>
struct Gauss2DSigmaXY {
    enum { K = TILT ? 8 : 6, NC = 0, CrlbAbs = 0 };
    const T* cst = nullptr;
    int R;
    T isx, isy;
    static constexpr int iS = TILT ? 6 : 4;
    static T StopLimit(int k) {
        if (k == 2) return T(1e-1f);
        if (k == 3 || (TILT && (k == 4 || k == 5))) return T(1e-4f);
        return T(1e-5f);
    }
    void CheckLimits(T* t) const {
        xy_clamp(t, R);
        if (t[2] < T(10.0f)) t[2] = T(10.0f);
        if (t[3] < T(0.0f)) t[3] = T(0.0f);
        t[iS] = clampf_<T>(t[iS], T(1.0f), T(R) * T(0.5f));
        t[iS + 1] = clampf_<T>(t[iS + 1], T(1.0f), T(R) * T(0.5f));
    }
    void Initial(const T com[3], T* t) const {
        t[0] = com[0]; t[1] = com[1]; t[2] = com[2]; t[3] = 0;
        if (TILT) { t[4] = 0; t[5] = 0; }
        t[iS] = isx; t[iS + 1] = isy;
    }
    template <typename CB> void Eval(const T* th, CB cb) const {
        const T tx = th[0], ty = th[1], tI = th[2], tbg = th[3];
        const float half = R * 0.5f;
        for (int y = 0; y < R; y++) {
            GAxis<T> gy = gaxis<T>(y, ty, th[iS + 1]);
            for (int x = 0; x < R; x++) {
                GAxis<T> gx = gaxis<T>(x, tx, th[iS]);
                T mu = tbg + tI * gx.D * gy.D;
                T J[8];
                J[0] = tI * gx.dE * gy.D;
                J[1] = tI * gy.dE * gx.D;
                J[2] = gx.D * gy.D;
                J[3] = T(1);
                if (TILT) { J[4] = T(x) - T(half); J[5] = T(y) - T(half); }
                J[iS] = tI * gx.dS * gy.D;
                J[iS + 1] = tI * gy.dS * gx.D;
                cb(y * R + x, mu, J);
            }
        }
    }
};

// IntensityBgModel (GaussianPSFModels.h:557-623): per-spot constants x, y, sigma.
template <typename T>
struct IntensityBg {
    enum { K = 2, NC = 3, CrlbAbs = 1 };
    const T* cst = nullptr;
    int R;
    static T StopLimit(int k) { return k == 0 ? T(1e-2f) : T(1e-4f); }
    void CheckLimits(T* t) const {
        if (t[0] < T(1e-5f)) t[0] = T(1e-5f);
        if (t[1] < T(1e-8f)) t[1] = T(1e-8f);
    }
    void Initial(const T*, T* t) const { t[0] = 1; t[1] = 0; }
    template <typename CB> void Eval(const T* th, CB cb) const {
        for (int y = 0; y < R; y++) {
            GAxis<T> gy = gaxis<T>(y, cst[1], cst[2]);
            for (int x = 0; x < R; x++) {
                GAxis<T> gx = gaxis<T>(x, cst[0], cst[2]);
                const T psf = gx.D * gy.D;
                const T J[2] = {psf, T(1)};
                cb(y * R + x, th[0] * psf + th[1], J);
            }
        }
    }
};

template <typename T>
bool cholesky(int d, const T* S, T* D) {
    for (int k = 0; k < d; ++k) {
        T sum{};
        for (int p = 0; p < k; ++p) sum += D[k * d + p] * D[k * d + p];
        T v = S[k * d + k] - sum;
        if (v == T(0)) return false;
        D[k * d + k] = std::sqrt(v);
        for (int i = k + 1; i < d; ++i) {
            T s2{};
            for (int p = 0; p < k; ++p) s2 += D[i * d + p] * D[k * d + p];
            T q = D[k * d + k];
            if (q == T(0)) return false;
            D[i * d + k] = (S[i * d + k] - s2) / q;
        }
    }
    return true;
}

template <typename T>
bool solve_cholesky(int d, const T* L, const T* b, T* x) {
    T y[8];
    for (int i = 0; i < d; ++i) {
        T sum{};
        for (int k = 0; k < i; ++k) sum += L[i * d + k] * y[k];
        T v = L[i * d + i];
        if (v == T(0)) return false;
        y[i] = (b[i] - sum) / v;
    }
    for (int i = d - 1; i >= 0; --i) {
        T sum{};
        for (int k = i + 1; k < d; ++k) sum += L[k * d + i] * x[k];
        x[i] = (y[i] - sum) / L[i * d + i];
    }
    return true;
}

// LU with partial pivoting; returns false when the largest pivot is < tol.
template <typename T>
bool lup_decompose(int N, T* A, T tol, int* P) {
    for (int i = 0; i <= N; i++) P[i] = i;
    for (int i = 0; i < N; i++) {
        T maxA = 0, absA;
        int imax = i;
        for (int k = i; k < N; k++)
            if ((absA = std::fabs(A[k * N + i])) > maxA) { maxA = absA; imax = k; }
        if (maxA < tol) return false;
        if (imax != i) {
            std::swap(P[i], P[imax]);
            for (int j = 0; j < N; j++) std::swap(A[i * N + j], A[imax * N + j]);
            P[N]++;
        }
        for (int j = i + 1; j < N; j++) {
            A[j * N + i] /= A[i * N + i];
            for (int k = i + 1; k < N; k++) A[j * N + k] -= A[j * N + i] * A[i * N + k];
        }
    }
    return true;
}

template <typename T>
void lup_invert(int N, const T* A, const int* P, T* IA) {
    for (int j = 0; j < N; j++) {
        for (int i = 0; i < N; i++) {
            IA[i * N + j] = (P[i] == j) ? T(1) : T(0);
            for (int k = 0; k < i; k++) IA[i * N + j] -= A[i * N + k] * IA[k * N + j];
        }
        for (int i = N - 1; i >= 0; i--) {
            for (int k = i + 1; k < N; k++) IA[i * N + j] -= A[i * N + k] * IA[k * N + j];
            IA[i * N + j] = IA[i * N + j] / A[i * N + i];
        }
    }
}

template <typename T>
void lup_solve(int N, const T* A, const int* P, const T* b, T* x) {
    for (int i = 0; i < N; i++) {
        x[i] = b[P[i]];
        for (int k = 0; k < i; k++) x[i] -= A[i * N + k] * x[k];
    }
    for (int i = N - 1; i >= 0; i--) {
        for (int k = i + 1; k < N; k++) x[i] -= A[i * N + k] * x[k];
        x[i] = x[i] / A[i * N + i];
    }
}

// ---------------------------------------------------------- estimation
template <typename T>
void com(const T* s, int R, T out[3]) {
    T sx = 0, sy = 0, sum = 0;
    for (int y = 0; y < R; y++)
        for (int x = 0; x < R; x++) {
            T v = s[y * R + x];
            sx += T(x) * v;
            sy += T(y) * v;
            sum += v;
        }
    out[0] = sx / sum;
    out[1] = sy / sum;
    out[2] = sum;
}

template <typename T, typename M>
int lm_optimize(const M& m, const T* smp, T* theta, int maxIter, T lambda, T* trace, int traceLen) {
    const int K = M::K;
    int i;
    for (i = 0; i < maxIter; i++) {
        if (i < traceLen)
            for (int k = 0; k < K; k++) trace[i * K + k] = theta[k];
        T alpha[K * K] = {};
        T beta[K] = {};
        T step[K];
        T L[K * K] = {};
        m.Eval(theta, [&](int idx, T mu, const T* J) {
            T xv = smp[idx];
            if (xv < T(1e-6f)) xv = T(1e-6f);
            T muc = mu > T(1e-6f) ? mu : T(1e-6f);
            T inv = T(1.0f) / muc;
            T xf2 = xv * inv * inv;
            for (int a = 0; a < K; a++)
                for (int b = a; b < K; b++) alpha[K * a + b] += J[a] * J[b] * xf2;
            T bf = T(1) - xv * inv;
            for (int a = 0; a < K; a++) beta[a] -= bf * J[a];
        });
        for (int a = 1; a < K; a++)
            for (int b = 0; b < a; b++) alpha[K * a + b] = alpha[K * b + a];
        for (int k = 0; k < K; k++) {
            if (lambda < T(0)) alpha[k * K + k] -= lambda;
            else {
                T s = 0;
                for (int j = 0; j < K; j++) s += alpha[j * K + k] * alpha[j * K + k];
                alpha[k * K + k] += s * lambda;
            }
        }
        if (!cholesky<T>(K, alpha, L)) break;
        if (!solve_cholesky<T>(K, L, beta, step)) break;
        T last[K];
        for (int k = 0; k < K; k++) last[k] = theta[k];
        for (int k = 0; k < K; k++) theta[k] += step[k];
        m.CheckLimits(theta);
        bool small = true;
        for (int k = 0; k < K; k++)
            if (std::fabs(last[k] - theta[k]) > M::StopLimit(k)) { small = false; break; }
        if (i > 0 && small) break;
    }
    return i + 1;
}

template <typename T, typename M>
void crlb_chisq(const M& m, const T* theta, const T* smp, T* crlb, T* chisq) {
    const int K = M::K;
    T fi[K * K] = {};
    T chi = 0;
    m.Eval(theta, [&](int idx, T mu, const T* J) {
        mu = std::fmax(T(1e-8f), mu);
        if (smp) {
            T e = smp[idx] - mu;
            chi += e * e / mu;
        }
        T inv = T(1.0f) / mu;
        for (int a = 0; a < K; a++)
            for (int b = a; b < K; b++) fi[K * a + b] += J[a] * J[b] * inv;
    });
    for (int a = 1; a < K; a++)
        for (int b = 0; b < a; b++) fi[K * a + b] = fi[K * b + a];
    T A[K * K], inv[K * K];
    int P[K + 1];
    std::memcpy(A, fi, sizeof(A));
    if (lup_decompose<T>(K, A, T(1e-9f), P)) {
        lup_invert<T>(K, A, P, inv);
        for (int k = 0; k < K; k++) crlb[k] = std::sqrt(M::CrlbAbs ? std::fabs(inv[k * K + k]) : inv[k * K + k]);
    } else {
        for (int k = 0; k < K; k++) crlb[k] = T(INFINITY);
    }
    *chisq = chi;
}

template <typename T, typename M>
T loglik(const M& m, const T* theta, const T* smp) {
    T LL = 0;
    m.Eval(theta, [&](int idx, T mu, const T*) {
        T xv = smp[idx];
        mu = std::fmax(mu, T(1e-8f));
        xv = std::fmax(xv, T(1e-8f));
        LL += xv * std::log(mu) - mu;
    });
    return LL;
}

template <typename T, int BG, typename F>
void spline_case(const OrModel* d, F&& fn) {
    Spline3D<T, BG> m{};
    m.R = d->roisize; m.Nz = d->shape[0]; m.Ny = d->shape[1]; m.Nx = d->shape[2];
    m.zmin = d->p[0]; m.zmax = d->p[1]; m.coefs = d->coefs;
    fn(m);
}

template <typename T, typename F>
void dispatch(const OrModel* d, F fn) {
    switch (d->kind) {
    case 0: { Gauss2DFixed<T> m{}; m.R = d->roisize; m.sx = d->p[0]; m.sy = d->p[1]; fn(m); break; }
    case 1: { Gauss2DSigma<T> m{}; m.R = d->roisize; m.initialSigma = d->p[0]; fn(m); break; }
    case 2: {
        Astig3D<T> m{};
        m.R = d->roisize;
        for (int i = 0; i < 4; i++) { m.cx[i] = d->p[i]; m.cy[i] = d->p[4 + i]; }
        m.minz = d->p[8]; m.maxz = d->p[9]; m.Bx = d->p[10]; m.By = d->p[11];
        fn(m);
        break;
    }
    case 3: spline_case<T, 0>(d, fn); break;
    case 9: spline_case<T, 1>(d, fn); break;
    case 10: spline_case<T, 2>(d, fn); break;
    case 4: { Gauss2DConstSigma<T> m{}; m.R = d->roisize; fn(m); break; }
    case 5: { Gauss2DTiltedBg<T> m{}; m.R = d->roisize; fn(m); break; }
    case 6: { Gauss2DSigmaXY<T, false> m{}; m.R = d->roisize; m.isx = d->p[0]; m.isy = d->p[1]; fn(m); break; }
    case 7: { Gauss2DSigmaXY<T, true> m{}; m.R = d->roisize; m.isx = d->p[0]; m.isy = d->p[1]; fn(m); break; }
    case 8: { IntensityBg<T> m{}; m.R = d->roisize; fn(m); break; }
    default: break;
    }
}

int num_params(int kind) {
    switch (kind) {
    case 0: case 4: return 4;
    case 5: case 6: return 6;
    case 7: return 8;
    case 8: return 2;
    case 9: return 7;
    default: return 5;
    }
}

int num_consts(int kind, int R) { return kind == 4 || kind == 5 ? 2 : kind == 8 ? 3 : kind == 10 ? R * R : 0; }

// The model for spot s: a copy that points at the spot's nc constants.
template <typename M, typename T>
M for_spot(const M& m, const T* cst, int s, int nc) {
    M c = m;
    if (nc > 0) c.cst = cst + (size_t)s * nc;
    return c;
}

template <typename T>
void expected_value(const OrModel* d, int n, const T* params, const T* cst, T* ev, int nth) {
    const int S = d->roisize * d->roisize, K = num_params(d->kind);
    dispatch<T>(d, [&](auto& m) {
        pool_for(n, nth, [&](int s) {
            for_spot(m, cst, s, num_consts(d->kind, d->roisize)).Eval(params + (size_t)s * K, [&](int idx, T mu, const T*) { ev[(size_t)s * S + idx] = mu; });
        });
    });
}

template <typename T>
void derivatives(const OrModel* d, int n, const T* params, const T* cst, T* deriv, T* ev, int nth) {
    const int S = d->roisize * d->roisize, K = num_params(d->kind);
    dispatch<T>(d, [&](auto& m) {
        pool_for(n, nth, [&](int s) {
            for_spot(m, cst, s, num_consts(d->kind, d->roisize)).Eval(params + (size_t)s * K, [&](int idx, T mu, const T* J) {
                for (int k = 0; k < K; k++) deriv[((size_t)s * K + k) * S + idx] = J[k];
                ev[(size_t)s * S + idx] = mu;
            });
        });
    });
}

template <typename T>
void estimate(const OrModel* d, int n, const T* smp, const T* cst, const T* initial, T* params, int* iters, T* trace,
              int traceLen, int maxIter, T lambda, int nth) {
    const int R = d->roisize, S = R * R, K = num_params(d->kind);
    dispatch<T>(d, [&](auto& m0) {
        pool_for(n, nth, [&](int s) {
            const auto m = for_spot(m0, cst, s, num_consts(d->kind, d->roisize));
            const T* sp = smp + (size_t)s * S;
            T* th = params + (size_t)s * K;
            if (initial) {
                for (int k = 0; k < K; k++) th[k] = initial[(size_t)s * K + k];
            } else {
                T c[3];
                com<T>(sp, R, c);
                m.Initial(c, th);
            }
            int it = lm_optimize<T>(m, sp, th, maxIter, lambda, trace ? trace + (size_t)s * traceLen * K : nullptr,
                                    trace ? traceLen : 0);
            if (iters) iters[s] = it;
        });
    });
}

template <typename T>
void crlb(const OrModel* d, int n, const T* params, const T* cst, const T* smp, T* out, T* chisq, int nth) {
    const int S = d->roisize * d->roisize, K = num_params(d->kind);
    dispatch<T>(d, [&](auto& m) {
        pool_for(n, nth, [&](int s) {
            T c;
            crlb_chisq<T>(for_spot(m, cst, s, num_consts(d->kind, d->roisize)), params + (size_t)s * K, smp ? smp + (size_t)s * S : nullptr,
                          out + (size_t)s * K, &c);
            if (chisq) chisq[s] = c;
        });
    });
}

// GLRT_PSF::Estimate diagnostics (GLRT_PSF.cu:103-129) of fitted params: per spot
// {ll_on, ll_off, mean}, mean over the first spot's samples (meanOfSpot0) or the spot's own.
template <typename T>
void glrt_all(const OrModel* d, int n, const T* params, const T* cst, const T* smp, int meanOfSpot0, T* diag,
              int nth) {
    const int S = d->roisize * d->roisize, K = num_params(d->kind);
    dispatch<T>(d, [&](auto& m) {
        pool_for(n, nth, [&](int s) {
            const T* own = smp + (size_t)s * S;
            const T* ms = meanOfSpot0 ? smp : own;
            T mean = 0;
            for (int i = 0; i < S; i++) mean += ms[i];
            mean /= T(S);
            std::vector<T> ev(S);
            for_spot(m, cst, s, num_consts(d->kind, d->roisize))
                .Eval(params + (size_t)s * K, [&](int idx, T mu, const T*) { ev[idx] = mu; });
            T on = 0, off = 0;
            for (int i = 0; i < S; i++) {
                const T mu = std::fmax(ev[i], T(1e-8f));
                const T x = std::fmax(own[i], T(1e-8f));
                on += x * std::log(mu) - mu;
                off += x * std::log(mean) - mean;
            }
            diag[3 * (size_t)s] = on;
            diag[3 * (size_t)s + 1] = off;
            diag[3 * (size_t)s + 2] = mean;
        });
    });
}

template <typename T>
void loglik_all(const OrModel* d, int n, const T* params, const T* cst, const T* smp, T* ll, int nth) {
    const int S = d->roisize * d->roisize, K = num_params(d->kind);
    dispatch<T>(d, [&](auto& m) {
        pool_for(n, nth, [&](int s) {
            ll[s] = loglik<T>(for_spot(m, cst, s, num_consts(d->kind, d->roisize)), params + (size_t)s * K, smp + (size_t)s * S);
        });
    });
}

}  // namespace

extern "C" {

int or_num_params(int kind) { return num_params(kind); }
void or_glrt_f32(const OrModel* d, int n, const float* p, const float* cst, const float* smp, int meanOfSpot0,
                 float* diag, int nth) { glrt_all<float>(d, n, p, cst, smp, meanOfSpot0, diag, nth); }
void or_glrt_f64(const OrModel* d, int n, const double* p, const double* cst, const double* smp, int meanOfSpot0,
                 double* diag, int nth) { glrt_all<double>(d, n, p, cst, smp, meanOfSpot0, diag, nth); }
int or_num_consts(int kind, int roisize) { return num_consts(kind, roisize); }

// cst: [n, NC] per-spot constants of the models that have them, else NULL.
void or_expected_value_f32(const OrModel* d, int n, const float* p, const float* cst, float* ev, int nth) { expected_value<float>(d, n, p, cst, ev, nth); }
void or_expected_value_f64(const OrModel* d, int n, const double* p, const double* cst, double* ev, int nth) { expected_value<double>(d, n, p, cst, ev, nth); }
void or_derivatives_f32(const OrModel* d, int n, const float* p, const float* cst, float* dv, float* ev, int nth) { derivatives<float>(d, n, p, cst, dv, ev, nth); }
void or_derivatives_f64(const OrModel* d, int n, const double* p, const double* cst, double* dv, double* ev, int nth) { derivatives<double>(d, n, p, cst, dv, ev, nth); }
void or_estimate_f32(const OrModel* d, int n, const float* smp, const float* cst, const float* initial, float* params,
                     int* iters, float* trace, int traceLen, int maxIter, float lambda, int nth) {
    estimate<float>(d, n, smp, cst, initial, params, iters, trace, traceLen, maxIter, lambda, nth);
}
void or_estimate_f64(const OrModel* d, int n, const double* smp, const double* cst, const double* initial,
                     double* params, int* iters, double* trace, int traceLen, int maxIter, double lambda, int nth) {
    estimate<double>(d, n, smp, cst, initial, params, iters, trace, traceLen, maxIter, lambda, nth);
}
void or_crlb_f32(const OrModel* d, int n, const float* p, const float* cst, const float* smp, float* out, float* chisq, int nth) { crlb<float>(d, n, p, cst, smp, out, chisq, nth); }
void or_crlb_f64(const OrModel* d, int n, const double* p, const double* cst, const double* smp, double* out, double* chisq, int nth) { crlb<double>(d, n, p, cst, smp, out, chisq, nth); }
void or_loglik_f32(const OrModel* d, int n, const float* p, const float* cst, const float* smp, float* ll, int nth) { loglik_all<float>(d, n, p, cst, smp, ll, nth); }
void or_loglik_f64(const OrModel* d, int n, const double* p, const double* cst, const double* smp, double* ll, int nth) { loglik_all<double>(d, n, p, cst, smp, ll, nth); }

// Spline coefficients for a z-stack [Nz,Ny,Nx]; output [(Nz-3+2b),(Ny-3+2b),(Nx-3+2b),64].
// Follows CSpline_Compute (CubicSplinePSF.cu:151-204): one 64x64 system whose row
// (a,b,c) holds (a-1)^l (b-1)^m (c-1)^n, solved per cell with FP64 LU.
void or_spline_compute(const int* shape, const float* data, float* out, int border) {
    static double A[64 * 64];
    static int P[65];
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++)
            for (int c = 0; c < 4; c++)
                for (int l = 0; l < 4; l++)
                    for (int m = 0; m < 4; m++)
                        for (int q = 0; q < 4; q++)
                            A[64 * (a * 16 + b * 4 + c) + (l * 16 + m * 4 + q)] =
                                std::pow(double(a - 1), l) * std::pow(double(b - 1), m) * std::pow(double(c - 1), q);
    lup_decompose<double>(64, A, 1e-10, P);
    const int o0 = shape[0] - 3 + 2 * border, o1 = shape[1] - 3 + 2 * border, o2 = shape[2] - 3 + 2 * border;
    auto px = [&](int z, int y, int x) {
        z = clampi_(z - border, 0, shape[0] - 1);
        y = clampi_(y - border, 0, shape[1] - 1);
        x = clampi_(x - border, 0, shape[2] - 1);
        return (double)data[(size_t)shape[1] * shape[2] * z + (size_t)shape[2] * y + x];
    };
    pool_for(o0, (int)std::thread::hardware_concurrency(), [&](int i) {
        double b[64], c[64];
        for (int j = 0; j < o1; j++)
            for (int k = 0; k < o2; k++) {
                for (int m = 0; m < 4; m++)
                    for (int n = 0; n < 4; n++)
                        for (int q = 0; q < 4; q++) b[m * 16 + n * 4 + q] = px(i + m, j + n, k + q);
                lup_solve<double>(64, A, P, b, c);
                float* dst = out + 64 * ((size_t)o1 * o2 * i + (size_t)o2 * j + k);
                for (int a = 0; a < 64; a++) dst[a] = (float)c[a];
            }
    });
}

}  // extern "C"
