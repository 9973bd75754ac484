// C ABI of libphotonpy_amd.so (declarations and reference citations in
// include/photonpy_amd.h).  Host-pointer calls stage through a per-estimator
// device workspace and run synchronously; device-pointer calls launch on the
// caller's stream.  Everything computes on the GPU: there is no CPU path.
#include "capi_internal.h"

#include <atomic>

namespace ppa {
hipError_t launch_spline_coefs(const int* shape, const float* d_data, float* d_out, int border, hipStream_t st);
hipError_t launch_spline_evaluate(const int* shape, const float* d_coefs, const float* shift, int border,
                                  float* d_out, hipStream_t st);
}

// .zstack file (CubicSplinePSF.cu:109-133, smlm/psf.py:256-276): a packed 24-byte header
// {int version = 1; int shape[3] (z, y, x); float zmin, zmax} and the float32 samples.
struct ZStackHeader {
    int version;
    int shape[3];
    float zmin, zmax;
};
static_assert(sizeof(ZStackHeader) == 24, "zstack header is 24 bytes");

static std::vector<float> load_zstack(const char* filename, ZStackHeader& h) {
    if (!filename) throw std::runtime_error("photonpy_amd: NULL file name");
    FILE* f = std::fopen(filename, "rb");
    if (!f) throw std::runtime_error(std::string("Can't open file ") + filename);
    std::vector<float> data;
    const bool ok = std::fread(&h, sizeof(h), 1, f) == 1;
    if (ok && h.version == 1 && h.shape[0] > 0 && h.shape[1] > 0 && h.shape[2] > 0) {
        data.resize((size_t)h.shape[0] * h.shape[1] * h.shape[2]);
        const size_t got = std::fread(data.data(), sizeof(float), data.size(), f);
        std::fclose(f);
        if (got != data.size()) throw std::runtime_error(std::string("truncated zstack file ") + filename);
        return data;
    }
    std::fclose(f);
    if (!ok) throw std::runtime_error(std::string("truncated zstack file ") + filename);
    if (h.version != 1) throw std::runtime_error("Wrong ZStack file version (should be 1)");
    throw std::runtime_error("photonpy_amd: invalid zstack shape");
}

using namespace ppa;


// Samples whose local tricubic interpolant a spline table is.  Along each axis
// cell c of a CSpline_Compute table interpolates samples g[c..c+3] at the local
// nodes t = -1, 0, 1, 2 (CubicSplinePSF.cu:151-204), so sample d is the value of
// cell clamp(d-1, 0, N-1) at node d - c - 1.  The samples are read off the
// polynomials in FP64; the table is accepted when every cell's coefficients,
// recomputed from the samples with the exact inverse of the 4x4 Vandermonde
// factor, agree with the given ones to 1e-5 of the largest coefficient (FP32
// rounding of coefficients and samples stays below 1e-6).  Tables made any other
// way (smoothing splines) fail the check and keep the coefficient form.
static bool spline_grid_from_coefs(int Nz, int Ny, int Nx, const float* c, std::vector<float>& grid) {
    const int D[3] = {Nz + 3, Ny + 3, Nx + 3}, N[3] = {Nz, Ny, Nx};
    // the fit kernels address the table and its padding by 32-bit byte offsets
    if (((size_t)D[0] * D[1] * D[2] + SplineGridModel::pad_floats(Nx)) * sizeof(float) >= ((size_t)1 << 31))
        return false;
    grid.assign((size_t)D[0] * D[1] * D[2], 0.0f);
    auto cellOf = [&](int axis, int d, int& cell, double& t) {
        cell = std::min(std::max(d - 1, 0), N[axis] - 1);
        t = (double)(d - cell - 1);
    };
    for (int dz = 0; dz < D[0]; dz++)
        for (int dy = 0; dy < D[1]; dy++)
            for (int dx = 0; dx < D[2]; dx++) {
                int cz, cy, cx;
                double tz, ty, tx;
                cellOf(0, dz, cz, tz);
                cellOf(1, dy, cy, ty);
                cellOf(2, dx, cx, tx);
                const float* cc = c + 64 * (((size_t)cz * Ny + cy) * Nx + cx);
                double v = 0.0, pz = 1.0;
                for (int m = 0; m < 4; m++, pz *= tz) {
                    double py = 1.0;
                    for (int n = 0; n < 4; n++, py *= ty) {
                        double px = 1.0;
                        for (int o = 0; o < 4; o++, px *= tx) v += (double)cc[16 * m + 4 * n + o] * pz * py * px;
                    }
                }
                grid[((size_t)dz * D[1] + dy) * D[2] + dx] = (float)v;
            }
    // inverse Vandermonde factor: coefficient p of the cubic through nodes -1, 0, 1, 2
    static const double Vi[4][4] = {{0.0, 1.0, 0.0, 0.0},
                                    {-1.0 / 3.0, -0.5, 1.0, -1.0 / 6.0},
                                    {0.5, -1.0, 0.5, 0.0},
                                    {-1.0 / 6.0, 0.5, -0.5, 1.0 / 6.0}};
    double cmax = 0.0;
    const size_t ncoef = (size_t)Nz * Ny * Nx * 64;
    for (size_t i = 0; i < ncoef; i++) cmax = std::max(cmax, std::fabs((double)c[i]));
    const double tol = 1e-5 * cmax;
    for (int cz = 0; cz < Nz; cz++)
        for (int cy = 0; cy < Ny; cy++)
            for (int cx = 0; cx < Nx; cx++) {
                double b[64], t1[64], t2[64];
                for (int m = 0; m < 4; m++)
                    for (int n = 0; n < 4; n++)
                        for (int q = 0; q < 4; q++)
                            b[16 * m + 4 * n + q] = grid[(((size_t)cz + m) * D[1] + cy + n) * D[2] + cx + q];
                for (int m = 0; m < 4; m++)
                    for (int n = 0; n < 4; n++)
                        for (int p = 0; p < 4; p++) {
                            double a = 0.0;
                            for (int q = 0; q < 4; q++) a += Vi[p][q] * b[16 * m + 4 * n + q];
                            t1[16 * m + 4 * n + p] = a;
                        }
                for (int m = 0; m < 4; m++)
                    for (int p = 0; p < 4; p++)
                        for (int q = 0; q < 4; q++) {
                            double a = 0.0;
                            for (int n = 0; n < 4; n++) a += Vi[p][n] * t1[16 * m + 4 * n + q];
                            t2[16 * m + 4 * p + q] = a;
                        }
                const float* cc = c + 64 * (((size_t)cz * Ny + cy) * Nx + cx);
                for (int p = 0; p < 4; p++)
                    for (int n = 0; n < 4; n++)
                        for (int q = 0; q < 4; q++) {
                            double a = 0.0;
                            for (int m = 0; m < 4; m++) a += Vi[p][m] * t2[16 * m + 4 * n + q];
                            if (!(std::fabs(a - (double)cc[16 * p + 4 * n + q]) <= tol)) return false;
                        }
            }
    return true;
}

static Estimator* make_gauss(int kind, int roisize, const GaussCalib& gc) {
    if (roisize < 1) throw std::runtime_error("photonpy_amd: roisize must be positive");
    auto* e = new Estimator();
    e->roisize = roisize;
    e->desc.kind = kind;
    e->desc.gc = gc;
    const float b = 1.5f, hi = (float)roisize - 1 - b;
    if (kind == kKindFixed) {
        e->fmt = "x,y,I,bg";
        e->limits = {{b, hi}, {b, hi}, {10.0f, 1e9f}, {1e-6f, 1e9f}};  // GaussianPSF.cu:92-97
    } else if (kind == kKindSigma) {
        e->fmt = "x,y,\tI,bg,sigma";  // literal tab as in GaussianPSFModels.h:316
        e->limits = {{b, hi}, {b, hi}, {10.0f, 1e9f}, {1e-6f, 1e9f}, {1.0f, roisize * 2.0f}};  // GaussianPSF.cu:141-147
    } else if (kind == kKindConstSigma) {
        e->fmt = "x,y,I,bg";
        e->limits = {{b, hi}, {b, hi}, {10.0f, 1e9f}, {1e-6f, 1e9f}};  // GaussianPSF.cu:113-118
    } else if (kind == kKindTiltedBg) {
        e->fmt = "x,y,I,bg,bgx,bgy";
        e->limits = {{b, hi}, {b, hi}, {10.0f, 1e9f}, {1e-6f, 1e9f}, {-1e9f, 1e9f}, {-1e9f, 1e9f}};  // :66-73
    } else if (kind == kKindSigmaXY) {
        e->fmt = "x,y,I,bg,sx,sy";
        e->limits = {{b, hi}, {b, hi}, {10.0f, 1e9f}, {1e-6f, 1e9f}, {1.0f, roisize * 2.0f},
                     {1.0f, roisize * 2.0f}};  // :169-176
    } else if (kind == kKindTiltedBgSigmaXY) {
        e->fmt = "x,y,I,bg,bgx,bgy,sx,sy";
        e->limits = {{b, hi}, {b, hi}, {10.0f, 1e9f}, {1e-6f, 1e9f}, {-1e9f, 1e9f}, {-1e9f, 1e9f},
                     {1.0f, roisize * 2.0f}, {1.0f, roisize * 2.0f}};  // :193-202
    } else {
        e->fmt = "x,y,z,I,bg";
        e->limits = {{b, hi}, {b, hi}, {gc.minz, gc.maxz}, {10.0f, 1e9f}, {1e-6f, 1e9f}};  // GaussianPSF.cu:42-48
    }
    return e;
}

// ---------------------------------------------------------------- launch helpers
int fit_group(Estimator* e, int groupHint) {
    for (int G = groupHint; G > e->G(); G /= 2)
        if (fit_compiled(e->desc, e->roisize, G)) return G;
    return e->G();
}

void run_fit(Estimator* e, int n, const float* d_sample, const float* d_constants, const float* d_initial,
             float* d_params, int* d_iter, float* d_trace, int traceLen, float* d_crlb, float* d_chisq, hipStream_t st,
             bool timing, int groupHint, int* counter) {
    if (counter) {
        if (!e->device_ready())
            throw std::runtime_error("photonpy_amd: the spline tables of this estimator are not on the device yet; "
                                     "an allocation-free call cannot upload them (any other fit or evaluation call "
                                     "on the estimator does)");
    } else {
        e->ensure_device();
    }
    // The fit kernel's work counter and the spot strides are 32-bit and run past n by
    // at most the grid's group count (< 2^21 lane groups on one device).
    constexpr int kMaxSpots = INT_MAX - (1 << 21);
    if (n > kMaxSpots)
        throw std::runtime_error("photonpy_amd: at most " + std::to_string(kMaxSpots) + " spots per fit call, got " +
                                 std::to_string(n));
    const int R = e->roisize;
    FitArgs a{};
    a.n = n;
    a.sample = d_sample;
    a.initial = d_initial;
    a.params = d_params;
    a.iterations = d_iter;
    a.trace = d_trace;
    a.traceLen = d_trace ? traceLen : 0;
    a.crlb = d_crlb;
    a.chisq = d_chisq;
    a.maxIter = e->lm.iterations;
    a.lambda = e->lm.lambda;
    a.refill = e->refill;
    a.counter = counter ? counter : e->next_counter();
    a.segShift = e->segShift;
    a.constants = d_constants;
    if (e->NC() > 0 && !d_constants)
        throw std::runtime_error("photonpy_amd: this estimator needs " + std::to_string(e->NC()) +
                                 " constants per spot");
    hipEvent_t* events = nullptr;
    if (e->phaseTiming && timing && !counter) {
        for (int i = 0; i < 3; i++) {
            hipEvent_t ev;
            hip_check(hipEventCreate(&ev), "hipEventCreate");
            e->phaseEvents.push_back(ev);
        }
        events = e->phaseEvents.data() + e->phaseEvents.size() - 3;
    }
    if (R < kMinFitR || R > kMaxFitR) {
        // no per-size kernel: the any-size kernel (one wavefront per spot)
        hip_check(launch_fit_large_any(e->desc, R, a, events, st), "fit kernel launch");
        return;
    }
    const int G = fit_group(e, groupHint);
    hipError_t err = launch_fit_any(e->desc, R, G, a, events, st);
    if (err == hipErrorInvalidValue)
        throw std::runtime_error("photonpy_amd: no fit kernel compiled for roisize " + std::to_string(R) + ", group " +
                                 std::to_string(G));
    hip_check(err, "fit kernel launch");
}

// ---------------------------------------------------------------- allocation-free fit
// Work counters of Estimator_EstimateDeviceNoAlloc: a fixed array in device memory
// (each device holds its own copy until the process ends), so that the call itself
// allocates nothing.  Each call takes the next of the kNoAllocSlots slots of
// kWorkSegs counters and resets it with a kernel on its stream before the fit reads it.
constexpr int kNoAllocSlots = 16;  // 16 x 8 counters 256 bytes apart: 32 KiB
static __device__ __align__(256) int scratch[kNoAllocSlots * kCounterInts];

// Device address of `scratch` on the current device, resolved once per device with
// that device current.  Host code passes this address, never the array's name.
static int* scratch_on_current_device() {
    constexpr int kMaxDevices = 64;
    static std::mutex mtx;
    static int* addr[kMaxDevices] = {};
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("photonpy_amd: device index out of range");
    std::lock_guard<std::mutex> l(mtx);
    if (!addr[dev]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        addr[dev] = static_cast<int*>(p);
    }
    return addr[dev];
}

static std::atomic<unsigned> g_noAllocSlot{0};

// workspace layout helper: consecutive 256-byte aligned slices
struct Slicer {
    char* base;
    size_t off = 0;
    template <class T>
    T* take(size_t count) {
        if (count == 0) return nullptr;
        T* p = reinterpret_cast<T*>(base + off);
        off += (count * sizeof(T) + 255) & ~size_t(255);
        return p;
    }
    static size_t bytes(std::initializer_list<size_t> sizes) {
        size_t t = 0;
        for (size_t s : sizes) t += (s + 255) & ~size_t(255);
        return t;
    }
};

// CSpline_CreatePSF (throws on failure).
static Estimator* create_spline(int roisize, const CSpline_Calibration* calib, const float* coefs, int mode, bool cuda,
                                Context* ctx) {
    if (!calib) throw std::runtime_error("photonpy_amd: calibration is NULL");
    // CSPLINE_FLATBG / TILTEDBG / BGIMAGE (CubicSplinePSF.cu:72-97)
    if (mode < 0 || mode > 2) throw std::runtime_error("Invalid model background mode");
    if (roisize < 1) throw std::runtime_error("photonpy_amd: roisize must be positive");
    const float* src = coefs ? coefs : calib->coefs;
    if (!src) throw std::runtime_error("photonpy_amd: spline coefficients are NULL");
    const int Nz = calib->shape[0], Ny = calib->shape[1], Nx = calib->shape[2];
    if (Nz < 1 || Ny < 1 || Nx < 1) throw std::runtime_error("photonpy_amd: invalid spline shape");
    auto* e = new Estimator();
    e->roisize = roisize;
    e->desc.kind = mode == 1 ? kKindSplineTilted : mode == 2 ? kKindSplineImage : kKindSpline;
    e->desc.sm.Nz = Nz;
    e->desc.sm.Ny = Ny;
    e->desc.sm.Nx = Nx;
    e->desc.sm.zmin = calib->z_min;
    e->desc.sm.zmax = calib->z_max;
    e->desc.sm.coefs = nullptr;
    e->hostCoefs.assign(src, src + (size_t)Nz * Ny * Nx * 64);  // copied, as CubicSplinePSF.cu:48-52
    e->desc.sg.Nz = Nz;
    e->desc.sg.Ny = Ny;
    e->desc.sg.Nx = Nx;
    e->desc.sg.zmin = calib->z_min;
    e->desc.sg.zmax = calib->z_max;
    e->desc.sg.grid = nullptr;
    e->gridForm = spline_grid_from_coefs(Nz, Ny, Nx, e->hostCoefs.data(), e->hostGrid);
    if (!e->gridForm) e->hostGrid.clear();
    e->desc.useGrid = e->gridForm;
    e->fmt = mode == 1 ? "x,y,z,I,bg,bgx,bgy" : "x,y,z,I,bg";
    const float b = 1.5f, hi = (float)roisize - 1 - b;
    e->limits = {{b, hi}, {b, hi}, {calib->z_min, calib->z_max}, {10.0f, 1e9f}, {1e-6f, 1e9f}};
    if (mode == 1) {
        e->limits.push_back({-1e6f, 1e6f});
        e->limits.push_back({-1e6f, 1e6f});
    }
    set_flavour(e, cuda);
    attach(e, ctx);
    return e;
}

// CSpline_Compute on the GPU (throws on failure).
static void compute_spline_coefs(const int* shape, const float* data, float* splineCoefs, int border) {
    if (!shape || !data || !splineCoefs) throw std::runtime_error("photonpy_amd: NULL argument");
    const size_t nin = (size_t)shape[0] * shape[1] * shape[2];
    const int o0 = shape[0] - 3 + 2 * border, o1 = shape[1] - 3 + 2 * border, o2 = shape[2] - 3 + 2 * border;
    if (o0 <= 0 || o1 <= 0 || o2 <= 0) throw std::runtime_error("photonpy_amd: z-stack too small for a spline");
    const size_t nout = (size_t)o0 * o1 * o2 * 64;
    float *din = nullptr, *dout = nullptr;
    hip_check(dev_alloc(&din, nin * sizeof(float)), "hipMalloc");
    hipError_t err = dev_alloc(&dout, nout * sizeof(float));
    if (err == hipSuccess) err = hipMemcpy(din, data, nin * sizeof(float), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = launch_spline_coefs(shape, din, dout, border, 0);
    if (err == hipSuccess) err = hipMemcpy(splineCoefs, dout, nout * sizeof(float), hipMemcpyDeviceToHost);
    (void)dev_release(din);
    (void)dev_release(dout);
    hip_check(err, "CSpline_Compute");
}

extern "C" {

const char* SMLM_GetLastError(void) { return g_lastError.c_str(); }
void SMLM_ClearError(void) { g_lastError.clear(); }

int SMLM_DeviceCount(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

Context* Context_Create(void) { return new Context(); }

void Context_Destroy(Context* ctx) {
    if (!ctx) return;
    std::vector<Estimator*> objs;
    std::vector<EstimationQueue*> queues;
    {
        std::lock_guard<std::mutex> l(ctx->mtx);
        objs.assign(ctx->objects.begin(), ctx->objects.end());
        queues.assign(ctx->queues.begin(), ctx->queues.end());
        ctx->objects.clear();
        ctx->queues.clear();
    }
    for (EstimationQueue* q : queues) queue_destroy(q);
    for (Estimator* e : objs) {
        e->ctx = nullptr;
        delete e;
    }
    delete ctx;
}

Estimator* Gauss2D_CreatePSF_XYIBg(int roisize, float sigmaX, float sigmaY, bool cuda, Context* ctx) {
    return guarded(
        [&]() -> Estimator* {
            GaussCalib gc{};
            gc.sx = sigmaX;
            gc.sy = sigmaY;
            Estimator* e = make_gauss(kKindFixed, roisize, gc);
            set_flavour(e, cuda);
            attach(e, ctx);
            return e;
        },
        nullptr);
}

Estimator* Gauss2D_CreatePSF_XYIBgSigma(int roisize, float initialSigma, bool cuda, Context* ctx) {
    return guarded(
        [&]() -> Estimator* {
            GaussCalib gc{};
            gc.initialSigma = initialSigma;
            Estimator* e = make_gauss(kKindSigma, roisize, gc);
            set_flavour(e, cuda);
            attach(e, ctx);
            return e;
        },
        nullptr);
}

// GaussianPSF.cu:108-129 -- x,y,I,bg with sigma x, y as per-spot constants
Estimator* Gauss2D_CreatePSF_XYIBgConstSigma(int roisize, bool cuda, Context* ctx) {
    return guarded(
        [&]() -> Estimator* {
            Estimator* e = make_gauss(kKindConstSigma, roisize, GaussCalib{});
            set_flavour(e, cuda);
            attach(e, ctx);
            return e;
        },
        nullptr);
}

// GaussianPSF.cu:61-85 -- x,y,I,bg,bgx,bgy with sigma x, y as per-spot constants
Estimator* Gauss2D_CreatePSF_XYITiltedBg(int roisize, bool cuda, Context* ctx) {
    return guarded(
        [&]() -> Estimator* {
            Estimator* e = make_gauss(kKindTiltedBg, roisize, GaussCalib{});
            set_flavour(e, cuda);
            attach(e, ctx);
            return e;
        },
        nullptr);
}

static Estimator* create_sigmaxy(int kind, int roisize, float isx, float isy, bool cuda, Context* ctx) {
    GaussCalib gc{};
    gc.initialSigma = isx;
    gc.initialSigmaY = isy;
    Estimator* e = make_gauss(kind, roisize, gc);
    set_flavour(e, cuda);
    attach(e, ctx);
    return e;
}

// GaussianPSF.cu:163-185 -- x,y,I,bg,sx,sy
Estimator* Gauss2D_CreatePSF_XYIBgSigmaXY(int roisize, float initialSigmaX, float initialSigmaY, bool cuda,
                                          Context* ctx) {
    return guarded([&]() -> Estimator* {
        return create_sigmaxy(kKindSigmaXY, roisize, initialSigmaX, initialSigmaY, cuda, ctx);
    }, nullptr);
}

// GaussianPSF.cu:187-212 -- x,y,I,bg,bgx,bgy,sx,sy
Estimator* Gauss2D_CreatePSF_XYITiltedBgSigmaXY(int roisize, float initialSigmaX, float initialSigmaY, bool cuda,
                                                Context* ctx) {
    return guarded([&]() -> Estimator* {
        return create_sigmaxy(kKindTiltedBgSigmaXY, roisize, initialSigmaX, initialSigmaY, cuda, ctx);
    }, nullptr);
}

// GaussianPSF.cu:216-240.  Per spot i: images [n, imgw, imgw], xy [n, 2] and sigma[i] (one
// float per spot: the i-th element of the array, as the reference reads it); a
// Levenberg-Marquardt fit of {I, bg} from {1, 0} with maxIterations steps and lambda 1e-15,
// then the CRLB of {I, bg}.  roipos is unused, as in the reference; the fit always runs on
// the GPU (the cuda flag chose the reference's CPU or CUDA loop).
void Gauss2D_EstimateIntensityBg(const float* imageData, float* IBg, float* IBg_crlb, int numspots, const float* xy,
                                 const int* roipos, const float* sigma, int imgw, int maxIterations, bool cuda) {
    (void)roipos;
    (void)cuda;
    if (numspots <= 0) return;
    guarded(
        [&]() -> int {
            if (!imageData || !xy || !sigma || !IBg) throw std::runtime_error("photonpy_amd: NULL argument");
            Estimator e;
            e.roisize = imgw;
            e.desc.kind = kKindIBg;
            e.lm.iterations = maxIterations;
            e.lm.lambda = 1e-15f;  // LevMarOptimize's default step (Estimation.h:194-196)
            const size_t n = numspots, S = (size_t)imgw * imgw;
            std::vector<float> cst(n * 3);
            for (size_t i = 0; i < n; i++) {
                cst[3 * i] = xy[2 * i];
                cst[3 * i + 1] = xy[2 * i + 1];
                cst[3 * i + 2] = sigma[i];
            }
            Slicer sl{(char*)e.workspace(Slicer::bytes({n * S * 4, n * 3 * 4, n * 2 * 4, n * 2 * 4, n * 4}))};
            float* ds = sl.take<float>(n * S);
            float* dc = sl.take<float>(n * 3);
            float* dp = sl.take<float>(n * 2);
            float* dcr = sl.take<float>(n * 2);
            int* dit = sl.take<int>(n);
            hip_check(hipMemcpy(ds, imageData, n * S * 4, hipMemcpyHostToDevice), "H2D images");
            hip_check(hipMemcpy(dc, cst.data(), n * 3 * 4, hipMemcpyHostToDevice), "H2D constants");
            run_fit(&e, numspots, ds, dc, nullptr, dp, dit, nullptr, 0, dcr, nullptr, 0);
            hip_check(hipMemcpy(IBg, dp, n * 2 * 4, hipMemcpyDeviceToHost), "D2H IBg");
            if (IBg_crlb) hip_check(hipMemcpy(IBg_crlb, dcr, n * 2 * 4, hipMemcpyDeviceToHost), "D2H IBg CRLB");
            return 0;
        },
        0);
}

static Estimator* create_astig(int roisize, const float* x, const float* y, float minz, float maxz, float Bx, float By,
                               bool cuda, Context* ctx) {
    GaussCalib gc{};
    for (int i = 0; i < 4; i++) {
        gc.ax[i] = x[i];
        gc.ay[i] = y[i];
    }
    gc.ax[4] = Bx;
    gc.ay[4] = By;
    gc.minz = minz;
    gc.maxz = maxz;
    Estimator* e = make_gauss(kKindAstig, roisize, gc);
    set_flavour(e, cuda);
    attach(e, ctx);
    return e;
}

Estimator* Gauss2D_CreatePSF_XYZIBg(int roisize, const Gauss3D_Calibration* c, bool cuda, Context* ctx) {
    return guarded(
        [&]() -> Estimator* {
            if (!c) throw std::runtime_error("photonpy_amd: calibration is NULL");
            return create_astig(roisize, c->x, c->y, c->minz, c->maxz, 0.0f, 0.0f, cuda, ctx);
        },
        nullptr);
}

Estimator* Gauss2D_CreatePSF_XYZIBgQ(int roisize, const Gauss3D_CalibrationQ* c, bool cuda, Context* ctx) {
    return guarded(
        [&]() -> Estimator* {
            if (!c) throw std::runtime_error("photonpy_amd: calibration is NULL");
            return create_astig(roisize, c->x, c->y, c->minz, c->maxz, c->Bx, c->By, cuda, ctx);
        },
        nullptr);
}

Estimator* CSpline_CreatePSF(int roisize, const CSpline_Calibration* calib, const float* coefs, int mode, bool cuda,
                             Context* ctx) {
    return guarded([&]() -> Estimator* { return create_spline(roisize, calib, coefs, mode, cuda, ctx); }, nullptr);
}

void CSpline_Compute(const int* shape, const float* data, float* splineCoefs, int border) {
    guarded(
        [&]() -> int {
            compute_spline_coefs(shape, data, splineCoefs, border);
            return 0;
        },
        0);
}

Estimator* CSpline_CreatePSF_FromFile(int roisize, const char* filename, int mode, bool cuda, Context* ctx) {
    return guarded(
        [&]() -> Estimator* {
            ZStackHeader h;
            std::vector<float> data = load_zstack(filename, h);
            const int cells[3] = {h.shape[0] - 3, h.shape[1] - 3, h.shape[2] - 3};
            if (cells[0] < 1 || cells[1] < 1 || cells[2] < 1)
                throw std::runtime_error("photonpy_amd: z-stack too small for a spline");
            std::vector<float> coefs((size_t)cells[0] * cells[1] * cells[2] * 64);
            compute_spline_coefs(h.shape, data.data(), coefs.data(), 0);
            // the coefficient table's shape (the reference passes the z-stack's own
            // shape here, CubicSplinePSF.cu:141, and so reads past its table)
            CSpline_Calibration calib;
            calib.shape[0] = cells[0];
            calib.shape[1] = cells[1];
            calib.shape[2] = cells[2];
            calib.z_min = h.zmin;
            calib.z_max = h.zmax;
            calib.coefs = coefs.data();
            return create_spline(roisize, &calib, nullptr, mode, cuda, ctx);
        },
        nullptr);
}

void CSpline_Evaluate(const int* shape, const float* splineCoefs, const float* shift, int border, float* values) {
    guarded(
        [&]() -> int {
            if (!shape || !splineCoefs || !shift || !values) throw std::runtime_error("photonpy_amd: NULL argument");
            const int o0 = shape[0] - 2 * border, o1 = shape[1] - 2 * border, o2 = shape[2] - 2 * border;
            if (shape[0] < 1 || shape[1] < 1 || shape[2] < 1 || o0 <= 0 || o1 <= 0 || o2 <= 0)
                throw std::runtime_error("photonpy_amd: invalid spline shape or border");
            const size_t ncoef = (size_t)shape[0] * shape[1] * shape[2] * 64;
            const size_t nout = (size_t)o0 * o1 * o2 * 4;
            float *dc = nullptr, *dout = nullptr;
            hip_check(dev_alloc(&dc, ncoef * sizeof(float)), "hipMalloc");
            hipError_t err = dev_alloc(&dout, nout * sizeof(float));
            if (err == hipSuccess) err = hipMemcpy(dc, splineCoefs, ncoef * sizeof(float), hipMemcpyHostToDevice);
            if (err == hipSuccess) err = launch_spline_evaluate(shape, dc, shift, border, dout, 0);
            if (err == hipSuccess) err = hipMemcpy(values, dout, nout * sizeof(float), hipMemcpyDeviceToHost);
            (void)dev_release(dc);
            if (dout) (void)dev_release(dout);
            hip_check(err, "CSpline_Evaluate");
            return 0;
        },
        0);
}

void Estimator_Delete(Estimator* e) {
    if (!e) return;
    detach(e);
    delete e;
}

const char* Estimator_ParamFormat(Estimator* e) { return e ? e->fmt.c_str() : ""; }

void Estimator_GetProperties(Estimator* e, EstimatorProperties* p) {
    if (!e || !p) return;
    p->numParams = e->K();
    p->sampleCount = e->S();
    p->numDiag = e->numDiag;  // 0 for the models, as the reference's GPU estimators (EstimatorImpl.h:34)
    p->numConst = e->NC();
    p->sampleIndexDims = 2;
}

void Estimator_SampleDims(Estimator* e, int* dims) {
    if (!e || !dims) return;
    dims[0] = e->roisize;
    dims[1] = e->roisize;
}

void Estimator_GetParamLimits(Estimator* e, ParamLimit* limits) {
    if (!e || !limits) return;
    for (size_t i = 0; i < e->limits.size(); i++) limits[i] = e->limits[i];
}

void Estimator_SetLMParams(Estimator* e, const LMParams* p) {
    if (e && p) e->lm = *p;
}

void Estimator_GetLMParams(Estimator* e, LMParams* p) {
    if (e && p) *p = e->lm;
}

int Estimator_SetGroupSize(Estimator* e, int group) {
    e = model_of(e);
    if (!e) return 1;
    if (group < 0 || group > 64 || (group & (group - 1)) != 0) return 1;
    e->group = group;
    return 0;
}

int Estimator_GetGroupSize(Estimator* e) {
    e = model_of(e);
    return e ? e->G() : 0;
}

int Estimator_SetRefill(Estimator* e, int refill) {
    e = model_of(e);
    if (!e || refill < 0 || refill > 64) return 1;
    e->refill = refill;
    return 0;
}

int Estimator_SetWorkSegments(Estimator* e, int segments) {
    e = model_of(e);
    if (!e || (segments != 1 && segments != kWorkSegs)) return 1;
    e->segShift = segments == 1 ? 0 : 3;
    return 0;
}

int Estimator_GetWorkSegments(Estimator* e) {
    e = model_of(e);
    return e ? 1 << e->segShift : 0;
}

int Estimator_SetSplineMode(Estimator* e, int mode) {
    e = model_of(e);
    if (!e || !is_spline(e->desc.kind) || mode < 0 || mode > 1) return 1;
    std::lock_guard<std::mutex> lk(e->mtx);
    e->splineMode = mode;
    e->desc.useGrid = e->gridForm && mode == 0;
    return 0;
}

int Estimator_SplineForm(Estimator* e) {
    e = model_of(e);
    if (!e || !is_spline(e->desc.kind)) return -1;
    return e->desc.useGrid ? 1 : 0;
}

int Estimator_SplineStripLength(Estimator* e) {
    e = model_of(e);
    if (!e || !is_spline(e->desc.kind) || !e->desc.useGrid) return -1;
    if (e->roisize < kMinFitR || e->roisize > kMaxFitR) return 0;
    return SplineGridModel::strip_len(e->roisize, e->G());
}

int Estimator_SetPhaseTiming(Estimator* e, int enable) {
    e = model_of(e);
    if (!e) return 1;
    std::lock_guard<std::mutex> lk(e->mtx);
    e->phaseTiming = enable != 0;
    if (!e->phaseTiming) e->release_phase_events();
    return 0;
}

int Estimator_GetPhaseTimes(Estimator* e, float* fitMs, float* crlbMs) {
    e = model_of(e);
    if (!e) return -1;
    return guarded(
        [&]() -> int {
            std::lock_guard<std::mutex> lk(e->mtx);
            const int launches = (int)(e->phaseEvents.size() / 3);
            double fit = 0.0, crlb = 0.0;
            for (int i = 0; i < launches; i++) {
                hipEvent_t* ev = e->phaseEvents.data() + 3 * i;
                float a = 0.0f, b = 0.0f;
                hip_check(hipEventSynchronize(ev[2]), "hipEventSynchronize");
                hip_check(hipEventElapsedTime(&a, ev[0], ev[1]), "hipEventElapsedTime");
                hip_check(hipEventElapsedTime(&b, ev[1], ev[2]), "hipEventElapsedTime");
                fit += a;
                crlb += b;
            }
            e->release_phase_events();
            if (fitMs) *fitMs = launches ? (float)(fit / launches) : 0.0f;
            if (crlbMs) *crlbMs = launches ? (float)(crlb / launches) : 0.0f;
            return launches;
        },
        -1);
}

// Uploads the per-spot constants of a host call (NULL when the model has none).
static const float* upload_constants(Estimator* e, Slicer& sl, int n, const float* constants) {
    const size_t C = e->NC();
    if (C == 0) return nullptr;
    if (!constants)
        throw std::runtime_error("photonpy_amd: this estimator needs " + std::to_string(C) + " constants per spot");
    float* dc = sl.take<float>(n * C);
    hip_check(hipMemcpy(dc, constants, n * C * 4, hipMemcpyHostToDevice), "H2D constants");
    return dc;
}

void Estimator_ComputeExpectedValue(Estimator* e, int n, const float* params, const float* constants,
                                    const int* spotpos, float* ev) {
    e = model_of(e);
    (void)spotpos;
    if (!e || n <= 0) return;
    guarded(
        [&]() -> int {
            std::lock_guard<std::mutex> l(e->mtx);
            e->ensure_device();
            const size_t K = e->K(), S = e->S(), C = e->NC();
            Slicer sl{(char*)e->workspace(Slicer::bytes({n * K * 4, n * S * 4, n * C * 4}))};
            float* dp = sl.take<float>(n * K);
            float* dev = sl.take<float>(n * S);
            const float* dcn = upload_constants(e, sl, n, constants);
            hip_check(hipMemcpy(dp, params, n * K * 4, hipMemcpyHostToDevice), "H2D params");
            hip_check(launch_ev_any(e->desc, e->roisize, n, dp, dcn, dev, nullptr, 0), "expected value kernel");
            hip_check(hipMemcpy(ev, dev, n * S * 4, hipMemcpyDeviceToHost), "D2H ev");
            return 0;
        },
        0);
}

void Estimator_ComputeDerivatives(Estimator* e, int n, const float* params, const float* constants,
                                  const int* spotpos, float* deriv, float* ev) {
    e = model_of(e);
    (void)spotpos;
    if (!e || n <= 0) return;
    guarded(
        [&]() -> int {
            std::lock_guard<std::mutex> l(e->mtx);
            e->ensure_device();
            const size_t K = e->K(), S = e->S(), C = e->NC();
            Slicer sl{(char*)e->workspace(Slicer::bytes({n * K * 4, n * S * 4, n * K * S * 4, n * C * 4}))};
            float* dp = sl.take<float>(n * K);
            float* dev = sl.take<float>(n * S);
            float* dd = sl.take<float>(n * K * S);
            const float* dcn = upload_constants(e, sl, n, constants);
            hip_check(hipMemcpy(dp, params, n * K * 4, hipMemcpyHostToDevice), "H2D params");
            hip_check(launch_ev_any(e->desc, e->roisize, n, dp, dcn, dev, dd, 0), "derivatives kernel");
            if (ev) hip_check(hipMemcpy(ev, dev, n * S * 4, hipMemcpyDeviceToHost), "D2H ev");
            if (deriv) hip_check(hipMemcpy(deriv, dd, n * K * S * 4, hipMemcpyDeviceToHost), "D2H deriv");
            return 0;
        },
        0);
}

void Estimator_Estimate(Estimator* e, int n, const float* sample, const float* constants, const int* spotpos,
                        const float* initial, float* params, float* diagnostics, int* iterations, float* trace,
                        int traceBufLen) {
    (void)spotpos;
    Estimator* outer = e;
    e = model_of(e);
    if (!e || n <= 0) return;
    guarded(
        [&]() -> int {
            std::lock_guard<std::mutex> l(e->mtx);
            const size_t K = e->K(), S = e->S(), C = e->NC();
            const size_t tl = (trace && traceBufLen > 0) ? (size_t)traceBufLen : 0;
            // diagnostics are computed for a GLRT wrapper; a cuda = false estimator's
            // column is left as the caller passed it (set_flavour)
            const size_t D = (diagnostics && outer->inner) ? (size_t)outer->numDiag : 0;
            Slicer sl{(char*)e->workspace(Slicer::bytes(
                {n * S * 4, initial ? n * K * 4 : 0, n * K * 4, (size_t)n * 4, n * tl * K * 4, n * C * 4, n * D * 4}))};
            float* ds = sl.take<float>(n * S);
            float* di = initial ? sl.take<float>(n * K) : nullptr;
            float* dp = sl.take<float>(n * K);
            int* dit = sl.take<int>(n);
            float* dt = tl ? sl.take<float>(n * tl * K) : nullptr;
            const float* dcn = upload_constants(e, sl, n, constants);
            float* dd = D ? sl.take<float>(n * D) : nullptr;
            if (di) hip_check(hipMemcpy(di, initial, n * K * 4, hipMemcpyHostToDevice), "H2D initial");
            // Large calls without trace or diagnostics upload the samples in chunks of
            // up to 32 MB: chunk k is copied by the process's copy threads into pinned
            // staging buffer k % 2, sent on the copy stream and fitted on the fit stream
            // while chunk k + 1 is being copied (each spot's fit is independent of the
            // batch it is in, so the results equal one fit over all spots).
            const int chunk = std::max(4096, (int)((32u << 20) / (S * 4)));
            if (!dt && !dd && n >= 2 * chunk) {
                e->ensure_upload((size_t)chunk * S);
                auto& u = e->up;
                for (int i0 = 0, k = 0; i0 < n; i0 += chunk, k++) {
                    const int b = k & 1, m = std::min(chunk, n - i0);
                    // staging[b] was last read by the H2D copy of chunk k - 2
                    if (k >= 2) hip_check(hipEventSynchronize(u.copied[b]), "hipEventSynchronize(staging)");
                    parallel_copy(u.staging[b], sample + (size_t)i0 * S, (size_t)m * S * 4);
                    hip_check(hipMemcpyAsync(ds + (size_t)i0 * S, u.staging[b], (size_t)m * S * 4, hipMemcpyHostToDevice,
                                             u.copy),
                              "H2D samples");
                    hip_check(hipEventRecord(u.copied[b], u.copy), "hipEventRecord");
                    hip_check(hipStreamWaitEvent(u.fit, u.copied[b], 0), "hipStreamWaitEvent");
                    run_fit(e, m, ds + (size_t)i0 * S, dcn ? dcn + (size_t)i0 * C : nullptr,
                            di ? di + (size_t)i0 * K : nullptr, dp + (size_t)i0 * K, dit + i0, nullptr, 0, nullptr,
                            nullptr, u.fit);
                }
                hip_check(hipStreamSynchronize(u.fit), "hipStreamSynchronize(fit)");
            } else {
                hip_check(hipMemcpy(ds, sample, n * S * 4, hipMemcpyHostToDevice), "H2D samples");
                if (dt) hip_check(hipMemset(dt, 0, n * tl * K * 4), "clear trace");
                run_fit(e, n, ds, dcn, di, dp, dit, dt, (int)tl, nullptr, nullptr, 0);
            }
            if (dd) {
                hip_check(launch_glrt_any(e->desc, e->roisize, n, dp, dcn, ds, outer->glrtMeanOfSpot0, dd, 0),
                          "glrt kernel");
                hip_check(hipMemcpy(diagnostics, dd, n * D * 4, hipMemcpyDeviceToHost), "D2H diagnostics");
            }
            hip_check(hipMemcpy(params, dp, n * K * 4, hipMemcpyDeviceToHost), "D2H params");
            if (iterations) hip_check(hipMemcpy(iterations, dit, (size_t)n * 4, hipMemcpyDeviceToHost), "D2H iterations");
            if (dt) hip_check(hipMemcpy(trace, dt, n * tl * K * 4, hipMemcpyDeviceToHost), "D2H trace");
            return 0;
        },
        0);
}

void Estimator_ChiSquareAndCRLB(Estimator* e, const float* params, const float* sample, const float* constants,
                                const int* spotpos, int n, float* crlb, float* chisq) {
    e = model_of(e);
    (void)spotpos;
    if (!e || n <= 0) return;
    guarded(
        [&]() -> int {
            std::lock_guard<std::mutex> l(e->mtx);
            e->ensure_device();
            const size_t K = e->K(), S = e->S(), C = e->NC();
            Slicer sl{(char*)e->workspace(
                Slicer::bytes({n * K * 4, sample ? n * S * 4 : 0, n * K * 4, (size_t)n * 4, n * C * 4}))};
            float* dp = sl.take<float>(n * K);
            float* ds = sample ? sl.take<float>(n * S) : nullptr;
            float* dc = sl.take<float>(n * K);
            float* dx = sl.take<float>(n);
            const float* dcn = upload_constants(e, sl, n, constants);
            hip_check(hipMemcpy(dp, params, n * K * 4, hipMemcpyHostToDevice), "H2D params");
            if (ds) hip_check(hipMemcpy(ds, sample, n * S * 4, hipMemcpyHostToDevice), "H2D samples");
            hip_check(launch_crlb_any(e->desc, e->roisize, n, dp, dcn, ds, dc, dx, 0), "crlb kernel");
            if (crlb) hip_check(hipMemcpy(crlb, dc, n * K * 4, hipMemcpyDeviceToHost), "D2H crlb");
            if (chisq) hip_check(hipMemcpy(chisq, dx, (size_t)n * 4, hipMemcpyDeviceToHost), "D2H chisq");
            return 0;
        },
        0);
}

// ---------------------------------------------------------------- device-pointer API
static void require_constants(Estimator* e, const float* d_constants) {
    if (e->NC() > 0 && !d_constants)
        throw std::runtime_error("photonpy_amd: this estimator needs " + std::to_string(e->NC()) +
                                 " constants per spot");
}

int Estimator_EstimateDevice(Estimator* e, int n, const float* d_sample, const float* d_constants,
                             const float* d_initial, float* d_params, int* d_iterations, float* d_trace, int traceLen,
                             float* d_crlb, float* d_chisq, void* stream) {
    e = model_of(e);
    if (!e) return 1;
    if (n <= 0) return 0;
    return guarded(
        [&]() -> int {
            std::lock_guard<std::mutex> l(e->mtx);
            run_fit(e, n, d_sample, d_constants, d_initial, d_params, d_iterations, d_trace, traceLen, d_crlb, d_chisq,
                    as_stream(stream));
            return 0;
        },
        2);
}

// Estimator_EstimateDevice without any allocation: the work counter comes from the
// library's fixed device array (see scratch_on_current_device).
int Estimator_EstimateDeviceNoAlloc(Estimator* e, int n, const float* d_sample, const float* d_constants,
                                    const float* d_initial, float* d_params, int* d_iterations, float* d_trace,
                                    int traceLen, float* d_crlb, float* d_chisq, void* stream) {
    e = model_of(e);
    if (!e) return 1;
    if (n <= 0) return 0;
    return guarded(
        [&]() -> int {
            std::lock_guard<std::mutex> l(e->mtx);
            int* counters = scratch_on_current_device();
            int* counter = counters + (size_t)(g_noAllocSlot.fetch_add(1) % kNoAllocSlots) * kCounterInts;
            run_fit(e, n, d_sample, d_constants, d_initial, d_params, d_iterations, d_trace, traceLen, d_crlb, d_chisq,
                    as_stream(stream), false, 0, counter);
            return 0;
        },
        2);
}

int Estimator_ExpectedValueDevice(Estimator* e, int n, const float* d_params, const float* d_constants, float* d_ev,
                                  void* stream) {
    e = model_of(e);
    if (!e) return 1;
    if (n <= 0) return 0;
    return guarded(
        [&]() -> int {
            std::lock_guard<std::mutex> l(e->mtx);
            e->ensure_device();
            require_constants(e, d_constants);
            hip_check(launch_ev_any(e->desc, e->roisize, n, d_params, d_constants, d_ev, nullptr, as_stream(stream)),
                      "ev kernel");
            return 0;
        },
        2);
}

int Estimator_DerivativesDevice(Estimator* e, int n, const float* d_params, const float* d_constants, float* d_deriv,
                                float* d_ev, void* stream) {
    e = model_of(e);
    if (!e) return 1;
    if (n <= 0) return 0;
    return guarded(
        [&]() -> int {
            std::lock_guard<std::mutex> l(e->mtx);
            e->ensure_device();
            require_constants(e, d_constants);
            hip_check(launch_ev_any(e->desc, e->roisize, n, d_params, d_constants, d_ev, d_deriv, as_stream(stream)),
                      "derivatives kernel");
            return 0;
        },
        2);
}

int Estimator_ChiSquareAndCRLBDevice(Estimator* e, int n, const float* d_params, const float* d_constants,
                                     const float* d_sample, float* d_crlb, float* d_chisq, void* stream) {
    e = model_of(e);
    if (!e) return 1;
    if (n <= 0) return 0;
    return guarded(
        [&]() -> int {
            std::lock_guard<std::mutex> l(e->mtx);
            e->ensure_device();
            require_constants(e, d_constants);
            hip_check(launch_crlb_any(e->desc, e->roisize, n, d_params, d_constants, d_sample, d_crlb, d_chisq,
                                      as_stream(stream)),
                      "crlb kernel");
            return 0;
        },
        2);
}

int Estimator_DiagnosticsDevice(Estimator* e, int n, const float* d_sample, const float* d_constants,
                                const float* d_params, float* d_diag, void* stream) {
    if (!e) return 1;
    if (n <= 0 || !e->inner) return 0;
    Estimator* outer = e;
    e = model_of(e);
    return guarded(
        [&]() -> int {
            std::lock_guard<std::mutex> l(e->mtx);
            e->ensure_device();
            require_constants(e, d_constants);
            if (!d_sample || !d_params || !d_diag) throw std::runtime_error("photonpy_amd: NULL argument");
            hip_check(launch_glrt_any(e->desc, e->roisize, n, d_params, d_constants, d_sample, outer->glrtMeanOfSpot0,
                                      d_diag, as_stream(stream)),
                      "glrt kernel");
            return 0;
        },
        2);
}

Estimator* GLRT_CreatePSF(Estimator* model, Context* ctx) {
    return guarded(
        [&]() -> Estimator* {
            if (!model) throw std::runtime_error("photonpy_amd: GLRT_CreatePSF needs an estimator");
            auto* g = new Estimator();
            g->inner = model;
            g->roisize = model->roisize;
            g->desc = model->desc;  // for the shapes (K, constants, samples); fits run on the model
            g->fmt = model->fmt;
            g->limits = model->limits;
            g->numDiag = 3;
            attach(g, ctx);
            return g;
        },
        nullptr);
}

int GLRT_SetMeanOfFirstSpot(Estimator* e, int enable) {
    if (!e || e->numDiag != 3) return 1;
    e->glrtMeanOfSpot0 = enable != 0;
    return 0;
}

int Estimator_NumShared(Estimator* e) {
    e = model_of(e);
    return e ? (int)e->num_shared() : -1;
}

int Estimator_GetShared(Estimator* e, float* out) {
    e = model_of(e);
    if (!e || !out) return 1;
    std::vector<float> v = e->shared_values();
    std::copy(v.begin(), v.end(), out);
    return 0;
}

int Estimator_NLLGradDevice(Estimator* e, int n, const float* d_params, const float* d_sample, const float* d_shared,
                            const float* d_gout, float* d_nll, float* d_grad_params, float* d_grad_shared,
                            void* stream) {
    e = model_of(e);
    if (!e) return 1;
    if (n <= 0) return 0;
    return guarded(
        [&]() -> int {
            std::lock_guard<std::mutex> l(e->mtx);
            if (e->desc.kind > kKindSpline)
                throw std::runtime_error("photonpy_amd: the adjoint covers the fixed-sigma, free-sigma, astigmatic and "
                                         "spline models");
            AdjArgs a{};
            a.n = n;
            a.R = e->roisize;
            a.params = d_params;
            a.sample = d_sample;
            a.gout = d_gout;
            a.shared = d_shared ? d_shared : e->device_shared();
            a.nll = d_nll;
            a.gparams = d_grad_params;
            a.gshared = d_grad_shared;
            void* ws = nullptr;
            bool sorted = false;
            if (e->desc.kind == kKindSpline && spline_adjoint_tiled(e->roisize)) {
                const long long nkeys = (long long)e->desc.sm.Nz * e->desc.sm.Ny * e->desc.sm.Nx;
                if (nkeys > (1LL << 30)) throw std::runtime_error("photonpy_amd: spline table too large");
                ws = e->adjoint_workspace(spline_adjoint_workspace(n, (int)nkeys));
                // a gradient call with upstream weights (the backward of an autograd step)
                // reuses the window sort of the previous call for as many spots on the same
                // stream (its forward); a call without them sorts (per-spot results do not
                // depend on the order, a stale sort only costs speed)
                sorted = (d_grad_params || d_grad_shared) && d_gout && e->adjSortN == n &&
                         e->adjSortStream == stream;
                e->adjSortN = -1;  // recorded again only once this call's launches succeeded
            }
            hip_check(launch_adjoint_any(e->desc, a, ws, sorted, as_stream(stream)), "adjoint kernel");
            if (ws) {
                e->adjSortN = n;
                e->adjSortStream = stream;
            }
            return 0;
        },
        2);
}

}  // extern "C"

