/* photonpy_amd — MI355X-native batched Poisson-MLE spot fitting.
 *
 * C ABI of libphotonpy_amd.so.  The first block keeps the names, argument
 * order and meaning of the reference's exported functions so that the
 * reference's ctypes bindings (photonpy/cpp/estimator.py, gaussian.py,
 * cspline.py, context.py) bind to it unchanged; every pointer there is host
 * memory and every call is synchronous.  Each declaration cites the reference
 * function it replaces.  The second block adds device-pointer entry points
 * (inputs already resident in HBM, asynchronous on a caller-supplied HIP
 * stream) that have no counterpart in the reference.
 *
 * Errors: every caught failure goes to the debug-print callback
 * (SetDebugPrintCallback) and is kept for SMLM_GetLastError(); creators also
 * return NULL and device-pointer calls return a nonzero code.  numspots == 0
 * is a no-op everywhere, as in the reference.
 *
 * The `cuda` flag of the creators: true gives the reference's GPU estimator
 * interface (NumDiag 0), false its CPU estimator interface (NumDiag 1, a
 * diagnostics column that Estimate leaves untouched, EstimatorImpl.h:108-185).
 * Both compute on the GPU; there is no CPU code path.
 */
#ifndef PHOTONPY_AMD_H
#define PHOTONPY_AMD_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct Context Context;
typedef struct Estimator Estimator;
typedef struct EstimationQueue EstimationQueue;

/* Estimators/Estimator.h:18-26 */
typedef struct LMParams {
    int iterations;
    float lambda;
    float factor; /* accepted and stored; unused by the fit, as in the reference */
} LMParams;

/* Estimators/Estimator.h:28-31 */
typedef struct ParamLimit {
    float min, max;
} ParamLimit;

/* Estimators/Estimator.h:151-154 */
typedef struct EstimatorProperties {
    int numParams, sampleCount, numDiag, numConst, sampleIndexDims;
} EstimatorProperties;

/* Estimators/Gaussian/GaussianPSF.h:27-33: x/y = {s0, gamma, d, A} */
typedef struct Gauss3D_Calibration {
    float x[4];
    float y[4];
    float minz, maxz;
} Gauss3D_Calibration;

/* Gauss3D_Calibration followed by the quartic coefficients B (per axis) of
 * sigma(z) = s0 sqrt(1 + t^2/d^2 + A t^3/d^2 + B t^4/d^2), t = z - gamma.
 * No reference counterpart (the reference model has B = 0). */
typedef struct Gauss3D_CalibrationQ {
    float x[4];
    float y[4];
    float minz, maxz;
    float Bx, By;
} Gauss3D_CalibrationQ;

/* Estimators/EstimationQueue.h:21-25 */
typedef struct EstimatorResult {
    int id;
    float chisq;
    int iterations;
} EstimatorResult;

/* Estimators/CSpline/CubicSplinePSF.h:12-23; shape = {Nz, Ny, Nx} */
typedef struct CSpline_Calibration {
    int shape[3];
    float z_min;
    float z_max;
    const float* coefs; /* [Nz, Ny, Nx, 64] */
} CSpline_Calibration;

/* ---------------------------------------------------------------------- */
/* Reference-compatible surface (host pointers, synchronous)                */

/* Context.cpp:65-73 */
Context* Context_Create(void);
void Context_Destroy(Context* ctx);

/* Estimators/Gaussian/GaussianPSF.cu:87-108 — params x,y,I,bg */
Estimator* Gauss2D_CreatePSF_XYIBg(int roisize, float sigmaX, float sigmaY, bool cuda, Context* ctx);
/* Estimators/Gaussian/GaussianPSF.cu:136-158 — params x,y,I,bg,sigma */
Estimator* Gauss2D_CreatePSF_XYIBgSigma(int roisize, float initialSigma, bool cuda, Context* ctx);
/* Estimators/Gaussian/GaussianPSF.cu:108-129 — params x,y,I,bg; constants sigma x,y per spot */
Estimator* Gauss2D_CreatePSF_XYIBgConstSigma(int roisize, bool cuda, Context* ctx);
/* Estimators/Gaussian/GaussianPSF.cu:61-85 — params x,y,I,bg,bgx,bgy; constants sigma x,y */
Estimator* Gauss2D_CreatePSF_XYITiltedBg(int roisize, bool cuda, Context* ctx);
/* Estimators/Gaussian/GaussianPSF.cu:163-185 — params x,y,I,bg,sx,sy */
Estimator* Gauss2D_CreatePSF_XYIBgSigmaXY(int roisize, float initialSigmaX, float initialSigmaY, bool cuda,
                                          Context* ctx);
/* Estimators/Gaussian/GaussianPSF.cu:187-212 — params x,y,I,bg,bgx,bgy,sx,sy */
Estimator* Gauss2D_CreatePSF_XYITiltedBgSigmaXY(int roisize, float initialSigmaX, float initialSigmaY, bool cuda,
                                                Context* ctx);
/* Estimators/Gaussian/GaussianPSF.cu:216-240 — per spot: I and bg of images[n,imgw,imgw]
 * at xy[n,2] with width sigma[i] (one float per spot i), IBg[n,2], IBg_crlb[n,2];
 * maxIterations LM steps from {1, 0}.  roipos unused; computed on the GPU whatever cuda says. */
void Gauss2D_EstimateIntensityBg(const float* imageData, float* IBg, float* IBg_crlb, int numspots, const float* xy,
                                 const int* roipos, const float* sigma, int imgw, int maxIterations, bool cuda);
/* Estimators/Gaussian/GaussianPSF.cu:36-59 — params x,y,z,I,bg */
Estimator* Gauss2D_CreatePSF_XYZIBg(int roisize, const Gauss3D_Calibration* calib, bool cuda, Context* ctx);
/* Estimators/CSpline/CubicSplinePSF.cu:55-107 — mode 0 (CSPLINE_FLATBG): params x,y,z,I,bg;
 * mode 1 (TILTEDBG): x,y,z,I,bg,bgx,bgy; mode 2 (BGIMAGE): x,y,z,I,bg with an R x R background
 * image per spot as its constants (NumConstants = roisize^2) */
Estimator* CSpline_CreatePSF(int roisize, const CSpline_Calibration* calib, const float* coefs, int mode, bool cuda,
                             Context* ctx);
/* Estimators/CSpline/CubicSplinePSF.cu:151-204 — shape {Nz,Ny,Nx};
 * splineCoefs [(Nz-3+2b),(Ny-3+2b),(Nx-3+2b),64] */
void CSpline_Compute(const int* shape, const float* data, float* splineCoefs, int border);
/* Estimators/CSpline/CubicSplinePSF.cu:135-146 — a .zstack file (24-byte header {version 1,
 * shape[3], zmin, zmax} + float32 samples) to a spline PSF (border 0).  The calibration takes
 * the coefficient table's shape (z-stack shape - 3); the reference passes the z-stack shape. */
Estimator* CSpline_CreatePSF_FromFile(int roisize, const char* filename, int mode, bool cuda, Context* ctx);
/* Estimators/CSpline/CubicSplinePSF.cu:208-271 — shape {Nz,Ny,Nx} of splineCoefs[Nz,Ny,Nx,64],
 * shift {z,y,x}; values[(Nz-2b)(Ny-2b)(Nx-2b), 4] = value, d/dx, d/dy, d/dz. */
void CSpline_Evaluate(const int* shape, const float* splineCoefs, const float* shift, int border, float* values);

/* Estimators/Estimator.cpp:45-100 */
void Estimator_Delete(Estimator* e);
const char* Estimator_ParamFormat(Estimator* e);
void Estimator_GetProperties(Estimator* e, EstimatorProperties* props);
void Estimator_SampleDims(Estimator* e, int* dims);
void Estimator_GetParamLimits(Estimator* e, ParamLimit* limits);
void Estimator_SetLMParams(Estimator* e, const LMParams* p);
void Estimator_GetLMParams(Estimator* e, LMParams* p);

/* Estimators/Estimator.cpp:102-113 — ev[numspots, H, W] */
void Estimator_ComputeExpectedValue(Estimator* e, int numspots, const float* params, const float* constants,
                                    const int* spotpos, float* ev);
/* Estimators/Estimator.cpp:115-128 — initial may be NULL (centre-of-mass start),
 * trace [numspots, traceBufLen, K] may be NULL; diagnostics [numspots, numDiag] (may be NULL):
 * numDiag is 0 for the models and 3 for a GLRT wrapper.  A call with at least two 32 MB
 * chunks of samples and no trace or diagnostics buffer uploads the chunks while earlier
 * ones are fitted (same results), through two 32 MB pinned buffers that the estimator
 * keeps until it is deleted. */
void Estimator_Estimate(Estimator* e, int numspots, const float* sample, const float* constants, const int* spotpos,
                        const float* initial, float* params, float* diagnostics, int* iterations, float* trace,
                        int traceBufLen);
/* Estimators/Estimator.cpp:131-144 — deriv[numspots, K, H, W], ev[numspots, H, W] */
void Estimator_ComputeDerivatives(Estimator* e, int numspots, const float* params, const float* constants,
                                  const int* spotpos, float* psf_deriv, float* ev);
/* Estimators/Estimator.cpp:146-157 — sample may be NULL (chisq then 0) */
void Estimator_ChiSquareAndCRLB(Estimator* e, const float* params, const float* sample, const float* constants,
                                const int* spotpos, int numspots, float* crlb, float* chisq);

/* Estimators/EstimationQueue.h:111-127, EstimationQueue.cpp:329-412.  Asynchronous
 * batches from host memory: Schedule copies spots into pinned batches of batchSize and
 * blocks while GetQueueLength() (spots in queued batches + batches in flight) is at
 * least maxQueueLen; a worker thread fits each batch (+ CRLB, chi^2) on one of
 * numStreams streams (< 0: 3); GetResults copies whole finished batches, newest first,
 * while they fit in maxresults and returns the number of spots copied (estim [n,K],
 * crlb [n,K], roipos [n,2], samples [n,H,W] when keepSamples, result [n]; each may be
 * NULL; diag [n, numDiag] for a GLRT wrapper).  h_constants [n, NumConstants] is required by the
 * models that have per-spot constants.  A queue created in a Context is deleted by
 * Context_Destroy. */
EstimationQueue* EstimQueue_Create(Estimator* psf, int batchSize, int maxQueueLen, bool keepSamples, int numStreams,
                                   Context* ctx);
void EstimQueue_Delete(EstimationQueue* q);
void EstimQueue_Schedule(EstimationQueue* q, int numspots, const int* ids, const float* h_samples,
                         const float* h_initial, const float* h_constants, const int* h_roipos);
void EstimQueue_Flush(EstimationQueue* q);
bool EstimQueue_IsIdle(EstimationQueue* q);
int EstimQueue_GetResultCount(EstimationQueue* q);
int EstimQueue_GetQueueLength(EstimationQueue* q);
int EstimQueue_GetResults(EstimationQueue* q, int maxresults, float* estim, float* diag, float* crlb, int* roipos,
                          float* samples, EstimatorResult* result);

/* Estimators/GLRT_PSF.cu:58-169 — wraps `model` (not owned: delete the wrapper first);
 * every call computes with the model and its LM parameters, and Estimate also writes
 * numDiag = 3 diagnostics per spot {ll_on, ll_off, mean}: the Poisson log-likelihoods
 * sum d log(mu) - mu of the fit and of a flat background at `mean` (mu, d clamped at 1e-8).
 * As in the reference, `mean` is the mean sample of the first spot of each call (each
 * batch for a queue); GLRT_SetMeanOfFirstSpot(e, 0) makes it each spot's own. */
Estimator* GLRT_CreatePSF(Estimator* model, Context* ctx);

/* ---------------------------------------------------------------------- */
/* Process-level functions the reference's Python loader binds at load time  */
/* (photonpy/cpp/lib.py:71-153, gaussian.py:76-84).                          */

/* CudaUtils.cu:12-16 — number of visible HIP devices */
int CudaGetNumDevices(void);
/* CudaUtils.cu:18-21 — makes device `index` current for the calling thread */
bool CudaSetDevice(int index);
/* CudaUtils.cu:23-32 — compute-unit count and device name (truncated to namelen - 1) */
bool CudaGetDeviceInfo(int index, int* numMultiprocessors, char* name, int namelen);
/* CudaUtils.cu:38-46 — bytes held by this library's device and pinned host allocations
 * (pitched allocations do not exist here: those two tallies are 0) */
void CudaGetMemoryUse(uint64_t* pinnedBytes, uint64_t* devicePitchedBytes, uint64_t* deviceBytes,
                      uint64_t* pitchedNumAllocs);
/* StringUtils.cpp:45-49 — receives every error message a C-ABI call catches (also kept for
 * SMLM_GetLastError); a callback returning nonzero lets the library echo it to stderr */
void SetDebugPrintCallback(int (*cb)(const char* msg));
/* DebugImageCallback.cpp:14-26 — ShowDebugImage passes its arguments to the callback */
void SetDebugImageCallback(void (*cb)(int width, int height, int numImages, const float* data, const char* title));
void ShowDebugImage(int w, int h, int numImages, const float* data, const char* title);
/* FFT.cpp:151-172, Rendering.cpp:12-59: bound by the loader but outside the spot-fitting
 * path; each reports "not provided" through the error channel and does nothing else */
void FFT(const void* src, void* dst, int batchsize, int siglen, int forward);
void FFT2(const void* src, void* dst, int batchsize, int sigw, int sigh, int forward);
void DrawROIs(float* image, int width, int height, const float* rois, int numrois, int roisize, const int* roiposYX);
void Gauss2D_Draw(double* image, int imgw, int imgh, const float* spotList, int nspots, float addSigma);

/* ---------------------------------------------------------------------- */
/* MI355X additions (no reference counterpart)                              */

/* Lanes per spot the queue's batch fits use: the estimator's own group size, or for
 * batches of a few thousand spots up to 16 (batchSize x lanes <= 65536) where a kernel is
 * compiled, so a small batch fills more of the GPU.  Results then differ from a direct fit
 * with the estimator's own group size in summation order only. */
int EstimQueue_GroupSize(EstimationQueue* q);
/* GLRT wrapper: 1 (default) = the reference's mean of the call's first spot, 0 = each
 * spot's own mean.  Returns 0, or 1 when e is not a GLRT wrapper. */
int GLRT_SetMeanOfFirstSpot(Estimator* e, int enable);
/* Diagnostics of fitted spots on the device: d_diag [n, numDiag] from d_sample [n,H,W] and
 * d_params [n,K] (a GLRT wrapper: {ll_on, ll_off, mean}); a no-op when numDiag = 0. */
int Estimator_DiagnosticsDevice(Estimator* e, int n, const float* d_sample, const float* d_constants,
                                const float* d_params, float* d_diag, void* stream);

/* Last error message of the calling thread ("" if none); cleared by SMLM_ClearError. */
const char* SMLM_GetLastError(void);
void SMLM_ClearError(void);
/* Number of visible HIP devices (0 without a GPU). */
int SMLM_DeviceCount(void);

/* Astigmatic model with the quartic sigma(z) term. */
Estimator* Gauss2D_CreatePSF_XYZIBgQ(int roisize, const Gauss3D_CalibrationQ* calib, bool cuda, Context* ctx);

/* Lane-group size (power of two) used by the fit kernel of this estimator;
 * 0 restores the built-in choice.  Returns nonzero when no kernel is compiled
 * for (roisize, group). */
int Estimator_SetGroupSize(Estimator* e, int group);
int Estimator_GetGroupSize(Estimator* e);
/* Number of stopped lane groups of a wavefront that are refilled together
 * (0 = one eighth of the wavefront's groups, at least 1). */
int Estimator_SetRefill(Estimator* e, int refill);
/* Work counters of a fit launch larger than the resident grid (no reference
 * counterpart; a tuning knob like the two above): 1 = one counter, 8 = the spots in 8
 * contiguous segments with a counter each, one per XCD.  Returns nonzero for other
 * values.  Results do not depend on it. */
int Estimator_SetWorkSegments(Estimator* e, int segments);
int Estimator_GetWorkSegments(Estimator* e);

/* Kernel timing of the fit (no reference counterpart; for benchmarks).  While
 * enabled, every fit launch records HIP events on its stream around the fit
 * kernel and the CRLB kernel.  Estimator_GetPhaseTimes waits for the recorded
 * launches, writes the mean duration in ms of each kernel and releases the
 * events; it returns the number of launches averaged (-1 on error). */
/* Spline evaluation form (no reference counterpart).  A coefficient table that CSpline_Compute
 * made is the local tricubic interpolant of a sample grid; the library recovers that grid at
 * creation (checked against every coefficient) and by default fits and evaluates from it (mode
 * 0).  Mode 1 keeps the coefficient form.  Estimator_SplineForm: 1 grid, 0 coefficients, -1 not
 * a spline estimator.  Results of the two forms agree to FP32 rounding of the model. */
int Estimator_SetSplineMode(Estimator* e, int mode);
int Estimator_SplineForm(Estimator* e);
/* Pixel layout of a grid-form spline fit (no reference counterpart; for FLOP accounting):
 * the strip length L of the strip layout at the estimator's ROI size and lane-group size
 * (each lane walks vertical strips of up to L pixels, contracting each window row once per
 * strip), 0 for the flat layout (every pixel contracts its own 4x4 block), -1 when the
 * estimator is not a grid-form spline. */
int Estimator_SplineStripLength(Estimator* e);

int Estimator_SetPhaseTiming(Estimator* e, int enable);
int Estimator_GetPhaseTimes(Estimator* e, float* fitMs, float* crlbMs);

/* Fused fit on device memory: Levenberg-Marquardt from the centre-of-mass start
 * (or d_initial), then CRLB and chi^2 when d_crlb / d_chisq are non-NULL.
 * d_sample [n,H,W], d_params [n,K], d_iterations [n], d_trace [n,traceLen,K]. */
int Estimator_EstimateDevice(Estimator* e, int n, const float* d_sample, const float* d_constants,
                             const float* d_initial, float* d_params, int* d_iterations, float* d_trace, int traceLen,
                             float* d_crlb, float* d_chisq, void* stream);
/* No counterpart in the reference.  For callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc): Estimator_EstimateDevice with the
 * same arguments and results, allocating nothing.  Its scratch -- the fit kernel's 8 work
 * counters -- comes from a fixed array of 16 slots of 8 counters 256 bytes apart (32 KiB) in
 * the library's device memory, one slot per call in turn, reset by a kernel on `stream`; it
 * does not grow with n (n is bounded as for Estimator_EstimateDevice).  Calls that overlap on
 * one GPU with 16 or more others in flight share a slot and race.  A spline estimator's tables must already be on
 * the device (any other fit or evaluation call on it uploads them); otherwise the call fails
 * with an error. */
int Estimator_EstimateDeviceNoAlloc(Estimator* e, int n, const float* d_sample, const float* d_constants,
                                    const float* d_initial, float* d_params, int* d_iterations, float* d_trace,
                                    int traceLen, float* d_crlb, float* d_chisq, void* stream);
/* d_constants [n, NumConstants]: the per-spot constants of the models that have them
 * (XYIBgConstSigma, XYITiltedBg: sigma x, y; spline mode 2: the background image), else NULL. */
int Estimator_ExpectedValueDevice(Estimator* e, int n, const float* d_params, const float* d_constants, float* d_ev,
                                  void* stream);
int Estimator_DerivativesDevice(Estimator* e, int n, const float* d_params, const float* d_constants, float* d_deriv,
                                float* d_ev, void* stream);
int Estimator_ChiSquareAndCRLBDevice(Estimator* e, int n, const float* d_params, const float* d_constants,
                                     const float* d_sample, float* d_crlb, float* d_chisq, void* stream);

/* Poisson negative log-likelihood per spot, sum_px mu' - x' log mu' with mu' = max(mu,1e-8),
 * x' = max(x,1e-8) (the negated ComputeLogLikelihood of Estimation.h:138-156), and its adjoint.
 * mu and the Jacobian are recomputed from d_params; nothing from a forward pass is stored.
 * d_shared: the shared PSF parameters (NULL = the estimator's calibration) in the layout
 *   fixed sigma {sigma_x, sigma_y}; astigmatic {s0, gamma, d, A, B} for x then y;
 *   spline the coefficient table [Nz,Ny,Nx,64]; free sigma: none.
 * d_gout [n]: upstream gradient of each spot's NLL (NULL = 1).  Outputs (each nullable):
 * d_nll [n], d_grad_params [n,K] (true derivatives), d_grad_shared (accumulated with +=).
 * The spline forms of R <= 28 sort the spots by spline window in a scratch buffer the
 * estimator keeps (grown to the largest n seen); a gradient call with d_gout (a backward pass)
 * reuses the sort of the previous call for as many spots on the same stream (its forward), a
 * call without d_gout sorts (any order gives the same per-spot results; a sort of other
 * parameters only costs speed). */
int Estimator_NLLGradDevice(Estimator* e, int n, const float* d_params, const float* d_sample, const float* d_shared,
                            const float* d_gout, float* d_nll, float* d_grad_params, float* d_grad_shared,
                            void* stream);
/* Mixed batch (no reference counterpart): bins of spots, one estimator (model + ROI size) per
 * bin.  The bins are spread longest-first over nstreams internal HIP streams (0 = 4, one per
 * hardware queue of the process) that run concurrently, so one bin's tail overlaps the next
 * bins' work and different models share the CUs; each bin runs the estimator's own persistent
 * fit + CRLB launches with work counters the batch owns.  With useGraph the fork, every launch
 * and the join are captured into one HIP graph on the first run and replayed with one
 * hipGraphLaunch on `stream`; a graph run with other buffers or changed estimator settings (LM
 * parameters, group size, refill batch, work segments, spline form) than the captured ones is refused with an
 * error (make a new MixedBatch for them; eager runs take any buffers).  Run takes per-bin
 * device pointers (arrays of nbins pointers; d_initial [n,K] per bin, or the whole array NULL
 * for the centre-of-mass start; d_iters/d_crlb/d_chisq may be NULL).  Runs of one batch must
 * not overlap each other (they share the batch's counters). */
typedef struct MixedBatch MixedBatch;
MixedBatch* MixedBatch_Create(int nbins, Estimator* const* estimators, const int* counts, int nstreams);
void MixedBatch_Delete(MixedBatch* mb);
int MixedBatch_NumStreams(MixedBatch* mb);
int MixedBatch_Run(MixedBatch* mb, const float* const* d_samples, const float* const* d_initial,
                   float* const* d_params, int* const* d_iters, float* const* d_crlb, float* const* d_chisq,
                   int useGraph, void* stream);
/* The same bins as one plain launch per bin, in bin order on `stream` (the comparison form). */
int MixedBatch_RunPlain(MixedBatch* mb, const float* const* d_samples, const float* const* d_initial,
                        float* const* d_params, int* const* d_iters, float* const* d_crlb, float* const* d_chisq,
                        void* stream);

/* Number of shared PSF parameters and a host copy of their current values. */
int Estimator_NumShared(Estimator* e);
int Estimator_GetShared(Estimator* e, float* out);

#ifdef __cplusplus
}
#endif

#endif /* PHOTONPY_AMD_H */
