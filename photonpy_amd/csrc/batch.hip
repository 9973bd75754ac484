// Mixed batch (MixedBatch_* in include/photonpy_amd.h): bins of spots on concurrent
// streams, optionally captured into one HIP graph.
#include "capi_internal.h"

using namespace ppa;

// ---------------------------------------------------------------- mixed batch
struct MixedBatch {
    std::vector<Estimator*> est;
    std::vector<int> counts;
    std::vector<int> branch;              // stream of each bin
    std::vector<std::vector<int>> order;  // bins of each stream, costliest first
    std::vector<hipStream_t> streams;     // one per branch
    std::vector<hipEvent_t> joins;        // one per branch
    hipEvent_t fork = nullptr;
    hipStream_t capStream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t lastLaunch = nullptr;    // stream of the last graph launch
    std::vector<uintptr_t> capturedKey;  // buffers and estimator settings the graph was captured with
    // Work counters of the batch's fit launches, kWorkSegs per bin, owned by the batch: a
    // replayed graph reuses them, and no other launch of the estimators touches them.
    int* counters = nullptr;

    void drop_graph() {
        if (exec) (void)hipStreamSynchronize(lastLaunch);  // the last replay has finished
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        capturedKey.clear();
    }

    ~MixedBatch() {
        drop_graph();
        for (hipStream_t s : streams) (void)hipStreamSynchronize(s);
        for (hipStream_t s : streams) (void)hipStreamDestroy(s);
        for (hipEvent_t ev : joins) (void)hipEventDestroy(ev);
        if (fork) (void)hipEventDestroy(fork);
        if (capStream) (void)hipStreamDestroy(capStream);
        (void)dev_release(counters);
    }
};

// Relative fit time per pixel and spot of each model, for the longest-first assignment
// of bins to branches: the mixed workload's bins measured alone on one MI355X gave
// about 0.023 ns per pixel for fixed sigma, 0.045 free sigma, 0.040 astigmatic and 0.16
// grid-form spline (ROI sizes 9 to 25; the coefficient form is about ten times the grid
// form).
static double bin_cost(const Estimator* e, int count) {
    double c = 1.0;
    switch (e->desc.kind) {
    case kKindFixed: c = 1.0; break;
    case kKindSigma: c = 2.0; break;
    case kKindAstig: c = 1.8; break;
    default: c = is_spline(e->desc.kind) ? (e->desc.useGrid ? 7.0 : 70.0) : 2.0; break;
    }
    return c * (double)count * e->roisize * e->roisize;
}

// Launches every bin: each branch waits for `fork` on its stream, runs its bins in
// order, and records its join event; the caller's stream waits for all joins.
static void launch_bins(MixedBatch* mb, const float* const* d_samples, const float* const* d_initial,
                        float* const* d_params, int* const* d_iters, float* const* d_crlb, float* const* d_chisq,
                        hipStream_t origin) {
    hip_check(hipEventRecord(mb->fork, origin), "hipEventRecord(fork)");
    for (size_t b = 0; b < mb->streams.size(); b++) {
        hipStream_t s = mb->streams[b];
        hip_check(hipStreamWaitEvent(s, mb->fork, 0), "hipStreamWaitEvent(fork)");
        for (int bin : mb->order[b]) {
            Estimator* e = mb->est[bin];
            std::lock_guard<std::mutex> l(e->mtx);
            run_fit(e, mb->counts[bin], d_samples[bin], nullptr, d_initial ? d_initial[bin] : nullptr, d_params[bin],
                    d_iters ? d_iters[bin] : nullptr, nullptr, 0, d_crlb ? d_crlb[bin] : nullptr,
                    d_chisq ? d_chisq[bin] : nullptr, s, false, 0, mb->counters + (size_t)bin * kCounterInts);
        }
        hip_check(hipEventRecord(mb->joins[b], s), "hipEventRecord(join)");
        hip_check(hipStreamWaitEvent(origin, mb->joins[b], 0), "hipStreamWaitEvent(join)");
    }
}

// Device tables of every bin's estimator exist before any launch or capture (no
// allocation inside a capture).
static void ensure_tables(MixedBatch* mb) {
    for (Estimator* e : mb->est) {
        std::lock_guard<std::mutex> l(e->mtx);
        e->ensure_device();
    }
}

extern "C" {

MixedBatch* MixedBatch_Create(int nbins, Estimator* const* ests, const int* counts, int nstreams) {
    return guarded(
        [&]() -> MixedBatch* {
            if (nbins <= 0 || !ests || !counts) throw std::runtime_error("photonpy_amd: empty mixed batch");
            if (nstreams <= 0) nstreams = 4;
            nstreams = std::min(nstreams, nbins);
            auto* mb = new MixedBatch();
            try {
                std::vector<std::pair<double, int>> cost;
                for (int b = 0; b < nbins; b++) {
                    Estimator* e = ests[b];
                    if (!e) throw std::runtime_error("photonpy_amd: NULL estimator in mixed batch");
                    if (e->inner)
                        throw std::runtime_error("photonpy_amd: a mixed batch returns no diagnostics; pass the "
                                                 "wrapped estimator instead of its GLRT wrapper");
                    if (e->NC() > 0)
                        throw std::runtime_error("photonpy_amd: models with per-spot constants are not batched");
                    if (counts[b] < 0) throw std::runtime_error("photonpy_amd: negative spot count");
                    mb->est.push_back(e);
                    mb->counts.push_back(counts[b]);
                    cost.push_back({bin_cost(e, counts[b]), b});
                }
                // longest-processing-time-first onto the least-loaded branch
                std::stable_sort(cost.begin(), cost.end(), [](auto& x, auto& y) { return x.first > y.first; });
                std::vector<double> load(nstreams, 0.0);
                mb->branch.assign(nbins, 0);
                mb->order.assign(nstreams, {});
                for (auto& c : cost) {
                    const int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                    load[s] += c.first;
                    mb->branch[c.second] = s;
                    mb->order[s].push_back(c.second);
                }
                for (int i = 0; i < nstreams; i++) {
                    hipStream_t st;
                    hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
                    mb->streams.push_back(st);
                    hipEvent_t ev;
                    hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
                    mb->joins.push_back(ev);
                }
                hip_check(hipEventCreateWithFlags(&mb->fork, hipEventDisableTiming), "hipEventCreate");
                hip_check(dev_alloc(&mb->counters, sizeof(int) * kCounterInts * (size_t)nbins), "hipMalloc(batch counters)");
            } catch (...) {
                delete mb;
                throw;
            }
            return mb;
        },
        nullptr);
}

void MixedBatch_Delete(MixedBatch* mb) { delete mb; }

int MixedBatch_NumStreams(MixedBatch* mb) { return mb ? (int)mb->streams.size() : -1; }

int MixedBatch_Run(MixedBatch* mb, const float* const* d_samples, const float* const* d_initial,
                   float* const* d_params, int* const* d_iters, float* const* d_crlb, float* const* d_chisq, int useGraph,
                   void* stream) {
    if (!mb) return 1;
    return guarded(
        [&]() -> int {
            hipStream_t st = as_stream(stream);
            const size_t nb = mb->est.size();
            for (size_t b = 0; b < nb; b++)
                if (mb->counts[b] > 0 && (!d_samples[b] || !d_params[b]))
                    throw std::runtime_error("photonpy_amd: NULL sample or params buffer");
            ensure_tables(mb);
            if (!useGraph) {
                launch_bins(mb, d_samples, d_initial, d_params, d_iters, d_crlb, d_chisq, st);
                return 0;
            }
            // the kernels' buffers and arguments (LM settings, lane groups, refill batch, work segments, spline
            // form) are baked into the graph: capture again when any of them changed
            std::vector<uintptr_t> key;
            auto ptr = [](const void* p) { return reinterpret_cast<uintptr_t>(p); };
            for (size_t b = 0; b < nb; b++) {
                const Estimator* e = mb->est[b];
                float lam = e->lm.lambda;
                uint32_t lamBits;
                std::memcpy(&lamBits, &lam, 4);
                for (uintptr_t v : {ptr(d_samples[b]), ptr(d_initial ? d_initial[b] : nullptr), ptr(d_params[b]),
                                    ptr(d_iters ? d_iters[b] : nullptr), ptr(d_crlb ? d_crlb[b] : nullptr),
                                    ptr(d_chisq ? d_chisq[b] : nullptr), (uintptr_t)e->G(), (uintptr_t)e->refill, (uintptr_t)e->segShift,
                                    (uintptr_t)e->desc.useGrid, (uintptr_t)e->lm.iterations, (uintptr_t)lamBits})
                    key.push_back(v);
            }
            // A second capture on one batch is refused: one run that captured again (a graph
            // with initial values replayed, then a capture without them, 4,000,000 ROIs) ended
            // in a host segmentation fault inside this call whose cause reading did not find.
            if (mb->exec && key != mb->capturedKey)
                throw std::runtime_error("photonpy_amd: this mixed batch's graph was captured with other buffers or "
                                         "estimator settings; create a new MixedBatch for them");
            if (!mb->exec) {
                if (!mb->capStream) hip_check(hipStreamCreateWithFlags(&mb->capStream, hipStreamNonBlocking),
                                              "hipStreamCreate");
                hip_check(hipStreamBeginCapture(mb->capStream, hipStreamCaptureModeRelaxed), "begin capture");
                try {
                    launch_bins(mb, d_samples, d_initial, d_params, d_iters, d_crlb, d_chisq, mb->capStream);
                } catch (...) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(mb->capStream, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                hip_check(hipStreamEndCapture(mb->capStream, &mb->graph), "end capture");
                hip_check(hipGraphInstantiate(&mb->exec, mb->graph, nullptr, nullptr, 0), "graph instantiate");
                mb->capturedKey = key;
            }
            hip_check(hipGraphLaunch(mb->exec, st), "graph launch");
            mb->lastLaunch = st;
            return 0;
        },
        2);
}

// The same bins as one plain launch per bin, one after another on the caller's stream.
int MixedBatch_RunPlain(MixedBatch* mb, const float* const* d_samples, const float* const* d_initial,
                        float* const* d_params, int* const* d_iters, float* const* d_crlb, float* const* d_chisq,
                        void* stream) {
    if (!mb) return 1;
    return guarded(
        [&]() -> int {
            ensure_tables(mb);
            for (size_t b = 0; b < mb->est.size(); b++) {
                Estimator* e = mb->est[b];
                std::lock_guard<std::mutex> l(e->mtx);
                run_fit(e, mb->counts[b], d_samples[b], nullptr, d_initial ? d_initial[b] : nullptr, d_params[b],
                        d_iters ? d_iters[b] : nullptr, nullptr, 0, d_crlb ? d_crlb[b] : nullptr,
                        d_chisq ? d_chisq[b] : nullptr, as_stream(stream), false, 0, mb->counters + (size_t)b * kCounterInts);
            }
            return 0;
        },
        2);
}

}  // extern "C"
