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
    hipStream_t lastLaunch = nullptr;       // stream of the last graph launch
    std::vector<const void*> capturedPtrs;  // buffers the graph was captured with

    void drop_graph() {
        if (exec) (void)hipStreamSynchronize(lastLaunch);  // the last replay has finished
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        capturedPtrs.clear();
    }

    ~MixedBatch() {
        drop_graph();
        for (hipStream_t s : streams) (void)hipStreamSynchronize(s);
        for (hipStream_t s : streams) (void)hipStreamDestroy(s);
        for (hipEvent_t ev : joins) (void)hipEventDestroy(ev);
        if (fork) (void)hipEventDestroy(fork);
        if (capStream) (void)hipStreamDestroy(capStream);
    }
};

// Relative cost of one pixel-iteration per model (grid spline included), for the
// longest-first assignment of bins to branches.
static double bin_cost(const Estimator* e, int count) {
    double c = 1.0;
    switch (e->desc.kind) {
    case kKindFixed: c = 1.0; break;
    case kKindSigma: c = 1.3; break;
    case kKindAstig: c = 1.4; break;
    default: c = is_spline(e->desc.kind) ? (e->desc.useGrid ? 4.0 : 40.0) : 1.3; break;
    }
    return c * (double)count * e->roisize * e->roisize;
}

// Launches every bin: each branch waits for `fork` on its stream, runs its bins in
// order, and records its join event; the caller's stream waits for all joins.
static void launch_bins(MixedBatch* mb, const float* const* d_samples, float* const* d_params, int* const* d_iters,
                        float* const* d_crlb, float* const* d_chisq, hipStream_t origin) {
    hip_check(hipEventRecord(mb->fork, origin), "hipEventRecord(fork)");
    for (size_t b = 0; b < mb->streams.size(); b++) {
        hipStream_t s = mb->streams[b];
        hip_check(hipStreamWaitEvent(s, mb->fork, 0), "hipStreamWaitEvent(fork)");
        for (int bin : mb->order[b]) {
            Estimator* e = mb->est[bin];
            std::lock_guard<std::mutex> l(e->mtx);
            run_fit(e, mb->counts[bin], d_samples[bin], nullptr, nullptr, d_params[bin], d_iters ? d_iters[bin] : nullptr,
                    nullptr, 0, d_crlb ? d_crlb[bin] : nullptr, d_chisq ? d_chisq[bin] : nullptr, s, false);
        }
        hip_check(hipEventRecord(mb->joins[b], s), "hipEventRecord(join)");
        hip_check(hipStreamWaitEvent(origin, mb->joins[b], 0), "hipStreamWaitEvent(join)");
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
                    const int R = e->roisize;
                    if (R < kMinFitR || R > kMaxFitR)
                        throw std::runtime_error("photonpy_amd: fit kernels are compiled for roisize " +
                                                 std::to_string(kMinFitR) + ".." + std::to_string(kMaxFitR) +
                                                 ", got " + std::to_string(R));
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

int MixedBatch_Run(MixedBatch* mb, const float* const* d_samples, float* const* d_params, int* const* d_iters,
                   float* const* d_crlb, float* const* d_chisq, int useGraph, void* stream) {
    if (!mb) return 1;
    return guarded(
        [&]() -> int {
            hipStream_t st = as_stream(stream);
            const size_t nb = mb->est.size();
            for (size_t b = 0; b < nb; b++)
                if (mb->counts[b] > 0 && (!d_samples[b] || !d_params[b]))
                    throw std::runtime_error("photonpy_amd: NULL sample or params buffer");
            for (Estimator* e : mb->est) {
                // device tables and counters exist before any capture (no allocation inside it)
                std::lock_guard<std::mutex> l(e->mtx);
                e->ensure_device();
                e->ensure_counters();
            }
            if (!useGraph) {
                launch_bins(mb, d_samples, d_params, d_iters, d_crlb, d_chisq, st);
                return 0;
            }
            std::vector<const void*> ptrs;
            for (size_t b = 0; b < nb; b++) {
                ptrs.push_back(d_samples[b]);
                ptrs.push_back(d_params[b]);
                ptrs.push_back(d_iters ? d_iters[b] : nullptr);
                ptrs.push_back(d_crlb ? d_crlb[b] : nullptr);
                ptrs.push_back(d_chisq ? d_chisq[b] : nullptr);
            }
            if (!mb->exec || ptrs != mb->capturedPtrs) {
                // the kernels' buffers are baked into the graph: capture again for new ones
                mb->drop_graph();
                if (!mb->capStream) hip_check(hipStreamCreateWithFlags(&mb->capStream, hipStreamNonBlocking),
                                              "hipStreamCreate");
                hip_check(hipStreamBeginCapture(mb->capStream, hipStreamCaptureModeRelaxed), "begin capture");
                try {
                    launch_bins(mb, d_samples, d_params, d_iters, d_crlb, d_chisq, mb->capStream);
                } catch (...) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(mb->capStream, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                hip_check(hipStreamEndCapture(mb->capStream, &mb->graph), "end capture");
                hip_check(hipGraphInstantiate(&mb->exec, mb->graph, nullptr, nullptr, 0), "graph instantiate");
                mb->capturedPtrs = ptrs;
            }
            hip_check(hipGraphLaunch(mb->exec, st), "graph launch");
            mb->lastLaunch = st;
            return 0;
        },
        2);
}

// The same bins as one plain launch per bin, one after another on the caller's stream.
int MixedBatch_RunPlain(MixedBatch* mb, const float* const* d_samples, float* const* d_params, int* const* d_iters,
                        float* const* d_crlb, float* const* d_chisq, void* stream) {
    if (!mb) return 1;
    return guarded(
        [&]() -> int {
            for (size_t b = 0; b < mb->est.size(); b++) {
                Estimator* e = mb->est[b];
                std::lock_guard<std::mutex> l(e->mtx);
                run_fit(e, mb->counts[b], d_samples[b], nullptr, nullptr, d_params[b], d_iters ? d_iters[b] : nullptr, nullptr,
                        0, d_crlb ? d_crlb[b] : nullptr, d_chisq ? d_chisq[b] : nullptr, as_stream(stream), false);
            }
            return 0;
        },
        2);
}

}  // extern "C"
