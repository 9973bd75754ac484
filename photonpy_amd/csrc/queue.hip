// EstimationQueue (EstimationQueue.h:27-108, EstimationQueue.cpp:11-412): spots are
// scheduled from host memory into pinned batches; a worker thread launches full
// (or flushed) batches round-robin on a few HIP streams -- copy in, fused fit + CRLB
// + chi^2, copy out -- and moves completed batches to a result list that
// GetResults drains.  Same entry points, argument meaning and queue semantics as
// the reference:
//   - Schedule blocks while GetQueueLength() >= maxQueueLen, where the queue length
//     is the spots in scheduled-but-unlaunched batches plus the number of batches in
//     flight (EstimationQueue.cpp:219-226, 160-171);
//   - a call with initial values switches the queue to initial values for every
//     later launch (:223);
//   - GetResults copies whole finished batches, newest first, while they fit in the
//     space left, and returns the number of spots copied (:117-158);
//   - IsIdle ignores the batch still being filled; Flush queues it (:175-205).
// Per-spot constants (models that have them) travel with the samples.
// Differences: batches are fitted by this library's fused kernels (CRLB from the
// fit's own second kernel), and maxQueueLen < 1 is treated as 1 (the reference never
// returns).  Diagnostics are those of the estimator: none for the models (as the
// reference's GPU path), {ll_on, ll_off, mean} per spot for a GLRT wrapper, whose
// mean comes from each batch's first spot as when the reference's queue calls it.
// Waits use condition variables instead of the reference's 1 ms / 5 ms sleeps:
// Schedule wakes when the worker starts or finishes a batch, and the worker
// sleeps on new work when nothing is in flight and polls its streams every 50 us
// when something is.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>

#include "capi_internal.h"

namespace {

template <class T>
T* pinned(size_t n) {
    void* p = nullptr;
    if (n == 0) return nullptr;
    hip_check(pinned_alloc(&p, n * sizeof(T)), "hipHostMalloc(queue batch)");
    return static_cast<T*>(p);
}

template <class T>
T* device_alloc(size_t n) {
    void* p = nullptr;
    if (n == 0) return nullptr;
    hip_check(dev_alloc(&p, n * sizeof(T)), "hipMalloc(queue stream buffers)");
    return static_cast<T*>(p);
}

// Copies large host buffers with a few helper threads (the calling thread helps):
// Schedule's copy of the caller's ROIs into pinned batches is bound by one core's
// memcpy rate otherwise.  One pool per process (copy_pool()); copies from several
// queues take turns; small copies stay on the caller.
class CopyPool {
public:
    explicit CopyPool(int helpers) {
        for (int i = 0; i < helpers; i++) threads_.emplace_back([this] { work(); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> l(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : threads_) t.join();
    }
    void copy(void* dst, const void* src, size_t bytes) {
        if (threads_.empty() || bytes < kMinParallel) {
            std::memcpy(dst, src, bytes);
            return;
        }
        std::lock_guard<std::mutex> one(turn_);
        {
            std::lock_guard<std::mutex> l(m_);
            dst_ = static_cast<char*>(dst);
            src_ = static_cast<const char*>(src);
            bytes_ = bytes;
            next_ = 0;
            busy_ = (int)threads_.size();
            gen_++;
        }
        cv_.notify_all();
        run_chunks();
        std::unique_lock<std::mutex> l(m_);
        done_.wait(l, [&] { return busy_ == 0; });
    }

private:
    static constexpr size_t kMinParallel = 1 << 20, kChunk = 256 << 10;
    void run_chunks() {
        for (;;) {
            const size_t o = next_.fetch_add(kChunk);
            if (o >= bytes_) return;
            std::memcpy(dst_ + o, src_ + o, std::min(kChunk, bytes_ - o));
        }
    }
    void work() {
        unsigned long long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> l(m_);
                cv_.wait(l, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            run_chunks();
            std::lock_guard<std::mutex> l(m_);
            if (--busy_ == 0) done_.notify_all();
        }
    }
    std::vector<std::thread> threads_;
    std::mutex turn_;  // one parallel copy at a time
    std::mutex m_;
    std::condition_variable cv_, done_;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t bytes_ = 0;
    std::atomic<size_t> next_{0};
    int busy_ = 0;
    unsigned long long gen_ = 0;
    bool stop_ = false;
};

CopyPool& copy_pool() {
    static CopyPool pool(std::max(0, std::min(7, (int)std::thread::hardware_concurrency() / 2 - 1)));
    return pool;
}

}  // namespace

void parallel_copy(void* dst, const void* src, size_t bytes) { copy_pool().copy(dst, src, bytes); }

struct EstimationQueue {
    struct Batch {
        std::vector<int> ids;
        float* samples = nullptr;  // pinned [batch, S]
        float* params = nullptr;   // pinned [batch, K]: initial values in, estimates out
        float* crlb = nullptr;     // pinned [batch, K]
        float* chisq = nullptr;    // pinned [batch]
        int* iterations = nullptr; // pinned [batch]
        float* constants = nullptr;  // pinned [batch, NC]
        float* diag = nullptr;       // pinned [batch, D]
        std::vector<int> roipos;   // [batch, 2], returned with the results
        int numspots = 0;

        Batch(int n, int S, int K, int NC, int D) : ids(n), roipos((size_t)n * 2) {
            constants = pinned<float>((size_t)n * NC);
            diag = pinned<float>((size_t)n * D);
            // written by the GLRT pass only; a cuda = false estimator's column stays 0
            if (diag) std::memset(diag, 0, sizeof(float) * (size_t)n * D);
            samples = pinned<float>((size_t)n * S);
            params = pinned<float>((size_t)n * K);
            crlb = pinned<float>((size_t)n * K);
            chisq = pinned<float>(n);
            iterations = pinned<int>(n);
        }
        ~Batch() {
            for (void* p : {(void*)samples, (void*)params, (void*)crlb, (void*)chisq, (void*)iterations,
                            (void*)constants, (void*)diag})
                if (p) (void)pinned_release(p);
        }
    };

    struct Lane {
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
        std::unique_ptr<Batch> current;
        float *samples = nullptr, *params = nullptr, *crlb = nullptr, *chisq = nullptr, *constants = nullptr;
        float* diag = nullptr;
        int* iterations = nullptr;
    };

    Estimator* outer;  // the estimator the queue was created with (a GLRT wrapper or a model)
    Estimator* est;    // the model that fits
    Context* ctx;
    int device = 0;
    int batchSize, maxQueueLen, K, S, NC, D;
    // Lanes per spot for the queue's batches: a batch of a few thousand spots fills a
    // small part of the GPU, so each spot gets up to 16 lanes while batch x lanes stays
    // within 64K (the fit falls back to the estimator's own size where none is compiled).
    int groupHint = 0;
    bool keepSamples;
    std::atomic<bool> useInitial{false};
    std::vector<Lane> lanes;

    std::deque<std::unique_ptr<Batch>> todo, finished, recycle;
    int numActive = 0;
    std::mutex todoMtx, finishedMtx, recycleMtx, scheduleMtx;
    std::condition_variable todoCv;  // todo or numActive changed, or stop (todoMtx)
    std::unique_ptr<Batch> next;

    std::thread worker;
    std::atomic<bool> stop{false};
    std::string workerError;  // first launch failure, reported by later calls

    EstimationQueue(Estimator* e, int batch, int maxLen, bool keep, int streams, Context* c)
        : outer(e), est(model_of(e)), ctx(c), batchSize(batch), maxQueueLen(std::max(1, maxLen)), K(e->K()),
          S(e->S()), NC(e->NC()), D(e->numDiag), keepSamples(keep) {
        if (streams < 0) streams = 3;  // EstimationQueue.cpp:16-17
        if (streams == 0) streams = 1;
        hip_check(hipGetDevice(&device), "hipGetDevice");
        {
            std::lock_guard<std::mutex> l(est->mtx);
            est->ensure_device();
            est->ensure_counters();
        }
        lanes.resize(streams);
        for (Lane& ln : lanes) {
            hip_check(hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipEventCreateWithFlags(&ln.done, hipEventDisableTiming), "hipEventCreate");
            ln.samples = device_alloc<float>((size_t)batchSize * S);
            ln.params = device_alloc<float>((size_t)batchSize * K);
            ln.crlb = device_alloc<float>((size_t)batchSize * K);
            ln.chisq = device_alloc<float>(batchSize);
            ln.iterations = device_alloc<int>(batchSize);
            ln.constants = device_alloc<float>((size_t)batchSize * NC);
            ln.diag = device_alloc<float>((size_t)batchSize * D);
        }
        next = std::make_unique<Batch>(batchSize, S, K, NC, D);
        for (groupHint = 16; groupHint > 1 && (long)batchSize * groupHint > 65536; groupHint /= 2) {
        }
        worker = std::thread([this]() { run(); });
    }

    ~EstimationQueue() {
        {
            std::lock_guard<std::mutex> l(todoMtx);
            stop = true;
        }
        todoCv.notify_all();
        if (worker.joinable()) worker.join();
        for (Lane& ln : lanes) {
            if (ln.stream) (void)hipStreamSynchronize(ln.stream);
            for (void* p : {(void*)ln.samples, (void*)ln.params, (void*)ln.crlb, (void*)ln.chisq, (void*)ln.iterations,
                            (void*)ln.constants, (void*)ln.diag})
                if (p) (void)dev_release(p);
            if (ln.done) (void)hipEventDestroy(ln.done);
            if (ln.stream) (void)hipStreamDestroy(ln.stream);
        }
    }

    void launch(Lane& ln, std::unique_ptr<Batch> b) {
        const int n = b->numspots;
        hip_check(hipMemcpyAsync(ln.samples, b->samples, sizeof(float) * n * S, hipMemcpyHostToDevice, ln.stream),
                  "queue: samples to device");
        if (NC > 0)
            hip_check(hipMemcpyAsync(ln.constants, b->constants, sizeof(float) * n * NC, hipMemcpyHostToDevice,
                                     ln.stream),
                      "queue: constants to device");
        const bool init = useInitial;
        if (init)
            hip_check(hipMemcpyAsync(ln.params, b->params, sizeof(float) * n * K, hipMemcpyHostToDevice, ln.stream),
                      "queue: initial values to device");
        {
            std::lock_guard<std::mutex> l(est->mtx);
            run_fit(est, n, ln.samples, NC > 0 ? ln.constants : nullptr, init ? ln.params : nullptr, ln.params,
                    ln.iterations, nullptr, 0, ln.crlb, ln.chisq, ln.stream, false, groupHint);
            if (D > 0 && outer->inner)
                hip_check(launch_glrt_any(est->desc, est->roisize, n, ln.params, NC > 0 ? ln.constants : nullptr,
                                          ln.samples, outer->glrtMeanOfSpot0, ln.diag, ln.stream),
                          "queue: glrt kernel");
        }
        if (D > 0 && outer->inner)
            hip_check(hipMemcpyAsync(b->diag, ln.diag, sizeof(float) * n * D, hipMemcpyDeviceToHost, ln.stream),
                      "queue: diagnostics to host");
        hip_check(hipMemcpyAsync(b->params, ln.params, sizeof(float) * n * K, hipMemcpyDeviceToHost, ln.stream),
                  "queue: estimates to host");
        hip_check(hipMemcpyAsync(b->crlb, ln.crlb, sizeof(float) * n * K, hipMemcpyDeviceToHost, ln.stream),
                  "queue: CRLB to host");
        hip_check(hipMemcpyAsync(b->chisq, ln.chisq, sizeof(float) * n, hipMemcpyDeviceToHost, ln.stream),
                  "queue: chi^2 to host");
        hip_check(hipMemcpyAsync(b->iterations, ln.iterations, sizeof(int) * n, hipMemcpyDeviceToHost, ln.stream),
                  "queue: iterations to host");
        hip_check(hipEventRecord(ln.done, ln.stream), "hipEventRecord");
        ln.current = std::move(b);
    }

    // Worker thread (EstimationQueue.cpp:54-99): start queued batches on idle
    // streams, collect completed ones; wait for work when nothing is in flight.
    void run() {
        (void)hipSetDevice(device);
        while (!stop) {
            bool changed = false, inFlight = false;
            for (Lane& ln : lanes) {
                if (!ln.current) {
                    std::unique_ptr<Batch> b;
                    {
                        std::lock_guard<std::mutex> l(todoMtx);
                        if (!todo.empty()) {
                            b = std::move(todo.front());
                            todo.pop_front();
                            numActive++;
                        }
                    }
                    if (b) {
                        try {
                            launch(ln, std::move(b));
                        } catch (const std::exception& ex) {
                            std::lock_guard<std::mutex> l(todoMtx);
                            if (workerError.empty()) workerError = ex.what();
                            numActive--;
                        }
                        changed = true;
                    }
                }
                if (ln.current && hipEventQuery(ln.done) == hipSuccess) {
                    {
                        std::lock_guard<std::mutex> l(finishedMtx);
                        finished.push_back(std::move(ln.current));
                    }
                    {
                        std::lock_guard<std::mutex> l(todoMtx);
                        numActive--;
                    }
                    changed = true;
                }
                inFlight |= (bool)ln.current;
            }
            if (changed) {
                todoCv.notify_all();
                continue;
            }
            if (inFlight) {
                std::this_thread::sleep_for(std::chrono::microseconds(50));
            } else {
                std::unique_lock<std::mutex> l(todoMtx);
                todoCv.wait_for(l, std::chrono::milliseconds(20), [&] { return stop || !todo.empty(); });
            }
        }
    }

    void check_worker() {
        std::lock_guard<std::mutex> l(todoMtx);
        if (!workerError.empty()) throw std::runtime_error(workerError);
    }

    int queue_length() {
        std::lock_guard<std::mutex> l(todoMtx);
        int sum = 0;
        for (auto& b : todo) sum += b->numspots;
        return sum + numActive;
    }

    bool is_idle() {
        std::lock_guard<std::mutex> l(todoMtx);
        return todo.empty() && numActive == 0;
    }

    // moves the batch being filled to the queue (scheduleMtx held by the caller)
    void flush_locked() {
        if (next->numspots == 0) return;
        {
            std::lock_guard<std::mutex> l(todoMtx);
            todo.push_back(std::move(next));
        }
        todoCv.notify_all();
        {
            std::lock_guard<std::mutex> l(recycleMtx);
            if (!recycle.empty()) {
                next = std::move(recycle.front());
                recycle.pop_front();
            }
        }
        if (!next) next = std::make_unique<Batch>(batchSize, S, K, NC, D);
    }

    void schedule(int count, const int* ids, const float* samples, const float* initial, const float* constants,
                  const int* roipos) {
        check_worker();
        if (NC > 0 && !constants)
            throw std::runtime_error("photonpy_amd: this estimator needs " + std::to_string(NC) + " constants per spot");
        {
            std::unique_lock<std::mutex> l(todoMtx);
            todoCv.wait(l, [&] {
                int len = numActive;
                for (auto& b : todo) len += b->numspots;
                return len < maxQueueLen || !workerError.empty();
            });
        }
        useInitial = initial != nullptr;
        std::lock_guard<std::mutex> l(scheduleMtx);
        // runs of spots that fit in the batch being filled, copied in bulk
        for (int i = 0; i < count;) {
            Batch& b = *next;
            const int j = b.numspots;
            const int m = std::min(count - i, batchSize - j);
            for (int u = 0; u < m; u++) {
                b.ids[j + u] = ids ? ids[i + u] : 0;
                b.roipos[2 * (j + u)] = roipos ? roipos[2 * (i + u)] : 0;
                b.roipos[2 * (j + u) + 1] = roipos ? roipos[2 * (i + u) + 1] : 0;
            }
            copy_pool().copy(b.samples + (size_t)j * S, samples + (size_t)i * S, sizeof(float) * S * m);
            if (initial) std::memcpy(b.params + (size_t)j * K, initial + (size_t)i * K, sizeof(float) * K * m);
            if (NC > 0) copy_pool().copy(b.constants + (size_t)j * NC, constants + (size_t)i * NC, sizeof(float) * NC * m);
            b.numspots = j + m;
            i += m;
            if (b.numspots == batchSize) flush_locked();
        }
    }

    int result_count() {
        std::lock_guard<std::mutex> l(finishedMtx);
        int total = 0;
        for (auto& b : finished) total += b->numspots;
        return total;
    }

    int results(int count, float* estim, float* diag, float* crlb, int* roipos, float* samples,
                EstimatorResult* res) {
        int copied = 0;
        while (copied < count) {
            const int space = count - copied;
            std::unique_ptr<Batch> b;
            {
                std::lock_guard<std::mutex> l(finishedMtx);
                if (finished.empty() || finished.back()->numspots > space) return copied;
                b = std::move(finished.back());
                finished.pop_back();
            }
            const int n = b->numspots;
            for (int i = 0; i < n; i++) {
                if (res) {
                    res[copied + i].id = b->ids[i];
                    res[copied + i].chisq = b->chisq[i];
                    res[copied + i].iterations = b->iterations[i];
                }
            }
            if (estim) std::memcpy(estim + (size_t)copied * K, b->params, sizeof(float) * n * K);
            if (crlb) std::memcpy(crlb + (size_t)copied * K, b->crlb, sizeof(float) * n * K);
            if (diag && D > 0) std::memcpy(diag + (size_t)copied * D, b->diag, sizeof(float) * n * D);
            if (roipos) std::memcpy(roipos + (size_t)copied * 2, b->roipos.data(), sizeof(int) * n * 2);
            if (samples && keepSamples) std::memcpy(samples + (size_t)copied * S, b->samples, sizeof(float) * n * S);
            copied += n;
            b->numspots = 0;
            std::lock_guard<std::mutex> l(recycleMtx);
            recycle.push_back(std::move(b));
        }
        return copied;
    }
};

void queue_destroy(EstimationQueue* q) { delete q; }

extern "C" {

EstimationQueue* EstimQueue_Create(Estimator* psf, int batchSize, int maxQueueLen, bool keepSamples, int numStreams,
                                   Context* ctx) {
    return guarded(
        [&]() -> EstimationQueue* {
            if (!psf) throw std::runtime_error("photonpy_amd: EstimationQueue needs an estimator");
            if (batchSize < 1) throw std::runtime_error("photonpy_amd: batchSize must be positive");
            auto* q = new EstimationQueue(psf, batchSize, maxQueueLen, keepSamples, numStreams, ctx);
            if (ctx) {
                std::lock_guard<std::mutex> l(ctx->mtx);
                ctx->queues.insert(q);
            }
            return q;
        },
        nullptr);
}

void EstimQueue_Delete(EstimationQueue* q) {
    if (!q) return;
    if (q->ctx) {
        std::lock_guard<std::mutex> l(q->ctx->mtx);
        q->ctx->queues.erase(q);
    }
    delete q;
}

void EstimQueue_Schedule(EstimationQueue* q, int numspots, const int* ids, const float* h_samples,
                         const float* h_initial, const float* h_constants, const int* h_roipos) {
    if (!q || numspots <= 0) return;
    guarded(
        [&]() -> int {
            if (!h_samples) throw std::runtime_error("photonpy_amd: NULL samples");
            q->schedule(numspots, ids, h_samples, h_initial, h_constants, h_roipos);
            return 0;
        },
        0);
}

void EstimQueue_Flush(EstimationQueue* q) {
    if (!q) return;
    std::lock_guard<std::mutex> l(q->scheduleMtx);
    q->flush_locked();
}

bool EstimQueue_IsIdle(EstimationQueue* q) { return q ? q->is_idle() : true; }

int EstimQueue_GetResultCount(EstimationQueue* q) { return q ? q->result_count() : 0; }

int EstimQueue_GetQueueLength(EstimationQueue* q) { return q ? q->queue_length() : 0; }

int EstimQueue_GroupSize(EstimationQueue* q) {
    if (!q) return 0;
    return guarded([&]() -> int { return fit_group(q->est, q->groupHint); }, 0);
}

int EstimQueue_GetResults(EstimationQueue* q, int maxresults, float* estim, float* diag, float* crlb, int* roipos,
                          float* samples, EstimatorResult* result) {
    if (!q) return 0;
    return guarded([&]() -> int { return q->results(maxresults, estim, diag, crlb, roipos, samples, result); }, 0);
}

}  // extern "C"
