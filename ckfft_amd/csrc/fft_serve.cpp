// fft_serve.cpp -- continuous-batching serving loop over the persistent batch kernel.
//
// The reference has no serving path: a caller runs one transform per synchronous call
// (src/ckfft/ckfft.cpp:36-114).  Here client threads submit transforms (device buffers)
// at any time; one dispatcher thread
//  * admits queued requests first-come-first-served while the bytes of admitted,
//    unfinished requests stay within the HBM budget (back-pressure, not failure);
//  * whenever one of `max_in_flight` slots is free, launches every admitted request
//    as one batch -- one persistent-kernel launch on the slot's own stream, descriptors
//    uploaded asynchronously from the slot's pinned buffers -- so batches formed while
//    others run overlap on the device;
//  * polls the slots' completion events, finishes their requests and frees budget.
// Requests above one workgroup's LDS (complex n >= 2^16, real n >= 2^17) run as ordinary
// multi-pass launches in the batch's stream, one after another, with their intermediate in
// the slot's resident buffer (up to 64 MiB; a larger one is allocated on the stream and
// charged to the budget with the request), so the serving path allocates nothing.
// Twiddles, descriptor buffers, intermediates and streams stay resident for the server's
// life.  A request's result does not depend on the batch it ran in: every kernel that
// computes a transform rounds the same way.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_set>
#include <vector>

#include "ckfft/ckfft_gpu.h"
#include "fft_batch.h"

using namespace ckfft_gpu;

namespace {

using Clock = std::chrono::steady_clock;

uint64_t now_ns()
{
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now().time_since_epoch()).count();
}

enum ReqState : int { kQueued = 0, kAdmitted = 1, kLaunched = 2, kDone = 3, kFailed = 4 };

// One entry of the request ring (all fields guarded by the server mutex).
struct Request {
    CkFftBatchDesc d{};
    long long bytes = 0;     // algorithmic bytes, charged against the budget
    uint64_t t_arrive = 0;   // arrival (submission or scheduled arrival) time
    uint64_t t_submit = 0;   // when the submission reached the server
    uint64_t t_launch = 0;   // launch of its batch
    long long ticket = -1;   // the submission this entry holds
    int state = kDone;
};

// Requests live in a ring indexed by ticket; an entry is reused once its request has
// finished (a submitter waits while the entry it needs is still pending).  Latencies are
// kept for the most recent completions only.
constexpr long long kRing = 1 << 18;
constexpr size_t kLatencyWindow = 1 << 20;

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    char* h_desc = nullptr;              // pinned staging of the batch's queue and transform table
    char* d_desc = nullptr;              // their device copy
    int* d_ctl = nullptr;
    // intermediate of the multi-pass transforms of the slot's batches: they run one after
    // another on the slot's stream, so one buffer serves them all and none allocates
    float2* d_scratch = nullptr;
    size_t scratch_bytes = 0;
    std::vector<long long> reqs;
    long long bytes = 0;
    bool busy = false;
};

constexpr int kMaxBatch = 2048;          // transforms per batch (slot buffer capacity)
// A slot's resident intermediate covers the multi-pass sizes up to this many bytes (8n for a
// complex transform, 4n for a real one: complex 2^23 or real 2^24); larger requests allocate
// theirs on the slot's stream and charge it to the budget.
constexpr size_t kSlotScratchCap = (size_t)64 << 20;
constexpr uint64_t kSpinNs = 2000000;    // idle dispatcher spins this long before sleeping
// queue entries first, the transform table after them at a 256-byte boundary, so one
// copy uploads both
constexpr size_t kDescBytes = sizeof(BatchItem) * kMaxBatch + 256 + sizeof(BatchTransform) * kMaxBatch;

}  // namespace

struct _CkFftServer {
    CkFftContext* ctx = nullptr;
    long long budget = 0;                // bytes available to requests
    long long resident = 0;              // bytes the server itself holds on the device
    int grid = 0;
    std::vector<Slot> slots;
    std::vector<Request> ring = std::vector<Request>(kRing);
    long long next_ticket = 0;
    std::mutex m;
    std::condition_variable cv;
    std::deque<long long> queued;        // submitted, not admitted
    std::vector<long long> admitted;     // admitted, not launched
    long long inflight_bytes = 0;
    long long peak_inflight = 0;
    long long outstanding = 0;
    std::atomic<bool> stop{false};
    std::thread worker;
    // statistics (guarded by m): latencies of the last kLatencyWindow completions
    std::vector<double> lat_us;
    size_t lat_next = 0;
    long long completed = 0, failed = 0, batches = 0, batched = 0, done_bytes = 0;
    uint64_t t_first = 0, t_last = 0;
    // the slowest request since the reset: its latency, scheduled arrival to submission
    // (the client's delay), submission to launch, launch to completion seen
    double max_us = 0, max_client_us = 0, max_wait_us = 0, max_device_us = 0;
    size_t slot_scratch = 0;             // bytes of each slot's resident intermediate
    // tickets of failed requests, kept past the reuse of their ring entries so that a query
    // of an old ticket still reports its failure (failures are rare: a launch error)
    std::unordered_set<long long> failed_tickets;
};

namespace {

thread_local char g_serve_error[256] = "no error";

bool serr(hipError_t e, const char* where)
{
    if (e == hipSuccess) return true;
    std::snprintf(g_serve_error, sizeof(g_serve_error), "%s: %s (%d)", where, hipGetErrorString(e), (int)e);
    return false;
}

void finish_slot(CkFftServer* S, Slot& sl, bool ok)
{
    const uint64_t t = now_ns();
    {
        std::lock_guard<std::mutex> lock(S->m);
        for (long long id : sl.reqs) {
            Request& r = S->ring[id % kRing];
            if (ok) {
                const double us = (double)(t - r.t_arrive) * 1e-3;
                if (us > S->max_us) {
                    S->max_us = us;
                    S->max_client_us = r.t_submit > r.t_arrive ? (double)(r.t_submit - r.t_arrive) * 1e-3 : 0.0;
                    S->max_wait_us = (double)(r.t_launch - std::max(r.t_arrive, r.t_submit)) * 1e-3;
                    S->max_device_us = (double)(t - r.t_launch) * 1e-3;
                }
                if (S->lat_us.size() < kLatencyWindow)
                    S->lat_us.push_back(us);
                else
                    S->lat_us[S->lat_next] = us;
                S->lat_next = (S->lat_next + 1) % kLatencyWindow;
                ++S->completed;
                S->done_bytes += r.bytes;
            } else {
                ++S->failed;
                S->failed_tickets.insert(id);
            }
            r.state = ok ? kDone : kFailed;
        }
        S->inflight_bytes -= sl.bytes;
        S->outstanding -= (long long)sl.reqs.size();
        S->t_last = t;
    }
    sl.reqs.clear();
    sl.bytes = 0;
    sl.busy = false;
    S->cv.notify_all();
}

// `descs` and `bytes` were copied from the requests `batch` under the lock (the request
// deque may grow concurrently, so it is only indexed under the lock).
bool launch_batch(CkFftServer* S, Slot& sl, std::vector<long long>& batch, const std::vector<CkFftBatchDesc>& descs,
                  long long bytes)
{
    BatchBuild b;
    // descriptors were validated at submission
    // requests arrive a few at a time: a complex 2^16 one is faster as its own multi-pass
    // launch than as one workgroup's item
    if (!batch_build(S->ctx, descs.data(), (int)descs.size(), -1, false, b, TwoPass::kExtra, false)) return false;
    const DeviceTables tb = context_tables(S->ctx);
    bool ok = true;
    if (!b.items.empty() && b.items.size() <= (size_t)kInlineQueue && b.trs.size() <= (size_t)kInlineQueue) {
        // small batch: the queue travels in the launch arguments
        ok = serr(launch_persistent_inline(b, sl.d_ctl, std::min(S->grid, (int)b.items.size()), tb, sl.stream),
                  "persistent launch");
    } else if (!b.items.empty()) {
        const size_t items_bytes = sizeof(BatchItem) * b.items.size();
        const size_t trs_off = (items_bytes + 255) & ~(size_t)255;
        const size_t total = trs_off + sizeof(BatchTransform) * b.trs.size();
        std::memcpy(sl.h_desc, b.items.data(), items_bytes);
        std::memcpy(sl.h_desc + trs_off, b.trs.data(), sizeof(BatchTransform) * b.trs.size());
        ok = serr(hipMemcpyAsync(sl.d_desc, sl.h_desc, total, hipMemcpyHostToDevice, sl.stream), "upload queue") &&
             serr(launch_persistent_raw((const BatchItem*)sl.d_desc, (int)b.items.size(),
                                        (const BatchTransform*)(sl.d_desc + trs_off), sl.d_ctl,
                                        std::min(S->grid, (int)b.items.size()), tb, sl.stream, false),
                  "persistent launch");
    }
    for (const auto& e : b.extras) {
        const bool fits = transform_scratch_bytes(e.kind, e.log2n, 1, tb) <= sl.scratch_bytes;
        if (ok)
            ok = serr(launch_transform(e.kind, e.log2n, 1, e.in, e.out, tb, sl.stream, fits ? sl.d_scratch : nullptr,
                                       fits ? sl.scratch_bytes : 0),
                      "extra launch");
    }
    if (ok) ok = serr(hipEventRecord(sl.done, sl.stream), "record");
    sl.reqs.swap(batch);
    sl.bytes = bytes;
    sl.busy = true;
    {
        std::lock_guard<std::mutex> lock(S->m);
        const uint64_t t = now_ns();
        for (long long id : sl.reqs) {
            S->ring[id % kRing].state = kLaunched;
            S->ring[id % kRing].t_launch = t;
        }
        ++S->batches;
        S->batched += (long long)sl.reqs.size();
    }
    if (!ok) {
        // work enqueued before the failing step may still be running on the slot's stream:
        // wait for it before the requests are reported failed and their budget returned,
        // so that no client reuses a buffer the device still writes
        (void)hipStreamSynchronize(sl.stream);
        (void)hipGetLastError();
        finish_slot(S, sl, false);
    }
    return ok;
}

void dispatcher(CkFftServer* S)
{
    (void)hipSetDevice(S->ctx->device);
    std::vector<long long> batch;
    std::vector<CkFftBatchDesc> descs;
    long long bytes = 0;
    batch.reserve(kMaxBatch);
    descs.reserve(kMaxBatch);
    uint64_t idle_since = 0;
    for (;;) {
        bool progress = false;
        for (Slot& sl : S->slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipSuccess) {
                finish_slot(S, sl, true);
                progress = true;
            } else if (q != hipErrorNotReady) {
                serr(q, "batch completion");
                finish_slot(S, sl, false);
                progress = true;
            }
        }
        Slot* free_slot = nullptr;
        for (Slot& sl : S->slots)
            if (!sl.busy) {
                free_slot = &sl;
                break;
            }
        {
            std::unique_lock<std::mutex> lock(S->m);
            while (!S->queued.empty()) {
                Request& r = S->ring[S->queued.front() % kRing];
                if (S->inflight_bytes + r.bytes > S->budget) break;      // back-pressure
                S->inflight_bytes += r.bytes;
                S->peak_inflight = std::max(S->peak_inflight, S->inflight_bytes);
                r.state = kAdmitted;
                S->admitted.push_back(S->queued.front());
                S->queued.pop_front();
            }
            if (free_slot && !S->admitted.empty()) {
                // take admitted requests in order, up to the slot's queue capacity
                const size_t take = std::min(S->admitted.size(), (size_t)kMaxBatch);
                batch.assign(S->admitted.begin(), S->admitted.begin() + take);
                S->admitted.erase(S->admitted.begin(), S->admitted.begin() + take);
                descs.clear();
                bytes = 0;
                for (long long id : batch) {
                    descs.push_back(S->ring[id % kRing].d);
                    bytes += S->ring[id % kRing].bytes;
                }
            }
            if (batch.empty() && !progress && S->outstanding == 0) {
                if (S->stop.load()) return;
                // nothing submitted: spin briefly (a request arriving now is batched at
                // once), then sleep until a submission or shutdown
                const uint64_t t = now_ns();
                if (!idle_since) idle_since = t;
                if (t - idle_since > kSpinNs) {
                    S->cv.wait(lock, [&] { return !S->queued.empty() || S->stop.load(); });
                    idle_since = 0;
                    continue;
                }
            } else {
                idle_since = 0;
            }
        }
        if (!batch.empty()) {
            launch_batch(S, *free_slot, batch, descs, bytes);
            batch.clear();
            progress = true;
        }
        if (!progress) std::this_thread::yield();
    }
}

void destroy_slots(CkFftServer* S)
{
    for (Slot& sl : S->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.h_desc) (void)hipHostFree(sl.h_desc);
        if (sl.d_desc) (void)hipFree(sl.d_desc);
        if (sl.d_ctl) (void)hipFree(sl.d_ctl);
        if (sl.d_scratch) (void)hipFree(sl.d_scratch);
    }
    S->slots.clear();
}

}  // namespace

extern "C" {

CkFftServer* CkFftServerCreate(CkFftContext* context, size_t hbmBudget, int maxInFlight)
{
    if (!context || context->magic != kContextMagic || maxInFlight < 1 || maxInFlight > 64) return nullptr;
    auto* S = new (std::nothrow) CkFftServer();
    if (!S) return nullptr;
    S->ctx = context;
    S->slots.resize(maxInFlight);
    // the largest multi-pass intermediate the context's sizes need (complex nMax >= 2^16:
    // 8 nMax bytes; real nMax >= 2^17: 4 nMax), up to the cap
    const size_t nmax = (size_t)context->max_count;
    S->slot_scratch = nmax >= ((size_t)1 << (kLogMaxLds + 1)) ? std::min(8 * nmax, kSlotScratchCap) : 0;
    for (Slot& sl : S->slots) {
        if (!serr(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking), "stream") ||
            !serr(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "event") ||
            !serr(hipHostMalloc((void**)&sl.h_desc, kDescBytes, hipHostMallocDefault), "pinned queue") ||
            !serr(hipMalloc((void**)&sl.d_desc, kDescBytes), "queue") ||
            !serr(hipMalloc(&sl.d_ctl, sizeof(int) * kCtlWords), "control") ||
            !serr(hipMemset(sl.d_ctl, 0, sizeof(int) * kCtlWords), "clear control")) {
            destroy_slots(S);
            delete S;
            return nullptr;
        }
        S->resident += (long long)(kDescBytes + sizeof(int) * kCtlWords);
        if (S->slot_scratch) {
            if (!serr(hipMalloc((void**)&sl.d_scratch, S->slot_scratch), "slot intermediate")) {
                destroy_slots(S);
                delete S;
                return nullptr;
            }
            sl.scratch_bytes = S->slot_scratch;
            S->resident += (long long)S->slot_scratch;
        }
    }
    S->budget = (long long)hbmBudget - S->resident;
    S->grid = persistent_grid(context, false);
    if (S->budget <= 0 || S->grid <= 0) {
        std::snprintf(g_serve_error, sizeof(g_serve_error), "budget below the server's resident %lld bytes",
                      S->resident);
        destroy_slots(S);
        delete S;
        return nullptr;
    }
    S->worker = std::thread(dispatcher, S);
    return S;
}

static long long submit_at(CkFftServer* S, const CkFftBatchDesc& d, uint64_t t_arrive)
{
    if (!batch_check_desc(S->ctx, d, 0, true)) {
        std::snprintf(g_serve_error, sizeof(g_serve_error), "invalid request (%s)", CkFftBatchLastError());
        return -1;
    }
    int log2n = 0;
    while ((1 << log2n) < d.n) ++log2n;
    // algorithmic bytes, plus the intermediate of a multi-pass launch when the slot's resident
    // one is too small for it (one half-size complex signal, 8 * n/2 bytes, for a real kind)
    const bool real = d.kind >= kRealForward;
    const bool multipass = real ? log2n > kLogMaxLds + 1 : log2n > kLogMaxLds;
    const bool mfma = S->ctx->precision == kCkFftPrecision_SplitBf16x3 && mfma_supported((Kind)d.kind, log2n);
    const long long inter = multipass && !mfma ? (real ? 4ll : 8ll) << log2n : 0;
    const long long bytes = transform_bytes(d.kind, log2n) + (inter > (long long)S->slot_scratch ? inter : 0);
    if (bytes > S->budget) {
        std::snprintf(g_serve_error, sizeof(g_serve_error), "request larger than the admission budget");
        return -1;
    }
    long long id;
    {
        std::unique_lock<std::mutex> lock(S->m);
        id = S->next_ticket++;
        Request& r = S->ring[id % kRing];
        // the entry's previous request (kRing submissions ago) must have finished
        S->cv.wait(lock, [&] { return r.state >= kDone; });
        r.d = d;
        r.bytes = bytes;
        r.t_arrive = t_arrive;
        r.t_submit = now_ns();
        r.ticket = id;
        r.state = kQueued;
        S->queued.push_back(id);
        ++S->outstanding;
        if (!S->t_first) S->t_first = t_arrive;
    }
    S->cv.notify_all();
    return id;
}

long long CkFftServerSubmit(CkFftServer* server, int kind, int n, const void* input, void* output)
{
    if (!server) return -1;
    return submit_at(server, CkFftBatchDesc{kind, n, input, output}, now_ns());
}

int CkFftServerReplay(CkFftServer* server, const CkFftBatchDesc* descs, const double* arrivalUs, int count,
                      long long* tickets)
{
    if (!server || count < 0 || (count && (!descs || !arrivalUs))) return 0;
    const uint64_t t0 = now_ns();
    int ok = 1;
    for (int i = 0; i < count; ++i) {
        const uint64_t t = t0 + (uint64_t)(arrivalUs[i] * 1e3);
        while (now_ns() < t) {
        }
        const long long id = submit_at(server, descs[i], t);
        if (tickets) tickets[i] = id;
        if (id < 0) ok = 0;
    }
    return ok;
}

int CkFftServerQuery(CkFftServer* server, long long ticket)
{
    if (!server || ticket < 0) return -1;
    std::lock_guard<std::mutex> lock(server->m);
    if (ticket >= server->next_ticket) return -1;
    const Request& r = server->ring[ticket % kRing];
    if (r.ticket != ticket)                    // older than the ring: finished long ago
        return server->failed_tickets.count(ticket) ? -1 : 1;
    return r.state == kDone ? 1 : (r.state == kFailed ? -1 : 0);
}

int CkFftServerWait(CkFftServer* server, long long ticket)
{
    if (!server || ticket < 0) return 0;
    std::unique_lock<std::mutex> lock(server->m);
    if (ticket >= server->next_ticket) return 0;
    const Request& r = server->ring[ticket % kRing];
    server->cv.wait(lock, [&] { return r.ticket != ticket || r.state >= kDone; });
    if (r.ticket != ticket) return server->failed_tickets.count(ticket) ? 0 : 1;
    return r.state == kDone ? 1 : 0;
}

int CkFftServerDrain(CkFftServer* server)
{
    if (!server) return 0;
    std::unique_lock<std::mutex> lock(server->m);
    server->cv.wait(lock, [&] { return server->outstanding == 0; });
    return server->failed == 0 ? 1 : 0;
}

int CkFftServerGetStats(CkFftServer* server, CkFftServerStats* out)
{
    if (!server || !out) return 0;
    std::lock_guard<std::mutex> lock(server->m);
    std::vector<double> l = server->lat_us;
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) { return l.empty() ? 0.0 : l[std::min(l.size() - 1, (size_t)(p * (l.size() - 1) + 0.5))]; };
    double mean = 0;
    for (double v : l) mean += v;
    *out = CkFftServerStats{};
    out->completed = server->completed;
    out->failed = server->failed;
    out->p50_us = pct(0.50);
    out->p99_us = pct(0.99);
    out->max_us = l.empty() ? 0.0 : l.back();
    out->mean_us = l.empty() ? 0.0 : mean / l.size();
    const double span = server->t_last > server->t_first ? (server->t_last - server->t_first) * 1e-9 : 0.0;
    out->requests_per_s = span > 0 ? server->completed / span : 0.0;
    out->GB_per_s = span > 0 ? server->done_bytes / span * 1e-9 : 0.0;
    out->peak_inflight_bytes = server->peak_inflight;
    out->budget_bytes = server->budget;
    out->resident_bytes = server->resident;
    out->max_client_us = server->max_client_us;
    out->max_wait_us = server->max_wait_us;
    out->max_device_us = server->max_device_us;
    out->batches = server->batches;
    out->mean_batch = server->batches ? (double)server->batched / server->batches : 0.0;
    return 1;
}

void CkFftServerResetStats(CkFftServer* server)
{
    if (!server) return;
    std::lock_guard<std::mutex> lock(server->m);
    server->lat_us.clear();
    server->lat_next = 0;
    server->completed = server->failed = server->batches = server->batched = server->done_bytes = 0;
    server->peak_inflight = server->inflight_bytes;
    server->t_first = server->t_last = 0;
    server->max_us = server->max_client_us = server->max_wait_us = server->max_device_us = 0;
}

void CkFftServerDestroy(CkFftServer* server)
{
    if (!server) return;
    {
        std::lock_guard<std::mutex> lock(server->m);
        server->stop.store(true);
    }
    server->cv.notify_all();
    if (server->worker.joinable()) server->worker.join();   // finishes outstanding requests first
    destroy_slots(server);
    delete server;
}

const char* CkFftServerLastError(void) { return g_serve_error; }

}  // extern "C"
