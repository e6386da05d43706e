// fft_staging.cpp -- host-pointer calls of the drop-in API: staging through pinned
// bounce buffers, with the copies pipelined.
//
// The reference's callers pass host memory (inc/ckfft/ckfft.h:85-150); every call is
// synchronous.  A call with a host buffer takes a staging slot from the context's pool
// (its own stream, device buffer and pinned host buffer) so that calls from several
// threads proceed concurrently: their uploads, transforms and downloads overlap.  Within a
// call:
//  * pageable input is copied into the pinned buffer chunk by chunk, each chunk's DMA to
//    the device starting as soon as it is copied (the CPU copy of chunk c+1 overlaps the
//    DMA of chunk c);
//  * the output comes back by DMA chunk by chunk into the pinned buffer and is copied out
//    behind it (chunk c's CPU copy overlaps the DMA of chunk c+1);
//  * pinned (hipHostMalloc) caller buffers are DMA'd directly, device buffers used in place;
//  * a transform that reads its input once and writes its output once (one workgroup per
//    signal: complex n <= 2^15, real n <= 2^16, FP32 mode) skips the DMA: the kernel reads
//    and writes host memory itself (pinned caller buffers in place, pageable ones through
//    the slot's pinned buffer).  A DMA costs a fixed ~10 us per copy on this path, more
//    than moving these sizes over the bus.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "ckfft_internal.h"

namespace ckfft_gpu {

namespace {

constexpr int kMaxSlots = 8;            // concurrent host-pointer calls per context
// A slot keeps its buffers between calls up to this size; a call that needed more frees
// them when it returns the slot, so a context does not hold the largest call's staging
// memory (up to 8 x 2 x 4 GiB at 2^28) until CkFftShutdown.
constexpr size_t kKeepBytes = (size_t)256 << 20;
constexpr size_t kMinChunk = 1 << 20;   // bytes per pipelined copy chunk (at least)
constexpr int kMaxChunks = 16;

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev[kMaxChunks] = {};
    char* d_buf = nullptr;   // device [in | out]
    char* h_buf = nullptr;   // pinned [in | out]
    size_t bytes = 0;
};

enum MemKind { kHostPageable, kHostPinned, kDevice };

MemKind classify(const void* p)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();   // plain pageable host memory is unknown to HIP
        return kHostPageable;
    }
    if (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged) return kDevice;
    if (a.type == hipMemoryTypeHost) return kHostPinned;
    return kHostPageable;
}

void free_slot(Slot* s)
{
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (hipEvent_t e : s->ev)
        if (e) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    if (s->d_buf) (void)hipFree(s->d_buf);
    if (s->h_buf) (void)hipHostFree(s->h_buf);
    delete s;
}

// Chunk layout of `bytes`: at most kMaxChunks chunks of at least kMinChunk, 256-aligned.
size_t chunk_size(size_t bytes)
{
    size_t c = std::max(kMinChunk, (bytes + kMaxChunks - 1) / kMaxChunks);
    return (c + 255) & ~size_t(255);
}

}  // namespace

struct StagingPool {
    std::mutex m;
    std::condition_variable cv;
    std::vector<Slot*> idle;
    int slots = 0;
};

StagingPool* staging_create() { return new (std::nothrow) StagingPool(); }

void staging_destroy(StagingPool* p)
{
    if (!p) return;
    for (Slot* s : p->idle) free_slot(s);   // every call has returned its slot by now
    delete p;
}

namespace {

Slot* acquire(StagingPool* p, hipError_t* err)
{
    std::unique_lock<std::mutex> lock(p->m);
    for (;;) {
        if (!p->idle.empty()) {
            Slot* s = p->idle.back();
            p->idle.pop_back();
            return s;
        }
        if (p->slots < kMaxSlots) {
            ++p->slots;
            lock.unlock();
            Slot* s = new (std::nothrow) Slot();
            hipError_t e = s ? hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) : hipErrorOutOfMemory;
            for (int i = 0; e == hipSuccess && i < kMaxChunks; ++i)
                e = hipEventCreateWithFlags(&s->ev[i], hipEventDisableTiming);
            if (e != hipSuccess) {
                if (s) free_slot(s);
                std::lock_guard<std::mutex> relock(p->m);
                --p->slots;
                p->cv.notify_one();
                *err = e;
                return nullptr;
            }
            return s;
        }
        p->cv.wait(lock);
    }
}

void release(StagingPool* p, Slot* s)
{
    if (s->bytes > kKeepBytes) {
        // the call synchronised its stream before returning the slot
        (void)hipFree(s->d_buf);
        (void)hipHostFree(s->h_buf);
        s->d_buf = nullptr;
        s->h_buf = nullptr;
        s->bytes = 0;
    }
    {
        std::lock_guard<std::mutex> lock(p->m);
        p->idle.push_back(s);
    }
    p->cv.notify_one();
}

hipError_t grow(Slot* s, size_t need)
{
    if (s->bytes >= need) return hipSuccess;
    if (s->d_buf) (void)hipFree(s->d_buf);
    if (s->h_buf) (void)hipHostFree(s->h_buf);
    s->d_buf = nullptr;
    s->h_buf = nullptr;
    s->bytes = 0;
    hipError_t e = hipMalloc((void**)&s->d_buf, need);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->h_buf, need, hipHostMallocDefault);
    if (e == hipSuccess) s->bytes = need;
    return e;
}

// One pass over the data: the kernel may read and write host memory in place.
bool one_pass(Kind kind, int log2n, const DeviceTables& tb)
{
    const bool real = kind == kRealForward || kind == kRealInverse;
    return tb.precision == kCkFftPrecision_Fp32 && log2n <= (real ? kLogMaxLds + 1 : kLogMaxLds);
}

hipError_t run_direct(Slot* s, Kind kind, int log2n, const void* in, size_t ib, MemKind kin, void* out, size_t ob,
                      MemKind kout, const DeviceTables& tb)
{
    const size_t ib_al = (ib + 255) & ~size_t(255);
    const void* src = in;
    void* dst = out;
    if (kin == kHostPageable) {
        std::memcpy(s->h_buf, in, ib);
        src = s->h_buf;
    }
    if (kout == kHostPageable) dst = s->h_buf + ib_al;
    hipError_t e = launch_transform(kind, log2n, 1, src, dst, tb, s->stream);
    const hipError_t es = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = es;
    if (e == hipSuccess && kout == kHostPageable) std::memcpy(out, dst, ob);
    return e;
}

hipError_t run_staged(Slot* s, Kind kind, int log2n, const void* in, size_t ib, MemKind kin, void* out, size_t ob,
                      MemKind kout, const DeviceTables& tb)
{
    const size_t ib_al = (ib + 255) & ~size_t(255);
    hipError_t e = grow(s, ib_al + ob);
    if (e != hipSuccess) return e;
    if (one_pass(kind, log2n, tb)) return run_direct(s, kind, log2n, in, ib, kin, out, ob, kout, tb);
    char* d_in = s->d_buf;
    char* d_out = s->d_buf + ib_al;
    // input
    const void* src = in;
    if (kin == kHostPinned) {
        e = hipMemcpyAsync(d_in, in, ib, hipMemcpyHostToDevice, s->stream);
        src = d_in;
    } else if (kin == kHostPageable) {
        const size_t cs = chunk_size(ib);
        for (size_t off = 0; off < ib && e == hipSuccess; off += cs) {
            const size_t len = std::min(cs, ib - off);
            std::memcpy(s->h_buf + off, (const char*)in + off, len);
            e = hipMemcpyAsync(d_in + off, s->h_buf + off, len, hipMemcpyHostToDevice, s->stream);
        }
        src = d_in;
    }
    void* dst = kout == kDevice ? out : (void*)d_out;
    if (e == hipSuccess) e = launch_transform(kind, log2n, 1, src, dst, tb, s->stream);
    // output
    if (e == hipSuccess && kout == kHostPinned) {
        e = hipMemcpyAsync(out, d_out, ob, hipMemcpyDeviceToHost, s->stream);
    } else if (e == hipSuccess && kout == kHostPageable) {
        char* h_out = s->h_buf + ib_al;
        const size_t cs = chunk_size(ob);
        int nc = 0;
        for (size_t off = 0; off < ob && e == hipSuccess; off += cs, ++nc) {
            e = hipMemcpyAsync(h_out + off, d_out + off, std::min(cs, ob - off), hipMemcpyDeviceToHost, s->stream);
            if (e == hipSuccess) e = hipEventRecord(s->ev[nc], s->stream);
        }
        size_t off = 0;
        for (int c = 0; c < nc && e == hipSuccess; ++c, off += cs) {
            e = hipEventSynchronize(s->ev[c]);
            if (e == hipSuccess) std::memcpy((char*)out + off, h_out + off, std::min(cs, ob - off));
        }
    }
    const hipError_t es = hipStreamSynchronize(s->stream);
    return e != hipSuccess ? e : es;
}

}  // namespace

hipError_t staged_transform(StagingPool* pool, Kind kind, int log2n, const void* in, size_t ib, void* out, size_t ob,
                            const DeviceTables& tb)
{
    const MemKind kin = classify(in), kout = classify(out);
    if (kin == kDevice && kout == kDevice) {
        // the calling thread's own default stream: calls from several threads sharing a
        // context (inc/ckfft/ckfft.h:39-41) overlap on the device instead of queueing on
        // the null stream, which they still follow as a blocking stream does
        hipError_t e = launch_transform(kind, log2n, 1, in, out, tb, hipStreamPerThread);
        return e != hipSuccess ? e : hipStreamSynchronize(hipStreamPerThread);
    }
    hipError_t e = hipSuccess;
    Slot* s = acquire(pool, &e);
    if (!s) return e;
    e = run_staged(s, kind, log2n, in, ib, kin, out, ob, kout, tb);
    release(pool, s);
    return e;
}

}  // namespace ckfft_gpu
