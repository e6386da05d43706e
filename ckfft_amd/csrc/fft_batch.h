// fft_batch.h -- host-side interface of the persistent batch kernel, shared by batch
// plans (fft_batch.hip) and the serving loop (fft_serve.cpp).
#pragma once
#include <vector>

#include "ckfft/ckfft_gpu.h"
#include "ckfft_internal.h"

namespace ckfft_gpu {

// One queue entry of the persistent kernel.
struct BatchItem {
    int cls;        // class id (see fft_batch.hip)
    int first;      // first transform (index into the BatchTransform table)
    int count;      // transforms in this item (single-workgroup classes)
    int pad;
};

// The buffers of one transform taken by the persistent kernel or the table passes.
struct BatchTransform {
    const float2* in;
    float2* out;
    float2* scratch;   // intermediate of a table-pass transform (else null)
};

// A transform the persistent kernel does not take: run with an ordinary launch.
struct BatchExtra {
    Kind kind;
    int log2n;
    const void* in;
    void* out;
};

constexpr int kPersistWG = 512;       // threads per persistent workgroup
constexpr int kPassLogN = 16;         // complex size run as two global passes (not by the persistent kernel)
constexpr int kCtlWords = 2;          // control words of one launch: queue head, finished count

// A queue small enough to travel in the kernel arguments (no upload before the launch).
constexpr int kInlineQueue = 32;
struct InlineQueue {
    BatchItem items[kInlineQueue];
    BatchTransform trs[kInlineQueue];
    int n_items;
};

// Host description of a batch.
struct BatchBuild {
    std::vector<BatchItem> items;     // the wide kernel's queue, then the narrow kernel's
    int n_wide = 0;                   // items of the wide kernel's queue
    std::vector<BatchTransform> trs;  // the persistent kernel's transforms, by class
    std::vector<BatchExtra> extras;
    // complex 2^16 transforms run by the table pass launches (TwoPass::kTable), forward
    // then inverse; scratch set by batch_assign_scratch
    std::vector<BatchTransform> pass_fwd, pass_inv;
    long long bytes = 0;              // algorithmic bytes of the whole batch
    int persistent = 0;               // transforms taken by the persistent kernel
};

// Validation of one descriptor (single-signal rules, plus device buffers when asked);
// false records the reason for CkFftBatchLastError.
bool batch_check_desc(const CkFftContext* ctx, const CkFftBatchDesc& d, int index, bool check_pointers);
// Where complex 2^16 transforms run.  Both compute the same two radix-2^8 passes:
//  kExtra  an ordinary multi-pass launch each (the lower latency for a lone transform,
//          used by the server);
//  kTable  two launches over all of them (both directions), tiles spread over the whole GPU
//          (the higher throughput for many, used by batch plans).
enum class TwoPass { kExtra, kTable };

// Validate and group a batch; var_override >= 0 forces one thread layout for every class.
// narrow: classes the narrow persistent kernel takes (two workgroups per CU) go to its
// queue; otherwise every item goes to the wide kernel (one workgroup per CU, any class).
bool batch_build(const CkFftContext* ctx, const CkFftBatchDesc* descs, int count, int var_override,
                 bool check_pointers, BatchBuild& out, TwoPass two_pass, bool narrow);
// Scratch of the table-pass transforms: (1 << kPassLogN) points each.
size_t batch_scratch_transforms(const BatchBuild& b);
void batch_assign_scratch(BatchBuild& b, float2* scratch);
// The table pass launches: pass 1 then pass 2 over the n_fwd forward transforms of d_trs
// followed by its n_inv inverse ones.
hipError_t launch_pass_table(const BatchTransform* d_trs, int n_fwd, int n_inv, const DeviceTables& tb, hipStream_t s);
// Workgroups that fill the GPU with the wide (narrow) persistent kernel.
int persistent_grid(const CkFftContext* ctx, bool narrow);
// One launch of the wide (narrow) persistent kernel over n_items items; d_ctl holds its
// kCtlWords control words, zero before the first launch (the kernel leaves them zero).
hipError_t launch_persistent_raw(const BatchItem* d_items, int n_items, const BatchTransform* d_trs, int* d_ctl,
                                 int grid, const DeviceTables& tb, hipStream_t s, bool narrow);
// The same launch with the queue of `b` passed by value; requires at most kInlineQueue
// items and transforms.
hipError_t launch_persistent_inline(const BatchBuild& b, int* d_ctl, int grid, const DeviceTables& tb, hipStream_t s);
long long transform_bytes(int kind, int log2n);
bool is_device_memory(const void* p);
bool batch_error(hipError_t e, const char* where);

}  // namespace ckfft_gpu
