// Host-side runtime shared by the kernel families: per-calling-thread HIP stream and
// scratch, status codes, the fixed-order tile merge, and result finishing.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "device_common.hpp"
#include "mcdx.h"

namespace mcdx {

// Per-thread engine state: its own non-blocking stream (default priority) and growable
// device/pinned scratch, so concurrent callers on different threads never share buffers
// or serialize on a lock while kernels run.
struct ThreadCtx {
    hipStream_t stream = nullptr;
    void* dev = nullptr;
    size_t dev_bytes = 0;
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    // Scratch of the enqueue-only entry points, which launch on the caller's streams: kept
    // apart from `dev` (used on `stream` by the synchronous entry points) and ordered across
    // caller streams by `enq_done`, recorded after every use (see enqueue_scratch).
    void* enq_dev = nullptr;
    size_t enq_bytes = 0;
    hipEvent_t enq_done = nullptr;
    hipStream_t enq_last = nullptr; // stream of the last recorded use
    bool enq_used = false;
};

// Returns MCDX_OK or an error status; on success *ctx is ready on the current device.
int thread_ctx(ThreadCtx** ctx);
// Ensure the device scratch holds at least `bytes`; contents are not preserved.
int ensure_dev(ThreadCtx* ctx, size_t bytes);
int ensure_pinned(ThreadCtx* ctx, size_t bytes);
// Enqueue scratch of at least `bytes` for work about to be launched on `s`: waits (on the
// device, through `s`) for the previous use on another stream, or (on the host) before
// growing the buffer. The caller launches its work on `s`, then calls enqueue_scratch_done.
int enqueue_scratch(ThreadCtx* ctx, hipStream_t s, size_t bytes, void** out);
int enqueue_scratch_done(ThreadCtx* ctx, hipStream_t s);
int status_of(hipError_t e);

// Device allocation and launch-status checks behind a test seam (mcdx_fault_inject): when the
// calling thread is armed, its Nth allocation reports out-of-memory and allocates nothing, or
// its Nth launch check reports a device error after consuming the real status, so every error
// path can run without a real fault.
bool inject_alloc_failure();
template <class T>
hipError_t dev_malloc(T** p, size_t bytes) {
    if (inject_alloc_failure()) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc((void**)p, bytes);
}
hipError_t launch_check();

// Stream to launch on: the caller's stream if given, else the thread's own.
inline hipStream_t pick_stream(ThreadCtx* ctx, void* stream) {
    return stream ? (hipStream_t)stream : ctx->stream;
}

// Launch the fixed-order merge of each contract's tile partials into `out[contract]`: contract
// c's tiles are tile_offsets[c] .. tile_offsets[c + 1] - 1, or, with tile_offsets null,
// c * n_tiles .. (c + 1) * n_tiles - 1.
hipError_t launch_final_merge(const Welford* partials, const uint64_t* tile_offsets,
                              uint64_t n_tiles, uint32_t n_contracts, Welford* out,
                              hipStream_t stream);

// price = mean, SE = sqrt(var/n) with var = m2/(n-1) (stats.hpp:35-41).
mcdx_result finish(const Welford& w);

inline uint64_t tiles_for(uint64_t paths) { return (paths + kTilePaths - 1) / kTilePaths; }

} // namespace mcdx
