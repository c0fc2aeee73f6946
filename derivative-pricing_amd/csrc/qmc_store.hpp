// Sobol + Brownian-bridge paths written into an LSM path store (QMC for American options).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mcdx {

// Bytes of device memory qmc_fill_store needs beside the stores for m dates and `runs` runs
// (bridge program and shift words); the caller provides them so the thread's scratch is not
// reallocated.
size_t qmc_store_scratch_bytes(int32_t m, uint32_t runs);

// store[d][p] (d < m, p < n; float, or double when fp64) = the spot at date d + 1 of the path
// built from Sobol index p + 1 by the Brownian bridge (brownian_bridge_path.cpp:20-77). With
// `shifted`, dimension k's integer is XORed with Philox4x32-10 word 0 at key shift_seed,
// counter {replicate, 0, k, 3}, and taken at its cell midpoint (x + 0.5) / 2^32; otherwise
// u = x / 2^32 (sobol.cpp:96). With runs > 1 (shifted only), run k uses replicate
// replicate + k and writes the store at store + k m n elements, all in one launch. Enqueued on
// s; `scratch` holds qmc_store_scratch_bytes(m, runs).
int qmc_fill_store(double spot, double rate, double carry, double vol, double time, int32_t m,
                   uint64_t n, bool fp64, bool shifted, uint64_t shift_seed, uint32_t replicate,
                   uint32_t runs, void* store, void* scratch, hipStream_t s);

} // namespace mcdx
