// VALU issue-rate probe (no reference counterpart): the wave64 instructions per second this
// GPU sustains for each instruction class the path kernels are built from, measured with eight
// independent dependency chains per lane at full occupancy, so the rate is the issue limit, not
// a latency. It gives the lane-issue roofline a measured peak per class: a kernel's
// issue-limited time is the sum over classes of its count of that class / the class's rate.
// Two mixed classes (alternate chains of two classes) show whether the classes overlap.
#include "runtime.hpp"

namespace mcdx {

namespace {

constexpr int kChains = 8;

// One class's chain step: one instruction of the class per chain. The operands depend on the
// previous value, so no step can be hoisted or dropped; the final values are stored, so
// nothing is dead. (The values need not stay meaningful: the issue cost of an instruction does
// not depend on its operands.)
template <int CLASS>
__device__ __forceinline__ void step(uint64_t (&q)[kChains], uint32_t (&u)[kChains],
                                     const uint32_t (&w)[kChains], float (&f)[kChains],
                                     double (&d)[kChains]) {
    if constexpr (CLASS == MCDX_ISSUE_XOR3) {
        // written out: from C the compiler shares pairs between chains and emits 2-way XORs
#pragma unroll
        for (int c = 0; c < kChains; ++c)
            asm("v_bitop3_b32 %0, %0, %1, %2 bitop3:0x96" : "+v"(u[c]) : "v"(w[c]), "v"(w[(c + 1) % kChains]));
        return;
    }
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
        if constexpr (CLASS == MCDX_ISSUE_FMA_F32) {
            f[c] = __builtin_fmaf(f[c], 1.0000001f, 0.5f);
        } else if constexpr (CLASS == MCDX_ISSUE_MAD_U64_U32) {
            q[c] = (uint64_t)(uint32_t)q[c] * 0xD2511F53u + q[c];
        } else if constexpr (CLASS == MCDX_ISSUE_MUL_HI_U32) {
            u[c] = __umulhi(u[c], 0xCD9E8D57u);
        } else if constexpr (CLASS == MCDX_ISSUE_EXP2_F32) {
            f[c] = __builtin_amdgcn_exp2f(-f[c]);
        } else if constexpr (CLASS == MCDX_ISSUE_MIX_FMA_MAD) {
            if (c & 1) q[c] = (uint64_t)(uint32_t)q[c] * 0xD2511F53u + q[c];
            else f[c] = __builtin_fmaf(f[c], 1.0000001f, 0.5f);
        } else if constexpr (CLASS == MCDX_ISSUE_MIX_FMA_EXP2) {
            if (c & 1) f[c] = __builtin_amdgcn_exp2f(-f[c]);
            else f[c] = __builtin_fmaf(f[c], 1.0000001f, 0.5f);
        } else {
            d[c] = __builtin_fma(d[c], 1.0000000000001, 0.5);
        }
    }
}

template <int CLASS>
__global__ __launch_bounds__(256) void issue_probe_kernel(uint32_t iters, uint32_t* out_u,
                                                          float* out_f, double* out_d) {
    uint64_t q[kChains];
    uint32_t u[kChains], w[kChains];
    float f[kChains];
    double d[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
        u[c] = threadIdx.x * 2654435761u + (uint32_t)c;
        q[c] = u[c];
        w[c] = u[c] * 0x9E3779B9u;
        f[c] = 0.25f + 0.01f * (float)c;
        d[c] = 0.25 + 0.01 * (double)c;
    }
    for (uint32_t k = 0; k < iters; ++k) step<CLASS>(q, u, w, f, d);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t su = 0;
    float sf = 0.f;
    double sd = 0.0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
        su ^= u[c] ^ (uint32_t)q[c] ^ (uint32_t)(q[c] >> 32);
        sf += f[c];
        sd += d[c];
    }
    out_u[i] = su;
    out_f[i] = sf;
    out_d[i] = sd;
}

template <int CLASS>
void launch(unsigned blocks, uint32_t iters, uint32_t* u, float* f, double* d, hipStream_t s) {
    hipLaunchKernelGGL(issue_probe_kernel<CLASS>, dim3(blocks), dim3(256), 0, s, iters, u, f, d);
}

} // namespace
} // namespace mcdx

using namespace mcdx;

extern "C" {

int mcdx_issue_probe(int instruction_class, uint32_t iterations, double* seconds,
                     uint64_t* wave_instructions) {
    if (!seconds || !wave_instructions || iterations == 0 || iterations > (1u << 24) ||
        instruction_class < MCDX_ISSUE_FMA_F32 || instruction_class > MCDX_ISSUE_MIX_FMA_EXP2)
        return MCDX_EINVAL;
    ThreadCtx* ctx;
    int st = thread_ctx(&ctx);
    if (st) return st;
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return status_of(e);
    // 8 blocks of 256 lanes per CU: 8 waves on each of its four SIMDs
    const unsigned blocks = 8u * (unsigned)cus;
    const size_t lanes = (size_t)blocks * 256;
    if ((st = ensure_dev(ctx, lanes * (4 + 4 + 8)))) return st;
    uint32_t* u = (uint32_t*)ctx->dev;
    float* f = (float*)(u + lanes);
    double* d = (double*)(f + lanes);
    hipEvent_t a, b;
    if ((e = hipEventCreate(&a)) != hipSuccess) return status_of(e);
    if ((e = hipEventCreate(&b)) != hipSuccess) {
        (void)hipEventDestroy(a);
        return status_of(e);
    }
    auto run = [&](uint32_t it) {
        switch (instruction_class) {
        case MCDX_ISSUE_FMA_F32: launch<MCDX_ISSUE_FMA_F32>(blocks, it, u, f, d, ctx->stream); break;
        case MCDX_ISSUE_MAD_U64_U32: launch<MCDX_ISSUE_MAD_U64_U32>(blocks, it, u, f, d, ctx->stream); break;
        case MCDX_ISSUE_MUL_HI_U32: launch<MCDX_ISSUE_MUL_HI_U32>(blocks, it, u, f, d, ctx->stream); break;
        case MCDX_ISSUE_XOR3: launch<MCDX_ISSUE_XOR3>(blocks, it, u, f, d, ctx->stream); break;
        case MCDX_ISSUE_EXP2_F32: launch<MCDX_ISSUE_EXP2_F32>(blocks, it, u, f, d, ctx->stream); break;
        case MCDX_ISSUE_MIX_FMA_MAD: launch<MCDX_ISSUE_MIX_FMA_MAD>(blocks, it, u, f, d, ctx->stream); break;
        case MCDX_ISSUE_MIX_FMA_EXP2: launch<MCDX_ISSUE_MIX_FMA_EXP2>(blocks, it, u, f, d, ctx->stream); break;
        default: launch<MCDX_ISSUE_FMA_F64>(blocks, it, u, f, d, ctx->stream); break;
        }
    };
    run(iterations); // warm-up
    e = hipEventRecord(a, ctx->stream);
    if (e == hipSuccess) {
        run(iterations);
        e = launch_check();
    }
    if (e == hipSuccess) e = hipEventRecord(b, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (e != hipSuccess) return status_of(e);
    *seconds = (double)ms * 1e-3;
    // chains x iterations per lane, per wave of 64 lanes
    *wave_instructions = (uint64_t)kChains * iterations * (lanes / 64);
    return MCDX_OK;
}

} // extern "C"
