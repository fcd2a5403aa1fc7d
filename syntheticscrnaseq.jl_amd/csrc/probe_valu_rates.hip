// VALU issue-rate probe for gfx950: sustained throughput of each instruction
// class the CLE stepper issues, expressed as a cost relative to full-rate
// FP32 FMA.
//
// Every lane runs 8 independent dependency chains of one operation, so the
// measurement is throughput-bound, not latency-bound; 2048 workgroups of 4
// waves (8 waves per SIMD on 256 CUs) keep every SIMD busy, and 16384
// iterations make each kernel run for milliseconds, so launch overhead is
// negligible.  A chain step that needs a helper instruction to stay alive
// (e.g. log2(x) + c) is corrected by the helper's own chain: add_f32 for
// floating-point steps, alignbit_b32 (a rotate by a run-time amount) for integer
// ones.
// All memory accesses are in bounds: each thread reads its seed from and
// writes one word to a buffer of grid size.
//
// Output: one JSON line per chain with its time, chain-steps per second and
// "cost" = time of the probed instruction alone / time of fma_f32.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#define CHECK(x)                                                                         \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) {                                                          \
            fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));               \
            exit(1);                                                                     \
        }                                                                                \
    } while (0)

constexpr int kChains = 8;
constexpr int kIters = 16384;

enum Op { FMA_F32, ADD_F32, ALIGN_B32, MAD_U64_U32, MUL_HI_U32, MUL_LO_U32, LOG_F32, EXP_F32, SQRT_F32, RCP_F32,
          SIN_F32, COS_F32, CVT_F32_U32, N_OPS };
// name, helper chain subtracted (-1: none)
struct OpInfo {
    const char* name;
    int helper;
};
static const OpInfo kOps[N_OPS] = {
    {"fma_f32", -1},           {"add_f32", -1},          {"alignbit_b32", -1},          {"mad_u64_u32 (+xor)", ALIGN_B32},
    {"mul_hi_u32 (+add)", ALIGN_B32}, {"mul_lo_u32 (+add)", ALIGN_B32}, {"log_f32 (+add)", ADD_F32},
    {"exp_f32 (+mul)", ADD_F32}, {"sqrt_f32 (+add)", ADD_F32}, {"rcp_f32 (+add)", ADD_F32},
    {"sin_f32 (+add)", ADD_F32}, {"cos_f32 (+add)", ADD_F32}, {"cvt_f32_u32 (+xor)", ALIGN_B32}};

template <int OP>
__global__ void __launch_bounds__(256) rate_kernel(const uint32_t* __restrict__ seed, uint32_t* __restrict__ out,
                                                   uint32_t k) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t u[kChains];
    float f[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
        u[c] = seed[tid] + 0x9E3779B9u * (uint32_t)c;
        f[c] = 1.0f + (float)(u[c] & 1023) * 0x1p-10f;
    }
    const float fk = (float)(k & 7) * 0.125f;
    for (int i = 0; i < kIters; ++i) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) {
            if (OP == FMA_F32) f[c] = fmaf(f[c], 0.999f, 0.001f);
            if (OP == ADD_F32) f[c] = f[c] + fk;
            if (OP == ALIGN_B32) u[c] = __builtin_amdgcn_alignbit(u[c], u[c], k);
            if (OP == MAD_U64_U32) {
                const uint64_t p = (uint64_t)u[c] * (uint64_t)0xD2511F53u + (uint64_t)k;
                u[c] = (uint32_t)(p >> 32) ^ (uint32_t)p;
            }
            if (OP == MUL_HI_U32) u[c] = __umulhi(u[c], 0xD2511F53u) + k;
            if (OP == MUL_LO_U32) u[c] = u[c] * 0xD2511F53u + k;
            if (OP == LOG_F32) f[c] = __builtin_amdgcn_logf(f[c]) + 2.0f;
            if (OP == EXP_F32) f[c] = __builtin_amdgcn_exp2f(f[c]) * 0.5f;
            if (OP == SQRT_F32) f[c] = __builtin_amdgcn_sqrtf(f[c]) + 0.5f;
            if (OP == RCP_F32) f[c] = __builtin_amdgcn_rcpf(f[c]) + 1.0f;
            if (OP == SIN_F32) f[c] = __builtin_amdgcn_sinf(f[c]) + 0.25f;
            if (OP == COS_F32) f[c] = __builtin_amdgcn_cosf(f[c]) + 0.25f;
            if (OP == CVT_F32_U32) {
                f[c] = (float)u[c];
                u[c] = __float_as_uint(f[c]) ^ k;
            }
        }
    }
    uint32_t acc = 0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) acc ^= u[c] ^ __float_as_uint(f[c]);
    out[tid] = acc;
}

typedef void (*KernelFn)(const uint32_t*, uint32_t*, uint32_t);

int main(int argc, char** argv) {
    const int blocks = argc > 1 ? atoi(argv[1]) : 2048;
    const int threads = 256;
    const size_t n = (size_t)blocks * threads;
    uint32_t *d_seed = nullptr, *d_out = nullptr;
    CHECK(hipMalloc(&d_seed, n * sizeof(uint32_t)));
    CHECK(hipMalloc(&d_out, n * sizeof(uint32_t)));
    uint32_t* h = (uint32_t*)malloc(n * sizeof(uint32_t));
    for (size_t i = 0; i < n; ++i) h[i] = (uint32_t)(i * 2654435761u);
    CHECK(hipMemcpy(d_seed, h, n * sizeof(uint32_t), hipMemcpyHostToDevice));
    const KernelFn fns[N_OPS] = {rate_kernel<FMA_F32>,     rate_kernel<ADD_F32>,     rate_kernel<ALIGN_B32>,
                                 rate_kernel<MAD_U64_U32>, rate_kernel<MUL_HI_U32>,  rate_kernel<MUL_LO_U32>,
                                 rate_kernel<LOG_F32>,     rate_kernel<EXP_F32>,     rate_kernel<SQRT_F32>,
                                 rate_kernel<RCP_F32>,     rate_kernel<SIN_F32>,     rate_kernel<COS_F32>,
                                 rate_kernel<CVT_F32_U32>};
    int dev = 0, clk_khz = 0, cus = 0;
    CHECK(hipGetDevice(&dev));
    CHECK(hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeClockRate, dev));
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    double t[N_OPS];
    const double lane_steps = (double)n * kIters * kChains;
    printf("{\"device\": {\"cus\": %d, \"clock_mhz\": %.0f}, \"lanes\": %zu, \"chains\": %d, \"iters\": %d}\n", cus,
           clk_khz / 1e3, n, kChains, kIters);
    for (int op = 0; op < N_OPS; ++op) {
        hipLaunchKernelGGL(fns[op], dim3(blocks), dim3(threads), 0, 0, d_seed, d_out, 7u);   // warm-up
        CHECK(hipGetLastError());
        float best = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
            CHECK(hipEventRecord(e0, 0));
            hipLaunchKernelGGL(fns[op], dim3(blocks), dim3(threads), 0, 0, d_seed, d_out, 7u);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e1, 0));
            CHECK(hipEventSynchronize(e1));
            float ms = 0.0f;
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) best = ms;
        }
        t[op] = best;
        const double alone = kOps[op].helper >= 0 ? t[op] - t[kOps[op].helper] : t[op];
        printf("{\"op\": \"%s\", \"ms\": %.4f, \"chain_steps_per_s\": %.4e, \"cost_vs_fma_f32\": %.3f}\n",
               kOps[op].name, best, lane_steps / (best * 1e-3), alone / t[FMA_F32]);
        fflush(stdout);
    }
    CHECK(hipFree(d_seed));
    CHECK(hipFree(d_out));
    free(h);
    return 0;
}
