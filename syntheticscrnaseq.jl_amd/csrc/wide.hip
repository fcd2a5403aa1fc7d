// Wide path: CLE networks above the fused steppers' 256 genes.
//
// Per step, one library GEMM forms the additive regulation of every cell,
//     reg[cells x G] = act[cells x G] * W^T      (W = A_pos - A_neg, network.jl:206-219)
// in FP32 (rocBLAS SGEMM), BF16 (gemm_ex, bf16 operands, FP32 accumulation) or
// BF16x3 (three gemm_ex: a_hi W_hi + a_lo W_hi + a_hi W_lo), and one fused
// elementwise kernel does everything else of the step: bias (basal rate +
// rowsum(A_neg), the rep = 1 - act form), AND/OR gates from the FP32 Hill
// values of their sources, telegraph switching, the Philox normals, the
// Euler-Maruyama update with its clamps (cle.jl:108-120), snapshots, and the
// Hill values of the new protein counts that the next step's GEMM reads.  So
// a step moves about 30 bytes per cell*gene through HBM (the GEMM's act and
// reg, the state in and out), against the reference structure's ~136 of
// eleven elementwise kernels and two SGEMMs (unfused.hip).
//
// Thread mapping of the elementwise kernel: a wave holds 256 consecutive cells
// (4 per lane, 16-byte accesses; 64 cells, one per lane, when the cell count
// is not a multiple of 4) of one gene pair (g, g+1), g even -- the pair of
// one normal block (cell, s, g>>1) -- so its blob reads are uniform and its
// state accesses coalesce along the [G][cells] layout of the state pool.  The
// draws, the switching rule and the update are those of the dense stepper;
// trajectories match the oracle within the FP32 or BF16 bars of the fused forms.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <stdint.h>

#include <cstring>

#include "blob.h"
#include "device_rng.h"
#include "hill.h"
#include "runtime_internal.h"

namespace scrna {

namespace {

__device__ __forceinline__ float w_as_f(uint32_t w) { return __uint_as_float(w); }
__device__ __forceinline__ uint64_t w_u64(const uint32_t* B, int i) {
    return (uint64_t)B[i] | ((uint64_t)B[i + 1] << 32);
}
__device__ __forceinline__ uint32_t w_pick4(const U4& u, int i) {
    return (i & 2) ? ((i & 1) ? u.w : u.z) : ((i & 1) ? u.y : u.x);
}

struct WideArgs {
    const uint32_t* B;     // the network's descriptor
    float* state_pool;
    float* snap_pool;
    uint8_t* prom_pool;
    const float* inj_pool;
    const float* reg;      // [G][cells] GEMM output
    void* act;             // [G][cells] GEMM input: f32 (FP32) or bf16 (BF16 forms, high plane)
    __bf16* act_lo;        // [G][cells] BF16x3 low plane
    float* gh;             // [2 buffers][2 S rows][cells] FP32 act / rep of gate sources
};

// MODE 1 BF16, 2 BF16X3, 3 FP32.  s < 0: write the initial state, promoters
// and Hill values; s >= 0: step s.  A lane holds V consecutive cells (V = 4 when
// the cell count is a multiple of 4: 16-byte state loads and stores).
template <int MODE, bool TELE, bool INJ, int V>
__global__ void __launch_bounds__(256) wide_step_kernel(WideArgs a, int s) {
    typedef float fv __attribute__((ext_vector_type(V)));
    typedef __bf16 bv __attribute__((ext_vector_type(V)));
    typedef uint8_t cv __attribute__((ext_vector_type(V)));
    const uint32_t* __restrict__ B = a.B;
    const int G = (int)B[H_G];
    const int pair = (int)blockIdx.y * 4 + (int)(threadIdx.x >> 6);
    const int g0 = 2 * pair;
    if (g0 >= G) return;                                    // whole waves
    const int C = (int)B[H_CELLS];
    const int S = (int)B[H_S];
    const int lane = threadIdx.x & 63;
    const int64_t c0l = ((int64_t)blockIdx.x * 64 + lane) * V;   // first cell of the lane; C % V == 0
    const int c0 = (int)(c0l < C ? c0l : C);
    const bool live = c0 < C;
    const int cb = live ? c0 : 0;                            // idle lanes shadow cells 0.. and store nothing
    const uint32_t flags = B[H_FLAGS];
    const uint32_t k0 = B[H_SEED_LO], k1 = B[H_SEED_HI];
    const uint32_t net = B[H_NETID];
    const float invK = w_as_f(B[H_INVK]);
    const float hill_n = w_as_f(B[H_HILLN]);
    const bool hill2 = (flags & F_HILL2) != 0;
    const size_t GC = (size_t)G * C;
    float* __restrict__ M = a.state_pool + (size_t)w_u64(B, H_OUT_FINAL_LO);
    float* __restrict__ P = M + GC;
    uint8_t* __restrict__ prom = a.prom_pool + (size_t)w_u64(B, H_OUT_PROM_LO);
    const int32_t* __restrict__ slot_of = (const int32_t*)(B + B[H_OFF_SLOT]);
    const uint32_t* __restrict__ telew = B + B[H_OFF_TELE];
    const int gbuf = s < 0 ? 0 : (s & 1);                     // Hill values this step reads
    const float* __restrict__ ghr = a.gh + (size_t)gbuf * 2 * S * C;
    float* __restrict__ ghw = a.gh + (size_t)(s < 0 ? 0 : 1 - gbuf) * 2 * S * C;

    // Hill values of the new protein counts -> the GEMM input (and gate rows).
    auto write_hill = [&](int g, const fv& p) {
        fv act, rep;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            float av, rv;
            if constexpr (MODE == 3) {
                if (hill2)
                    hill_pair<true>(p[v], invK, 2.0f, av, rv);
                else
                    hill_pair<false>(p[v], invK, hill_n, av, rv);
            } else {
                // act = 1 - rep from one fma, as in the dense BF16 forms
                const float x = p[v] * invK;
                const float d = hill2 ? fmaf(x, x, 1.0f)
                                      : 1.0f + __builtin_amdgcn_exp2f(hill_n * __builtin_amdgcn_logf(x));
                rv = __builtin_amdgcn_rcpf(d);
                av = 1.0f - rv;
            }
            act[v] = av;
            rep[v] = rv;
        }
        if (!live) return;
        const size_t i = (size_t)g * C + c0;
        if constexpr (MODE == 3) {
            *(fv*)((float*)a.act + i) = act;
        } else {
            bv hi, lo;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                hi[v] = (__bf16)act[v];
                lo[v] = (__bf16)(act[v] - (float)hi[v]);
            }
            *(bv*)((__bf16*)a.act + i) = hi;
            if constexpr (MODE == 2) *(bv*)(a.act_lo + i) = lo;
        }
        const int sl = slot_of[g];
        if (sl >= 0) {
            *(fv*)(ghw + (size_t)(2 * sl) * C + c0) = act;
            *(fv*)(ghw + (size_t)(2 * sl + 1) * C + c0) = rep;
        }
    };

    if (s < 0) {
        const uint32_t* __restrict__ initw = B + B[H_OFF_INIT];
        for (int j = 0; j < 2 && g0 + j < G; ++j) {
            const int g = g0 + j;
            fv m, p;
            cv o;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                m[v] = w_as_f(initw[2 * g]);
                p[v] = w_as_f(initw[2 * g + 1]);
                bool on = true;
                if (TELE && (telew[g * 8 + T_FLAGS] & 2u)) {   // tauleap_binomial.jl:64-76
                    const U4 ui = philox4x32_10((uint32_t)(cb + v), 0u, (uint32_t)(g >> 2), (net << 4) | kPromInit,
                                                k0, k1);
                    on = (uint64_t)w_pick4(ui, g) < w_u64(telew, g * 8 + T_INIT_LO);
                }
                o[v] = (uint8_t)on;
            }
            if (live) {
                *(fv*)(M + (size_t)g * C + c0) = m;
                *(fv*)(P + (size_t)g * C + c0) = p;
                if (TELE) *(cv*)(prom + (size_t)g * C + c0) = o;
            }
            write_hill(g, p);
        }
        return;
    }

    const float dt = w_as_f(B[H_DT]);
    const float neg2ln2dt = w_as_f(B[H_NEG2LN2DT]);
    const float k_t = w_as_f(B[H_KT]);
    const float mu_m = w_as_f(B[H_MUM]);
    const float mu_p = w_as_f(B[H_MUP]);
    const int snap_every = (int)B[H_SNAP_EVERY];
    const int32_t* __restrict__ gate_off = (const int32_t*)(B + B[H_OFF_GATEOFF]);
    const uint32_t* __restrict__ gates = B + B[H_OFF_GATES];
    // The normal blocks (cell, s, g>>1) of the lane's cells: genes g0 (words 0, 1) and g0+1 (words 2, 3).
    U4 u[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        u[v] = U4{0u, 0u, 0u, 0u};
        if (!INJ || TELE)   // the switching bits come from the normal block also with injected noise
            u[v] = philox4x32_10((uint32_t)(cb + v), (uint32_t)s, (uint32_t)pair, (net << 4) | kNormal, k0, k1);
    }
    for (int j = 0; j < 2 && g0 + j < G; ++j) {
        const int g = g0 + j;
        const size_t i = (size_t)g * C + cb;
        const fv regv = *(const fv*)(a.reg + i);
        const fv mv = *(const fv*)(M + i), pv = *(const fv*)(P + i);
        cv ov;
        if constexpr (TELE) ov = *(const cv*)(prom + i);
        fv injm, injp;
        if constexpr (INJ) {
            const float* __restrict__ inj = a.inj_pool + (size_t)w_u64(B, H_INJ_LO);
            const size_t base = ((size_t)s * G + g) * 2;
            injm = *(const fv*)(inj + base * C + cb);
            injp = *(const fv*)(inj + (base + 1) * C + cb);
        }
        const float bias = w_as_f(B[B[H_OFF_BIAS] + g]) + w_as_f(B[B[H_OFF_BETA] + g]);
        const int q0 = gate_off[g], q1 = gate_off[g + 1];
        fv mo, po;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int cell = cb + v;
            const uint32_t wa = j ? u[v].z : u[v].x, wb = j ? u[v].w : u[v].y;
            float r2, c, sn;
            if constexpr (INJ) {
                r2 = dt;
                c = injm[v];
                sn = injp[v];
            } else {
                normal_pair_r2(wa, wb, neg2ln2dt, r2, c, sn);
            }
            // Regulation: the GEMM's W act, then bias, basal rate and gates (cle.jl:84-106).
            float reg = regv[v] + bias;
            for (int q = q0; q < q1;) {
                const uint32_t hd = gates[q];
                const uint32_t type = hd >> 24;
                const int ns = (int)(hd & 0xffffu);
                float prod = 1.0f;
                for (int k = 0; k < ns; ++k) prod *= ghr[(size_t)(2 * (int)gates[q + 2 + k] + (int)type) * C + cell];
                reg = fmaf(w_as_f(gates[q + 1]), type == GATE_AND ? prod : 1.0f - prod, reg);
                q += 2 + ns;
            }
            bool on = true;
            if constexpr (TELE) {
                // Telegraph switching (tauleap_binomial.jl:110-121): X = q << 14 | e flips
                // iff below the current state's threshold; e only matters when q == T.
                const uint32_t tf = telew[g * 8 + T_FLAGS];
                const bool cur = ov[v] != 0;
                const uint32_t thr_on = telew[g * 8 + T_ON_HI] != 0u ? 1u << 18 : telew[g * 8 + T_ON_LO] >> 14;
                const uint32_t thr_off = telew[g * 8 + T_OFF_HI] != 0u ? 1u << 18 : telew[g * 8 + T_OFF_LO] >> 14;
                const uint32_t qq = ((wa & 0x1FFu) << 9) | (wb & 0x1FFu);
                const uint32_t T = cur ? thr_off : thr_on;
                const bool sw = (tf & 1u) != 0;
                bool flip = sw && qq < T;
                if (sw && qq == T) {
                    const U4 ue = philox4x32_10((uint32_t)cell, (uint32_t)s, (uint32_t)(g >> 2), (net << 4) | kSwitch,
                                                k0, k1);
                    const uint32_t e = w_pick4(ue, g) >> 18;
                    flip = (((uint64_t)qq << 14) | e) < w_u64(telew, g * 8 + (cur ? T_OFF_LO : T_ON_LO));
                }
                on = cur != flip;
                ov[v] = (uint8_t)on;
            }
            // Euler-Maruyama update (cle.jl:108-120), with the clamps of negative start counts.
            const float tx = on ? reg : 0.0f;
            const float mm = mv[v], pp = pv[v];
            const float var_m = fmaxf(fmaf(mu_m, fmaxf(mm, 0.0f), tx), 0.0f);
            float mn = fmaf(fmaf(-mu_m, mm, tx), dt, mm);
            mn = fmaxf(fmaf(c, __builtin_amdgcn_sqrtf(r2 * var_m), mn), 0.0f);
            const float var_p = fmaxf(fmaf(k_t, mn, mu_p * fmaxf(pp, 0.0f)), 0.0f);
            float pn = fmaf(fmaf(k_t, mn, -mu_p * pp), dt, pp);
            pn = fmaxf(fmaf(sn, __builtin_amdgcn_sqrtf(r2 * var_p), pn), 0.0f);
            mo[v] = mn;
            po[v] = pn;
        }
        if (live) {
            *(fv*)(M + i) = mo;
            *(fv*)(P + i) = po;
            if constexpr (TELE) *(cv*)(prom + i) = ov;
            if (snap_every > 0 && (s + 1) % snap_every == 0) {
                float* __restrict__ snp = a.snap_pool + (size_t)w_u64(B, H_OUT_SNAP_LO);
                const size_t snap = (size_t)((s + 1) / snap_every - 1);
                *(fv*)(snp + (2 * snap * G + g) * (size_t)C + c0) = mo;
                *(fv*)(snp + ((2 * snap + 1) * G + g) * (size_t)C + c0) = po;
            }
        }
        write_hill(g, po);
    }
}

template <int MODE>
const void* wide_fn(bool tele, bool inj, bool vec4) {
    const void* f[8] = {(const void*)wide_step_kernel<MODE, false, false, 1>,
                        (const void*)wide_step_kernel<MODE, false, true, 1>,
                        (const void*)wide_step_kernel<MODE, true, false, 1>,
                        (const void*)wide_step_kernel<MODE, true, true, 1>,
                        (const void*)wide_step_kernel<MODE, false, false, 4>,
                        (const void*)wide_step_kernel<MODE, false, true, 4>,
                        (const void*)wide_step_kernel<MODE, true, false, 4>,
                        (const void*)wide_step_kernel<MODE, true, true, 4>};
    return f[(vec4 ? 4 : 0) | (tele ? 2 : 0) | (inj ? 1 : 0)];
}

uint16_t host_f2bf(float f) {     // round to nearest even
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
float host_bf2f(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

}  // namespace

int wide_prepare(scrna_plan* P, WideNet& w) {
    const CompiledNet& cn = P->nets[w.net];
    const size_t G = (size_t)cn.G, GC = G * (size_t)cn.cells;
    const size_t act_bytes = w.mode == 3 ? 4 : 2;
    void* p = nullptr;
    int rc;
    if ((rc = plan_alloc(P, &p, GC * sizeof(float)))) return rc;
    w.reg = (float*)p;
    if ((rc = plan_alloc(P, &p, GC * act_bytes))) return rc;
    w.act = p;
    if (w.mode == 2) {
        if ((rc = plan_alloc(P, &p, GC * 2))) return rc;
        w.act_lo = p;
    }
    if (cn.S > 0) {
        if ((rc = plan_alloc(P, &p, (size_t)2 * 2 * cn.S * cn.cells * sizeof(float)))) return rc;
        w.gh = (float*)p;
    }
    // W column-major (element (g, k) at g + k G): FP32, or bf16 high (and low) planes.
    const size_t GG = G * G;
    if (w.mode == 3) {
        if ((rc = plan_alloc(P, &p, GG * sizeof(float)))) return rc;
        w.W = p;
        if (hipMemcpy(w.W, w.host_W.data(), GG * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            return set_error(SCRNA_ERR_DEVICE, "upload of W failed");
    } else {
        std::vector<uint16_t> hi(GG), lo(w.mode == 2 ? GG : 0);
        for (size_t k = 0; k < GG; ++k) {
            hi[k] = host_f2bf(w.host_W[k]);
            if (w.mode == 2) lo[k] = host_f2bf(w.host_W[k] - host_bf2f(hi[k]));
        }
        if ((rc = plan_alloc(P, &p, GG * 2))) return rc;
        w.W = p;
        if (hipMemcpy(w.W, hi.data(), GG * 2, hipMemcpyHostToDevice) != hipSuccess)
            return set_error(SCRNA_ERR_DEVICE, "upload of W failed");
        if (w.mode == 2) {
            if ((rc = plan_alloc(P, &p, GG * 2))) return rc;
            w.W_lo = p;
            if (hipMemcpy(w.W_lo, lo.data(), GG * 2, hipMemcpyHostToDevice) != hipSuccess)
                return set_error(SCRNA_ERR_DEVICE, "upload of W failed");
        }
    }
    w.host_W.clear();
    w.host_W.shrink_to_fit();
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success)
        return set_error(SCRNA_ERR_DEVICE, "rocblas_create_handle failed");
    w.blas = h;
    // Deterministic products: the same run gives the same trajectories.
    if (rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed) != rocblas_status_success)
        return set_error(SCRNA_ERR_DEVICE, "rocblas_set_atomics_mode failed");
    return SCRNA_OK;
}

void wide_release(WideNet& w) {
    if (w.blas) rocblas_destroy_handle((rocblas_handle)w.blas);
    w.blas = nullptr;
}

int wide_launch(scrna_plan* P, WideNet& w, hipStream_t st) {
    const CompiledNet& cn = P->nets[w.net];
    const uint32_t flags = cn.blob[H_FLAGS];
    const bool tele = (flags & F_TELE) != 0, inj = (flags & F_INJECT) != 0;
    const int G = cn.G, C = (int)cn.cells;
    // 16-byte aligned 4-cell groups: the cell count and the network's pool offsets are multiples of 4
    const bool vec4 = C % 4 == 0 && cn.off_state % 4 == 0 && cn.off_snap % 4 == 0 && cn.off_prom % 4 == 0 &&
                      cn.off_inj % 4 == 0;
    const void* f = w.mode == 1 ? wide_fn<1>(tele, inj, vec4)
                  : w.mode == 2 ? wide_fn<2>(tele, inj, vec4)
                                : wide_fn<3>(tele, inj, vec4);
    const int per_wave = vec4 ? 256 : 64;    // cells of one wave
    const dim3 grid((unsigned)(((int64_t)C + per_wave - 1) / per_wave), (unsigned)(((G + 1) / 2 + 3) / 4)), block(256);
    if (grid.y > 65535u) return set_error(SCRNA_ERR_INVALID, "wide path: G=%d needs more than 65535 block rows", G);
    WideArgs a{P->d_blobs + cn.blob_off, P->d_state, P->d_snap, P->d_prom, P->d_inj, w.reg, w.act,
               (__bf16*)w.act_lo, w.gh};
    rocblas_handle h = (rocblas_handle)w.blas;
    if (rocblas_set_stream(h, st) != rocblas_status_success) return set_error(SCRNA_ERR_DEVICE, "rocblas_set_stream");
    auto step = [&](int s) -> int {
        void* args[] = {(void*)&a, (void*)&s};
        const hipError_t e = hipLaunchKernel(f, grid, block, args, 0, st);
        return e == hipSuccess ? SCRNA_OK
                               : set_error(SCRNA_ERR_DEVICE, "wide step kernel: %s", hipGetErrorString(e));
    };
    const float one = 1.0f, zero = 0.0f;
    auto gemm_bf16 = [&](const void* A, const void* Bm, const float* beta) {
        return rocblas_gemm_ex(h, rocblas_operation_none, rocblas_operation_transpose, C, G, G, &one, A,
                               rocblas_datatype_bf16_r, C, Bm, rocblas_datatype_bf16_r, G, beta, w.reg,
                               rocblas_datatype_f32_r, C, w.reg, rocblas_datatype_f32_r, C, rocblas_datatype_f32_r,
                               rocblas_gemm_algo_standard, 0, 0);
    };
    int rc = step(-1);
    if (rc) return rc;
    for (int64_t s = 0; s < cn.n_steps; ++s) {
        // reg[C x G] = act[C x G] * W^T (both column-major; W's element (g, k) at g + k G)
        rocblas_status bs;
        if (w.mode == 3) {
            bs = rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_transpose, C, G, G, &one,
                               (const float*)w.act, C, (const float*)w.W, G, &zero, w.reg, C);
        } else {
            bs = gemm_bf16(w.act, w.W, &zero);
            if (bs == rocblas_status_success && w.mode == 2) bs = gemm_bf16(w.act_lo, w.W, &one);
            if (bs == rocblas_status_success && w.mode == 2) bs = gemm_bf16(w.act, w.W_lo, &one);
        }
        if (bs != rocblas_status_success)
            return set_error(SCRNA_ERR_DEVICE, "wide path GEMM failed (rocblas status %d)", (int)bs);
        if ((rc = step((int)s))) return rc;
    }
    return SCRNA_OK;
}

}  // namespace scrna
