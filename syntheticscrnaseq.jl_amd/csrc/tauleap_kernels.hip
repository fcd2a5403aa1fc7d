// gfx950 binomial tau-leap stepper (src/algorithms/tauleap_binomial.jl:31-180,
// regulation mode :transcription; GPU form ext/SyntheticscRNAseqCUDAExt.jl:230-311
// and 317-399, whose Poisson inverse-CDF truncation and Poisson-for-binomial
// approximation, ext:21-39, are replaced by exact draws).
//
// Work mapping as in the FP32 CLE stepper: a workgroup owns up to 64 cells of
// one network (lane = cell), wave w owns genes [w*CH, w*CH+CH), and each lane
// keeps its genes' integer mRNA and protein counts on chip (in LDS) for the
// whole horizon.  Per step:
//   1. telegraph switching (tauleap_binomial.jl:110-121), the same Philox
//      switch stream as the CLE;
//   2. the Hill values of every source gene, in FP64, from the protein counts
//      at the start of the step (:81), into LDS rows [2*slot][64] (act) and
//      [2*slot+1][64] (rep); one workgroup barrier;
//   3. per gene: regulation in FP64 in the reference's loop order
//      (network.jl:158-193: additive edges in source order, then AND, then
//      OR gates), then four exact draws from one 52-bit uniform stream per
//      (cell, step, gene): transcription Poisson(max(reg+beta,0) dt), mRNA
//      decay Binomial(m, 1-exp(-mu_m dt)), translation Poisson(k_t m' dt)
//      with the updated m', protein decay Binomial(p, 1-exp(-mu_p dt))
//      (:128-163).
// All FP64 arithmetic mirrors the CPU restatement operation for operation
// (FP contraction off), so for n = 2, where every Hill value is an exact
// integer square, the counts agree bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "blob.h"
#include "cle_kernels.h"
#include "device_binomial.h"
#include "device_rng.h"
#include "segment_schedule.h"

namespace scrna {

__device__ __forceinline__ uint64_t tl_u64(const uint32_t* B, int i) {
    return (uint64_t)B[i] | ((uint64_t)B[i + 1] << 32);
}

// The four draws of one gene in one step (tauleap_binomial.jl:140-163), from
// the gene's own uniform stream.
// tab: the three start-value tables of the network (in LDS), each kTlTable
// long: exp(-k_t n dt), exp(n log q_m), exp(n log q_p); counts beyond them
// evaluate exp.
__device__ __forceinline__ void tl_gene_step(int& m, int& p, double lam_tx, double k_t, double dt, double pdm,
                                             const double* lqm, double pdp, const double* lqp, const double* tab,
                                             uint32_t cell, uint32_t step, uint32_t g, uint32_t c3t, uint32_t k0,
                                             uint32_t k1) {
#pragma clang fp contract(off)
    UStream st(cell, step, g << 16, c3t, k0, k1);
    const int n_tx = (int)poisson_sample(lam_tx, st);                                // :140
    const int n_md = m > 0 ? (int)binomial_sample(m, pdm, st, lqm, m < kTlTable ? tab[kTlTable + m] : -1.0)
                           : 0;                                                      // :143-149
    m += n_tx - n_md;                                                                // :150
    const int mpos = m > 0 ? m : 0;
    const int n_tl = (int)poisson_sample(k_t * (double)mpos * dt, st,
                                         mpos < kTlTable ? tab[mpos] : -1.0);       // :153
    const int n_pd = p > 0 ? (int)binomial_sample(p, pdp, st, lqp, p < kTlTable ? tab[2 * kTlTable + p] : -1.0)
                           : 0;                                                      // :156-161
    p += n_tl - n_pd;                                                                // :163
}

// Dynamic LDS: Hill rows [nbuf][2S][64] doubles, then the sampler start-value
// tables [3][kTlTable] doubles, then the lanes' counts [NW][CH][2][64] int32.  Keeping the counts in LDS lets the gene loop stay
// rolled, so the exact samplers are inlined once rather than once per gene,
// and CH (genes per wave, <= 32) is a run-time value.  At most 8 waves per
// workgroup leave the FP64 samplers 256 VGPRs.
constexpr int kTauLeapMaxWaves = 8;

// Four waves per SIMD (128 VGPRs; the samplers' rarely used values spill to
// scratch): a CU then holds four 4-wave workgroups instead of two, and the
// extra ones fill the barrier and sampler-loop stalls of the others.  On the
// G=20 benchmark network: 2 waves per SIMD (the compiler's own 199 VGPRs)
// 205 ms, 3 waves 159.6 ms, 4 waves 153.6 ms.
constexpr int kTauLeapWavesPerSimd = 4;

__global__ void __launch_bounds__(64 * kTauLeapMaxWaves) __attribute__((amdgpu_waves_per_eu(kTauLeapWavesPerSimd)))
tauleap_stepper_kernel(const uint32_t* __restrict__ blobs, const Tile* __restrict__ tiles,
                       float* __restrict__ state_pool, float* __restrict__ snap_pool,
                       uint8_t* __restrict__ prom_pool, double* __restrict__ vol_pool, int s_begin, int s_end,
                       int nbuf, int CH, int* __restrict__ sched, int n_tiles, int n_seg, long long wait_ticks) {
#pragma clang fp contract(off)
    extern __shared__ double tl_lds[];
    __shared__ int sh_ticket, sh_fail;
    const SegmentSlot sl = segment_take(sched, n_tiles, &sh_ticket);
    const Tile tile = tiles[sl.tile];
    const uint32_t* __restrict__ B = blobs + tile.blob;
    const int G = (int)B[H_G];
    const int n_steps = (int)B[H_NSTEPS];
    const int r_lo = max(s_begin, tile.s0);
    const int r_hi = max(r_lo, min(min(s_end, tile.s1), n_steps));
    int s_start, s_stop;
    segment_range(r_lo, r_hi, sl.seg, n_seg, &s_start, &s_stop);
    if (!segment_wait(sched, sl, wait_ticks, &sh_fail)) return;
    const bool resume = s_start > 0;
    const int S = (int)B[H_S];
    const uint32_t flags = B[H_FLAGS];
    const int snap_every = (int)B[H_SNAP_EVERY];
    const int C = (int)B[H_CELLS];
    const uint32_t k0 = B[H_SEED_LO], k1 = B[H_SEED_HI];
    const uint32_t c3s = (B[H_NETID] << 4) | kSwitch;
    const uint32_t c3t = (B[H_NETID] << 4) | kTauLeap;
    const bool hill2 = (flags & F_HILL2) != 0;
    const bool tele = (flags & F_TELE) != 0;
    const bool pop = (flags & F_POP) != 0;
    const double growth = __longlong_as_double((long long)tl_u64(B, H_GROWTH_LO));
    const double* __restrict__ D = (const double*)(B + B[H_OFF_TL]);
    const double g_rate = D[TL_GROWTH_RATE], g_noise = D[TL_GROWTH_NOISE];
    const double Kn = D[TL_KN], hill_n = D[TL_HILLN], k_t = D[TL_KT], dt = D[TL_DT];
    const double pdm = D[TL_PDM], pdp = D[TL_PDP];
    const double lqm[2] = {D[TL_LQM], D[TL_LPM]};
    const double lqp[2] = {D[TL_LQP], D[TL_LPP]};
    const double* __restrict__ beta_d = D + TL_HEAD;
    const double* __restrict__ w_edge = beta_d + G;
    const int n_edges = (int)B[H_TL_NEDGES];
    const double* __restrict__ w_gate = w_edge + n_edges;
    const int n_gate_words = (int)B[H_TL_NGATEW];
    const double* __restrict__ init_d = w_gate + n_gate_words;
    const int32_t* __restrict__ slot_of = (const int32_t*)(B + B[H_OFF_SLOT]);
    const int32_t* __restrict__ in_off = (const int32_t*)(B + B[H_OFF_INOFF]);
    const uint32_t* __restrict__ edges = B + B[H_OFF_EDGES];
    const int32_t* __restrict__ gate_off = (const int32_t*)(B + B[H_OFF_GATEOFF]);
    const uint32_t* __restrict__ gates = B + B[H_OFF_GATES];
    const uint32_t* __restrict__ telew = B + B[H_OFF_TELE];
    const size_t off_state = (size_t)tl_u64(B, H_OUT_FINAL_LO);
    const size_t off_snap = (size_t)tl_u64(B, H_OUT_SNAP_LO);
    const size_t off_prom = (size_t)tl_u64(B, H_OUT_PROM_LO);
    const size_t off_vol = (size_t)tl_u64(B, H_OUT_V_LO);

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const bool live = lane < tile.ncell;
    const int cell = tile.cell0 + (live ? lane : 0);   // inactive lanes shadow the tile's first cell
    const int g0 = __builtin_amdgcn_readfirstlane(wave * CH);
    const int ng = __builtin_amdgcn_readfirstlane(max(0, min(CH, G - g0)));   // genes of this wave
    const int rows = 2 * S;
    double* __restrict__ tab = tl_lds + (size_t)nbuf * rows * 64;
    int* __restrict__ cnt = (int*)(tab + 3 * kTlTable) + (size_t)wave * CH * 128;
    {   // start-value tables -> LDS (read after the first step's barrier)
        const double* __restrict__ src = init_d + 2 * G;
        for (int i = threadIdx.x; i < 3 * kTlTable; i += blockDim.x) tab[i] = src[i];
    }
    auto M = [&](int k) -> int& { return cnt[k * 128 + lane]; };
    auto P = [&](int k) -> int& { return cnt[k * 128 + 64 + lane]; };

    uint32_t on = 0, sw_mask = 0;
#pragma unroll 1
    for (int k = 0; k < ng; ++k) {
        const int g = g0 + k;
        const uint32_t tf = telew[g * 8 + T_FLAGS];
        sw_mask |= (tf & 1u) << k;
        if (resume) {
            M(k) = ((const int32_t*)state_pool)[off_state + (size_t)g * C + cell];     // int32 counts
            P(k) = ((const int32_t*)state_pool)[off_state + (size_t)(G + g) * C + cell];
            on |= (uint32_t)(prom_pool[off_prom + (size_t)g * C + cell] != 0) << k;
        } else {
            M(k) = (int)init_d[2 * g];                 // round(Int, steady state), tauleap_binomial.jl:61-62
            P(k) = (int)init_d[2 * g + 1];
            bool o = true;                             // constitutive genes are always ON
            if (tele && (tf & 2u)) {                   // tauleap_binomial.jl:64-76
                const U4 u = philox4x32_10((uint32_t)cell, 0u, (uint32_t)(g >> 2),
                                           (B[H_NETID] << 4) | kPromInit, k0, k1);
                const uint32_t uu[4] = {u.x, u.y, u.z, u.w};
                o = (uint64_t)uu[g & 3] < tl_u64(telew, g * 8 + T_INIT_LO);
            }
            on |= (uint32_t)o << k;
        }
    }
    if (!tele) sw_mask = 0;
    // cell volume (population runs): written by the population init kernel or a previous phase
    double V = pop ? vol_pool[off_vol + cell] : 1.0;

    for (int s = s_start; s < s_stop; ++s) {
        double* __restrict__ H = tl_lds + (nbuf == 2 ? (s & 1) * rows * 64 : 0);

        // Telegraph switching first (tauleap_binomial.jl:110-121).
        if (sw_mask != 0) {
            // Switch block (cell, s, g>>2) carries genes 4q..4q+3; a wave's genes
            // need not start on a block boundary (genes per wave is any value),
            // so a new block is drawn whenever g>>2 changes.
            int have = -1;
            U4 u{0u, 0u, 0u, 0u};
#pragma unroll 1
            for (int k = 0; k < ng; ++k) {
                if (!((sw_mask >> k) & 1u)) continue;
                const int g = g0 + k;
                if ((g >> 2) != have) {
                    have = g >> 2;
                    u = philox4x32_10_uk((uint32_t)cell, (uint32_t)s, (uint32_t)have, c3s, k0, k1);
                }
                const uint32_t* t = telew + g * 8;
                const uint32_t on_lo = __builtin_amdgcn_readfirstlane(t[T_ON_LO]);
                const uint32_t on_hi = __builtin_amdgcn_readfirstlane(t[T_ON_HI]);
                const uint32_t off_lo = __builtin_amdgcn_readfirstlane(t[T_OFF_LO]);
                const uint32_t off_hi = __builtin_amdgcn_readfirstlane(t[T_OFF_HI]);
                const int w = g & 3;
                const uint32_t x = (w & 2) ? ((w & 1) ? u.w : u.z) : ((w & 1) ? u.y : u.x);
                const bool cur = (on >> k) & 1u;
                const uint32_t lo = cur ? off_lo : on_lo;
                const uint32_t hi = cur ? off_hi : on_hi;
                on ^= (uint32_t)((hi != 0u) | (x < lo)) << k;
            }
        }

        // Hill values (network.jl:112-125) from the step-start protein counts.
#pragma unroll 1
        for (int k = 0; k < ng; ++k) {
            const int slot = slot_of[g0 + k];
            if (slot >= 0) {
                const double pf = (double)P(k);
                const double pn = hill2 ? pf * pf : pow(pf, hill_n);
                const double den = pn + Kn;
                H[(2 * slot) * 64 + lane] = pn / den;
                H[(2 * slot + 1) * 64 + lane] = Kn / den;
            }
        }
        __syncthreads();

#pragma unroll 1
        for (int k = 0; k < ng; ++k) {
            const int g = g0 + k;
            // regulatory_input (network.jl:158-193): additive edges in source
            // order, then the AND and OR gates of this target.
            double reg = 0.0;
            const int e1 = in_off[g + 1];
            for (int e = in_off[g]; e < e1; ++e) reg += w_edge[e] * H[edges[2 * e] * 64 + lane];
            const int q1 = gate_off[g + 1];
            for (int q = gate_off[g]; q < q1;) {
                const uint32_t hd = gates[q];
                const int ns = (int)(hd & 0xffffu);
                double prod = 1.0;
                if ((hd >> 24) == GATE_AND) {
                    for (int i = 0; i < ns; ++i) prod *= H[(2 * (int)gates[q + 2 + i]) * 64 + lane];
                    reg += w_gate[q] * prod;
                } else {
                    for (int i = 0; i < ns; ++i) prod *= (1.0 - H[(2 * (int)gates[q + 2 + i]) * 64 + lane]);
                    reg += w_gate[q] * (1.0 - prod);
                }
                q += 2 + ns;
            }
            double prop = reg + beta_d[g];                                       // :129
            if (!(prop > 0.0)) prop = 0.0;
            if (pop) prop = prop * V;                                            // :221
            if (tele && !((on >> k) & 1u)) prop = 0.0;                           // :134-137
            int mk = M(k), pk = P(k);
            tl_gene_step(mk, pk, prop * dt, k_t, dt, pdm, lqm, pdp, lqp, tab, (uint32_t)cell, (uint32_t)s,
                         (uint32_t)g, c3t, k0, k1);
            M(k) = mk;
            P(k) = pk;
        }
        if (pop) {                                                               // :253
            if (g_noise > 0.0) {
                // population.jl:130-134: V *= exp((lambda + sigma xi) dt), xi ~ N(0,1) per cell and
                // step (FP64 Box-Muller on two 52-bit uniforms of the cell's growth block)
                const U4 u = philox4x32_10((uint32_t)cell, (uint32_t)s, 0u, (B[H_NETID] << 4) | kGrowth, k0, k1);
                const double xi = sqrt(-2.0 * log(u52(u.x, u.y))) * cos(6.283185307179586 * u52(u.z, u.w));
                V *= exp((g_rate + g_noise * xi) * dt);
            } else {
                V *= growth;                                                     // population.jl:136-137
            }
        }
        if (nbuf == 1) __syncthreads();

        if (snap_every > 0 && ((s + 1) % snap_every) == 0 && live) {
            const int snap = (s + 1) / snap_every - 1;
#pragma unroll 1
            for (int k = 0; k < ng; ++k) {
                const int g = g0 + k;
                snap_pool[off_snap + ((size_t)(2 * snap) * G + g) * C + cell] = (float)M(k);
                snap_pool[off_snap + ((size_t)(2 * snap + 1) * G + g) * C + cell] = (float)P(k);
            }
        }
    }
    if (live) {
#pragma unroll 1
        for (int k = 0; k < ng; ++k) {
            const int g = g0 + k;
            ((int32_t*)state_pool)[off_state + (size_t)g * C + cell] = M(k);         // exact past 2^24
            ((int32_t*)state_pool)[off_state + (size_t)(G + g) * C + cell] = P(k);
            prom_pool[off_prom + (size_t)g * C + cell] = (uint8_t)((on >> k) & 1u);
        }
        if (pop && wave == 0) vol_pool[off_vol + cell] = V;
    }
    segment_done(sched, sl);
}

size_t tauleap_lds_bytes(int S, int nbuf, int NW, int CH) {
    return (size_t)nbuf * 2 * (size_t)S * 64 * sizeof(double) + (size_t)3 * kTlTable * sizeof(double) +
           (size_t)NW * CH * 2 * 64 * sizeof(int);
}

hipError_t set_tauleap_lds_limit(size_t bytes) {
    bytes = bytes > kStepperStaticLds ? bytes - kStepperStaticLds : bytes;
    return hipFuncSetAttribute((const void*)tauleap_stepper_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
}

int tauleap_max_waves() { return kTauLeapMaxWaves; }

int tauleap_blocks_per_cu(int NW, size_t lds_bytes) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)tauleap_stepper_kernel, 64 * NW, lds_bytes) !=
        hipSuccess)
        return 0;
    return nb;
}

hipError_t launch_tauleap(int CH, int NW, int n_tiles, const uint32_t* blobs, const Tile* tiles, float* state_pool,
                          float* snap_pool, uint8_t* prom_pool, double* vol_pool, int s_begin, int s_end, int nbuf,
                          size_t lds_bytes, int* sched, int n_seg, long long wait_ticks, hipStream_t stream) {
    if (n_tiles <= 0) return hipSuccess;
    if (sched == nullptr) n_seg = 1;
    if (n_seg < 1 || (long long)n_tiles * n_seg > INT32_MAX || NW < 1 || NW > kTauLeapMaxWaves || CH < 1 || CH > 32)
        return hipErrorInvalidValue;
    const dim3 grid((unsigned)(n_tiles * n_seg)), block(64 * NW);
    hipLaunchKernelGGL(tauleap_stepper_kernel, grid, block, lds_bytes, stream, blobs, tiles, state_pool, snap_pool,
                       prom_pool, vol_pool, s_begin, s_end, nbuf, CH, sched, n_tiles, n_seg, wait_ticks);
    return hipGetLastError();
}

}  // namespace scrna
