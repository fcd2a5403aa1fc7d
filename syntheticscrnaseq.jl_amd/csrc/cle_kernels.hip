// gfx950 kernels for the Chemical Langevin Equation path.
//
//   cle_stepper_kernel<CH>  persistent multi-step Euler-Maruyama CLE
//                           (reference update rule: src/algorithms/cle.jl:78-121;
//                           GPU form ext/SyntheticscRNAseqCUDAExt.jl:87-108)
//   capture_kernel          LogNormal-Binomial capture (src/capture.jl:25-47)
//   moments_kernel          per-gene FP64 sums over cells
//
// Work mapping of the stepper.  A workgroup owns up to 64 cells of one
// network (lane = cell) and NW waves; wave w owns genes [w*CH, w*CH+CH).
// Each lane keeps the mRNA and protein of its CH genes in VGPRs for the whole
// horizon.  Regulation needs other genes' Hill values, so every step the
// owner of a "source" gene writes act = x^n/(1+x^n) and rep = 1/(1+x^n)
// (x = p/K) into LDS rows [2*slot][64 lanes] and [2*slot+1][64 lanes]; after
// one workgroup barrier each wave gathers its genes' in-edges from LDS.
// Edge lists are uniform per wave (every lane of a wave works on the same
// gene), so they stream through the scalar cache.  Only snapshots and the
// final state touch HBM.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "blob.h"
#include "device_binomial.h"
#include "device_rng.h"
#include "cle_kernels.h"
#include "segment_schedule.h"

namespace scrna {

__device__ __forceinline__ float as_f(uint32_t w) { return __uint_as_float(w); }
__device__ __forceinline__ uint64_t u64_at(const uint32_t* B, int i) {
    return (uint64_t)B[i] | ((uint64_t)B[i + 1] << 32);
}

// Hill pair for one protein value: act = x^n/(1+x^n), rep = 1/(1+x^n) with
// x = p/K (network.jl:112-125; same quantities as cle.jl:79-82 divided
// through by K^n, which keeps x^n finite for large p).
// HILL2 selects n = 2 at compile time; callers branch on the (uniform) flag
// outside their gene loops so that the n = 2 path never evaluates log/exp.
template <bool HILL2>
__device__ __forceinline__ void hill_pair(float p, float invK, float hill_n, float& act, float& rep) {
    const float x = p * invK;
    float xn;
    if (HILL2) {
        xn = x * x;
    } else {
        xn = __builtin_amdgcn_exp2f(hill_n * __builtin_amdgcn_logf(x));
    }
    xn = fminf(xn, 1e30f);
    rep = __builtin_amdgcn_rcpf(1.0f + xn);
    act = xn * rep;
}

// Hill values of a wave's CH genes that are regulation sources -> LDS rows
// [2*slot] (act) and [2*slot+1] (rep), one column per lane.
template <int CH, bool HILL2>
__device__ __forceinline__ void hill_to_lds(const float (&p)[CH], int g0, int G, const int32_t* __restrict__ slot_of,
                                            float invK, float hill_n, float* __restrict__ H, int lane) {
#pragma unroll
    for (int k = 0; k < CH; ++k) {
        const int g = g0 + k;
        if (g < G) {
            const int slot = slot_of[g];
            if (slot >= 0) {
                float act, rep;
                hill_pair<HILL2>(p[k], invK, hill_n, act, rep);
                H[(2 * slot) * 64 + lane] = act;
                H[(2 * slot + 1) * 64 + lane] = rep;
            }
        }
    }
}

// Up to 4 genes per lane the kernel fits 8 waves per SIMD (the SGPR budget is
// the limit); asking for that occupancy measured 10% faster on the G=20
// benchmark (33.5 vs 37.1 ms).  Wider tiles keep the compiler's choice, since
// their registers would spill.
template <int CH>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(CH <= 4 ? 8 : 1)))
cle_stepper_kernel(const uint32_t* __restrict__ blobs, const Tile* __restrict__ tiles,
                   float* __restrict__ state_pool, float* __restrict__ snap_pool,
                   uint8_t* __restrict__ prom_pool, double* __restrict__ vol_pool,
                   const float* __restrict__ inj_pool, int s_begin, int s_end,
                   int nbuf, int* __restrict__ sched, int n_tiles, int n_seg, long long wait_ticks) {
    extern __shared__ float lds[];
    __shared__ int sh_ticket, sh_fail;
    const SegmentSlot sl = segment_take(sched, n_tiles, &sh_ticket);
    const Tile tile = tiles[sl.tile];
    const uint32_t* __restrict__ B = blobs + tile.blob;
    const int G = (int)B[H_G];
    const int n_steps = (int)B[H_NSTEPS];
    // Step range of this launch, clamped to the tile and to the network's
    // horizon (networks of a batch may differ in horizon), then this segment's share.
    const int r_lo = max(s_begin, tile.s0);
    const int r_hi = max(r_lo, min(min(s_end, tile.s1), n_steps));
    int s_start, s_stop;
    segment_range(r_lo, r_hi, sl.seg, n_seg, &s_start, &s_stop);
    if (!segment_wait(sched, sl, wait_ticks, &sh_fail)) return;
    const bool resume = s_start > 0;       // later chunks and segments continue from the state pool
    const int S = (int)B[H_S];
    const uint32_t flags = B[H_FLAGS];
    const int snap_every = (int)B[H_SNAP_EVERY];
    const int C = (int)B[H_CELLS];
    const uint32_t k0 = B[H_SEED_LO], k1 = B[H_SEED_HI];
    const uint32_t c3n = (B[H_NETID] << 4) | kNormal;
    const uint32_t c3s = (B[H_NETID] << 4) | kSwitch;
    const float dt = as_f(B[H_DT]);
    const float neg2ln2dt = as_f(B[H_NEG2LN2DT]);
    const float k_t = as_f(B[H_KT]);
    const float mu_m = as_f(B[H_MUM]);
    const float mu_p = as_f(B[H_MUP]);
    const float invK = as_f(B[H_INVK]);
    const float hill_n = as_f(B[H_HILLN]);
    const float sqrt_dt = as_f(B[H_SQRTDT]);
    const double growth = __longlong_as_double((long long)u64_at(B, H_GROWTH_LO));
    const bool hill2 = (flags & F_HILL2) != 0;
    const bool tele = (flags & F_TELE) != 0;
    const bool inject = (flags & F_INJECT) != 0;
    const bool pop = (flags & F_POP) != 0;
    const uint32_t* __restrict__ beta_w = B + B[H_OFF_BETA];
    const int32_t* __restrict__ slot_of = (const int32_t*)(B + B[H_OFF_SLOT]);
    const int32_t* __restrict__ in_off = (const int32_t*)(B + B[H_OFF_INOFF]);
    const uint32_t* __restrict__ edges = B + B[H_OFF_EDGES];
    const int32_t* __restrict__ gate_off = (const int32_t*)(B + B[H_OFF_GATEOFF]);
    const uint32_t* __restrict__ gates = B + B[H_OFF_GATES];
    const uint32_t* __restrict__ telew = B + B[H_OFF_TELE];
    const uint32_t* __restrict__ initw = B + B[H_OFF_INIT];
    const size_t off_state = (size_t)u64_at(B, H_OUT_FINAL_LO);
    const size_t off_snap = (size_t)u64_at(B, H_OUT_SNAP_LO);
    const size_t off_prom = (size_t)u64_at(B, H_OUT_PROM_LO);
    const size_t off_vol = (size_t)u64_at(B, H_OUT_V_LO);
    const size_t off_inj = (size_t)u64_at(B, H_INJ_LO);

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const bool live = lane < tile.ncell;
    const int cell = tile.cell0 + (live ? lane : 0);   // inactive lanes shadow the tile's first cell
    const int g0 = __builtin_amdgcn_readfirstlane(wave * CH);

    float m[CH], p[CH];
    uint32_t on = 0;                                       // promoter bit k = gene g0+k
    uint32_t sw_mask = 0;                                  // bit k: gene g0+k switches
    double V = 1.0;                                       // cell volume (population runs)
#pragma unroll
    for (int k = 0; k < CH; ++k) {
        const int g = g0 + k;
        m[k] = 0.0f;
        p[k] = 0.0f;
        if (g < G) {
            const uint32_t tf = telew[g * 8 + T_FLAGS];
            sw_mask |= (tf & 1u) << k;
            if (resume) {
                m[k] = state_pool[off_state + (size_t)g * C + cell];
                p[k] = state_pool[off_state + (size_t)(G + g) * C + cell];
                on |= (uint32_t)(prom_pool[off_prom + (size_t)g * C + cell] != 0) << k;
            } else {
                m[k] = as_f(initw[2 * g]);
                p[k] = as_f(initw[2 * g + 1]);
                bool o = true;   // constitutive genes are always ON
                if (tele && (tf & 2u)) {
                    const U4 u = philox4x32_10((uint32_t)cell, 0u, (uint32_t)(g >> 2),
                                               (B[H_NETID] << 4) | kPromInit, k0, k1);
                    const uint32_t uu[4] = {u.x, u.y, u.z, u.w};
                    o = (uint64_t)uu[g & 3] < u64_at(telew, g * 8 + T_INIT_LO);
                }
                on |= (uint32_t)o << k;
            }
        }
    }
    if (pop) V = vol_pool[off_vol + cell];   // written by the population init kernel or a previous phase
    if (!tele) sw_mask = 0;
    const bool any_switch = sw_mask != 0;

    const int rows = 2 * S;
    for (int s = s_start; s < s_stop; ++s) {
        float* __restrict__ H = lds + (nbuf == 2 ? (s & 1) * rows * 64 : 0);

        // Telegraph switching first (tauleap_binomial.jl:110-121).
        if (any_switch) {
#pragma unroll
            for (int k = 0; k < CH; k += 4) {
                const int g = g0 + k;
                if (g < G) {
                    const U4 u = philox4x32_10_uk((uint32_t)cell, (uint32_t)s, (uint32_t)(g >> 2), c3s, k0, k1);
                    const uint32_t uu[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
                    for (int j = 0; j < 4 && k + j < CH; ++j) {
                        if ((sw_mask >> (k + j)) & 1u) {
                            // Both thresholds are uniform scalar loads; the lane's promoter
                            // state picks one (readfirstlane keeps the selection on values,
                            // not on a per-lane address).  thr <= 2^32, so x < thr is
                            // hi != 0 or x < lo.
                            const uint32_t* t = telew + (g + j) * 8;
                            const uint32_t on_lo = __builtin_amdgcn_readfirstlane(t[T_ON_LO]);
                            const uint32_t on_hi = __builtin_amdgcn_readfirstlane(t[T_ON_HI]);
                            const uint32_t off_lo = __builtin_amdgcn_readfirstlane(t[T_OFF_LO]);
                            const uint32_t off_hi = __builtin_amdgcn_readfirstlane(t[T_OFF_HI]);
                            const uint32_t x = uu[(g & 3) + j];
                            const bool cur = (on >> (k + j)) & 1u;
                            const uint32_t lo = cur ? off_lo : on_lo;
                            const uint32_t hi = cur ? off_hi : on_hi;
                            const bool flip = (hi != 0u) | (x < lo);
                            on ^= (uint32_t)flip << (k + j);
                        }
                    }
                }
            }
        }

        // Hill values of this wave's source genes -> LDS (cle.jl:79-82).
        if (hill2) {
            hill_to_lds<CH, true>(p, g0, G, slot_of, invK, hill_n, H, lane);
        } else {
            hill_to_lds<CH, false>(p, g0, G, slot_of, invK, hill_n, H, lane);
        }
        __syncthreads();

        // Regulation + Euler-Maruyama update, two genes per Philox block.
#pragma unroll
        for (int k = 0; k < CH; k += 2) {
            const int g = g0 + k;
            if (g < G) {
                float xm[2], xp[2];
                if (inject) {
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        const int gg = (g + j < G) ? g + j : g;
                        const size_t base = ((size_t)s * G + gg) * 2;
                        xm[j] = inj_pool[off_inj + base * C + cell] * sqrt_dt;
                        xp[j] = inj_pool[off_inj + (base + 1) * C + cell] * sqrt_dt;
                    }
                } else {
                    const U4 u = philox4x32_10_uk((uint32_t)cell, (uint32_t)s, (uint32_t)(g >> 1), c3n, k0, k1);
                    normal_pair_scaled(u.x, u.y, neg2ln2dt, xm[0], xp[0]);
                    normal_pair_scaled(u.z, u.w, neg2ln2dt, xm[1], xp[1]);
                }
#pragma unroll
                for (int j = 0; j < 2 && k + j < CH; ++j) {
                    const int gg = g + j;
                    if (gg < G) {
                        // Additive Hill regulation (cle.jl:84-86).
                        float reg = 0.0f;
                        const int e1 = in_off[gg + 1];
                        for (int e = in_off[gg]; e < e1; ++e)
                            reg = fmaf(as_f(edges[2 * e + 1]), H[edges[2 * e] * 64 + lane], reg);
                        // AND / OR gates (cle.jl:89-106); OR uses rep = 1 - act.
                        if (flags & F_GATES) {
                            const int q1 = gate_off[gg + 1];
                            for (int q = gate_off[gg]; q < q1;) {
                                const uint32_t hd = gates[q];
                                const uint32_t type = hd >> 24;
                                const int ns = (int)(hd & 0xffffu);
                                float prod = 1.0f;
                                for (int i = 0; i < ns; ++i)
                                    prod *= H[(2 * (int)gates[q + 2 + i] + (int)type) * 64 + lane];
                                reg = fmaf(as_f(gates[q + 1]), type == GATE_AND ? prod : 1.0f - prod, reg);
                                q += 2 + ns;
                            }
                        }
                        const float beta = as_f(beta_w[gg]);
                        const bool o = (on >> (k + j)) & 1u;
                        // telegraph gating of the transcription propensity
                        // (tauleap_binomial.jl:134-137); volume scaling (cle.jl:224-225)
                        const float tx = o ? (reg + beta) * (float)V : 0.0f;
                        // mRNA (cle.jl:110-113)
                        const float mm = m[k + j];
                        const float var_m = fmaxf(fmaf(mu_m, fmaxf(mm, 0.0f), tx), 0.0f);
                        float mn = fmaf(fmaf(-mu_m, mm, tx), dt, mm);
                        mn = fmaxf(fmaf(xm[j], __builtin_amdgcn_sqrtf(var_m), mn), 0.0f);
                        // protein with the updated mRNA (cle.jl:116-120)
                        const float pp = p[k + j];
                        const float var_p = fmaxf(fmaf(k_t, mn, mu_p * fmaxf(pp, 0.0f)), 0.0f);
                        float pn = fmaf(fmaf(k_t, mn, -mu_p * pp), dt, pp);
                        pn = fmaxf(fmaf(xp[j], __builtin_amdgcn_sqrtf(var_p), pn), 0.0f);
                        m[k + j] = mn;
                        p[k + j] = pn;
                    }
                }
            }
        }
        if (pop) V *= growth;                                 // cle.jl:237
        if (nbuf == 1) __syncthreads();

        if (snap_every > 0 && ((s + 1) % snap_every) == 0 && live) {
            const int snap = (s + 1) / snap_every - 1;
#pragma unroll
            for (int k = 0; k < CH; ++k) {
                const int g = g0 + k;
                if (g < G) {
                    snap_pool[off_snap + ((size_t)(2 * snap) * G + g) * C + cell] = m[k];
                    snap_pool[off_snap + ((size_t)(2 * snap + 1) * G + g) * C + cell] = p[k];
                }
            }
        }
    }
    if (live) {
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            const int g = g0 + k;
            if (g < G) {
                state_pool[off_state + (size_t)g * C + cell] = m[k];
                state_pool[off_state + (size_t)(G + g) * C + cell] = p[k];
                prom_pool[off_prom + (size_t)g * C + cell] = (uint8_t)((on >> k) & 1u);
            }
        }
        if (pop && wave == 0) vol_pool[off_vol + cell] = V;
    }
    segment_done(sched, sl);
}

// Resets one segmented-schedule block (ticket, error word, per-tile done
// flags) before a launch; a kernel rather than a memset so that the launch
// sequence can be captured in a graph.
__global__ void __launch_bounds__(256) sched_reset_kernel(int* __restrict__ sched, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) sched[i] = 0;
}

// ---------------------------------------------------------------------------
// Capture: one thread per (cell, gene) of one job (capture.jl:25-47).
// The per-cell efficiency is recomputed by every gene of the cell from the
// same counter, which is cheaper than a separate pass.
template <typename T>
__global__ void __launch_bounds__(256)
capture_kernel(const CaptureJob* __restrict__ jobs, const T* __restrict__ src_pool,
               int32_t* __restrict__ dst_pool) {
#pragma clang fp contract(off)
    const CaptureJob J = jobs[blockIdx.y];
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)J.cells * J.G) return;
    const int64_t g = idx / J.cells;
    const int64_t c = idx - g * J.cells;
    const U4 e = philox4x32_10((uint32_t)c, 0u, 0u, (J.net_id << 4) | kCaptureEff, J.k0, J.k1);
    const double U1 = u52(e.x, e.y), U2 = u52(e.z, e.w);
    const double z = sqrt(-2.0 * log(U1)) * cos(2.0 * 3.141592653589793 * U2);
    double pc = exp(J.mu_log + J.sigma_log * z);        // capture.jl:36
    pc = pc < 0.0 ? 0.0 : (pc > 1.0 ? 1.0 : pc);
    const double xv = (double)src_pool[J.src_off + (size_t)g * J.src_ld + c];
    int64_t n = (int64_t)rint(xv);                       // round half to even (capture.jl:38)
    if (n < 0) n = 0;
    int64_t y = 0;
    if (n != 0 && pc != 0.0) {                           // capture.jl:39-43
        UStream st((uint32_t)c, (uint32_t)g, 0u, (J.net_id << 4) | kCaptureBin, J.k0, J.k1);
        y = binomial_sample(n, pc, st);
    }
    dst_pool[J.dst_off + (size_t)g * J.dst_ld + c] = (int32_t)y;
}

// ---------------------------------------------------------------------------
// FP64 moment sums per gene: one workgroup per (job, gene); fixed-order tree
// reduction, so results are bitwise reproducible.
__global__ void __launch_bounds__(256)
moments_kernel(const MomentJob* __restrict__ jobs, const float* __restrict__ pool,
               double* __restrict__ out) {
    const MomentJob J = jobs[blockIdx.y];
    const int g = blockIdx.x;
    if (g >= J.G) return;
    const float* __restrict__ mrow = pool + J.m_off + (size_t)g * J.cells;
    const float* __restrict__ prow = pool + J.p_off + (size_t)g * J.cells;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0;
    for (int64_t c = threadIdx.x; c < J.cells; c += blockDim.x) {
        const double a = (double)mrow[c], b = (double)prow[c];
        s0 += a;
        s1 += a * a;
        s2 += b;
        s3 += b * b;
        s4 += a * b;
    }
    __shared__ double red[5][256];
    red[0][threadIdx.x] = s0;
    red[1][threadIdx.x] = s1;
    red[2][threadIdx.x] = s2;
    red[3][threadIdx.x] = s3;
    red[4][threadIdx.x] = s4;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int q = 0; q < 5; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) out[J.out_off + (size_t)g * 5 + threadIdx.x] = red[threadIdx.x][0];
}

// ---------------------------------------------------------------------------
// Population dynamics (Moran process), one thread per cell of a tile.
//
// pop_init_kernel: V ~ U(V_init_lo, V_init_hi) (cle.jl:158-159, population.jl:364-365).
__global__ void __launch_bounds__(64)
pop_init_kernel(const uint32_t* __restrict__ blobs, const Tile* __restrict__ tiles, double* __restrict__ vol_pool,
                const double* __restrict__ vinit /* [2*n_nets] lo, hi */) {
#pragma clang fp contract(off)
    const Tile tile = tiles[blockIdx.x];
    if ((int)threadIdx.x >= tile.ncell) return;
    const uint32_t* __restrict__ B = blobs + tile.blob;
    const int cell = tile.cell0 + threadIdx.x;
    const U4 u = philox4x32_10((uint32_t)cell, 0u, 0u, (B[H_NETID] << 4) | kVInit, B[H_SEED_LO], B[H_SEED_HI]);
    const double lo = vinit[2 * tile.net], hi = vinit[2 * tile.net + 1];
    vol_pool[(size_t)u64_at(B, H_OUT_V_LO) + cell] = lo + (hi - lo) * u52(u.x, u.y);
}

// Division phase 1 (population.jl:157-192 / cle.jl:241-261): every cell with
// V > V_div after step s_div splits its molecules with its daughter -- binomially
// (round half-even, then Binomial(n, 1/2)) or by exact halving -- halves its
// volume, buffers the daughter, and claims a uniformly drawn replacement slot.
// Later dividers (higher cell index) win a contested slot, which is what the
// reference's sequential phase-2 loop produces.
__global__ void __launch_bounds__(64)
pop_divide_kernel(const uint32_t* __restrict__ blobs, const Tile* __restrict__ tiles,
                  float* __restrict__ state_pool, uint8_t* __restrict__ prom_pool, double* __restrict__ vol_pool,
                  float* __restrict__ dstate, uint8_t* __restrict__ dprom, double* __restrict__ dvol,
                  int32_t* __restrict__ claim, int s_div) {
#pragma clang fp contract(off)
    const Tile tile = tiles[blockIdx.x];
    if ((int)threadIdx.x >= tile.ncell) return;
    const uint32_t* __restrict__ B = blobs + tile.blob;
    if (s_div >= (int)B[H_NSTEPS]) return;
    const int G = (int)B[H_G];
    const int C = (int)B[H_CELLS];
    const int cell = tile.cell0 + threadIdx.x;
    const size_t off_state = (size_t)u64_at(B, H_OUT_FINAL_LO);
    const size_t off_prom = (size_t)u64_at(B, H_OUT_PROM_LO);
    const size_t off_vol = (size_t)u64_at(B, H_OUT_V_LO);
    const double vdiv = __longlong_as_double((long long)u64_at(B, H_VDIV_LO));
    double V = vol_pool[off_vol + cell];
    if (!(V > vdiv)) return;
    const uint32_t mode = B[H_POPMODE];
    const uint32_t k0 = B[H_SEED_LO], k1 = B[H_SEED_HI];
    const uint32_t net = B[H_NETID];
    for (int g = 0; g < G; ++g) {
        for (int sp = 0; sp < 2; ++sp) {
            const size_t i = off_state + (size_t)(sp * G + g) * C + cell;
            const float x = state_pool[i];
            float mother, daughter;
            if (mode & 1u) {
                int64_t n = (int64_t)rint((double)x);
                if (n < 0) n = 0;
                int64_t d = 0;
                if (n > 0) {
                    UStream st((uint32_t)cell, (uint32_t)s_div, ((uint32_t)(g * 2 + sp)) << 16,
                               (net << 4) | kPartition, k0, k1);
                    d = binomial_sample(n, 0.5, st);
                }
                mother = (float)(n - d);
                daughter = (float)d;
            } else {
                mother = daughter = x / 2.0f;
            }
            state_pool[i] = mother;
            dstate[i] = daughter;
        }
        dprom[off_prom + (size_t)g * C + cell] = prom_pool[off_prom + (size_t)g * C + cell];
    }
    V /= 2.0;
    vol_pool[off_vol + cell] = V;
    dvol[off_vol + cell] = V;
    const U4 u = philox4x32_10((uint32_t)cell, (uint32_t)s_div, 0u, (net << 4) | kReplace, k0, k1);
    const uint64_t x = ((uint64_t)u.x << 32) | (uint64_t)u.y;
    int64_t r;
    if ((mode & 2u) && C > 1) {                        // cle.jl:265-268
        r = (int64_t)__umul64hi(x, (uint64_t)(C - 1));
        if (r >= cell) r += 1;
    } else {                                           // population.jl:196
        r = (int64_t)__umul64hi(x, (uint64_t)C);
    }
    atomicMax(&claim[off_vol + r], cell + 1);
}

// Division phase 2: each claimed slot takes its winning daughter.
__global__ void __launch_bounds__(64)
pop_place_kernel(const uint32_t* __restrict__ blobs, const Tile* __restrict__ tiles,
                 float* __restrict__ state_pool, uint8_t* __restrict__ prom_pool, double* __restrict__ vol_pool,
                 const float* __restrict__ dstate, const uint8_t* __restrict__ dprom,
                 const double* __restrict__ dvol, int32_t* __restrict__ claim) {
    const Tile tile = tiles[blockIdx.x];
    if ((int)threadIdx.x >= tile.ncell) return;
    const uint32_t* __restrict__ B = blobs + tile.blob;
    const int G = (int)B[H_G];
    const int C = (int)B[H_CELLS];
    const int r = tile.cell0 + threadIdx.x;
    const size_t off_state = (size_t)u64_at(B, H_OUT_FINAL_LO);
    const size_t off_prom = (size_t)u64_at(B, H_OUT_PROM_LO);
    const size_t off_vol = (size_t)u64_at(B, H_OUT_V_LO);
    const int w = claim[off_vol + r];
    if (w == 0) return;
    const int mother = w - 1;
    for (int g = 0; g < G; ++g) {
        state_pool[off_state + (size_t)g * C + r] = dstate[off_state + (size_t)g * C + mother];
        state_pool[off_state + (size_t)(G + g) * C + r] = dstate[off_state + (size_t)(G + g) * C + mother];
        prom_pool[off_prom + (size_t)g * C + r] = dprom[off_prom + (size_t)g * C + mother];
    }
    vol_pool[off_vol + r] = dvol[off_vol + mother];
    claim[off_vol + r] = 0;
}

// ---------------------------------------------------------------------------
// Host-side launchers.

hipError_t launch_pop_init(int n_tiles, const uint32_t* blobs, const Tile* tiles, double* vol, const double* vinit,
                           hipStream_t stream) {
    if (n_tiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(pop_init_kernel, dim3(n_tiles), dim3(64), 0, stream, blobs, tiles, vol, vinit);
    return hipGetLastError();
}

hipError_t launch_pop_divide(int n_tiles, const uint32_t* blobs, const Tile* tiles, float* state, uint8_t* prom,
                             double* vol, float* dstate, uint8_t* dprom, double* dvol, int32_t* claim, int s_div,
                             hipStream_t stream) {
    if (n_tiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(pop_divide_kernel, dim3(n_tiles), dim3(64), 0, stream, blobs, tiles, state, prom, vol, dstate,
                       dprom, dvol, claim, s_div);
    return hipGetLastError();
}

hipError_t launch_pop_place(int n_tiles, const uint32_t* blobs, const Tile* tiles, float* state, uint8_t* prom,
                            double* vol, const float* dstate, const uint8_t* dprom, const double* dvol, int32_t* claim,
                            hipStream_t stream) {
    if (n_tiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(pop_place_kernel, dim3(n_tiles), dim3(64), 0, stream, blobs, tiles, state, prom, vol, dstate,
                       dprom, dvol, claim);
    return hipGetLastError();
}

size_t stepper_lds_bytes(int S, int nbuf) { return (size_t)nbuf * 2 * (size_t)S * 64 * sizeof(float); }

hipError_t launch_stepper(int CH, int NW, int n_tiles, const uint32_t* blobs, const Tile* tiles,
                          float* state_pool, float* snap_pool, uint8_t* prom_pool, double* vol_pool,
                          const float* inj_pool, int s_begin, int s_end, int nbuf,
                          size_t lds_bytes, int* sched, int n_seg, long long wait_ticks, hipStream_t stream) {
    if (n_tiles <= 0) return hipSuccess;
    if (sched == nullptr) n_seg = 1;
    if (n_seg < 1 || (long long)n_tiles * n_seg > INT32_MAX) return hipErrorInvalidValue;
    const dim3 grid((unsigned)(n_tiles * n_seg)), block(64 * NW);
#define SCRNA_STEPPER_CASE(N)                                                                                  \
    case N:                                                                                                    \
        hipLaunchKernelGGL(cle_stepper_kernel<N>, grid, block, lds_bytes, stream, blobs, tiles, state_pool,   \
                           snap_pool, prom_pool, vol_pool, inj_pool, s_begin, s_end, nbuf, sched, n_tiles, n_seg, \
                           wait_ticks);                                                                        \
        break;
    switch (CH) {
        SCRNA_STEPPER_CASE(2)
        SCRNA_STEPPER_CASE(4)
        SCRNA_STEPPER_CASE(8)
        SCRNA_STEPPER_CASE(16)
        default:
            return hipErrorInvalidValue;
    }
#undef SCRNA_STEPPER_CASE
    return hipGetLastError();
}

hipError_t launch_sched_reset(int* sched, int n_words, hipStream_t stream) {
    if (n_words <= 0) return hipSuccess;
    hipLaunchKernelGGL(sched_reset_kernel, dim3((unsigned)((n_words + 255) / 256)), dim3(256), 0, stream, sched,
                       n_words);
    return hipGetLastError();
}

int stepper_blocks_per_cu(int CH, int NW, size_t lds_bytes) {
    const void* f = nullptr;
    switch (CH) {
        case 2: f = (const void*)cle_stepper_kernel<2>; break;
        case 4: f = (const void*)cle_stepper_kernel<4>; break;
        case 8: f = (const void*)cle_stepper_kernel<8>; break;
        case 16: f = (const void*)cle_stepper_kernel<16>; break;
        default: return 0;
    }
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * NW, lds_bytes) != hipSuccess) return 0;
    return nb;
}

hipError_t set_stepper_lds_limit(size_t bytes) {
    // The kernel's own static LDS (schedule ticket and flag) shares the budget.
    bytes = bytes > kStepperStaticLds ? bytes - kStepperStaticLds : bytes;
    hipError_t e = hipSuccess;
    const void* fns[] = {(const void*)cle_stepper_kernel<2>, (const void*)cle_stepper_kernel<4>,
                         (const void*)cle_stepper_kernel<8>, (const void*)cle_stepper_kernel<16>};
    for (const void* f : fns) {
        hipError_t r = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (r != hipSuccess) e = r;
    }
    return e;
}

hipError_t launch_capture_f32(int n_jobs, int64_t max_elems, const CaptureJob* jobs, const float* src,
                              int32_t* dst, hipStream_t stream) {
    if (n_jobs <= 0 || max_elems <= 0) return hipSuccess;
    const dim3 grid((unsigned)((max_elems + 255) / 256), (unsigned)n_jobs), block(256);
    hipLaunchKernelGGL(capture_kernel<float>, grid, block, 0, stream, jobs, src, dst);
    return hipGetLastError();
}

hipError_t launch_capture_f64(int n_jobs, int64_t max_elems, const CaptureJob* jobs, const double* src,
                              int32_t* dst, hipStream_t stream) {
    if (n_jobs <= 0 || max_elems <= 0) return hipSuccess;
    const dim3 grid((unsigned)((max_elems + 255) / 256), (unsigned)n_jobs), block(256);
    hipLaunchKernelGGL(capture_kernel<double>, grid, block, 0, stream, jobs, src, dst);
    return hipGetLastError();
}

hipError_t launch_moments(int n_jobs, int max_G, const MomentJob* jobs, const float* pool, double* out,
                          hipStream_t stream) {
    if (n_jobs <= 0 || max_G <= 0) return hipSuccess;
    const dim3 grid((unsigned)max_G, (unsigned)n_jobs), block(256);
    hipLaunchKernelGGL(moments_kernel, grid, block, 0, stream, jobs, pool, out);
    return hipGetLastError();
}

}  // namespace scrna
