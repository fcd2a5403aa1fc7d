// gfx950 kernels for the Chemical Langevin Equation path.
//
//   cle_stepper_kernel<CH,INJ,POP>  persistent multi-step Euler-Maruyama CLE
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
// Everything about a wave's genes is uniform across the wave (every lane
// works on the same genes): switch thresholds, basal rates, the Hill rows they
// write and their first in-edges are set up once in VGPRs; further edges and
// gates stream through the scalar cache.  Only snapshots and the final state
// touch HBM.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "blob.h"
#include "device_binomial.h"
#include "device_rng.h"
#include "cle_kernels.h"
#include "hill.h"
#include "segment_schedule.h"

namespace scrna {

__device__ __forceinline__ float as_f(uint32_t w) { return __uint_as_float(w); }
__device__ __forceinline__ uint64_t u64_at(const uint32_t* B, int i) {
    return (uint64_t)B[i] | ((uint64_t)B[i + 1] << 32);
}

// Hill values of a wave's source genes (bit k of srcm: gene g0+k) -> the LDS
// row pair at byte address haddr[k] + hbase: act in the first row, rep in the
// second (64 floats further), one column per lane.
template <int CH, bool HILL2>
__device__ __forceinline__ void hill_to_lds(const float (&p)[CH], uint32_t srcm, const uint32_t (&haddr)[CH],
                                            uint32_t hbase, float invK, float hill_n, float* __restrict__ lds) {
#pragma unroll
    for (int k = 0; k < CH; ++k) {
        if ((srcm >> k) & 1u) {
            float act, rep;
            hill_pair<HILL2>(p[k], invK, hill_n, act, rep);
            float* row = (float*)((char*)lds + (haddr[k] + hbase));
            row[0] = act;
            row[64] = rep;
        }
    }
}

// Telegraph switch of one gene in the 64 cells of a wave
// (tauleap_binomial.jl:110-121): a cell flips iff its 32-bit switching
// uniform X is below the threshold of its current state (ON cells use the
// ON->OFF one; thresholds are ceil(prob * 2^32) <= 2^32).  X = q << 14 | e:
// q = the low 9 bits of the gene's radius word a and angle word b of this
// step's normal block (Box-Muller reads only their top 23 bits), e = the top
// 14 bits of the gene's word of the switch block (cell, step, gene >> 2).
// With T = floor(threshold / 2^14) (t_on / t_off), q < T flips and q > T
// does not, whatever e is; e is drawn only in the rare waves where some lane
// has q == T (probability 2^-18 per cell).
__device__ __forceinline__ uint64_t cle_switch(bool cur, uint32_t a, uint32_t b, uint32_t t_on, uint32_t t_off,
                                               uint32_t cell, uint32_t step, int gene, uint32_t c3, uint32_t k0,
                                               uint32_t k1, const uint32_t* __restrict__ blobs, uint32_t tw_off) {
    const uint32_t q = ((a & 0x1FFu) << 9) | (b & 0x1FFu);
    const uint32_t T = cur ? t_off : t_on;
    uint64_t flip = __builtin_amdgcn_ballot_w64(q < T);
    const uint64_t amb = __builtin_amdgcn_ballot_w64(q == T);
    if (amb != 0) {
        // Here T < 2^18, so the threshold is below 2^32 and X fits in 32 bits.
        // The threshold words' offset is made opaque inside the branch, and
        // the keys too (round keys recomputed here), so that nothing of this
        // path is hoisted into the step and kept live there.
        uint32_t off = tw_off;
        asm("" : "+s"(off) : "s"((uint32_t)amb));
        const U4 u = philox4x32_10_uk(cell, step, (uint32_t)(gene >> 2), c3, k0, k1);
        const int w = gene & 3;
        const uint32_t e = (w == 0 ? u.x : w == 1 ? u.y : w == 2 ? u.z : u.w) >> 18;
        const uint32_t* __restrict__ tw = blobs + off;
        const uint32_t thr = cur ? tw[T_OFF_LO] : tw[T_ON_LO];
        flip |= amb & __builtin_amdgcn_ballot_w64(((q << 14) | e) < thr);
    }
    return flip;                                   // lanes whose promoter flips
}

// Promoter states of a lane's CH genes.  Up to 4 genes per lane: one 64-bit
// wave lane mask per gene (bit `lane` = that lane's cell), in scalar
// registers, so the switch tests produce masks straight from v_cmp, a flip is
// one s_xor_b64 and the gating of transcription one v_cndmask on the mask.
// Wider tiles keep one bit per gene in a per-lane register instead: 2*CH
// mask registers would not fit in the scalar file beside the step's other
// uniforms, and their spills cost a v_readlane at every use.
template <int CH>
struct Promoters {
    static constexpr bool kMasks = CH <= 4;
    uint64_t mask[kMasks ? CH : 1];
    uint32_t bits;
    __device__ __forceinline__ void set(int k, bool o) {
        if constexpr (kMasks) mask[k] = __builtin_amdgcn_ballot_w64(o);
        else bits = (bits & ~(1u << k)) | ((uint32_t)o << k);
    }
    __device__ __forceinline__ bool get(int k) const {
        if constexpr (kMasks) return __builtin_amdgcn_inverse_ballot_w64(mask[k]);
        else return (bits >> k) & 1u;
    }
    __device__ __forceinline__ void flip(int k, uint64_t f) {
        if constexpr (kMasks) mask[k] ^= f;
        else bits ^= (uint32_t)__builtin_amdgcn_inverse_ballot_w64(f) << k;
    }
};

// Occupancy and register slots for up to 4 genes per lane.  The kernel asks
// for 7 waves per SIMD (72 VGPRs, 94 SGPRs), which holds two register slots
// per gene for in-edges without spilling.  On the G=20 benchmark network
// (kernel ms): 2 slots at 7 waves 23.6, 2 slots at 8 waves (VGPRs spill) 23.7,
// 1 slot at 8 waves 24.4, 1 slot at 7 waves 24.6.  Wider tiles keep the
// compiler's choice of occupancy and use no register slots, since their
// registers would spill.
constexpr int kEdgeSlots4 = 2;
constexpr int kWavesPerSimd4 = 7;
// 8 genes per lane: 6 waves per SIMD (G=100, 50,000 cells, 2000 steps: 24.7
// against 25.2 ms at the compiler's 5); 16 genes per lane keep the
// compiler's choice (7 waves per SIMD spilled: G=200 telegraph 68 against 42 ms).
constexpr int kWavesPerSimd8 = 6;
//
// INJ, POP and TELE say whether a launch may hold networks with injected
// noise, a population or telegraph promoters; the flags in each blob still
// decide per network.  Launches without them get a kernel without those
// paths, whose pointers, offsets and flags would otherwise occupy registers
// for the whole horizon (a launch holding a telegraph network must use
// TELE = true: without it every promoter stays ON).
//
// CLAMP keeps the update's max(m, 0) and max(p, 0) and the protein
// variance's max(., 0) (cle.jl:111, 118).  Every count is >= 0 at the start
// of a step (each step ends with max(., 0); resumed, divided and placed states
// come from such steps) unless the initial values are negative (negative
// basal rates, F_NEG_INIT), and k_t, mu_p >= 0, so launches without such a
// network drop them: 7 VALU instructions fewer per 4 genes, same results.
//
// Scalar-register economy.  The step loop is long and every uniform value the
// compiler keeps live across it takes an SGPR (a uniform boolean takes two, as
// a lane mask); past the budget they spill into VGPR lanes and each reload is
// a VALU v_readlane.  So the blob offset and the per-wave bit masks are made
// opaque at the top of every step: header words are re-read through the
// scalar cache and per-gene tests recomputed on the scalar unit instead of
// being hoisted (booleans as lane masks) out of the loop.
template <int CH, bool INJ, bool POP, bool TELE, bool CLAMP>
__global__ void __launch_bounds__(1024)
__attribute__((amdgpu_waves_per_eu(CH <= 4 ? kWavesPerSimd4 : CH == 8 ? kWavesPerSimd8 : 1)))
cle_stepper_kernel(const uint32_t* __restrict__ blobs, const Tile* __restrict__ tiles,
                   float* __restrict__ state_pool, float* __restrict__ snap_pool,
                   uint8_t* __restrict__ prom_pool, double* __restrict__ vol_pool,
                   const float* __restrict__ inj_pool, int s_begin, int s_end,
                   int nbuf, int* __restrict__ sched, int n_tiles, int n_seg, long long wait_ticks) {
    extern __shared__ float lds[];
    __shared__ int sh_ticket, sh_fail;
    const SegmentSlot sl = segment_take(sched, n_tiles, &sh_ticket);
    const Tile tile = tiles[sl.tile];
    const uint32_t blob_off = __builtin_amdgcn_readfirstlane(tile.blob);
    const uint32_t* __restrict__ B = blobs + blob_off;
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
    // Values used before and after the step loop only; the loop re-reads
    // what it needs from the blob.
    const uint32_t flags = B[H_FLAGS];
    const int snap_every = (int)B[H_SNAP_EVERY];
    const int C = (int)B[H_CELLS];
    const bool tele = TELE && (flags & F_TELE) != 0;
    const bool pop = POP && (flags & F_POP) != 0;
    const uint32_t* __restrict__ telew = B + B[H_OFF_TELE];
    const uint32_t* __restrict__ initw = B + B[H_OFF_INIT];
    const size_t off_state = (size_t)u64_at(B, H_OUT_FINAL_LO);
    const size_t off_prom = (size_t)u64_at(B, H_OUT_PROM_LO);
    const size_t off_vol = (size_t)u64_at(B, H_OUT_V_LO);

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const bool live = lane < tile.ncell;
    const int cell = tile.cell0 + (live ? lane : 0);   // inactive lanes shadow the tile's first cell
    const int g0 = __builtin_amdgcn_readfirstlane(wave * CH);

    float m[CH], p[CH];
    Promoters<CH> on;                      // promoter state of gene g0+k in this lane's cell
    on.bits = 0u;
    // Per-gene constants of the step loop kept in VGPRs (uniform values; VALU
    // operands at no cost, and no scalar registers or reloads across the
    // loop): the OFF->ON and ON->OFF thresholds in units of 2^14,
    // floor(threshold / 2^14) <= 2^18 (0 for genes that never switch, which
    // then never flip; see cle_switch), and the LDS byte address of the
    // gene's Hill row pair in this lane (source genes only).
    uint32_t thr_on[CH], thr_off[CH], haddr[CH];
    constexpr bool kVTab = CH <= 4;                        // set up from the wave record
    float beta_v[CH];                                      // basal rate of each gene
    // The first EK additive in-edges of each gene as (LDS byte address in this
    // lane's column, weight); unused slots read the buffer's zero row with
    // weight 0.  Further edges and gates go through the blob (bits of
    // extra_mask: 0..15 more edges, 16..31 gates).
    constexpr int EK = CH <= 4 ? kEdgeSlots4 : 0;
    uint32_t eaddr[CH][EK > 0 ? EK : 1];
    float ew[CH][EK > 0 ? EK : 1];
    uint32_t sw_mask = 0;                                  // bit k: gene g0+k switches
    uint32_t src_mask = 0;                                 // bit k: gene g0+k is a regulation source
    uint32_t extra_mask = 0;
    double V = 1.0;                                       // cell volume (population runs)
    const int S = (int)B[H_S];
    const uint32_t zaddr = (uint32_t)((2 * S * 64 + lane) * 4);   // zero row of a buffer, this lane
    static_assert(kEdgeSlots4 == kRecEdgeSlots, "wave records hold the register edge slots");
    if constexpr (kVTab) {
        // The host laid these out per wave (RecWord): a few wide scalar loads.
        const uint32_t* __restrict__ rec = B + B[H_OFF_WREC] + wave * rec_stride(CH);
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            const uint32_t* __restrict__ rg = rec + k * R_GENE_WORDS;
            thr_on[k] = rg[R_TON];
            thr_off[k] = rg[R_TOFF];
            haddr[k] = rg[R_HROW] + 4u * (uint32_t)lane;
            beta_v[k] = as_f(rg[R_BETA]);
#pragma unroll
            for (int i = 0; i < EK; ++i) {
                eaddr[k][i] = rg[R_EDGE + 2 * i] + 4u * (uint32_t)lane;
                ew[k][i] = as_f(rg[R_EDGE + 2 * i + 1]);
            }
            asm volatile("" : "+v"(thr_on[k]), "+v"(thr_off[k]), "+v"(haddr[k]), "+v"(beta_v[k]));
#pragma unroll
            for (int i = 0; i < EK; ++i) asm volatile("" : "+v"(eaddr[k][i]), "+v"(ew[k][i]));
        }
        sw_mask = rec[R_GENE_WORDS * CH + RM_SWITCH];
        src_mask = rec[R_GENE_WORDS * CH + RM_SOURCE];
        extra_mask = rec[R_GENE_WORDS * CH + RM_EXTRA];
    } else {
        const int32_t* __restrict__ slot_of = (const int32_t*)(B + B[H_OFF_SLOT]);
        const uint32_t* __restrict__ beta_w = B + B[H_OFF_BETA];
        const int32_t* __restrict__ in_off = (const int32_t*)(B + B[H_OFF_INOFF]);
        const uint32_t* __restrict__ edges = B + B[H_OFF_EDGES];
        const int32_t* __restrict__ gate_off = (const int32_t*)(B + B[H_OFF_GATEOFF]);
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            const int g = g0 + k;
            thr_on[k] = thr_off[k] = haddr[k] = 0u;
            beta_v[k] = 0.0f;
#pragma unroll
            for (int i = 0; i < EK; ++i) {
                eaddr[k][i] = zaddr;
                ew[k][i] = 0.0f;
            }
            if (g < G) {
                beta_v[k] = as_f(beta_w[g]);
                const int e0 = in_off[g], e1 = in_off[g + 1];
#pragma unroll
                for (int i = 0; i < EK; ++i) {
                    if (e0 + i < e1) {
                        eaddr[k][i] = (uint32_t)(((int)edges[2 * (e0 + i)] * 64 + lane) * 4);
                        ew[k][i] = as_f(edges[2 * (e0 + i) + 1]);
                    }
                }
                if (e1 - e0 > EK) extra_mask |= 1u << k;
                if ((flags & F_GATES) && gate_off[g + 1] > gate_off[g]) extra_mask |= 1u << (16 + k);
                const uint32_t* t = telew + g * 8;
                if (tele && (t[T_FLAGS] & 1u)) {
                    thr_on[k] = t[T_ON_HI] != 0u ? 1u << 18 : t[T_ON_LO] >> 14;
                    thr_off[k] = t[T_OFF_HI] != 0u ? 1u << 18 : t[T_OFF_LO] >> 14;
                }
                const int slot = slot_of[g];
                if (slot >= 0) {
                    src_mask |= 1u << k;
                    haddr[k] = (uint32_t)((2 * slot * 64 + lane) * 4);
                }
            }
            asm volatile("" : "+v"(thr_on[k]), "+v"(thr_off[k]), "+v"(haddr[k]), "+v"(beta_v[k]));
#pragma unroll
            for (int i = 0; i < EK; ++i) asm volatile("" : "+v"(eaddr[k][i]), "+v"(ew[k][i]));
        }
    }
    // The zero rows (one per buffer) are never written by the steps.
    if (wave == 0) {
        lds[2 * S * 64 + lane] = 0.0f;
        if (nbuf == 2) lds[(2 * S + 1 + 2 * S) * 64 + lane] = 0.0f;
    }
#pragma unroll
    for (int k = 0; k < CH; ++k) {
        const int g = g0 + k;
        m[k] = 0.0f;
        p[k] = 0.0f;
        on.set(k, true);
        if (g < G) {
            const uint32_t tf = telew[g * 8 + T_FLAGS];
            if (!kVTab) sw_mask |= (tf & 1u) << k;
            bool o = true;   // constitutive genes are always ON
            if (resume) {
                m[k] = state_pool[off_state + (size_t)g * C + cell];
                p[k] = state_pool[off_state + (size_t)(G + g) * C + cell];
                o = prom_pool[off_prom + (size_t)g * C + cell] != 0;
            } else {
                m[k] = as_f(initw[2 * g]);
                p[k] = as_f(initw[2 * g + 1]);
                if (tele && (tf & 2u)) {
                    const U4 u = philox4x32_10((uint32_t)cell, 0u, (uint32_t)(g >> 2),
                                               (B[H_NETID] << 4) | kPromInit, B[H_SEED_LO], B[H_SEED_HI]);
                    const uint32_t uu[4] = {u.x, u.y, u.z, u.w};
                    o = (uint64_t)uu[g & 3] < u64_at(telew, g * 8 + T_INIT_LO);
                }
            }
            on.set(k, o);
        }
    }
    if (pop) V = vol_pool[off_vol + cell];   // written by the population init kernel or a previous phase
    // The lane's step-invariant part of its normal blocks (cell, s, pair): three
    // VGPRs that save two lane products per block (G=20 benchmark, 4 genes per
    // lane: 20.0 -> 19.3 ms).  Wider tiles, whose registers are already short,
    // and population launches, whose short phases pay for the extra spills at
    // every start (4096-network batch at T=20: 1.97 -> 2.13 s), draw the
    // blocks whole.
    constexpr bool kLaneRng = CH <= 4 && !POP;
    PhiloxLane plane{0u, 0u, 0u};
    if constexpr (kLaneRng)
        plane = philox_lane((uint32_t)cell, (B[H_NETID] << 4) | kNormal, B[H_SEED_LO], B[H_SEED_HI]);
    sw_mask = tele ? __builtin_amdgcn_readfirstlane(sw_mask) : 0u;
    src_mask = __builtin_amdgcn_readfirstlane(src_mask);
    extra_mask = __builtin_amdgcn_readfirstlane(extra_mask);

    // Next step after which a snapshot is due (a countdown instead of a
    // per-step division).
    int next_snap = snap_every > 0 ? (s_start / snap_every + 1) * snap_every - 1 : INT32_MAX;
    for (int s = s_start; s < s_stop; ++s) {
        // The blob offset is made opaque at every step, so the header words
        // and per-gene addresses below are re-read through the scalar cache
        // inside the step instead of being kept live (and spilled) across it.
        // The asm is not volatile (a volatile one would count as a memory
        // write and turn every later blob read into a vector load) and takes s
        // as an input so that it stays inside the loop; it works on the
        // offset, not on the pointer, which keeps the address space global.
        uint32_t boff = blob_off, swm = sw_mask, srcm = src_mask, xtra = extra_mask;
        asm("" : "+s"(boff) : "s"(s));
        asm("" : "+s"(swm) : "s"(s));
        asm("" : "+s"(srcm) : "s"(s));
        asm("" : "+s"(xtra) : "s"(s));
        const uint32_t* __restrict__ Bs = blobs + boff;
        const uint32_t fl = Bs[H_FLAGS];
        const int nv = min(CH, (int)Bs[H_G] - g0);            // genes of this wave
        const uint32_t sk0 = Bs[H_SEED_LO], sk1 = Bs[H_SEED_HI];
        const uint32_t netc = Bs[H_NETID] << 4;
        // Buffer of this step: 2S Hill rows and a zero row.
        float* __restrict__ H = lds + (nbuf == 2 ? (s & 1) * (2 * (int)Bs[H_S] + 1) * 64 : 0);
        const uint32_t hbase = (uint32_t)((H - lds) * 4);

        // Hill values of this wave's source genes -> LDS (cle.jl:79-82).
        {
            const float invK = as_f(Bs[H_INVK]);
            if (fl & F_HILL2)
                hill_to_lds<CH, true>(p, srcm, haddr, hbase, invK, 2.0f, lds);
            else
                hill_to_lds<CH, false>(p, srcm, haddr, hbase, invK, as_f(Bs[H_HILLN]), lds);
        }
        __syncthreads();

        // Regulation + Euler-Maruyama update, two genes per Philox block.
        {
            const float dt = as_f(Bs[H_DT]);
            const float neg2ln2dt = as_f(Bs[H_NEG2LN2DT]);
            const float k_t = as_f(Bs[H_KT]);
            const float mu_m = as_f(Bs[H_MUM]);
            const float mu_p = as_f(Bs[H_MUP]);
            const int32_t* __restrict__ in_off = (const int32_t*)(Bs + Bs[H_OFF_INOFF]);
            const uint32_t* __restrict__ edges = Bs + Bs[H_OFF_EDGES];
            const int32_t* __restrict__ gate_off = (const int32_t*)(Bs + Bs[H_OFF_GATEOFF]);
            const uint32_t* __restrict__ gates = Bs + Bs[H_OFF_GATES];
#pragma unroll
            for (int k = 0; k < CH; k += 2) {
                const int g = g0 + k;
                if (k < nv) {
                    // Noise of gene j: xi_m = c[j]*sqrt(r2[j]), xi_p = sn[j]*sqrt(r2[j]) (Box-Muller
                    // radius squared, scaled by dt, and the two unit directions), so that
                    // xi*sqrt(var) below costs one square root, sqrt(r2*var).
                    float r2[2], c[2], sn[2];
                    // The normal block also carries the genes' switching bits, so
                    // injected-noise runs draw it for telegraph networks too.
                    const bool inj = INJ && (fl & F_INJECT);
                    U4 u = {0u, 0u, 0u, 0u};
                    if (!inj || (TELE && swm != 0u)) {
                        if constexpr (kLaneRng)
                            u = philox4x32_10_lane(plane, (uint32_t)s, (uint32_t)(g >> 1), sk0, sk1);
                        else
                            u = philox4x32_10((uint32_t)cell, (uint32_t)s, (uint32_t)(g >> 1), netc | kNormal, sk0,
                                              sk1);
                    }
                    // Telegraph switching of the two genes (tauleap_binomial.jl:110-121),
                    // ahead of their transcription below; the Hill values written
                    // above do not depend on promoter states.
#pragma unroll
                    for (int j = 0; j < 2 && k + j < CH; ++j) {
                        if (TELE && k + j < nv && ((swm >> (k + j)) & 1u))
                            on.flip(k + j, cle_switch(on.get(k + j), j ? u.z : u.x, j ? u.w : u.y, thr_on[k + j],
                                                      thr_off[k + j], (uint32_t)cell, (uint32_t)s, g + j,
                                                      netc | kSwitch, sk0, sk1, blobs,
                                                      boff + Bs[H_OFF_TELE] + (uint32_t)(g + j) * 8u));
                    }
                    if (inj) {
                        const int Gs = (int)Bs[H_G];
                        const size_t Cs = (size_t)Bs[H_CELLS];
                        const float* __restrict__ inj = inj_pool + (size_t)u64_at(Bs, H_INJ_LO);
#pragma unroll
                        for (int j = 0; j < 2; ++j) {
                            const int gg = (k + j < nv) ? g + j : g;
                            const size_t base = ((size_t)s * Gs + gg) * 2;
                            r2[j] = dt;
                            c[j] = inj[base * Cs + cell];
                            sn[j] = inj[(base + 1) * Cs + cell];
                        }
                    } else {
                        normal_pair_r2(u.x, u.y, neg2ln2dt, r2[0], c[0], sn[0]);
                        normal_pair_r2(u.z, u.w, neg2ln2dt, r2[1], c[1], sn[1]);
                    }
#pragma unroll
                    for (int j = 0; j < 2 && k + j < CH; ++j) {
                        const int gg = g + j;
                        if (k + j < nv) {
                            // Additive Hill regulation (cle.jl:84-86): the register slots,
                            // then any further (LDS row, weight) pairs of the blob.
                            // basal rate first: the sum is the whole transcription drive
                            float reg = beta_v[k + j];
#pragma unroll
                            for (int i = 0; i < EK; ++i)
                                reg = fmaf(ew[k + j][i], *(const float*)((const char*)lds + (eaddr[k + j][i] + hbase)),
                                           reg);
                            if ((xtra >> (k + j)) & 1u) {
                                const uint32_t* __restrict__ ep = edges + 2 * (in_off[gg] + EK);
                                int n = in_off[gg + 1] - in_off[gg] - EK;
                                // Wide tiles (dense networks): eight edges per round, their
                                // records in one scalar load and their eight Hill reads in
                                // flight together, then the fmas in edge order (the same sum
                                // as one at a time).  G=256 precision workload in FP32 on
                                // this stepper: 667 -> 279 ms.  Narrow tiles keep the
                                // one-edge loop (their extra edges are few; the G=20
                                // workload measured 19.3 -> 19.8 ms with the eight-edge code).
                                if constexpr (CH > 4) {
#pragma unroll 1
                                    for (; n >= 8; n -= 8, ep += 16) {
                                        float hv[8];
#pragma unroll
                                        for (int i = 0; i < 8; ++i) hv[i] = H[ep[2 * i] * 64 + lane];
#pragma unroll
                                        for (int i = 0; i < 8; ++i) reg = fmaf(as_f(ep[2 * i + 1]), hv[i], reg);
                                    }
                                }
#pragma unroll 1
                                for (; n > 0; --n, ep += 2)
                                    reg = fmaf(as_f(ep[1]), H[ep[0] * 64 + lane], reg);
                            }
                            // AND / OR gates (cle.jl:89-106); OR uses rep = 1 - act.
                            if ((xtra >> (16 + k + j)) & 1u) {
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
                            const bool o = on.get(k + j);
                            // telegraph gating of the transcription propensity
                            // (tauleap_binomial.jl:134-137); volume scaling (cle.jl:224-225)
                            const float drive = POP ? reg * (float)V : reg;
                            const float tx = o ? drive : 0.0f;
                            // mRNA (cle.jl:110-113)
                            const float mm = m[k + j];
                            const float var_m = fmaxf(fmaf(mu_m, CLAMP ? fmaxf(mm, 0.0f) : mm, tx), 0.0f);
                            float mn = fmaf(fmaf(-mu_m, mm, tx), dt, mm);
                            mn = fmaxf(fmaf(c[j], __builtin_amdgcn_sqrtf(r2[j] * var_m), mn), 0.0f);
                            // protein with the updated mRNA (cle.jl:116-120)
                            const float pp = p[k + j];
                            float var_p, pn;
                            if (CLAMP) {
                                var_p = fmaxf(fmaf(k_t, mn, mu_p * fmaxf(pp, 0.0f)), 0.0f);
                                pn = fmaf(fmaf(k_t, mn, -mu_p * pp), dt, pp);
                            } else {
                                const float dp = mu_p * pp;        // protein decay propensity
                                var_p = fmaf(k_t, mn, dp);
                                pn = fmaf(fmaf(k_t, mn, -dp), dt, pp);
                            }
                            pn = fmaxf(fmaf(sn[j], __builtin_amdgcn_sqrtf(r2[j] * var_p), pn), 0.0f);
                            m[k + j] = mn;
                            p[k + j] = pn;
                        }
                    }
                }
            }
        }
        if (POP && (fl & F_POP)) V *= __longlong_as_double((long long)u64_at(Bs, H_GROWTH_LO));   // cle.jl:237
        if (nbuf == 1) __syncthreads();

        if (s == next_snap) {
            const int se = (int)Bs[H_SNAP_EVERY];
            next_snap += se;
            if (live) {
                const int snap = (s + 1) / se - 1;
                const int Gs = (int)Bs[H_G];
                const size_t Cs = (size_t)Bs[H_CELLS];
                float* __restrict__ out = snap_pool + (size_t)u64_at(Bs, H_OUT_SNAP_LO);
#pragma unroll
                for (int k = 0; k < CH; ++k) {
                    if (k < nv) {
                        out[((size_t)(2 * snap) * Gs + g0 + k) * Cs + cell] = m[k];
                        out[((size_t)(2 * snap + 1) * Gs + g0 + k) * Cs + cell] = p[k];
                    }
                }
            }
        }
    }
    if (live) {
        // Output offsets re-read through an opaque blob offset rather than
        // kept in scalar registers across the step loop.
        uint32_t boff = blob_off;
        asm("" : "+s"(boff) : "s"(s_stop));
        const uint32_t* __restrict__ Be = blobs + boff;
        const int Ge = (int)Be[H_G];
        const size_t Ce = (size_t)Be[H_CELLS];
        const size_t out_state = (size_t)u64_at(Be, H_OUT_FINAL_LO);
        const size_t out_prom = (size_t)u64_at(Be, H_OUT_PROM_LO);
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            const int g = g0 + k;
            if (g < Ge) {
                state_pool[out_state + (size_t)g * Ce + cell] = m[k];
                state_pool[out_state + (size_t)(Ge + g) * Ce + cell] = p[k];
                prom_pool[out_prom + (size_t)g * Ce + cell] = (uint8_t)on.get(k);
            }
        }
        if (pop && wave == 0) vol_pool[(size_t)u64_at(Be, H_OUT_V_LO) + cell] = V;
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
    const size_t si = J.src_off + (size_t)g * J.src_ld + c;
    const double xv = J.src_int ? (double)((const int32_t*)src_pool)[si] : (double)src_pool[si];
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
        const double a = J.src_int ? (double)((const int32_t*)mrow)[c] : (double)mrow[c];
        const double b = J.src_int ? (double)((const int32_t*)prow)[c] : (double)prow[c];
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
// pop_init_kernel: V ~ U(V_init_lo, V_init_hi) (cle.jl:158-159, population.jl:112-113).
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
    const bool counts = B[H_ALG] == 1u;   // binomial tau-leap (scrna_algorithm 1): int32 counts in the pool
    int32_t* __restrict__ state_i = (int32_t*)state_pool;
    int32_t* __restrict__ dstate_i = (int32_t*)dstate;
    for (int g = 0; g < G; ++g) {
        for (int sp = 0; sp < 2; ++sp) {
            const size_t i = off_state + (size_t)(sp * G + g) * C + cell;
            if (mode & 1u) {
                int64_t n = counts ? (int64_t)state_i[i] : (int64_t)rint((double)state_pool[i]);
                if (n < 0) n = 0;
                int64_t d = 0;
                if (n > 0) {
                    UStream st((uint32_t)cell, (uint32_t)s_div, ((uint32_t)(g * 2 + sp)) << 16,
                               (net << 4) | kPartition, k0, k1);
                    d = binomial_sample(n, 0.5, st);
                }
                if (counts) {
                    state_i[i] = (int32_t)(n - d);
                    dstate_i[i] = (int32_t)d;
                } else {
                    state_pool[i] = (float)(n - d);
                    dstate[i] = (float)d;
                }
            } else {
                const float x = state_pool[i];
                state_pool[i] = x / 2.0f;
                dstate[i] = x / 2.0f;
            }
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
    // Raw 32-bit words: FP32 molecules of the CLE, int32 counts of the tau-leap.
    uint32_t* __restrict__ sw = (uint32_t*)state_pool;
    const uint32_t* __restrict__ dw = (const uint32_t*)dstate;
    for (int g = 0; g < G; ++g) {
        sw[off_state + (size_t)g * C + r] = dw[off_state + (size_t)g * C + mother];
        sw[off_state + (size_t)(G + g) * C + r] = dw[off_state + (size_t)(G + g) * C + mother];
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

// Per buffer: 2S Hill rows (act, rep of each source slot) and one zero row.
size_t stepper_lds_bytes(int S, int nbuf) { return (size_t)nbuf * (2 * (size_t)S + 1) * 64 * sizeof(float); }

template <int CH>
static const void* stepper_fn(bool inj, bool pop, bool tele, bool clamp) {
#define SCRNA_STEPPER_VARIANTS(CL)                                                                       \
    (const void*)cle_stepper_kernel<CH, false, false, false, CL>,                                        \
        (const void*)cle_stepper_kernel<CH, true, false, false, CL>,                                     \
        (const void*)cle_stepper_kernel<CH, false, true, false, CL>,                                     \
        (const void*)cle_stepper_kernel<CH, true, true, false, CL>,                                      \
        (const void*)cle_stepper_kernel<CH, false, false, true, CL>,                                     \
        (const void*)cle_stepper_kernel<CH, true, false, true, CL>,                                      \
        (const void*)cle_stepper_kernel<CH, false, true, true, CL>,                                      \
        (const void*)cle_stepper_kernel<CH, true, true, true, CL>
    const void* fns[16] = {SCRNA_STEPPER_VARIANTS(false), SCRNA_STEPPER_VARIANTS(true)};
#undef SCRNA_STEPPER_VARIANTS
    return fns[(inj ? 1 : 0) | (pop ? 2 : 0) | (tele ? 4 : 0) | (clamp ? 8 : 0)];
}
static const void* stepper_fn(int CH, bool inj, bool pop, bool tele, bool clamp) {
    switch (CH) {
        case 2: return stepper_fn<2>(inj, pop, tele, clamp);
        case 4: return stepper_fn<4>(inj, pop, tele, clamp);
        case 8: return stepper_fn<8>(inj, pop, tele, clamp);
        case 16: return stepper_fn<16>(inj, pop, tele, clamp);
        default: return nullptr;
    }
}

// A launch with no injected-noise pool (no network of it injects noise), no
// volume pool (no population), tele = false (no telegraph network) or clamp =
// false (no negative initial value) runs the kernel variant without that path.
hipError_t launch_stepper(int CH, int NW, int n_tiles, const uint32_t* blobs, const Tile* tiles,
                          float* state_pool, float* snap_pool, uint8_t* prom_pool, double* vol_pool,
                          const float* inj_pool, int s_begin, int s_end, int nbuf,
                          size_t lds_bytes, int* sched, int n_seg, long long wait_ticks, bool tele,
                          bool clamp, hipStream_t stream) {
    if (n_tiles <= 0) return hipSuccess;
    if (sched == nullptr) n_seg = 1;
    if (n_seg < 1 || (long long)n_tiles * n_seg > INT32_MAX) return hipErrorInvalidValue;
    const void* f = stepper_fn(CH, inj_pool != nullptr, vol_pool != nullptr, tele, clamp);
    if (!f) return hipErrorInvalidValue;
    void* args[] = {(void*)&blobs, (void*)&tiles, (void*)&state_pool, (void*)&snap_pool, (void*)&prom_pool,
                    (void*)&vol_pool, (void*)&inj_pool, (void*)&s_begin, (void*)&s_end, (void*)&nbuf,
                    (void*)&sched, (void*)&n_tiles, (void*)&n_seg, (void*)&wait_ticks};
    return hipLaunchKernel(f, dim3((unsigned)(n_tiles * n_seg)), dim3(64 * NW), args, lds_bytes, stream);
}

hipError_t launch_sched_reset(int* sched, int n_words, hipStream_t stream) {
    if (n_words <= 0) return hipSuccess;
    hipLaunchKernelGGL(sched_reset_kernel, dim3((unsigned)((n_words + 255) / 256)), dim3(256), 0, stream, sched,
                       n_words);
    return hipGetLastError();
}

int stepper_blocks_per_cu(int CH, int NW, size_t lds_bytes) {
    const void* f = stepper_fn(CH, false, false, true, false);
    if (!f) return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * NW, lds_bytes) != hipSuccess) return 0;
    return nb;
}

hipError_t set_stepper_lds_limit(size_t bytes) {
    // The kernel's own static LDS (schedule ticket and flag) shares the budget.
    bytes = bytes > kStepperStaticLds ? bytes - kStepperStaticLds : bytes;
    hipError_t e = hipSuccess;
    for (int CH : {2, 4, 8, 16})
        for (int v = 0; v < 16; ++v) {
            hipError_t r = hipFuncSetAttribute(stepper_fn(CH, v & 1, v & 2, v & 4, v & 8),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
            if (r != hipSuccess) e = r;
        }
    return e;
}

// The job of a capture or moment workgroup is blockIdx.y, whose range is
// at most 65535: plans with more jobs launch them in chunks.
constexpr int kMaxGridY = 65535;

template <typename T>
static hipError_t launch_capture(int n_jobs, int64_t max_elems, const CaptureJob* jobs, const T* src,
                                 int32_t* dst, hipStream_t stream) {
    if (n_jobs <= 0 || max_elems <= 0) return hipSuccess;
    const int64_t gx = (max_elems + 255) / 256;
    if (gx > INT32_MAX) return hipErrorInvalidConfiguration;
    for (int j0 = 0; j0 < n_jobs; j0 += kMaxGridY) {
        const dim3 grid((unsigned)gx, (unsigned)std::min(kMaxGridY, n_jobs - j0)), block(256);
        hipLaunchKernelGGL(capture_kernel<T>, grid, block, 0, stream, jobs + j0, src, dst);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_capture_f32(int n_jobs, int64_t max_elems, const CaptureJob* jobs, const float* src,
                              int32_t* dst, hipStream_t stream) {
    return launch_capture<float>(n_jobs, max_elems, jobs, src, dst, stream);
}

hipError_t launch_capture_f64(int n_jobs, int64_t max_elems, const CaptureJob* jobs, const double* src,
                              int32_t* dst, hipStream_t stream) {
    return launch_capture<double>(n_jobs, max_elems, jobs, src, dst, stream);
}

hipError_t launch_moments(int n_jobs, int max_G, const MomentJob* jobs, const float* pool, double* out,
                          hipStream_t stream) {
    if (n_jobs <= 0 || max_G <= 0) return hipSuccess;
    for (int j0 = 0; j0 < n_jobs; j0 += kMaxGridY) {
        const dim3 grid((unsigned)max_G, (unsigned)std::min(kMaxGridY, n_jobs - j0)), block(256);
        hipLaunchKernelGGL(moments_kernel, grid, block, 0, stream, jobs + j0, pool, out);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace scrna
