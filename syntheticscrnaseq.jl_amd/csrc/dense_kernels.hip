// Dense-regulation CLE stepper for large networks, additive regulation on
// BF16 MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulate).
//
// Regulation (cle.jl:79-86) is A_pos*act + A_neg*rep.  With rep = 1 - act it
// equals W*act + b, W = A_pos - A_neg (the signed A_eff of network.jl:206-219)
// and b = rowsum(A_neg): one signed contraction per step,
//     reg[cells x G] = act[cells x G] * W^T[G x G] + b.
// MODE 1 (BF16): act and W rounded to bf16.
// MODE 2 (BF16X3): act = a_hi + a_lo, W = W_hi + W_lo (bf16 each) and
//   reg = a_hi W_hi + a_lo W_hi + a_hi W_lo, recovering ~16 mantissa bits.
// MODE 3 (FP32): act and W in FP32 on v_mfma_f32_16x16x4_f32 (the FP32
//   matrix rate equals the FP32 vector rate on gfx950, and the matrix cores
//   run beside the VALU's elementwise work).  The act tile is FP32 in LDS;
//   K is permuted so that lane group q = l >> 4 supplies k = q*Gp/4 + t in
//   the MFMA numbered t, and a lane reads four consecutive t with one
//   16-byte LDS load.
// AND / OR gates stay in FP32 on the VALU from an FP32 copy of their
// sources' Hill values.
//
// Work mapping.  A workgroup owns 16 cells and NW = Gp/(16*NB) waves (Gp = G
// rounded up to 32); wave w owns NB blocks of 16 genes, block b being genes
// 16*(w*NB+b) + [0,16).  The contraction runs transposed, reg^T = W act^T:
// W is the MFMA's A operand and the act tile its B operand, so in the 16x16
// result lane l holds genes 16*(w*NB+b) + 4*(l>>4) + v, v = 0..3, of the
// cell l & 15, and those four genes' (m, p) stay in the lane's VGPRs for the
// whole horizon, in accumulator layout.  W for the wave's genes is
// register-resident (4 VGPRs per 32-wide K chunk, gene block and bf16 plane;
// a lane holds row l & 15 of the block, sources 8*(l>>4) + [0,8) of each
// chunk).  NB = 2 halves the waves per workgroup, which gives each wave the
// 256-VGPR budget of 2 waves per SIMD that BF16X3's two weight planes need.
// Each step every lane writes act of its four genes for its cell into an LDS
// tile laid out [cell][gene] (one 8-byte store per bf16 plane), the
// workgroup synchronises once (double-buffered tile), and each wave issues
// NB*Gp/32 MFMAs (x3 for BF16X3).  A lane's genes come in pairs (g, g+1),
// g even, which is exactly the gene pair of one normal block (cell, s, g>>1).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "blob.h"
#include "cle_kernels.h"
#include "device_rng.h"

namespace scrna {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float d_as_f(uint32_t w) { return __uint_as_float(w); }
__device__ __forceinline__ uint64_t d_u64(const uint32_t* B, int i) {
    return (uint64_t)B[i] | ((uint64_t)B[i + 1] << 32);
}
// Word (i & 3) of a Philox block without a dynamically indexed array.
__device__ __forceinline__ uint32_t pick4(const U4& u, int i) {
    return (i & 2) ? ((i & 1) ? u.w : u.z) : ((i & 1) ? u.y : u.x);
}

constexpr int kDenseMaxKC = 8;   // Gp <= 256

// CLAMP keeps max(m, 0), max(p, 0) and the protein variance's max(., 0) of
// the update, which only a network starting from negative counts
// (F_NEG_INIT) needs: as in the FP32 stepper, counts are never negative at
// the start of a later step.
template <int MODE, int NB, bool TELE, bool INJECT, bool CLAMP>
__global__ void __launch_bounds__(1024 / NB)
cle_dense_kernel(const uint32_t* __restrict__ blobs, const Tile* __restrict__ tiles,
                 const bf16x8* __restrict__ wpool, float* __restrict__ state_pool, float* __restrict__ snap_pool,
                 uint8_t* __restrict__ prom_pool, const float* __restrict__ inj_pool, int s_begin, int s_end) {
    extern __shared__ __attribute__((aligned(16))) unsigned char dlds[];
    const Tile tile = tiles[blockIdx.x];
    const uint32_t* __restrict__ B = blobs + tile.blob;
    const int G = (int)B[H_G];
    const int s_start = max(s_begin, tile.s0);
    const bool resume = s_start > 0;       // later chunks continue from the state pool
    const int S = (int)B[H_S];             // gate-source slots
    const int Gp = (int)B[H_GP];           // G rounded up to 32
    const int KC = Gp / 32;
    const uint32_t flags = B[H_FLAGS];
    const int n_steps = (int)B[H_NSTEPS];
    const int snap_every = (int)B[H_SNAP_EVERY];
    const int C = (int)B[H_CELLS];
    const uint32_t k0 = B[H_SEED_LO], k1 = B[H_SEED_HI];
    const uint32_t c3n = (B[H_NETID] << 4) | kNormal;
    const uint32_t c3s = (B[H_NETID] << 4) | kSwitch;
    const float dt = d_as_f(B[H_DT]);
    const float neg2ln2dt = d_as_f(B[H_NEG2LN2DT]);
    const float k_t = d_as_f(B[H_KT]);
    const float mu_m = d_as_f(B[H_MUM]);
    const float mu_p = d_as_f(B[H_MUP]);
    const float invK = d_as_f(B[H_INVK]);
    const float hill_n = d_as_f(B[H_HILLN]);
    const bool hill2 = (flags & F_HILL2) != 0;
    const bool tele = TELE;                 // specialised: telegraph and injected-noise paths cost registers
    const bool inject = INJECT;
    const size_t off_state = (size_t)d_u64(B, H_OUT_FINAL_LO);
    const size_t off_snap = (size_t)d_u64(B, H_OUT_SNAP_LO);
    const size_t off_prom = (size_t)d_u64(B, H_OUT_PROM_LO);
    const size_t off_inj = (size_t)d_u64(B, H_INJ_LO);
    const size_t off_w = (size_t)d_u64(B, H_DENSE_LO);   // in bf16x8 units
    const int NBLK = Gp / 16;              // gene blocks in the network

    const int l = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int cl = l & 15;                 // the lane's cell: row cl of the tile
    const int qd = l >> 4;                 // its gene quad: genes 4*qd .. 4*qd+3 of each block

    // LDS: act tiles [plane][buffer][16 rows][Gp + 16] bf16 (FP32: [buffer][16 rows][Gp + 4] f32), then
    // FP32 gate Hill values [buffer][2 S][16 cells].  In the BF16 tile, rows whose cell has bit 2 set
    // hold their 16-byte chunks pairwise swapped (column c at c ^ 8).  Measured on gfx950 with the
    // bank-conflict counter (SQ_LDS_BANK_CONFLICT), for these exact lane patterns: rows 8 dwords
    // (mod 16) apart make the B-fragment ds_read_b128 conflict-free (rows 4 dwords apart, Gp + 8,
    // conflict 2-way on every read, with or without a swap), and the swap brings the act write
    // (ds_write_b64, served 16 consecutive lanes at a time over 32 banks) from 4-way to 2-way, the
    // least for 16 lanes writing the same 8-byte half of their chunks.
    constexpr bool F32 = MODE == 3;
    const int rowstride = F32 ? Gp + 4 : Gp + 16;
    // BF16 tiles: the row's chunk swap, in bf16 columns.  BF16X3 leaves it out: at its 256-VGPR
    // budget the swap's address register spills (36 B per lane), and its two planes' writes cost
    // less than that (4-way on 2 writes per step against 16 conflict-free reads).
    const int swz = MODE == 2 ? 0 : (cl & 4) << 1;
    const int tile_elems = 16 * rowstride;
    const int KT = Gp / 4;                 // FP32 mode: MFMAs per gene block and step
    __bf16* act_t = (__bf16*)dlds;
    float* act_f = (float*)dlds;
    float* ghill = (float*)(dlds + (F32 ? (size_t)2 * tile_elems * sizeof(float)
                                        : (size_t)(MODE == 2 ? 4 : 2) * tile_elems * sizeof(__bf16)));

    const int32_t* __restrict__ gate_off = (const int32_t*)(B + B[H_OFF_GATEOFF]);
    const uint32_t* __restrict__ gates = B + B[H_OFF_GATES];
    const uint32_t* __restrict__ telew = B + B[H_OFF_TELE];

    // Per gene (b, v): weights, constants, state.  A gene's gate, when it has
    // exactly one and it has two sources (the form of every gate the prior
    // draws), lives in registers as a signed weight and the LDS offsets of its
    // sources' Hill rows (AND reads act rows, OR rep rows): an AND gate adds
    // w*act_a*act_b, an OR gate w*(1 - rep_a*rep_b) = w - w*rep_a*rep_b, whose
    // constant w joins the bias.  Other gate layouts take the loop over the
    // blob's records (gslow).
    bf16x8 whi[NB][F32 ? 1 : kDenseMaxKC], wlo[NB][F32 ? 1 : kDenseMaxKC];
    float wf[NB][F32 ? 4 * kDenseMaxKC * 2 : 1];      // FP32: W[16 blk + (l & 15)][(l >> 4) KT + t]
    int slot[NB][4];
    float gw[NB][4];                       // signed gate weight (0: no register gate)
    uint32_t gab[NB][4];                   // LDS float offsets of the two source rows, 16 bits each
    uint32_t gslow_m[NB];                  // bit v: gene v takes the record loop
    bool wave_gate[NB], wave_slow[NB];
    f32x4 bias[NB];                        // basal rate + rowsum(A_neg) (+ OR weights): the MFMA's start value
    uint32_t tflags[NB][4];
    uint32_t thr_on[NB][4], thr_off[NB][4];   // floor(threshold / 2^14), see cle_switch
    float m[NB][4], p[NB][4];
    uint32_t on = 0;                       // promoter bit 4*b + v

    const bool live = cl < tile.ncell;
    const int cell = tile.cell0 + (live ? cl : 0);
    auto gene = [&](int b, int v) { return 16 * (w * NB + b) + 4 * qd + v; };
    // The step-invariant first Philox round of the lane's normal blocks (cell, s, g>>1).
    PhiloxFixed pfx[NB][2];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int h = 0; h < 2; ++h)
            pfx[b][h] = philox_fixed((uint32_t)cell, (uint32_t)(gene(b, 2 * h) >> 1), c3n, k0, k1);

#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int blk = w * NB + b;
        if constexpr (F32) {
            const float* __restrict__ wp = (const float*)wpool + off_w * 4;
#pragma unroll
            for (int t = 0; t < 8 * kDenseMaxKC; ++t)
                if (t < KT) wf[b][t] = wp[((size_t)blk * KT + t) * 64 + l];
        } else {
#pragma unroll
            for (int kc = 0; kc < kDenseMaxKC; ++kc) {
                if (kc < KC) {
                    whi[b][kc] = wpool[off_w + ((size_t)blk * KC + kc) * 64 + l];
                    if (MODE == 2) wlo[b][kc] = wpool[off_w + ((size_t)NBLK * KC + (size_t)blk * KC + kc) * 64 + l];
                }
            }
        }
        bool any_gate = false;
        gslow_m[b] = 0u;
        U4 uinit = U4{0u, 0u, 0u, 0u};
        if (!resume && tele)   // the promoter-init block (cell, 0, g>>2) covers the lane's four genes
            uinit = philox4x32_10((uint32_t)cell, 0u, (uint32_t)(gene(b, 0) >> 2), (B[H_NETID] << 4) | kPromInit,
                                  k0, k1);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int g = gene(b, v);
            const bool gl = g < G;
            const int gc = gl ? g : 0;
            // Padding genes (g >= G) get beta = bias = 0, zero weight rows and
            // columns, an OFF promoter and zero state, so every step leaves them at
            // exactly 0 and the update needs no mask.
            float bv = gl ? d_as_f(B[B[H_OFF_BIAS] + gc]) + d_as_f(B[B[H_OFF_BETA] + gc]) : 0.0f;
            slot[b][v] = gl ? ((const int32_t*)(B + B[H_OFF_SLOT]))[g] : -1;
            tflags[b][v] = gl ? telew[g * 8 + T_FLAGS] : 0u;
            thr_on[b][v] = telew[gc * 8 + T_ON_HI] != 0u ? 1u << 18 : telew[gc * 8 + T_ON_LO] >> 14;
            thr_off[b][v] = telew[gc * 8 + T_OFF_HI] != 0u ? 1u << 18 : telew[gc * 8 + T_OFF_LO] >> 14;
            const int q0 = gl ? gate_off[g] : 0;
            const int q1 = gl ? gate_off[g + 1] : 0;
            gw[b][v] = 0.0f;
            gab[b][v] = 0u;
            if (q1 > q0) {
                const uint32_t hd = gates[q0];
                const int type = (int)(hd >> 24), ns = (int)(hd & 0xffffu);
                if (ns == 2 && q1 - q0 == 2 + ns) {
                    any_gate = true;
                    const float gwt = d_as_f(gates[q0 + 1]);
                    const bool is_or = type == (int)GATE_OR;
                    gw[b][v] = is_or ? -gwt : gwt;
                    if (is_or) bv += gwt;
                    gab[b][v] = (uint32_t)((2 * (int)gates[q0 + 2] + type) * 16) |
                                ((uint32_t)((2 * (int)gates[q0 + 3] + type) * 16) << 16);
                } else {
                    gslow_m[b] |= 1u << v;
                }
            }
            bias[b][v] = bv;
            m[b][v] = 0.0f;
            p[b][v] = 0.0f;
            if (gl) {
                if (resume) {
                    m[b][v] = state_pool[off_state + (size_t)g * C + cell];
                    p[b][v] = state_pool[off_state + (size_t)(G + g) * C + cell];
                    on |= (uint32_t)(prom_pool[off_prom + (size_t)g * C + cell] != 0) << (4 * b + v);
                } else {
                    const uint32_t* initw = B + B[H_OFF_INIT];
                    m[b][v] = d_as_f(initw[2 * g]);
                    p[b][v] = d_as_f(initw[2 * g + 1]);
                    bool o = true;
                    if (tele && (tflags[b][v] & 2u))   // tauleap_binomial.jl:64-76
                        o = (uint64_t)pick4(uinit, g) < d_u64(telew, g * 8 + T_INIT_LO);
                    on |= (uint32_t)o << (4 * b + v);
                }
            }
        }
        wave_gate[b] = __builtin_amdgcn_ballot_w64(any_gate) != 0;
        wave_slow[b] = __builtin_amdgcn_ballot_w64(gslow_m[b] != 0u) != 0;
    }

    const int s_stop = min(min(s_end, tile.s1), n_steps);
    // Next step after which a snapshot is due (a countdown, not a division per step).
    int next_snap = snap_every > 0 ? (s_start / snap_every + 1) * snap_every - 1 : INT32_MAX;
    for (int s = s_start; s < s_stop; ++s) {
        const int buf = s & 1;
        __bf16* T0 = act_t + (size_t)buf * tile_elems;
        __bf16* T1 = act_t + (size_t)(2 + buf) * tile_elems;    // low plane (BF16X3)
        float* TF = act_f + (size_t)buf * tile_elems;           // FP32 mode
        float* GH = ghill + (size_t)buf * 2 * S * 16;

#pragma unroll
        for (int b = 0; b < NB; ++b) {
            // Hill values of the lane's four genes for its cell -> LDS, one
            // 8-byte (BF16) or 16-byte (FP32) store per plane.  The n = 2 test
            // sits outside the gene loop so that the n = 2 path never evaluates
            // log/exp (as a select, both sides would run).
            float act[4], rep[4];
            auto hill_phase = [&](auto h2) {
                constexpr bool H2 = decltype(h2)::value;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    // Padding genes (g >= G) keep p = 0, so act = 0 and rep = 1.
                    const float x = p[b][v] * invK;
                    if constexpr (MODE != 3) {
                        // act = 1 - rep: one fma forms 1 + x^2, and x^n = inf gives
                        // rep = 0, act = 1 without a clamp.  Its absolute error
                        // (2^-24) is far below the bf16 rounding of act.
                        const float d = H2 ? fmaf(x, x, 1.0f)
                                           : 1.0f + __builtin_amdgcn_exp2f(hill_n * __builtin_amdgcn_logf(x));
                        rep[v] = __builtin_amdgcn_rcpf(d);
                        act[v] = 1.0f - rep[v];
                        continue;
                    }
                    float xn = H2 ? x * x : __builtin_amdgcn_exp2f(hill_n * __builtin_amdgcn_logf(x));
                    xn = fminf(xn, 1e30f);
                    rep[v] = __builtin_amdgcn_rcpf(1.0f + xn);
                    act[v] = xn * rep[v];
                }
            };
            if (hill2)
                hill_phase(std::true_type{});
            else
                hill_phase(std::false_type{});
            const int col = 16 * (w * NB + b) + 4 * qd;
            if constexpr (F32) {
                *(f32x4*)(TF + cl * rowstride + col) = f32x4{act[0], act[1], act[2], act[3]};
            } else {
                typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
                bf16x4 hi;
#pragma unroll
                for (int v = 0; v < 4; ++v) hi[v] = (__bf16)act[v];
                *(bf16x4*)(T0 + cl * rowstride + (col ^ swz)) = hi;
                if (MODE == 2) {
                    bf16x4 lo;
#pragma unroll
                    for (int v = 0; v < 4; ++v) lo[v] = (__bf16)(act[v] - (float)hi[v]);
                    *(bf16x4*)(T1 + cl * rowstride + (col ^ swz)) = lo;
                }
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if (slot[b][v] >= 0) {
                    GH[(2 * slot[b][v]) * 16 + cl] = act[v];
                    GH[(2 * slot[b][v] + 1) * 16 + cl] = rep[v];
                }
            }
        }
        __syncthreads();

        // The step's normals: the block (cell, s, g>>1) carries genes g (words
        // 0,1) and g+1 (words 2,3) for even g, two blocks per gene block.
        // With one gene block per wave they are drawn here, between the MFMA
        // issue and the first use of its result; with two (BF16X3's register
        // budget) each pair is drawn just before its update.
        constexpr bool EARLY = NB == 1 && !INJECT;
        float nr2[NB][4], ncs[NB][4], nsn[NB][4];
        uint32_t nwa[NB][4], nwb[NB][4];
        auto draw = [&](const int b, const int h) {
            // Round keys: recomputed on the scalar unit at each call in the
            // one-block forms, kept in SGPRs by BF16X3 (whose 2-wave budget has
            // the scalar registers; 95 -> 91 ms, while BF16 measured 61 -> 65 ms).
            const U4 u = philox4x32_10_fixed<MODE != 2>(pfx[b][h], (uint32_t)s, k0, k1);
            normal_pair_r2(u.x, u.y, neg2ln2dt, nr2[b][2 * h], ncs[b][2 * h], nsn[b][2 * h]);
            normal_pair_r2(u.z, u.w, neg2ln2dt, nr2[b][2 * h + 1], ncs[b][2 * h + 1], nsn[b][2 * h + 1]);
            nwa[b][2 * h] = u.x;
            nwb[b][2 * h] = u.y;
            nwa[b][2 * h + 1] = u.z;
            nwb[b][2 * h + 1] = u.w;
        };
        auto draw_early = [&]() {
            if constexpr (EARLY) {
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    draw(b, 0);
                    draw(b, 1);
                }
            }
        };

        // reg^T = W * act^T on the matrix cores: W (register-resident) is the
        // A operand, the act tile the B operand, and the bias the start value,
        // so lane l receives genes 4*(l >> 4) + v of its block for cell l & 15.
        // The BF16 forms accumulate one chain per gene block: the normals drawn
        // between the MFMA issue and the result's first use (and, with two
        // blocks, the other block's chain) cover the chain's latency, and no
        // partial sums are added up afterwards.  The FP32 form, whose 16x16x4
        // MFMAs are 8 times as many, keeps two chains (even and odd groups of
        // four), summed after the normals are drawn.
        f32x4 acc[NB], acc2[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            acc[b] = bias[b];
            acc2[b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        }
        if constexpr (F32) {
            const float* arow = TF + cl * rowstride + qd * KT;
#pragma unroll
            for (int t = 0; t < 8 * kDenseMaxKC; t += 4) {
                if (t < KT) {
                    const f32x4 a4 = *(const f32x4*)(arow + t);
#pragma unroll
                    for (int b = 0; b < NB; ++b) {
                        acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(wf[b][t], a4[0], acc[b], 0, 0, 0);
                        acc2[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(wf[b][t + 1], a4[1], acc2[b], 0, 0, 0);
                        acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(wf[b][t + 2], a4[2], acc[b], 0, 0, 0);
                        acc2[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(wf[b][t + 3], a4[3], acc2[b], 0, 0, 0);
                    }
                }
            }
            draw_early();
        }
        if constexpr (!F32) {
            // With the chunk count a compile-time constant, the act reads of all
            // chunks issue ahead of the MFMAs.
            const __bf16* arow0 = T0 + cl * rowstride + ((8 * qd) ^ swz);   // + 32 kc keeps bit 3
            const __bf16* arow1 = T1 + cl * rowstride + ((8 * qd) ^ swz);
            auto contract = [&](auto kcn) {
                constexpr int KCN = decltype(kcn)::value;
#pragma unroll
                for (int kc = 0; kc < KCN; ++kc) {
                    const bf16x8 a = *(const bf16x8*)(arow0 + 32 * kc);
                    bf16x8 alo;
                    if (MODE == 2) alo = *(const bf16x8*)(arow1 + 32 * kc);
#pragma unroll
                    for (int b = 0; b < NB; ++b) {
                        f32x4& c = acc[b];
                        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(whi[b][kc], a, c, 0, 0, 0);
                        if (MODE == 2) {
                            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(whi[b][kc], alo, c, 0, 0, 0);
                            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wlo[b][kc], a, c, 0, 0, 0);
                        }
                    }
                }
                draw_early();   // in the MFMAs' basic block, so that the scheduler can overlap them
            };
            switch (KC) {
                case 8: contract(std::integral_constant<int, 8>{}); break;
                case 7: contract(std::integral_constant<int, 7>{}); break;
                case 6: contract(std::integral_constant<int, 6>{}); break;
                case 5: contract(std::integral_constant<int, 5>{}); break;
                case 4: contract(std::integral_constant<int, 4>{}); break;
                case 3: contract(std::integral_constant<int, 3>{}); break;
                case 2: contract(std::integral_constant<int, 2>{}); break;
                default: contract(std::integral_constant<int, 1>{}); break;
            }
        }

        if (F32) {
#pragma unroll
            for (int b = 0; b < NB; ++b) acc[b] += acc2[b];
        }

        // Elementwise update of gene (b, v) (cle.jl:108-120).
        // Noise enters as xi_m = c*sqrt(r2), xi_p = sn*sqrt(r2) (Box-Muller radius
        // squared, scaled by dt), so xi*sqrt(var) costs one square root, sqrt(r2*var).
        // wa, wb: the radius and angle words of the gene's normal block,
        // whose low 9 bits are its switching bits.
        auto update = [&](const int b, const int v, float r2, float c, float sn, uint32_t wa, uint32_t wb) {
            float reg = acc[b][v];
            // AND / OR (cle.jl:89-106).  Lanes without a register gate read a
            // valid Hill row with weight 0.
            if (wave_gate[b]) {
                const float prod = GH[(gab[b][v] & 0xffffu) + cl] * GH[(gab[b][v] >> 16) + cl];
                reg = fmaf(gw[b][v], prod, reg);
            }
            if (wave_slow[b] && ((gslow_m[b] >> v) & 1u)) {
                const int g = gene(b, v);
                const int q1 = gate_off[g + 1];
                for (int q = gate_off[g]; q < q1;) {
                    const uint32_t hd = gates[q];
                    const uint32_t type = hd >> 24;
                    const int ns = (int)(hd & 0xffffu);
                    float prod = 1.0f;
                    for (int i = 0; i < ns; ++i) prod *= GH[(2 * (int)gates[q + 2 + i] + (int)type) * 16 + cl];
                    reg = fmaf(d_as_f(gates[q + 1]), type == GATE_AND ? prod : 1.0f - prod, reg);
                    q += 2 + ns;
                }
            }
            // Telegraph switching first (tauleap_binomial.jl:110-121), as in
            // cle_switch: X = q << 14 | e flips iff below the threshold of the
            // current state; e (the top 14 bits of the gene's word of the switch
            // block) only matters when q == T.
            if (tele) {
                const int bit = 4 * b + v;
                const bool cur = (on >> bit) & 1u;
                const bool sw = tflags[b][v] & 1u;
                const uint32_t q = ((wa & 0x1FFu) << 9) | (wb & 0x1FFu);
                const uint32_t T = cur ? thr_off[b][v] : thr_on[b][v];
                bool flip = sw && q < T;
                const bool amb = sw && q == T;
                if (__builtin_amdgcn_ballot_w64(amb) != 0) {
                    const int g = gene(b, v);
                    const U4 u = philox4x32_10_uk((uint32_t)cell, (uint32_t)s, (uint32_t)(g >> 2), c3s, k0, k1);
                    const uint32_t e = pick4(u, g) >> 18;
                    const int gc = g < G ? g : 0;
                    const uint64_t thr = d_u64(telew, gc * 8 + (cur ? T_OFF_LO : T_ON_LO));
                    if (amb) flip = ((((uint64_t)q << 14) | e) < thr);
                }
                on ^= (uint32_t)flip << bit;
            }
            // Without telegraph genes every live gene is ON, and padding genes have
            // zero drive (reg = 0), so the gate on the promoter bit is dropped.
            const bool o = !tele || ((on >> (4 * b + v)) & 1u);
            const float tx = o ? reg : 0.0f;
            const float mm = m[b][v];
            const float var_m = fmaxf(fmaf(mu_m, CLAMP ? fmaxf(mm, 0.0f) : mm, tx), 0.0f);
            float mn = fmaf(fmaf(-mu_m, mm, tx), dt, mm);
            mn = fmaxf(fmaf(c, __builtin_amdgcn_sqrtf(r2 * var_m), mn), 0.0f);
            const float pp = p[b][v];
            float var_p, pn;
            if (CLAMP) {
                var_p = fmaxf(fmaf(k_t, mn, mu_p * fmaxf(pp, 0.0f)), 0.0f);
                pn = fmaf(fmaf(k_t, mn, -mu_p * pp), dt, pp);
            } else {
                const float dp = mu_p * pp;
                var_p = fmaf(k_t, mn, dp);
                pn = fmaf(fmaf(k_t, mn, -dp), dt, pp);
            }
            pn = fmaxf(fmaf(sn, __builtin_amdgcn_sqrtf(r2 * var_p), pn), 0.0f);
            m[b][v] = mn;
            p[b][v] = pn;
        };

#pragma unroll
        for (int b = 0; b < NB; ++b) {
            if (inject) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int g = gene(b, v);
                    const size_t base = ((size_t)s * G + (g < G ? g : 0)) * 2;
                    // the switching bits still come from the normal block
                    uint32_t wa = 0u, wb = 0u;
                    if (tele) {
                        const U4 u = philox4x32_10_uk((uint32_t)cell, (uint32_t)s, (uint32_t)(g >> 1), c3n, k0, k1);
                        wa = (g & 1) ? u.z : u.x;
                        wb = (g & 1) ? u.w : u.y;
                    }
                    update(b, v, dt, inj_pool[off_inj + base * C + cell], inj_pool[off_inj + (base + 1) * C + cell],
                           wa, wb);
                }
            } else {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if constexpr (!EARLY) draw(b, h);
#pragma unroll
                    for (int v = 2 * h; v < 2 * h + 2; ++v)
                        update(b, v, nr2[b][v], ncs[b][v], nsn[b][v], nwa[b][v], nwb[b][v]);
                }
            }
        }

        if (s == next_snap) {
            next_snap += snap_every;
            const int snap = (s + 1) / snap_every - 1;
            if (live) {
#pragma unroll
                for (int b = 0; b < NB; ++b) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int g = gene(b, v);
                        if (g < G) {
                            snap_pool[off_snap + ((size_t)(2 * snap) * G + g) * C + cell] = m[b][v];
                            snap_pool[off_snap + ((size_t)(2 * snap + 1) * G + g) * C + cell] = p[b][v];
                        }
                    }
                }
            }
        }
    }

    if (live) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int g = gene(b, v);
                if (g < G) {
                    state_pool[off_state + (size_t)g * C + cell] = m[b][v];
                    state_pool[off_state + (size_t)(G + g) * C + cell] = p[b][v];
                    prom_pool[off_prom + (size_t)g * C + cell] = (uint8_t)((on >> (4 * b + v)) & 1u);
                }
            }
        }
    }
}

// Gene blocks per wave for a precision mode (BF16X3 needs the larger budget).
int dense_nb(int mode) { return mode == 2 ? 2 : 1; }

size_t dense_lds_bytes(int Gp, int S, int mode) {
    const size_t tiles = mode == 3 ? (size_t)2 * 16 * (Gp + 4) * sizeof(float)
                                   : (size_t)(mode == 2 ? 4 : 2) * 16 * (Gp + 16) * 2;
    return tiles + (size_t)2 * 2 * S * 16 * sizeof(float);
}

template <int MODE, int NB>
static const void* dense_fn(bool tele, bool inject, bool clamp) {
#define SCRNA_DENSE_VARIANTS(CL)                                                                         \
    (const void*)cle_dense_kernel<MODE, NB, false, false, CL>,                                          \
        (const void*)cle_dense_kernel<MODE, NB, false, true, CL>,                                       \
        (const void*)cle_dense_kernel<MODE, NB, true, false, CL>,                                       \
        (const void*)cle_dense_kernel<MODE, NB, true, true, CL>
    const void* fns[8] = {SCRNA_DENSE_VARIANTS(false), SCRNA_DENSE_VARIANTS(true)};
#undef SCRNA_DENSE_VARIANTS
    return fns[(inject ? 1 : 0) | (tele ? 2 : 0) | (clamp ? 4 : 0)];
}

template <int MODE, int NB>
static hipError_t dense_attr(size_t bytes) {
    hipError_t e = hipSuccess;
    for (int v = 0; v < 8; ++v) {
        hipError_t r = hipFuncSetAttribute(dense_fn<MODE, NB>(v & 2, v & 1, v & 4),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (r != hipSuccess) e = r;
    }
    return e;
}

hipError_t set_dense_lds_limit(size_t bytes) {
    hipError_t e1 = dense_attr<1, 1>(bytes);
    hipError_t e2 = dense_attr<2, 2>(bytes);
    hipError_t e3 = dense_attr<3, 1>(bytes);
    return e1 != hipSuccess ? e1 : e2 != hipSuccess ? e2 : e3;
}

hipError_t launch_dense(int mode, bool tele, bool inject, bool clamp, int NW, int n_tiles, const uint32_t* blobs,
                        const Tile* tiles, const void* wpool, float* state_pool, float* snap_pool, uint8_t* prom_pool,
                        const float* inj_pool, int s_begin, int s_end, size_t lds_bytes,
                        hipStream_t stream) {
    if (n_tiles <= 0) return hipSuccess;
    const void* f = mode == 1 ? dense_fn<1, 1>(tele, inject, clamp)
                  : mode == 2 ? dense_fn<2, 2>(tele, inject, clamp)
                  : mode == 3 ? dense_fn<3, 1>(tele, inject, clamp)
                              : nullptr;
    if (!f) return hipErrorInvalidValue;
    const bf16x8* w = (const bf16x8*)wpool;
    void* args[] = {(void*)&blobs, (void*)&tiles, (void*)&w, (void*)&state_pool, (void*)&snap_pool,
                    (void*)&prom_pool, (void*)&inj_pool, (void*)&s_begin, (void*)&s_end};
    return hipLaunchKernel(f, dim3(n_tiles), dim3(64 * NW), args, lds_bytes, stream);
}

}  // namespace scrna
