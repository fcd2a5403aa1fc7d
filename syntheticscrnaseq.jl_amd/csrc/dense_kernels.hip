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
// rounded up to 32); wave w owns NB blocks of 16 output genes, block b being
// genes 16*(w*NB+b) + [0,16).  In a 16x16 MFMA result lane l holds gene
// 16*(w*NB+b) + (l & 15) for the 4 cells 4*(l>>4) + v, v = 0..3, so those
// cells' (m, p) stay in the lane's VGPRs for the whole horizon, in
// accumulator layout.  W^T for the wave's output genes is register-resident
// as the B operand (4 VGPRs per 32-wide K chunk, gene block and bf16 plane).
// NB = 2 halves the waves per workgroup, which gives each wave the 256-VGPR
// budget of 2 waves per SIMD that BF16X3's two weight planes need.  Each step
// every lane writes act of its genes and cells into an LDS tile laid out
// [cell][gene] -- the A operand's layout -- the workgroup synchronises once
// (double-buffered tile), and each wave issues NB*Gp/32 MFMAs (x3 for BF16X3).
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

// Cross-lane move inside a quad of lanes as a DPP operand of a v_mov (no
// LDS round trip as with ds_bpermute): lane j receives lane j ^ 1's value.
__device__ __forceinline__ uint32_t dpp_xor1(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
}

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
    const int rg = l >> 4;                 // row group: cells 4*rg .. 4*rg+3 of the tile
    const int par = l & 1;

    // LDS: act tiles [plane][buffer][16 rows][Gp + 8] bf16 (FP32: [buffer][16 rows][Gp + 4] f32, rows
    // 16 bytes apart in bank order so that the 16-byte reads of 16 lanes hit distinct banks), then FP32
    // gate Hill values.
    constexpr bool F32 = MODE == 3;
    const int rowstride = F32 ? Gp + 4 : Gp + 8;
    const int tile_elems = 16 * rowstride;
    const int KT = Gp / 4;                 // FP32 mode: MFMAs per gene block and step
    __bf16* act_t = (__bf16*)dlds;
    float* act_f = (float*)dlds;
    float* ghill = (float*)(dlds + (F32 ? (size_t)2 * tile_elems * sizeof(float)
                                        : (size_t)(MODE == 2 ? 4 : 2) * tile_elems * sizeof(__bf16)));

    const int32_t* __restrict__ gate_off = (const int32_t*)(B + B[H_OFF_GATEOFF]);
    const uint32_t* __restrict__ gates = B + B[H_OFF_GATES];
    const uint32_t* __restrict__ telew = B + B[H_OFF_TELE];

    // Per gene block: weights, constants, state.  A gene's gate, when it has
    // exactly one and it has two sources (the form of every gate the prior
    // draws), lives in registers: strength*beta, OR flag and the offsets of
    // its sources' Hill rows (AND reads act rows, OR rep rows).  Other gate
    // layouts take the loop over the blob's records (gslow).
    bf16x8 whi[NB][F32 ? 1 : kDenseMaxKC], wlo[NB][F32 ? 1 : kDenseMaxKC];
    float wf[NB][F32 ? 4 * kDenseMaxKC * 2 : 1];      // FP32: W[16 blk + (l & 15)][(l >> 4) KT + t]
    int g[NB], slot[NB], q0[NB], q1[NB];
    bool gl[NB];
    float gw[NB];
    int ga[NB], gb[NB];
    bool gor[NB], gslow[NB], wave_gate[NB], wave_slow[NB];
    float bias[NB];                        // basal rate + rowsum(A_neg) bias: the drive's constant part
    uint32_t tflags[NB];
    uint32_t thr_on[NB], thr_off[NB];     // floor(threshold / 2^14), see cle_switch
    float m[NB][4], p[NB][4];
    uint32_t on = 0;                       // promoter bit 4*b + v

    auto row = [&](int v) { return 4 * rg + v; };
    auto live = [&](int v) { return row(v) < tile.ncell; };
    auto cell = [&](int v) { return tile.cell0 + (live(v) ? row(v) : 0); };

#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int blk = w * NB + b;
        g[b] = blk * 16 + (l & 15);
        gl[b] = g[b] < G;
        const int gc = gl[b] ? g[b] : 0;
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
        // Padding genes (g >= G) get beta = bias = 0, zero weight rows and
        // columns, an OFF promoter and zero state, so every step leaves them at
        // exactly 0 and the update needs no mask.
        bias[b] = gl[b] ? d_as_f(B[B[H_OFF_BIAS] + gc]) + d_as_f(B[B[H_OFF_BETA] + gc]) : 0.0f;
        slot[b] = gl[b] ? ((const int32_t*)(B + B[H_OFF_SLOT]))[g[b]] : -1;
        tflags[b] = gl[b] ? telew[g[b] * 8 + T_FLAGS] : 0u;
        thr_on[b] = telew[gc * 8 + T_ON_HI] != 0u ? 1u << 18 : telew[gc * 8 + T_ON_LO] >> 14;
        thr_off[b] = telew[gc * 8 + T_OFF_HI] != 0u ? 1u << 18 : telew[gc * 8 + T_OFF_LO] >> 14;
        q0[b] = gl[b] ? gate_off[g[b]] : 0;
        q1[b] = gl[b] ? gate_off[g[b] + 1] : 0;
        gw[b] = 0.0f;
        ga[b] = gb[b] = 0;
        gor[b] = gslow[b] = false;
        bool greg = false;
        if (q1[b] > q0[b]) {
            const uint32_t hd = gates[q0[b]];
            const int type = (int)(hd >> 24), ns = (int)(hd & 0xffffu);
            if (ns == 2 && q1[b] - q0[b] == 2 + ns) {
                greg = true;
                gw[b] = d_as_f(gates[q0[b] + 1]);
                gor[b] = type == (int)GATE_OR;
                ga[b] = (2 * (int)gates[q0[b] + 2] + type) * 16;
                gb[b] = (2 * (int)gates[q0[b] + 3] + type) * 16;
            } else {
                gslow[b] = true;
            }
        }
        wave_gate[b] = __builtin_amdgcn_ballot_w64(greg) != 0;
        wave_slow[b] = __builtin_amdgcn_ballot_w64(gslow[b]) != 0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            m[b][v] = 0.0f;
            p[b][v] = 0.0f;
            if (gl[b]) {
                if (resume) {
                    m[b][v] = state_pool[off_state + (size_t)g[b] * C + cell(v)];
                    p[b][v] = state_pool[off_state + (size_t)(G + g[b]) * C + cell(v)];
                    on |= (uint32_t)(prom_pool[off_prom + (size_t)g[b] * C + cell(v)] != 0) << (4 * b + v);
                } else {
                    const uint32_t* initw = B + B[H_OFF_INIT];
                    m[b][v] = d_as_f(initw[2 * g[b]]);
                    p[b][v] = d_as_f(initw[2 * g[b] + 1]);
                    bool o = true;
                    if (tele && (tflags[b] & 2u)) {
                        const U4 u = philox4x32_10((uint32_t)cell(v), 0u, (uint32_t)(g[b] >> 2),
                                                   (B[H_NETID] << 4) | kPromInit, k0, k1);
                        o = (uint64_t)pick4(u, g[b]) < d_u64(telew, g[b] * 8 + T_INIT_LO);
                    }
                    on |= (uint32_t)o << (4 * b + v);
                }
            }
        }
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
            // Hill values of the lane's gene for its 4 cells -> LDS.  The n = 2
            // test sits outside the cell loop so that the n = 2 path never
            // evaluates log/exp (as a select, both sides would run).
            auto hill_phase = [&](auto h2) {
                constexpr bool H2 = decltype(h2)::value;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    // Padding genes (g >= G) keep p = 0, so act = 0 and rep = 1.
                    const float x = p[b][v] * invK;
                    float xn = H2 ? x * x : __builtin_amdgcn_exp2f(hill_n * __builtin_amdgcn_logf(x));
                    xn = fminf(xn, 1e30f);
                    const float rep = __builtin_amdgcn_rcpf(1.0f + xn);
                    const float act = xn * rep;
                    if constexpr (F32) {
                        TF[row(v) * rowstride + g[b]] = act;
                    } else {
                        const __bf16 hi = (__bf16)act;
                        T0[row(v) * rowstride + g[b]] = hi;
                        if (MODE == 2) T1[row(v) * rowstride + g[b]] = (__bf16)(act - (float)hi);
                    }
                    if (slot[b] >= 0) {
                        GH[(2 * slot[b]) * 16 + row(v)] = act;
                        GH[(2 * slot[b] + 1) * 16 + row(v)] = rep;
                    }
                }
            };
            if (hill2)
                hill_phase(std::true_type{});
            else
                hill_phase(std::false_type{});
        }
        __syncthreads();

        // reg = act * W^T on the matrix cores.
        f32x4 acc[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        if constexpr (F32) {
            // Two accumulation chains per block, summed after the loop.
            f32x4 acc2[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) acc2[b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            const float* arow = TF + (l & 15) * rowstride + rg * KT;
#pragma unroll
            for (int t = 0; t < 8 * kDenseMaxKC; t += 4) {
                if (t < KT) {
                    const f32x4 a4 = *(const f32x4*)(arow + t);
#pragma unroll
                    for (int b = 0; b < NB; ++b) {
                        acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[0], wf[b][t], acc[b], 0, 0, 0);
                        acc2[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[1], wf[b][t + 1], acc2[b], 0, 0, 0);
                        acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[2], wf[b][t + 2], acc[b], 0, 0, 0);
                        acc2[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[3], wf[b][t + 3], acc2[b], 0, 0, 0);
                    }
                }
            }
#pragma unroll
            for (int b = 0; b < NB; ++b) acc[b] += acc2[b];
        }
        if constexpr (!F32) {
            // With the chunk count a compile-time constant, the A reads of all
            // chunks issue ahead of the MFMAs, and even and odd chunks
            // accumulate in two independent chains.
            const __bf16* arow0 = T0 + (l & 15) * rowstride + 8 * rg;
            const __bf16* arow1 = T1 + (l & 15) * rowstride + 8 * rg;
            auto contract = [&](auto kcn) {
                constexpr int KCN = decltype(kcn)::value;
                f32x4 acc2[NB];
#pragma unroll
                for (int b = 0; b < NB; ++b) acc2[b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int kc = 0; kc < KCN; ++kc) {
                    const bf16x8 a = *(const bf16x8*)(arow0 + 32 * kc);
                    bf16x8 alo;
                    if (MODE == 2) alo = *(const bf16x8*)(arow1 + 32 * kc);
#pragma unroll
                    for (int b = 0; b < NB; ++b) {
                        f32x4& c = (kc & 1) ? acc2[b] : acc[b];
                        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, whi[b][kc], c, 0, 0, 0);
                        if (MODE == 2) {
                            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, whi[b][kc], c, 0, 0, 0);
                            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, wlo[b][kc], c, 0, 0, 0);
                        }
                    }
                }
#pragma unroll
                for (int b = 0; b < NB; ++b) acc[b] += acc2[b];
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

        // Elementwise update of element (b, v) (cle.jl:108-120).
        // Noise enters as xi_m = c*sqrt(r2), xi_p = sn*sqrt(r2) (Box-Muller radius
        // squared, scaled by dt), so xi*sqrt(var) costs one square root, sqrt(r2*var).
        // wa, wb: the radius and angle words of element (b, v)'s normal block,
        // whose low 9 bits are its switching bits.
        auto update = [&](const int b, const int v, float r2, float c, float sn, uint32_t wa, uint32_t wb) {
            float reg = acc[b][v] + bias[b];
            // AND / OR (cle.jl:89-106).  Lanes without a register gate read a
            // valid Hill row with weight 0.
            if (wave_gate[b]) {
                const float prod = GH[ga[b] + row(v)] * GH[gb[b] + row(v)];
                reg = fmaf(gw[b], gor[b] ? 1.0f - prod : prod, reg);
            }
            if (wave_slow[b] && gslow[b]) {
                for (int q = q0[b]; q < q1[b];) {
                    const uint32_t hd = gates[q];
                    const uint32_t type = hd >> 24;
                    const int ns = (int)(hd & 0xffffu);
                    float prod = 1.0f;
                    for (int i = 0; i < ns; ++i) prod *= GH[(2 * (int)gates[q + 2 + i] + (int)type) * 16 + row(v)];
                    reg = fmaf(d_as_f(gates[q + 1]), type == GATE_AND ? prod : 1.0f - prod, reg);
                    q += 2 + ns;
                }
            }
            // Telegraph switching first (tauleap_binomial.jl:110-121), as in
            // cle_switch with one gene per lane: X = q << 14 | e flips iff below
            // the threshold of the current state; e (the top 14 bits of the
            // gene's word of the switch block) only matters when q == T.
            if (tele) {
                const int bit = 4 * b + v;
                const bool cur = (on >> bit) & 1u;
                const bool sw = tflags[b] & 1u;
                const uint32_t q = ((wa & 0x1FFu) << 9) | (wb & 0x1FFu);
                const uint32_t T = cur ? thr_off[b] : thr_on[b];
                bool flip = sw && q < T;
                const bool amb = sw && q == T;
                if (__builtin_amdgcn_ballot_w64(amb) != 0) {
                    const U4 u = philox4x32_10_uk((uint32_t)cell(v), (uint32_t)s, (uint32_t)(g[b] >> 2), c3s, k0,
                                                  k1);
                    const uint32_t e = pick4(u, g[b]) >> 18;
                    const int gc = gl[b] ? g[b] : 0;
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
                const size_t base = ((size_t)s * G + (gl[b] ? g[b] : 0)) * 2;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    // the switching bits still come from the normal block
                    uint32_t wa = 0u, wb = 0u;
                    if (tele) {
                        const U4 u = philox4x32_10_uk((uint32_t)cell(v), (uint32_t)s, (uint32_t)(g[b] >> 1), c3n, k0,
                                                      k1);
                        wa = (g[b] & 1) ? u.z : u.x;
                        wb = (g[b] & 1) ? u.w : u.y;
                    }
                    update(b, v, dt, inj_pool[off_inj + base * C + cell(v)], inj_pool[off_inj + (base + 1) * C + cell(v)],
                           wa, wb);
                }
            } else {
                // The normal block (cell, s, g>>1) carries genes g (words 0,1) and
                // g^1 (words 2,3), i.e. lanes l and l^1, which own the same cells.
                // The even lane draws element t's block, the odd lane element
                // t+2's, and each hands its partner the partner's half.
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    const int cm = par ? cell(t + 2) : cell(t);
                    const U4 u = philox4x32_10_uk((uint32_t)cm, (uint32_t)s, (uint32_t)(g[b] >> 1), c3n, k0, k1);
                    const uint32_t mine_a = par ? u.z : u.x, mine_b = par ? u.w : u.y;
                    const uint32_t got_a = dpp_xor1(par ? u.x : u.z);
                    const uint32_t got_b = dpp_xor1(par ? u.y : u.w);
                    float r2, c, sn;
                    const uint32_t ta = par ? got_a : mine_a, tb = par ? got_b : mine_b;
                    normal_pair_r2(ta, tb, neg2ln2dt, r2, c, sn);
                    update(b, t, r2, c, sn, ta, tb);
                    const uint32_t ua = par ? mine_a : got_a, ub = par ? mine_b : got_b;
                    normal_pair_r2(ua, ub, neg2ln2dt, r2, c, sn);
                    update(b, t + 2, r2, c, sn, ua, ub);
                }
            }
        }

        if (s == next_snap) {
            next_snap += snap_every;
            const int snap = (s + 1) / snap_every - 1;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                if (!gl[b]) continue;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    if (live(v)) {
                        snap_pool[off_snap + ((size_t)(2 * snap) * G + g[b]) * C + cell(v)] = m[b][v];
                        snap_pool[off_snap + ((size_t)(2 * snap + 1) * G + g[b]) * C + cell(v)] = p[b][v];
                    }
                }
            }
        }
    }

#pragma unroll
    for (int b = 0; b < NB; ++b) {
        if (!gl[b]) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            if (live(v)) {
                state_pool[off_state + (size_t)g[b] * C + cell(v)] = m[b][v];
                state_pool[off_state + (size_t)(G + g[b]) * C + cell(v)] = p[b][v];
                prom_pool[off_prom + (size_t)g[b] * C + cell(v)] = (uint8_t)((on >> (4 * b + v)) & 1u);
            }
        }
    }
}

// Gene blocks per wave for a precision mode (BF16X3 needs the larger budget).
int dense_nb(int mode) { return mode == 2 ? 2 : 1; }

size_t dense_lds_bytes(int Gp, int S, int mode) {
    const size_t tiles = mode == 3 ? (size_t)2 * 16 * (Gp + 4) * sizeof(float)
                                   : (size_t)(mode == 2 ? 4 : 2) * 16 * (Gp + 8) * 2;
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
