// Counter-based random numbers for the gfx950 CLE kernels.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11).  Key = 64-bit run seed; counter words:
//   c0 = cell index inside its network, c1 = step (or gene / 0),
//   c2 = sub-index (gene pair, gene quad, sampler block),
//   c3 = (net_id << 4) | purpose.
// Purposes are fixed below; the CPU oracle uses the same assignments, so it
// can replay any cell's exact uniform stream.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace scrna {

enum Purpose : uint32_t {
    kNormal = 0,       // (cell, step, gene>>1): xi_m, xi_p for genes 2q and 2q+1 (top 23 bits of
                       // each word); the low 9 bits of a gene's two words are its CLE switching bits
    kSwitch = 1,       // (cell, step, gene>>2): one u32 per gene (tau-leap switching; the CLE
                       // steppers take 14 more switching bits from it when 18 do not decide)
    kPromInit = 2,     // (cell, 0, gene>>2)
    kCaptureEff = 3,   // (cell, 0, 0)
    kCaptureBin = 4,   // (cell, gene, block)
    kPartition = 5,    // (mother, step, ((2*gene+species)<<16) + block)
    kReplace = 6,      // (mother, step, 0)
    kVInit = 7,        // (cell, 0, 0)
    kTauLeap = 8,      // (cell, step, (gene << 16) + block): the tau-leap draws of one gene
    kGrowth = 9        // (cell, step, 0): the volume growth noise of one cell and step
};

struct U4 {
    uint32_t x, y, z, w;
};

// Three-input xor in one gfx950 VALU instruction (v_bitop3_b32, truth table
// 0x96); the compiler does not form it from two xors by itself.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// The kernels call it with c0 per lane and c1..c3 wave-uniform.  Then the
// first rounds' xors are scalar (c1 ^ k0 and c3 ^ k1 are uniform until the
// per-lane word has reached them), and a VALU bitop3 there would pull uniform
// values into VGPRs; from round 2 (n0) and round 3 (n2) on every operand but
// the key is per lane and one bitop3 replaces two v_xor_b32.
__device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                             uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        const uint32_t n0 = r >= 2 ? xor3((uint32_t)(p1 >> 32), c1, k0) : (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = r >= 3 ? xor3((uint32_t)(p0 >> 32), c3, k1) : (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return U4{c0, c1, c2, c3};
}

// The same generator for a key that is wave-uniform (the run seed).  The empty
// asm makes the key opaque at each call, so the ten round keys are recomputed
// on the scalar unit inside the call instead of being hoisted out of a long
// loop, where 20 live SGPRs push other uniforms into VGPR lanes and every
// reload costs a VALU v_readlane.
__device__ __forceinline__ U4 philox4x32_10_uk(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                uint32_t k0, uint32_t k1) {
    asm volatile("" : "+s"(k0), "+s"(k1));
    return philox4x32_10(c0, c1, c2, c3, k0, k1);
}

// A counter block whose c0, c2 and c3 stay fixed while c1 (the step) varies.
// Only the first round's c0 word depends on c1, so, writing p0 = M0*c0 and
// p1 = M1*c2 for each round's products, the following are step-invariant:
// round 0 entirely; in round 1 the product M1*c2 (c2 is round 0's
// hi(M0*c0) ^ c3 ^ k), hence the new c0 and c1; in round 2 the product
// M0*c0 of that fixed c0.  Their values, with the round keys they are xored
// with, are computed once (five words per block), and each block then costs
// one product in round 1, one in round 2 and two in every later round: 16
// products instead of 20 (18 after round 0).  The result is the same
// Philox4x32-10 block as philox4x32_10(c0, c1, c2, c3, k0, k1).
struct PhiloxFixed {
    uint32_t a;   // hi(M1*c2) ^ k0            round 1: c0 = a ^ c1
    uint32_t b;   // lo(M0*c0) ^ k1_1          round 1: c2 = hi(M0*c0_1) ^ b
    uint32_t c;   // lo(M1*c2_1) ^ k0_2        round 2: c0 = hi(M1*c2_2) ^ c
    uint32_t d;   // hi(M0*c0_2) ^ k1_2        round 2: c2 = lo(M0*c0_1) ^ d
    uint32_t e;   // lo(M0*c0_2) ^ k1_3        round 3: c2 = hi(M0*c0_3) ^ e
};
constexpr uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;
__device__ __forceinline__ PhiloxFixed philox_fixed(uint32_t c0, uint32_t c2, uint32_t c3, uint32_t k0,
                                                    uint32_t k1) {
    const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;                 // round 0
    const uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
    const uint32_t c2_1 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    const uint64_t q1 = (uint64_t)kPhiloxM1 * c2_1;               // round 1, product of c2
    const uint32_t c0_2 = (uint32_t)(q1 >> 32) ^ (uint32_t)p1 ^ (k0 + kPhiloxW0);
    const uint64_t q2 = (uint64_t)kPhiloxM0 * c0_2;               // round 2, product of c0
    return PhiloxFixed{(uint32_t)(p1 >> 32) ^ k0, (uint32_t)p0 ^ (k1 + kPhiloxW1),
                       (uint32_t)q1 ^ (k0 + 2u * kPhiloxW0), (uint32_t)(q2 >> 32) ^ (k1 + 2u * kPhiloxW1),
                       (uint32_t)q2 ^ (k1 + 3u * kPhiloxW1)};
}
// Block (c0, c1, c2, c3) from its fixed part, key wave-uniform.  OPAQUE keeps
// the key opaque as in philox4x32_10_uk (round keys recomputed on the scalar
// unit at each call); without it the compiler may keep them in SGPRs.
template <bool OPAQUE = true>
__device__ __forceinline__ U4 philox4x32_10_fixed(const PhiloxFixed& f, uint32_t c1, uint32_t k0, uint32_t k1) {
    if constexpr (OPAQUE) asm volatile("" : "+s"(k0), "+s"(k1));
    const uint64_t r1 = (uint64_t)kPhiloxM0 * (f.a ^ c1);        // round 1
    const uint64_t r2 = (uint64_t)kPhiloxM1 * ((uint32_t)(r1 >> 32) ^ f.b);   // round 2
    uint32_t c0 = (uint32_t)(r2 >> 32) ^ f.c, c1r = (uint32_t)r2;
    uint32_t c2 = (uint32_t)r1 ^ f.d, c3;
    k0 += 3u * kPhiloxW0;                                            // round 3
    k1 += 3u * kPhiloxW1;
    {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
        c0 = xor3((uint32_t)(p1 >> 32), c1r, k0);
        c2 = (uint32_t)(p0 >> 32) ^ f.e;
        c1r = (uint32_t)p1;
        c3 = (uint32_t)p0;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
#pragma unroll
    for (int r = 4; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
        const uint32_t n0 = xor3((uint32_t)(p1 >> 32), c1r, k0);
        const uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, k1);
        c1r = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
    return U4{c0, c1r, c2, c3};
}

// Blocks whose c0 is a lane's fixed value (the cell) and c3 fixed, with c1
// (the step) and c2 (the gene pair) the same in every lane of the wave, as in
// the FP32 stepper, where a wave works on the same genes of 64 cells.  Round
// 0's lane product M0*c0 and round 1's M1*c2 (c2 then being hi(M0*c0) ^ c3 ^ k)
// are the same for every step and pair; with the round keys they are xored
// with, three words per lane.  The uniform products (M1*c2, then M0*c0 of
// round 1, whose c0 is hi(M1*c2) ^ c1 ^ k) go to the scalar unit, so a block
// costs 16 lane products.  Same block as philox4x32_10(c0, c1, c2, c3, k0, k1).
struct PhiloxLane {
    uint32_t b;   // lo(M0*c0) ^ k1_1          round 1: c2 = hi(M0*c0_1) ^ b
    uint32_t x;   // hi(M1*c2_1) ^ k0_1        round 1: c0 = x ^ lo(M1*c2)
    uint32_t y;   // lo(M1*c2_1) ^ k0_2        round 2: c0 = hi(M1*c2_2) ^ y
};
__device__ __forceinline__ PhiloxLane philox_lane(uint32_t c0, uint32_t c3, uint32_t k0, uint32_t k1) {
    const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
    const uint64_t q1 = (uint64_t)kPhiloxM1 * ((uint32_t)(p0 >> 32) ^ c3 ^ k1);
    return PhiloxLane{(uint32_t)p0 ^ (k1 + kPhiloxW1), (uint32_t)(q1 >> 32) ^ (k0 + kPhiloxW0),
                      (uint32_t)q1 ^ (k0 + 2u * kPhiloxW0)};
}
// Block (c0, c1, c2, c3) from the lane part; c1, c2 and the key wave-uniform.
__device__ __forceinline__ U4 philox4x32_10_lane(const PhiloxLane& f, uint32_t c1u, uint32_t c2u, uint32_t k0,
                                                 uint32_t k1) {
    const uint64_t u1 = (uint64_t)kPhiloxM1 * c2u;                               // round 0, uniform
    const uint64_t u2 = (uint64_t)kPhiloxM0 * ((uint32_t)(u1 >> 32) ^ c1u ^ k0);   // round 1, uniform
    const uint32_t c0_2 = f.x ^ (uint32_t)u1;
    const uint32_t c2_2 = (uint32_t)(u2 >> 32) ^ f.b;
    const uint64_t p0 = (uint64_t)kPhiloxM0 * c0_2;                              // round 2
    const uint64_t p1 = (uint64_t)kPhiloxM1 * c2_2;
    uint32_t c0 = (uint32_t)(p1 >> 32) ^ f.y;
    uint32_t c2 = (uint32_t)(p0 >> 32) ^ ((uint32_t)u2 ^ (k1 + 2u * kPhiloxW1));
    uint32_t c1r = (uint32_t)p1, c3 = (uint32_t)p0;
    k0 += 3u * kPhiloxW0;
    k1 += 3u * kPhiloxW1;
#pragma unroll
    for (int r = 3; r < 10; ++r) {
        const uint64_t q0 = (uint64_t)kPhiloxM0 * c0;
        const uint64_t q1 = (uint64_t)kPhiloxM1 * c2;
        const uint32_t n0 = xor3((uint32_t)(q1 >> 32), c1r, k0);
        const uint32_t n2 = xor3((uint32_t)(q0 >> 32), c3, k1);
        c1r = (uint32_t)q1;
        c3 = (uint32_t)q0;
        c0 = n0;
        c2 = n2;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
    return U4{c0, c1r, c2, c3};
}

// 1 + (w >> 9) * 2^-23 as a float in [1, 2): the top 23 bits of w placed
// under the exponent of 1.0 by one v_alignbit_b32 ((0x7f:w) >> 9).
__device__ __forceinline__ float one_plus_u23(uint32_t w) {
    return __uint_as_float(__builtin_amdgcn_alignbit(0x7Fu, w, 9u));
}

// Box-Muller on two u32.  The radius uniform u1 = ((a>>9)+1/2)*2^-23 and the
// angle (b>>9)*2^-23 (revolutions) are exact in float, so the CPU oracle's
// double transform sees the same uniforms.  u1 is formed as
// (1 + (a>>9)*2^-23) - (1 - 2^-24), which is exact (the difference has at most
// 24 significant bits); the angle is passed as 1 + (b>>9)*2^-23, one whole
// revolution more, which v_sin/v_cos (periodic in revolutions, valid on
// [-256, 256]) do not see.  neg2ln2s2 = -2*ln(2)*scale^2 folds the log2->ln
// conversion and a scale (sqrt(dt)) into the radius.
__device__ __forceinline__ void normal_pair_scaled(uint32_t a, uint32_t b, float neg2ln2s2,
                                                   float& z0, float& z1) {
    const float u1 = one_plus_u23(a) - (1.0f - 0x1p-24f);
    const float r = __builtin_amdgcn_sqrtf(neg2ln2s2 * __builtin_amdgcn_logf(u1));
    const float th = one_plus_u23(b);
    z0 = r * __builtin_amdgcn_cosf(th);             // cos(2*pi*th)
    z1 = r * __builtin_amdgcn_sinf(th);             // sin(2*pi*th)
}

// The same pair as radius squared and direction: z0 = c*sqrt(r2), z1 = sn*sqrt(r2),
// for callers that fold sqrt(r2) into another square root.
__device__ __forceinline__ void normal_pair_r2(uint32_t a, uint32_t b, float neg2ln2s2, float& r2, float& c,
                                               float& sn) {
    const float u1 = one_plus_u23(a) - (1.0f - 0x1p-24f);
    r2 = neg2ln2s2 * __builtin_amdgcn_logf(u1);
    const float th = one_plus_u23(b);
    c = __builtin_amdgcn_cosf(th);
    sn = __builtin_amdgcn_sinf(th);
}

// 52-bit uniform in (0,1) from two u32 (capture and population draws).
// Computed as a * 2^20 + (b >> 12) + 1/2 in doubles: every intermediate is an
// integer or half-integer below 2^52, so each operation is exact and the value
// is that of the 52-bit integer x = (a << 20) | (b >> 12), without a 64-bit
// integer conversion.
__device__ __forceinline__ double u52(uint32_t a, uint32_t b) {
    return fma((double)a, 0x1p20, (double)(b >> 12) + 0.5) * 0x1p-52;
}

}  // namespace scrna
