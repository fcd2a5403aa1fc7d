// Exact Binomial(n, p) sampling on the GPU in double precision.
//
//   n*min(p,1-p) < 10 : sequential inversion from 0 (BINV, Kachitvichyanukul
//                       and Schmeiser 1988);
//   otherwise         : BTRD, transformed rejection with decomposition
//                       (W. Hormann, "The generation of binomial random
//                       variates", J. Statist. Comput. Simul. 46, 1993).
// Uniforms come from consecutive Philox blocks (two 52-bit uniforms per
// block) so the draw count may vary per element without losing
// reproducibility.  FP contraction is off so the arithmetic rounds like the
// CPU restatement and accept/reject decisions agree.
#pragma once
#include "device_rng.h"

namespace scrna {

struct UStream {
    uint32_t c0, c1, c3, k0, k1;
    uint32_t block;
    int pos;
    U4 u;
    __device__ __forceinline__ UStream(uint32_t c0_, uint32_t c1_, uint32_t block0, uint32_t c3_,
                                       uint32_t k0_, uint32_t k1_)
        : c0(c0_), c1(c1_), c3(c3_), k0(k0_), k1(k1_), block(block0), pos(2), u{0, 0, 0, 0} {}
    __device__ __forceinline__ double next() {
        if (pos >= 2) {
            u = philox4x32_10(c0, c1, block, c3, k0, k1);
            ++block;
            pos = 0;
        }
        const double v = pos == 0 ? u52(u.x, u.y) : u52(u.z, u.w);
        ++pos;
        return v;
    }
};

// log(k!) - [log(sqrt(2 pi)) + (k+1/2) log(k+1) - (k+1)]
__device__ __forceinline__ double stirling_fc(int64_t k) {
    switch (k) {
        case 0: return 0.08106146679532726;
        case 1: return 0.04134069595540929;
        case 2: return 0.02767792568499834;
        case 3: return 0.02079067210376509;
        case 4: return 0.01664469118982119;
        case 5: return 0.01387612882307075;
        case 6: return 0.01189670994589177;
        case 7: return 0.01041126526197209;
        case 8: return 0.009255462182712733;
        case 9: return 0.008330563433362871;
        default: break;
    }
    const double kp = (double)k + 1.0;
    const double kp2 = kp * kp;
    return (1.0 / 12.0 - (1.0 / 360.0 - 1.0 / 1260.0 / kp2) / kp2) / kp;
}

// lq: nullptr, or {log(1-p), log(p)} precomputed by the caller, in which case
// the inversion start q^n is exp(n * log q) instead of pow(q, n) (a fixed
// decay probability drawn every step).  r0_known >= 0 is that start value
// itself, exp(n * log q) computed ahead (a table over n).
// I: the count's integer type (int for the tau-leap's int32 counts, whose
// conversion to double is one instruction; int64_t for capture and division).
// (double)n + 1.0 equals (double)(n + 1) for every count below 2^53.
template <typename I>
__device__ inline int64_t binomial_sample(I n_in, double p, UStream& st, const double* lq = nullptr,
                                          double r0_known = -1.0) {
#pragma clang fp contract(off)
    const int64_t n = n_in;
    if (n_in <= 0 || !(p > 0.0)) return 0;
    if (p >= 1.0) return n;
    const bool flip = p > 0.5;
    const double pp = flip ? 1.0 - p : p;
    const double q = 1.0 - pp;
    const double nd = (double)n_in;
    int64_t x = 0;
    if (nd * pp < 10.0) {
        const double s = pp / q, a = (nd + 1.0) * s;
        const double r0 = r0_known >= 0.0 ? r0_known
                          : lq            ? exp(nd * (flip ? lq[1] : lq[0]))
                                          : pow(q, nd);
        // 32-bit counter: the loop ends long before 2^31 (k exceeding n redraws)
        const int nk = n > (int64_t)INT32_MAX ? INT32_MAX : (int)n;
        for (;;) {
            double u = st.next(), r = r0;
            int k = 0;
            bool ok = true;
            while (u > r) {
                u -= r;
                ++k;
                if (k > nk) { ok = false; break; }
                r *= (a / (double)k - s);
            }
            if (ok) { x = k; break; }
        }
    } else {
        const double m = floor((double)(n + 1) * pp);
        const double r = pp / q, nr = (double)(n + 1) * r, npq = (double)n * pp * q;
        const double spq = sqrt(npq);
        const double b = 1.15 + 2.53 * spq;
        const double a = -0.0873 + 0.0248 * b + 0.01 * pp;
        const double c = (double)n * pp + 0.5;
        const double alpha = (2.83 + 5.1 / b) * spq;
        const double vr = 0.92 - 4.2 / b;
        const double urvr = 0.86 * vr;
        const int64_t mi = (int64_t)m;
        for (;;) {
            double v = st.next(), u;
            if (v <= urvr) {
                u = v / vr - 0.43;
                x = (int64_t)floor((2.0 * a / (0.5 - fabs(u)) + b) * u + c);
                break;
            }
            if (v >= vr) {
                u = st.next() - 0.5;
            } else {
                u = v / vr - 0.93;
                u = (u < 0.0 ? -0.5 : 0.5) - u;
                v = st.next() * vr;
            }
            const double us = 0.5 - fabs(u);
            const double kf = floor((2.0 * a / us + b) * u + c);
            if (kf < 0.0 || kf > (double)n) continue;
            const int64_t k = (int64_t)kf;
            v = v * alpha / (a / (us * us) + b);
            const int64_t km = k > mi ? k - mi : mi - k;
            if (km <= 15) {
                double f = 1.0;
                if (mi < k) {
                    for (int64_t i = mi + 1; i <= k; ++i) f *= (nr / (double)i - r);
                } else if (mi > k) {
                    for (int64_t i = k + 1; i <= mi; ++i) v *= (nr / (double)i - r);
                }
                if (v <= f) { x = k; break; }
                continue;
            }
            const double lv = log(v);
            const double kmd = (double)km;
            const double rho = (kmd / npq) * (((kmd / 3.0 + 0.625) * kmd + 1.0 / 6.0) / npq + 0.5);
            const double t = -kmd * kmd / (2.0 * npq);
            if (lv < t - rho) { x = k; break; }
            if (lv > t + rho) continue;
            const double nm = (double)(n - mi + 1);
            const double h = (m + 0.5) * log((m + 1.0) / (r * nm)) + stirling_fc(mi) + stirling_fc(n - mi);
            const double nk = (double)(n - k + 1);
            if (lv <= h + (double)(n + 1) * log(nm / nk) + (kf + 0.5) * log(nk * r / (kf + 1.0)) -
                          stirling_fc(k) - stirling_fc(n - k)) {
                x = k;
                break;
            }
        }
    }
    return flip ? n - x : x;
}

// log(k!) from the Stirling series and its correction table.
__device__ __forceinline__ double log_factorial(int64_t k) {
#pragma clang fp contract(off)
    const double kd = (double)k;
    return 0.91893853320467274178 + (kd + 0.5) * log(kd + 1.0) - (kd + 1.0) + stirling_fc(k);
}

// Exact Poisson(lam): sequential inversion from 0 for lam < 10, otherwise
// PTRS (W. Hormann, "The transformed rejection method for generating Poisson
// random variables", Insurance: Mathematics and Economics 12, 1993).  lam <= 0
// draws nothing and returns 0.  p0_known >= 0 is exp(-lam) computed ahead.
__device__ inline int64_t poisson_sample(double lam, UStream& st, double p0_known = -1.0) {
#pragma clang fp contract(off)
    if (!(lam > 0.0)) return 0;
    if (lam < 10.0) {
        const double p0 = p0_known >= 0.0 ? p0_known : exp(-lam);
        for (;;) {
            double u = st.next(), p = p0;
            int k = 0;
            while (u > p) {
                u -= p;
                ++k;
                if (k > 200) break;   // u beyond the rounded total mass: redraw
                p *= lam / (double)k;
            }
            if (k <= 200) return k;
        }
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double alpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (;;) {
        const double u = st.next() - 0.5;
        const double v = st.next();
        const double us = 0.5 - fabs(u);
        const double kf = floor((2.0 * a / us + b) * u + lam + 0.43);
        if (us >= 0.07 && v <= vr) return (int64_t)kf;
        if (kf < 0.0 || (us < 0.013 && v > us)) continue;
        const double w = v * alpha / (a / (us * us) + b);
        if (log(w) <= -lam + kf * loglam - log_factorial((int64_t)kf)) return (int64_t)kf;
    }
}

}  // namespace scrna
