// Hill function of the CLE steppers (network.jl:112-125), shared by the
// fused FP32 stepper and the wide path's elementwise kernel.
#pragma once
#include <hip/hip_runtime.h>

namespace scrna {

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

}  // namespace scrna
