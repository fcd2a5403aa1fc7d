// Population dynamics (Moran process) host driver: phases of
// div_check_interval steps in the persistent stepper, each followed by the
// device-side division and replacement kernels.  No host round trip: the
// reference GPU path copies the whole state to the host at every division
// check (ext/SyntheticscRNAseqCUDAExt.jl:481-519).
#include <algorithm>
#include <vector>

#include "runtime_internal.h"

namespace scrna {

#define HIP_TRY_P(expr)                                                                          \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess)                                                                    \
            return set_error(SCRNA_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(_e));  \
    } while (0)

int population_prepare(scrna_plan* P, size_t state_elems, size_t prom_elems, size_t vol_elems) {
    int interval = 0;
    for (auto& cn : P->nets) {
        if (!cn.population) continue;
        if (interval == 0) interval = cn.pop.div_check_interval;
        if (cn.pop.div_check_interval != interval)
            return set_error(SCRNA_ERR_UNSUPPORTED, "population networks in one plan must share div_check_interval");
        P->pop_max_steps = std::max<int64_t>(P->pop_max_steps, cn.n_steps);
    }
    P->pop_interval = interval;
    void* p = nullptr;
    int rc;
    if ((rc = plan_alloc(P, &p, state_elems * sizeof(float)))) return rc;
    P->d_dstate = (float*)p;
    if ((rc = plan_alloc(P, &p, prom_elems))) return rc;
    P->d_dprom = (uint8_t*)p;
    if ((rc = plan_alloc(P, &p, vol_elems * sizeof(double)))) return rc;
    P->d_dvol = (double*)p;
    if ((rc = plan_alloc(P, &p, vol_elems * sizeof(int32_t)))) return rc;
    P->d_claim = (int32_t*)p;
    if ((rc = plan_alloc(P, &p, P->nets.size() * 2 * sizeof(double)))) return rc;
    P->d_vinit = (double*)p;
    std::vector<double> vinit(P->nets.size() * 2);
    for (size_t i = 0; i < P->nets.size(); ++i) {
        vinit[2 * i] = P->nets[i].pop.V_init_lo;
        vinit[2 * i + 1] = P->nets[i].pop.V_init_hi;
    }
    HIP_TRY_P(hipMemcpy(P->d_vinit, vinit.data(), vinit.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY_P(hipMemset(P->d_claim, 0, vol_elems * sizeof(int32_t)));   // claims are cleared by the place kernel
    return SCRNA_OK;
}

// Each stream lane owns whole networks (their tiles contiguous in d_tiles)
// and runs their phases and divisions on its own stream; a network's
// division only ever reads and writes that network's cells, so the lanes
// need to meet only at the end.
int population_launch(scrna_plan* P, hipStream_t st) {
    const int I = P->pop_interval;
    const int64_t n_max = P->pop_max_steps;
    int n_tiles = 0;
    for (auto& lg : P->groups) n_tiles += lg.n_tiles;
    HIP_TRY_P(launch_pop_init(n_tiles, P->d_blobs, P->d_tiles, P->d_vol, P->d_vinit, st));
    const int L = std::max(P->n_lanes, 1);   // lanes and their tile ranges are set by plan_build
    auto lane_stream = [&](int l) { return l == 0 ? st : P->side[l - 1]; };
    if (L > 1) {
        HIP_TRY_P(hipEventRecord(P->ev_fork, st));
        for (int i = 0; i + 1 < L; ++i) HIP_TRY_P(hipStreamWaitEvent(P->side[i], P->ev_fork, 0));
    }
    for (int64_t s0 = 0; s0 < n_max || s0 == 0; s0 += I) {
        const int64_t s1 = std::min<int64_t>(s0 + I, n_max);
        for (int l = 0; l < L; ++l) {
            hipStream_t gs = lane_stream(l);
            for (const LaunchGroup& lg : P->groups) {
                if (lg.lane != l) continue;
                int rc = launch_group_steps(P, lg, gs, (int)s0, (int)s1, false);
                if (rc) return rc;
            }
            if (s1 > s0 && s1 % I == 0 && P->lane_nt[l] > 0) {   // cle.jl:240 (step % interval == 0, 1-based step = s1)
                const Tile* tiles = P->d_tiles + P->lane_t0[l];
                HIP_TRY_P(launch_pop_divide(P->lane_nt[l], P->d_blobs, tiles, P->d_state, P->d_prom, P->d_vol,
                                            P->d_dstate, P->d_dprom, P->d_dvol, P->d_claim, (int)(s1 - 1), gs));
                HIP_TRY_P(launch_pop_place(P->lane_nt[l], P->d_blobs, tiles, P->d_state, P->d_prom, P->d_vol,
                                           P->d_dstate, P->d_dprom, P->d_dvol, P->d_claim, gs));
            }
        }
        if (n_max == 0) break;
    }
    for (int i = 0; i + 1 < L; ++i) {
        HIP_TRY_P(hipEventRecord(P->ev_join[i], P->side[i]));
        HIP_TRY_P(hipStreamWaitEvent(st, P->ev_join[i], 0));
    }
    return SCRNA_OK;
}

}  // namespace scrna
