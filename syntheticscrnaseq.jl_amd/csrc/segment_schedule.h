// Device side of the segmented schedule shared by the FP32 CLE stepper and
// the tau-leap stepper.
//
// A launch over n_tiles tiles with n_seg segments runs n_tiles * n_seg
// workgroups.  Each takes a ticket in start order; ticket t is segment
// t / n_tiles of tile t % n_tiles, so the segment it continues holds an
// earlier ticket and is resident or finished: waiting on it always ends.
// Hand-over is a release store / acquire load at agent scope (the tile's
// state crosses XCD L2s through the state pool).  A wall-clock bound turns a
// scheduling bug into an error word instead of a hang.
#pragma once
#include <hip/hip_runtime.h>

#include "cle_kernels.h"

namespace scrna {

struct SegmentSlot {
    int tile;   // tile index inside the launch group
    int seg;    // segment of that tile's step range
};

// Ticket -> (tile, segment); without a schedule the workgroup is its tile.
// The ticket comes back through LDS; readfirstlane states that it is uniform,
// so the tile descriptor and the blob behind it are read with scalar loads.
__device__ __forceinline__ SegmentSlot segment_take(int* __restrict__ sched, int n_tiles, int* sh_ticket) {
    if (sched == nullptr) return SegmentSlot{(int)blockIdx.x, 0};
    if (threadIdx.x == 0) *sh_ticket = atomicAdd(&sched[SCHED_TICKET], 1);
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(*sh_ticket);
    const int seg = t / n_tiles;
    return SegmentSlot{t - seg * n_tiles, seg};
}

// Waits for segment seg-1 of the tile; returns false (after releasing this
// segment so later ones do not wait) if the bound expired.
__device__ __forceinline__ bool segment_wait(int* __restrict__ sched, SegmentSlot sl, long long wait_ticks,
                                             int* sh_fail) {
    if (sched == nullptr || sl.seg == 0) return true;
    if (threadIdx.x == 0) {
        int fail = 0;
        const long long t0 = wall_clock64();
        while (__hip_atomic_load(&sched[SCHED_DONE + sl.tile], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < sl.seg) {
            __builtin_amdgcn_s_sleep(16);
            if (wall_clock64() - t0 > wait_ticks) {
                fail = 1;
                atomicOr(&sched[SCHED_ERROR], 1);
                break;
            }
        }
        *sh_fail = fail;
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (*sh_fail) {
        if (threadIdx.x == 0)
            __hip_atomic_store(&sched[SCHED_DONE + sl.tile], sl.seg + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        return false;
    }
    return true;
}

// Publishes this segment's state: every wave's stores complete at agent
// scope before the flag that releases the next segment.
__device__ __forceinline__ void segment_done(int* __restrict__ sched, SegmentSlot sl) {
    if (sched == nullptr) return;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_store(&sched[SCHED_DONE + sl.tile], sl.seg + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// This segment's share [first, last) of the step range [r_lo, r_hi).
__device__ __forceinline__ void segment_range(int r_lo, int r_hi, int seg, int n_seg, int* first, int* last) {
    *first = r_lo + (int)(((long long)(r_hi - r_lo) * seg) / n_seg);
    *last = r_lo + (int)(((long long)(r_hi - r_lo) * (seg + 1)) / n_seg);
}

}  // namespace scrna
