// Launch interface of the gfx950 CLE kernels (host C++ side).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "blob.h"

namespace scrna {

// One capture job = one cells x G matrix (element (c, g) at g*ld + c).
struct CaptureJob {
    uint64_t src_off;    // element offset into the source pool
    uint64_t dst_off;    // element offset into the int32 count pool
    int64_t src_ld, dst_ld;
    int32_t cells, G;
    double mu_log, sigma_log;   // log(eff) - sigma^2/2, sigma (capture.jl:30-31)
    uint32_t k0, k1, net_id;
    uint32_t src_int;           // 1: the source holds int32 counts (tau-leap state), 0: floating point
};

// One moment job = one (m, p) pair of G x cells float matrices.
struct MomentJob {
    uint64_t m_off, p_off;      // element offsets into the float pool
    uint64_t out_off;           // element offset into the double output ([G][5])
    int32_t cells, G;
    int32_t src_int, pad;       // src_int 1: the rows hold int32 counts (tau-leap state)
};

// Segmented schedule of the FP32 stepper: a launch over n_tiles tiles with
// n_seg > 1 runs n_tiles * n_seg workgroups; one int block per launch group
// holds the ticket counter, an error word and one done-flag per tile.
enum SchedWord : int { SCHED_TICKET = 0, SCHED_ERROR = 1, SCHED_DONE = 2 };
inline size_t sched_words(int n_tiles) { return (size_t)SCHED_DONE + (size_t)n_tiles; }

// Static LDS of the stepper kernel, beside its dynamic Hill buffers.
constexpr size_t kStepperStaticLds = 64;
size_t stepper_lds_bytes(int S, int nbuf);
hipError_t set_stepper_lds_limit(size_t bytes);
// Resident workgroups per CU for the stepper at this shape (0 on error).
int stepper_blocks_per_cu(int CH, int NW, size_t lds_bytes);
// sched == nullptr: one workgroup per tile over the whole step range.
// sched != nullptr: n_seg segments per tile, ordered by tickets; sched must be
// reset (launch_sched_reset) before every such launch.  wait_ticks bounds a
// segment's wait for its predecessor, in wall-clock ticks.  tele: some network
// of the launch has telegraph promoters (selects the kernel with switching);
// clamp: some network starts from negative counts (F_NEG_INIT).
hipError_t launch_stepper(int CH, int NW, int n_tiles, const uint32_t* blobs, const Tile* tiles,
                          float* state_pool, float* snap_pool, uint8_t* prom_pool, double* vol_pool,
                          const float* inj_pool, int s_begin, int s_end, int nbuf,
                          size_t lds_bytes, int* sched, int n_seg, long long wait_ticks, bool tele,
                          bool clamp, hipStream_t stream);
hipError_t launch_sched_reset(int* sched, int n_words, hipStream_t stream);
// Binomial tau-leap stepper (tauleap_kernels.hip); same tiles and schedule.
size_t tauleap_lds_bytes(int S, int nbuf, int NW, int CH);
hipError_t set_tauleap_lds_limit(size_t bytes);
int tauleap_blocks_per_cu(int NW, size_t lds_bytes);
int tauleap_max_waves();
hipError_t launch_tauleap(int CH, int NW, int n_tiles, const uint32_t* blobs, const Tile* tiles, float* state_pool,
                          float* snap_pool, uint8_t* prom_pool, double* vol_pool, int s_begin, int s_end, int nbuf,
                          size_t lds_bytes, int* sched, int n_seg, long long wait_ticks, hipStream_t stream);
hipError_t launch_capture_f32(int n_jobs, int64_t max_elems, const CaptureJob* jobs, const float* src,
                              int32_t* dst, hipStream_t stream);
hipError_t launch_capture_f64(int n_jobs, int64_t max_elems, const CaptureJob* jobs, const double* src,
                              int32_t* dst, hipStream_t stream);
hipError_t launch_pop_init(int n_tiles, const uint32_t* blobs, const Tile* tiles, double* vol, const double* vinit,
                           hipStream_t stream);
hipError_t launch_pop_divide(int n_tiles, const uint32_t* blobs, const Tile* tiles, float* state, uint8_t* prom,
                             double* vol, float* dstate, uint8_t* dprom, double* dvol, int32_t* claim, int s_div,
                             hipStream_t stream);
hipError_t launch_pop_place(int n_tiles, const uint32_t* blobs, const Tile* tiles, float* state, uint8_t* prom,
                            double* vol, const float* dstate, const uint8_t* dprom, const double* dvol, int32_t* claim,
                            hipStream_t stream);
size_t dense_lds_bytes(int Gp, int S, int mode);
hipError_t set_dense_lds_limit(size_t bytes);
int dense_nb(int mode);
hipError_t launch_dense(int mode, bool tele, bool inject, bool clamp, int NW, int n_tiles, const uint32_t* blobs,
                        const Tile* tiles, const void* wpool, float* state_pool, float* snap_pool, uint8_t* prom_pool,
                        const float* inj_pool, int s_begin, int s_end, size_t lds_bytes,
                        hipStream_t stream);
hipError_t launch_moments(int n_jobs, int max_G, const MomentJob* jobs, const float* pool, double* out,
                          hipStream_t stream);

}  // namespace scrna
