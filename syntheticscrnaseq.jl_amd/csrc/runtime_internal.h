// Internal types shared by the libscrna host runtime translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/scrna.h"
#include "blob.h"
#include "cle_kernels.h"

// Genes-per-lane choice for networks with G in [G_lo, G_hi] (autotuner table).
struct TileRule {
    int G_lo, G_hi, ch;
};

struct scrna_ctx {
    std::vector<TileRule> tiles;
    std::mutex tile_mu;
    int device = 0;
    uint64_t budget = 0;
    std::atomic<uint64_t> used{0};
    size_t lds_max = 65536;          // per-workgroup LDS limit of the device
    size_t lds_limit_soft = 65536;   // keep double buffers under this to fit 2+ groups per CU
    int num_cu = 0;
    int wall_khz = 100000;           // device wall-clock rate (wall_clock64 ticks per ms)
    int force_ch = 0;                // tile override (autotuner / tests); 0 = automatic
    hipStream_t stream = nullptr;
    std::mutex mu;
};

namespace scrna {

constexpr int kMaxGenes = 256;            // fused steppers (registers / LDS hold a network's genes)
constexpr int kMaxGenesUnfused = 16384;    // wide and reference-structured paths (rocBLAS products)
constexpr uint64_t kSnapPoolBit = 1ull << 63;

struct CompiledNet {
    std::vector<uint32_t> blob;
    size_t blob_off = 0;
    int G = 0, S = 0, CH = 2, NW = 1, n_edges = 0;
    int64_t cells = 0, n_steps = 0, snap_every = 0, n_snap = 0;
    int readout = 1, precision = 0, algorithm = 0;
    bool capture = false, population = false;
    scrna_capture cap{};
    scrna_population pop{};
    uint64_t seed = 0;
    uint32_t net_id = 0;
    size_t off_state = 0, off_snap = 0, off_prom = 0, off_vol = 0, off_inj = 0, off_counts = 0, off_mom = 0;
    int n_mom = 1;
    // dense (MFMA) regulation path
    bool dense = false;
    int mode = 0, Gp = 0;
    std::vector<uint16_t> wpack;     // bf16 bits
    size_t off_w = 0;                // in 8-element (16-byte) units
    bool wide = false;               // G > kMaxGenes: library GEMM + fused elementwise kernel (wide.hip)
};

// A network on the wide path (wide.hip): G above the fused steppers' limit,
// one library GEMM and one fused elementwise kernel per step.
struct WideNet {
    int net = 0;
    int mode = 3;                     // 1 BF16, 2 BF16X3, 3 FP32
    std::vector<float> host_W;        // W = A_pos - A_neg, column-major G x G (until uploaded)
    float* reg = nullptr;
    float* gh = nullptr;
    void *act = nullptr, *act_lo = nullptr, *W = nullptr, *W_lo = nullptr;
    void* blas = nullptr;             // rocblas_handle
};

// A network on the reference-structured FP32 path (unfused.hip).
struct UnfusedNet {
    int net = 0;
    float hill_n = 2.0f, Kn = 1.0f;
    bool has_gates = false;
    std::vector<float> host_A;        // A_pos then A_neg, column-major G x G
    float *Pn = nullptr, *den = nullptr, *act = nullptr, *rep = nullptr, *reg = nullptr, *tmp = nullptr,
          *nz = nullptr, *Apos = nullptr, *Aneg = nullptr;
    void* blas = nullptr;             // rocblas_handle
};

struct LaunchGroup {
    int kind = 0;                    // 0: sparse VALU stepper, 1: dense MFMA stepper, 2: tau-leap stepper
    int CH = 2, NW = 1, tile0 = 0, n_tiles = 0, nbuf = 1;
    int mode = 0;                    // dense precision mode
    bool tele = false, inject = false;
    bool clamp = false;              // FP32 stepper: some network starts from negative counts
    size_t lds = 0;
    int nseg = 1;                    // FP32 stepper: segments per tile on whole-horizon launches
    size_t sched_off = 0;            // its schedule block in scrna_plan::d_sched
    int lane = 0;                    // stream lane: 0 = the launch stream, k > 0 = scrna_plan::side[k-1]
};

}  // namespace scrna

namespace scrna {
// A fixed device region that a plan carves its buffers from instead of
// allocating (bump order, 256-byte aligned).  Used by the allocation-free
// entry point; such a plan owns no memory, events or streams.
struct Arena {
    char* base = nullptr;
    size_t size = 0, used = 0;
};
}  // namespace scrna

struct scrna_plan {
    scrna_ctx* ctx = nullptr;
    scrna::Arena* arena = nullptr;        // non-null: buffers come from this region
    std::vector<scrna::CompiledNet> nets;
    std::vector<scrna::LaunchGroup> groups;
    std::vector<scrna::UnfusedNet> unfused;
    std::vector<scrna::WideNet> wide;
    std::vector<std::pair<void*, size_t>> allocs;
    uint64_t bytes = 0;
    double cell_gene_steps = 0.0;
    bool launched = false, has_population = false;
    uint32_t* d_blobs = nullptr;
    scrna::Tile* d_tiles = nullptr;
    float* d_state = nullptr;
    float* d_snap = nullptr;
    uint8_t* d_prom = nullptr;
    double* d_vol = nullptr;
    float* d_inj = nullptr;
    int32_t* d_counts = nullptr;
    uint16_t* d_wpool = nullptr;
    double* d_mom = nullptr;
    int* d_sched = nullptr;               // segmented-schedule blocks of the FP32 stepper groups
    size_t sched_words = 0;
    long long wait_ticks = 0;             // bound on a segment's wait for its predecessor
    scrna::CaptureJob* d_cjobs = nullptr;
    scrna::MomentJob* d_mjobs_snap = nullptr;
    scrna::MomentJob* d_mjobs_state = nullptr;
    std::vector<scrna::MomentJob> mjobs_snap, mjobs_state;
    int n_cjobs = 0;
    int64_t max_cap_elems = 0;
    int max_mom_G = 0;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    hipEvent_t ev_done = nullptr;         // after the last kernel of a launch (sync and destroy wait on it)
    // Launch groups run side by side: lanes > 0 fork from the launch stream
    // at ev_fork and join back through ev_join (also inside graph capture).
    static constexpr int kSideStreams = 3;
    int n_lanes = 1;
    size_t n_tiles_total = 0;        // tiles of every launch group (d_tiles)
    // Population plans: each lane owns whole networks, whose tiles are
    // contiguous in d_tiles ([lane_t0, lane_t0 + lane_nt)), and advances its
    // phases and divisions on its own, joining the others only at the end.
    int lane_t0[1 + kSideStreams] = {}, lane_nt[1 + kSideStreams] = {};
    hipStream_t side[kSideStreams] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kSideStreams] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // population buffers: daughters, claims, initial volume ranges
    float* d_dstate = nullptr;
    uint8_t* d_dprom = nullptr;
    double* d_dvol = nullptr;
    int32_t* d_claim = nullptr;
    double* d_vinit = nullptr;
    int pop_interval = 0;
    int64_t pop_max_steps = 0;
};

namespace scrna {
int set_error(int code, const char* fmt, ...);
int64_t steps_for(double T, double dt);
int ctx_alloc(scrna_ctx* ctx, void** p, size_t bytes);
void ctx_free(scrna_ctx* ctx, void* p, size_t bytes);
int validate_inputs(const scrna_network* net, const scrna_kinetics* kin, const scrna_run* run,
                    const scrna_population* pop, const scrna_capture* cap);
int compile_network(const scrna_network* net, const scrna_kinetics* kin, const scrna_run* run,
                    const scrna_population* pop, int force_ch, CompiledNet* out);
int plan_build(scrna_ctx* ctx, int32_t n, const scrna_network* nets, const scrna_kinetics* kins,
               const scrna_run* runs, const scrna_population* pops, const scrna_capture* caps, int force_ch,
               scrna_plan** out, Arena* arena = nullptr);
int plan_launch(scrna_plan* P, hipStream_t st);
int plan_enqueue(scrna_plan* P, hipStream_t st, bool events);
int plan_check_schedule(scrna_plan* P);
void noalloc_scratch_init(int device);
int plan_download(scrna_plan* P, int32_t i, scrna_output* out);
void plan_release(scrna_plan* P);
int population_launch(scrna_plan* P, hipStream_t st);
int launch_steps(scrna_plan* P, hipStream_t st, int s_begin, int s_end, bool segmented);
// One launch group's stepper launch on stream gs (checked on the host first).
int launch_group_steps(scrna_plan* P, const LaunchGroup& lg, hipStream_t gs, int s_begin, int s_end, bool segmented);
int choose_segments(int n_tiles, long long resident, int64_t steps, int max_k);
int plan_make_graph(scrna_plan* P);
int unfused_prepare(scrna_plan* P, UnfusedNet& u);
int unfused_launch(scrna_plan* P, UnfusedNet& u, hipStream_t st);
void unfused_release(UnfusedNet& u);
int wide_prepare(scrna_plan* P, WideNet& w);
int wide_launch(scrna_plan* P, WideNet& w, hipStream_t st);
void wide_release(WideNet& w);
int ctx_tile_for(scrna_ctx* ctx, int G);
int population_prepare(scrna_plan* P, size_t state_elems, size_t prom_elems, size_t vol_elems);
int plan_alloc(scrna_plan* p, void** ptr, size_t bytes);
}  // namespace scrna
