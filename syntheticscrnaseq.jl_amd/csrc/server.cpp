// Serving loop: continuous batching of simulate() requests on one device.
//
// Requests are compiled on submit (host), admitted into a device arena when
// it has room (the HBM budget), and advanced together: every round builds one
// tile table covering all admitted requests, each tile carrying its request's
// next steps, and launches the stepper once per kernel shape.  Launches of
// different shapes go to separate streams so that they run side by side;
// a launch whose tiles fill the device runs the segmented schedule, so its
// last wave of workgroups is short.  A request whose per-step cost is a small
// fraction of the costliest active request's advances several chunks per
// round (in a launch of its own), so small requests are not held to one
// chunk per round at the pace of the largest.  Requests that arrive while
// others are running join at the next round; finished requests get their
// moment and capture kernels, are copied out and release their arena ranges.
// Results do not depend on the mix: random streams are keyed by (seed,
// net_id, cell, step), and a chunk or segment boundary only stores and
// reloads the FP32 state and the promoter bits.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "runtime_internal.h"

using namespace scrna;

namespace {

using Clock = std::chrono::steady_clock;

// Budget left beside the arenas for the server's scratch: tile tables,
// schedule blocks and job tables (a few KB per admitted request).
constexpr uint64_t kScratchReserve = 32ull << 20;

// First-fit range allocator over [0, size) in 256-byte granules.
class RangeAlloc {
   public:
    explicit RangeAlloc(size_t size) { free_[0] = size; }
    bool alloc(size_t bytes, size_t* off) {
        bytes = (bytes + 255) & ~(size_t)255;
        if (bytes == 0) bytes = 256;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes, start = it->first + bytes;
                free_.erase(it);
                if (rest) free_[start] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        if (bytes == 0) bytes = 256;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

   private:
    std::map<size_t, size_t> free_;
};

enum class St { Pending, Active, Done, Failed, Cancelled };

struct Req {
    int64_t id = 0;
    CompiledNet cn;
    bool capture = false;
    scrna_capture cap{};
    int readout = 1;
    St st = St::Pending;
    bool in_worker = false;                // admitted and not yet dropped by the worker
    int rc = SCRNA_OK;
    std::string err;
    int64_t s_done = 0;
    // arena placement (bytes) and blob placement (words)
    size_t off_state = 0, off_prom = 0, off_counts = 0, off_mom = 0, off_blob = 0;
    size_t b_state = 0, b_prom = 0, b_counts = 0, b_mom = 0, w_blob = 0;
    double cost = 0.0;                     // per-step cost estimate for the round planner
    Clock::time_point t_submit, t_done;
    std::vector<float> y;                  // readout rows, [col][cell], converted to double on wait
    std::vector<int32_t> counts;
    std::vector<double> moments;
};

}  // namespace

struct scrna_server {
    scrna_ctx* ctx = nullptr;
    int max_slots = 64;
    int chunk = 256;                       // steps per round of the costliest requests
    void* arena = nullptr;
    size_t arena_bytes = 0;
    uint32_t* blob_arena = nullptr;
    size_t blob_words = 0;
    Tile* d_tiles = nullptr;
    size_t tiles_cap = 0;
    CaptureJob* d_cjobs = nullptr;
    MomentJob* d_mjobs = nullptr;
    size_t jobs_cap = 0;
    std::unique_ptr<RangeAlloc> arena_alloc, blob_alloc;
    hipStream_t stream = nullptr;          // uploads, the first launch group and finishing work
    static constexpr int kStreams = 4;     // launch groups run side by side on these (default priority)
    hipStream_t side[kStreams - 1] = {};
    hipEvent_t ev_tiles = nullptr;         // tile table uploaded, on `stream`
    int* d_sched = nullptr;                // segmented-schedule blocks of one round's launches
    size_t sched_cap = 0;
    long long wait_ticks = 0;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> pending;
    std::vector<int64_t> active;
    std::unordered_map<int64_t, std::unique_ptr<Req>> reqs;
    int64_t next_id = 1;
    bool stop = false;
    std::thread worker;

    // statistics
    std::vector<double> latencies_ms;
    double cgs_done = 0.0, busy_s = 0.0;
    int64_t completed = 0, rejected = 0, cancelled = 0;
};

namespace {

void fail_req(Req& r, int rc, const std::string& msg) {
    r.st = St::Failed;
    r.rc = rc;
    r.err = msg;
    r.t_done = Clock::now();
}

void release_req(scrna_server* S, Req& r) {
    if (r.b_state) S->arena_alloc->release(r.off_state, r.b_state);
    if (r.b_prom) S->arena_alloc->release(r.off_prom, r.b_prom);
    if (r.b_counts) S->arena_alloc->release(r.off_counts, r.b_counts);
    if (r.b_mom) S->arena_alloc->release(r.off_mom, r.b_mom);
    if (r.w_blob) S->blob_alloc->release(r.off_blob * 4, r.w_blob * 4);
    r.b_state = r.b_prom = r.b_counts = r.b_mom = r.w_blob = 0;
}

// Try to place a request in the arena; upload its blob.
bool admit(scrna_server* S, Req& r) {
    const size_t GC = (size_t)r.cn.G * (size_t)r.cn.cells;
    size_t o1 = 0, o2 = 0, o3 = 0, o4 = 0, ob = 0;
    const size_t b1 = 2 * GC * sizeof(float), b2 = GC, b3 = r.capture ? GC * sizeof(int32_t) : 0,
                 b4 = (size_t)r.cn.G * 5 * sizeof(double), wb = r.cn.blob.size() * 4;
    auto& A = *S->arena_alloc;
    if (!A.alloc(b1, &o1)) return false;
    if (!A.alloc(b2, &o2)) { A.release(o1, b1); return false; }
    if (b3 && !A.alloc(b3, &o3)) { A.release(o1, b1); A.release(o2, b2); return false; }
    if (!A.alloc(b4, &o4)) {
        A.release(o1, b1); A.release(o2, b2); if (b3) A.release(o3, b3);
        return false;
    }
    if (!S->blob_alloc->alloc(wb, &ob)) {
        A.release(o1, b1); A.release(o2, b2); if (b3) A.release(o3, b3); A.release(o4, b4);
        return false;
    }
    r.off_state = o1; r.b_state = b1;
    r.off_prom = o2; r.b_prom = b2;
    r.off_counts = o3; r.b_counts = b3;
    r.off_mom = o4; r.b_mom = b4;
    r.off_blob = ob / 4; r.w_blob = wb / 4;
    // Per-step cost for the round planner: G x cells, with the cells floored
    // at what fills the device's resident workgroup slots once (a request
    // with fewer tiles takes about as long per step as one that fills them).
    const int Smax = std::max(1, r.cn.S);
    const int nbuf = stepper_lds_bytes(Smax, 2) <= S->ctx->lds_limit_soft ? 2 : 1;
    const int per_cu = std::max(1, stepper_blocks_per_cu(r.cn.CH, r.cn.NW, stepper_lds_bytes(Smax, nbuf)));
    const double fill = 64.0 * per_cu * std::max(1, S->ctx->num_cu);
    r.cost = (double)r.cn.G * std::max((double)r.cn.cells, fill);
    std::vector<uint32_t>& b = r.cn.blob;
    auto put = [&](int idx, uint64_t v) { b[idx] = (uint32_t)v; b[idx + 1] = (uint32_t)(v >> 32); };
    put(H_OUT_FINAL_LO, o1 / sizeof(float));
    put(H_OUT_SNAP_LO, 0);
    put(H_OUT_PROM_LO, o2);
    put(H_OUT_V_LO, 0);
    put(H_INJ_LO, 0);
    if (hipMemcpyAsync(S->blob_arena + r.off_blob, b.data(), wb, hipMemcpyHostToDevice, S->stream) != hipSuccess) {
        release_req(S, r);
        fail_req(r, SCRNA_ERR_DEVICE, "blob upload failed");
        return true;
    }
    r.st = St::Active;
    return true;
}

// Grow a scratch buffer (tiles, schedule blocks, job tables) to hold `need`
// elements.  Scratch is charged to the context's HBM budget like the arenas;
// scrna_server_create leaves kScratchReserve of the budget for it.
bool ensure_dev(scrna_server* S, void** p, size_t* cap, size_t need, size_t elem) {
    if (need <= *cap) return true;
    size_t n = std::max(need, *cap * 2);
    ctx_free(S->ctx, *p, *cap * elem);
    *p = nullptr;
    *cap = 0;
    if (ctx_alloc(S->ctx, p, n * elem) != SCRNA_OK) return false;
    *cap = n;
    return true;
}

void finish(scrna_server* S, std::vector<Req*>& done) {
    if (done.empty()) return;
    std::vector<CaptureJob> cj;
    std::vector<MomentJob> mj;
    int maxG = 0;
    int64_t max_cap = 0;
    for (Req* r : done) {
        const size_t GC = (size_t)r->cn.G * (size_t)r->cn.cells;
        MomentJob m{};
        m.m_off = r->off_state / sizeof(float);
        m.p_off = m.m_off + GC;
        m.out_off = r->off_mom / sizeof(double);
        m.cells = (int32_t)r->cn.cells;
        m.G = r->cn.G;
        mj.push_back(m);
        maxG = std::max(maxG, r->cn.G);
        if (r->capture) {
            CaptureJob c{};
            c.src_off = r->off_state / sizeof(float) + (r->cap.readout == SCRNA_READOUT_PROTEIN ? GC : 0);
            c.dst_off = r->off_counts / sizeof(int32_t);
            c.src_ld = c.dst_ld = r->cn.cells;
            c.cells = (int32_t)r->cn.cells;
            c.G = r->cn.G;
            c.mu_log = std::log(r->cap.efficiency) - r->cap.efficiency_std * r->cap.efficiency_std / 2.0;
            c.sigma_log = r->cap.efficiency_std;
            c.k0 = (uint32_t)r->cn.seed;
            c.k1 = (uint32_t)(r->cn.seed >> 32);
            c.net_id = r->cn.net_id;
            cj.push_back(c);
            max_cap = std::max<int64_t>(max_cap, (int64_t)GC);
        }
    }
    hipError_t e = hipSuccess;
    // job buffers are sized for jobs_cap entries of the larger struct
    if (mj.size() > S->jobs_cap || cj.size() > S->jobs_cap) {
        ctx_free(S->ctx, S->d_mjobs, S->jobs_cap * sizeof(MomentJob));
        ctx_free(S->ctx, S->d_cjobs, S->jobs_cap * sizeof(CaptureJob));
        S->d_mjobs = nullptr;
        S->d_cjobs = nullptr;
        const size_t cap = std::max(mj.size(), cj.size()) * 2;
        S->jobs_cap = 0;
        if (ctx_alloc(S->ctx, (void**)&S->d_mjobs, cap * sizeof(MomentJob)) != SCRNA_OK ||
            ctx_alloc(S->ctx, (void**)&S->d_cjobs, cap * sizeof(CaptureJob)) != SCRNA_OK) {
            ctx_free(S->ctx, S->d_mjobs, cap * sizeof(MomentJob));
            S->d_mjobs = nullptr;
            e = hipErrorOutOfMemory;
        } else {
            S->jobs_cap = cap;
        }
    }
    float* fpool = (float*)S->arena;
    if (e == hipSuccess)
        e = hipMemcpyAsync(S->d_mjobs, mj.data(), mj.size() * sizeof(MomentJob), hipMemcpyHostToDevice, S->stream);
    if (e == hipSuccess && !cj.empty())
        e = hipMemcpyAsync(S->d_cjobs, cj.data(), cj.size() * sizeof(CaptureJob), hipMemcpyHostToDevice, S->stream);
    if (e == hipSuccess) e = launch_moments((int)mj.size(), maxG, S->d_mjobs, fpool, (double*)S->arena, S->stream);
    if (e == hipSuccess && !cj.empty())
        e = launch_capture_f32((int)cj.size(), max_cap, S->d_cjobs, fpool, (int32_t*)S->arena, S->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(S->stream);
    for (Req* r : done) {
        const size_t C = (size_t)r->cn.cells, G = (size_t)r->cn.G, GC = G * C;
        if (e == hipSuccess) {
            // Only the rows the readout asks for; wait() converts them to double.
            const size_t ncol = r->readout == SCRNA_READOUT_BOTH ? 2 * G : G;
            const size_t row0 = r->readout == SCRNA_READOUT_PROTEIN ? G : 0;
            r->y.resize(ncol * C);
            e = hipMemcpy(r->y.data(), (char*)S->arena + r->off_state + row0 * C * sizeof(float),
                          ncol * C * sizeof(float), hipMemcpyDeviceToHost);
            if (e == hipSuccess && r->capture) {
                r->counts.resize(GC);
                e = hipMemcpy(r->counts.data(), (char*)S->arena + r->off_counts, GC * sizeof(int32_t),
                              hipMemcpyDeviceToHost);
            }
            if (e == hipSuccess) {
                r->moments.resize(G * 5);
                e = hipMemcpy(r->moments.data(), (char*)S->arena + r->off_mom, G * 5 * sizeof(double),
                              hipMemcpyDeviceToHost);
            }
        }
        std::lock_guard<std::mutex> lk(S->mu);
        release_req(S, *r);
        if (r->st == St::Cancelled) continue;      // cancelled during its last round: no result
        if (e != hipSuccess) {
            fail_req(*r, SCRNA_ERR_DEVICE, std::string("device error: ") + hipGetErrorString(e));
        } else {
            r->st = St::Done;
            r->t_done = Clock::now();
            S->latencies_ms.push_back(std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count());
            S->cgs_done += (double)r->cn.G * (double)r->cn.cells * (double)r->cn.n_steps;
            S->completed++;
        }
    }
    S->cv_done.notify_all();
}

// One launch of the stepper within a round: tiles of one kernel shape.
struct RoundGroup {
    int CH = 0, NW = 0, Smax = 1;
    bool multi = false;                    // requests advancing several chunks this round
    bool tele = false;                     // some request has telegraph promoters
    bool clamp = false;                    // some request starts from negative counts
    int64_t min_steps = INT64_MAX;         // fewest steps any of its tiles takes
    std::vector<Tile> tiles;
};

// Advances every request in `run` by its steps for this round and waits for
// the device.  Updates s_done; returns the first device error.
constexpr double kSegmentBelowFill = 0.85;

hipError_t run_round(scrna_server* S, const std::vector<Req*>& run) {
    constexpr int kMaxChunks = 16;
    double cmax = 0.0;
    for (Req* r : run) cmax = std::max(cmax, r->cost);
    std::map<std::tuple<bool, int, int>, RoundGroup> groups;   // multi-chunk groups sort first
    for (Req* r : run) {
        const int k = r->cost > 0.0 ? (int)std::min<double>(kMaxChunks, std::max(1.0, std::floor(cmax / r->cost))) : 1;
        const int64_t s0 = r->s_done, s1 = std::min<int64_t>(r->cn.n_steps, s0 + (int64_t)S->chunk * k);
        RoundGroup& g = groups[std::make_tuple(k == 1, r->cn.CH, r->cn.NW)];
        g.CH = r->cn.CH;
        g.NW = r->cn.NW;
        g.multi = k > 1;
        g.Smax = std::max(g.Smax, r->cn.S);
        g.tele = g.tele || (r->cn.blob[H_FLAGS] & F_TELE) != 0;
        g.clamp = g.clamp || (r->cn.blob[H_FLAGS] & F_NEG_INIT) != 0;
        g.min_steps = std::min<int64_t>(g.min_steps, s1 - s0);
        for (int64_t c0 = 0; c0 < r->cn.cells; c0 += 64) {
            Tile t{};
            t.blob = (int32_t)r->off_blob;
            t.cell0 = (int32_t)c0;
            t.ncell = (int32_t)std::min<int64_t>(64, r->cn.cells - c0);
            t.net = 0;
            t.s0 = (int32_t)s0;
            t.s1 = (int32_t)s1;
            g.tiles.push_back(t);
        }
        r->s_done = s1;
    }
    // One tile table and one schedule area for the round.
    std::vector<Tile> all;
    std::vector<size_t> tile0, sched0;
    std::vector<int> nseg;
    size_t sched_total = 0;
    for (auto& kv : groups) {
        RoundGroup& g = kv.second;
        tile0.push_back(all.size());
        all.insert(all.end(), g.tiles.begin(), g.tiles.end());
        const int n_tiles = (int)g.tiles.size();
        const int nbuf = stepper_lds_bytes(g.Smax, 2) <= S->ctx->lds_limit_soft ? 2 : 1;
        const long long resident =
            (long long)stepper_blocks_per_cu(g.CH, g.NW, stepper_lds_bytes(g.Smax, nbuf)) * S->ctx->num_cu;
        // Segment only when the unsegmented launch would leave its last wave
        // of resident workgroups mostly empty: rounds are short, so a
        // segment's own start and hand-over cost more than in a whole-horizon plan.
        int k = 1;
        if (!g.multi && resident > 0) {
            const long long waves = ((long long)n_tiles + resident - 1) / resident;
            if ((double)n_tiles < kSegmentBelowFill * (double)(waves * resident))
                k = choose_segments(n_tiles, resident, g.min_steps, 16);
        }
        nseg.push_back(k);
        sched0.push_back(sched_total);
        if (k > 1) sched_total += sched_words(n_tiles);
    }
    hipError_t e = hipSuccess;
    if (!ensure_dev(S, (void**)&S->d_tiles, &S->tiles_cap, all.size(), sizeof(Tile))) return hipErrorOutOfMemory;
    if (sched_total && !ensure_dev(S, (void**)&S->d_sched, &S->sched_cap, sched_total, sizeof(int)))
        return hipErrorOutOfMemory;
    e = hipMemcpyAsync(S->d_tiles, all.data(), all.size() * sizeof(Tile), hipMemcpyHostToDevice, S->stream);
    if (e == hipSuccess) e = hipEventRecord(S->ev_tiles, S->stream);
    int gi = 0;
    for (auto& kv : groups) {
        if (e != hipSuccess) break;
        const RoundGroup& g = kv.second;
        hipStream_t st = gi % S->kStreams == 0 ? S->stream : S->side[gi % S->kStreams - 1];
        if (st != S->stream) e = hipStreamWaitEvent(st, S->ev_tiles, 0);
        const int nbuf = stepper_lds_bytes(g.Smax, 2) <= S->ctx->lds_limit_soft ? 2 : 1;
        int* sched = nullptr;
        if (e == hipSuccess && nseg[gi] > 1) {
            sched = S->d_sched + sched0[gi];
            e = launch_sched_reset(sched, (int)sched_words((int)g.tiles.size()), st);
        }
        if (e == hipSuccess)
            e = launch_stepper(g.CH, g.NW, (int)g.tiles.size(), S->blob_arena, S->d_tiles + tile0[gi],
                               (float*)S->arena, nullptr, (uint8_t*)S->arena, nullptr, nullptr, 0, INT32_MAX, nbuf,
                               stepper_lds_bytes(g.Smax, nbuf), sched, nseg[gi], S->wait_ticks, g.tele, g.clamp, st);
        ++gi;
    }
    for (int i = 0; i < S->kStreams - 1; ++i) {
        const hipError_t es = hipStreamSynchronize(S->side[i]);
        if (e == hipSuccess) e = es;
    }
    {
        const hipError_t es = hipStreamSynchronize(S->stream);
        if (e == hipSuccess) e = es;
    }
    for (size_t i = 0; e == hipSuccess && i < nseg.size(); ++i) {
        if (nseg[i] <= 1) continue;
        int err = 0;
        e = hipMemcpy(&err, S->d_sched + sched0[i] + SCHED_ERROR, sizeof(int), hipMemcpyDeviceToHost);
        if (e == hipSuccess && err) e = hipErrorLaunchTimeOut;
    }
    return e;
}

void worker_main(scrna_server* S) {
    (void)hipSetDevice(S->ctx->device);
    for (;;) {
        std::vector<Req*> run;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return S->stop || !S->pending.empty() || !S->active.empty(); });
            if (S->stop) return;
            // drop cancelled actives
            std::vector<int64_t> keep;
            for (int64_t id : S->active) {
                Req& r = *S->reqs.at(id);
                if (r.st == St::Cancelled) {
                    release_req(S, r);
                    r.in_worker = false;
                    S->cv_done.notify_all();
                } else {
                    keep.push_back(id);
                }
            }
            S->active.swap(keep);
            // admission, FIFO with skipping over requests that do not fit yet
            for (auto it = S->pending.begin(); it != S->pending.end() && (int)S->active.size() < S->max_slots;) {
                Req& r = *S->reqs.at(*it);
                if (r.st == St::Cancelled) { it = S->pending.erase(it); continue; }
                if (admit(S, r)) {
                    if (r.st == St::Active) {
                        r.in_worker = true;
                        S->active.push_back(r.id);
                    }
                    it = S->pending.erase(it);
                } else {
                    ++it;
                }
            }
            for (int64_t id : S->active) run.push_back(S->reqs.at(id).get());
        }
        if (run.empty()) {
            // pending requests exist but none fits: wait for a slot to free (should not happen
            // with nothing active unless a request exceeds the arena; those are rejected on submit)
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait_for(lk, std::chrono::milliseconds(5));
            continue;
        }
        const auto t0 = Clock::now();
        hipError_t e = run_round(S, run);
        std::vector<Req*> finished;
        for (Req* r : run)
            if (r->s_done >= r->cn.n_steps) finished.push_back(r);
        if (e != hipSuccess) {
            std::lock_guard<std::mutex> lk(S->mu);
            for (Req* r : run) {
                release_req(S, *r);
                fail_req(*r, SCRNA_ERR_DEVICE, std::string("device error: ") + hipGetErrorString(e));
                r->in_worker = false;
            }
            S->active.clear();
            S->cv_done.notify_all();
            continue;
        }
        finish(S, finished);
        {
            std::lock_guard<std::mutex> lk(S->mu);
            std::vector<int64_t> keep;
            for (int64_t id : S->active) {
                Req& r = *S->reqs.at(id);
                if (r.st == St::Active) {
                    keep.push_back(id);
                } else {
                    release_req(S, r);
                    r.in_worker = false;
                }
            }
            S->active.swap(keep);
            S->cv_done.notify_all();
            S->busy_s += std::chrono::duration<double>(Clock::now() - t0).count();
        }
    }
}

void destroy_streams(scrna_server* S) {
    if (S->stream) {
        (void)hipStreamSynchronize(S->stream);
        (void)hipStreamDestroy(S->stream);
        S->stream = nullptr;
    }
    for (hipStream_t& st : S->side) {
        if (!st) continue;
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
        st = nullptr;
    }
    if (S->ev_tiles) {
        (void)hipEventDestroy(S->ev_tiles);
        S->ev_tiles = nullptr;
    }
}

}  // namespace

extern "C" {

int scrna_server_create(scrna_ctx* ctx, int32_t max_slots, scrna_server** out) {
    if (!ctx || !out) return set_error(SCRNA_ERR_INVALID, "null argument");
    *out = nullptr;
    if (max_slots <= 0) return set_error(SCRNA_ERR_INVALID, "max_slots must be positive");
    std::unique_ptr<scrna_server> S(new scrna_server());
    S->ctx = ctx;
    S->max_slots = max_slots;
    // Plans and servers of this context size themselves from the budget under ctx->mu.
    std::lock_guard<std::mutex> lk(ctx->mu);
    const uint64_t used = ctx->used.load();
    const uint64_t room = ctx->budget > used ? ctx->budget - used : 0;
    S->blob_words = 16u << 20;   // 64 MiB of network descriptors
    const uint64_t fixed = (uint64_t)S->blob_words * 4 + kScratchReserve;
    S->arena_bytes = (size_t)std::min<uint64_t>(room > fixed ? room - fixed : 0, (uint64_t)64 << 30);
    S->arena_bytes &= ~(size_t)255;
    if (S->arena_bytes < (1u << 20)) return set_error(SCRNA_ERR_BUDGET, "HBM budget leaves no room for a serving arena");
    int rc = ctx_alloc(ctx, &S->arena, S->arena_bytes);
    if (rc) return rc;
    rc = ctx_alloc(ctx, (void**)&S->blob_arena, S->blob_words * 4);
    if (rc) {
        ctx_free(ctx, S->arena, S->arena_bytes);
        return rc;
    }
    S->arena_alloc.reset(new RangeAlloc(S->arena_bytes));
    S->blob_alloc.reset(new RangeAlloc(S->blob_words * 4));
    bool ok = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) == hipSuccess;
    for (int i = 0; ok && i < S->kStreams - 1; ++i)
        ok = hipStreamCreateWithFlags(&S->side[i], hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&S->ev_tiles, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        destroy_streams(S.get());
        ctx_free(ctx, S->arena, S->arena_bytes);
        ctx_free(ctx, S->blob_arena, S->blob_words * 4);
        return set_error(SCRNA_ERR_DEVICE, "hipStreamCreate failed");
    }
    S->wait_ticks = 60000LL * ctx->wall_khz;   // 60 s bound on a segment's wait for its predecessor
    scrna_server* raw = S.get();
    S->worker = std::thread(worker_main, raw);
    *out = S.release();
    return SCRNA_OK;
}

int scrna_submit(scrna_server* S, const scrna_request* req, int64_t* ticket) {
    if (!S || !req || !ticket) return set_error(SCRNA_ERR_INVALID, "null argument");
    const scrna_capture* cap = req->capture.enabled ? &req->capture : nullptr;
    int rc = validate_inputs(&req->net, &req->kin, &req->run, nullptr, cap);
    if (rc) return rc;
    if (req->net.G > kMaxGenes)
        return set_error(SCRNA_ERR_UNSUPPORTED, "served requests run the fused FP32 stepper: G must be <= %d", kMaxGenes);
    if ((req->run.precision != SCRNA_PREC_FP32 && req->run.precision != SCRNA_PREC_FP32_VALU) ||
        req->run.snapshot_every != 0 || req->run.injected_noise ||
        req->run.algorithm != SCRNA_ALG_CLE)
        return set_error(SCRNA_ERR_UNSUPPORTED,
                         "served requests run the FP32 CLE without snapshots or injected noise");
    std::unique_ptr<Req> r(new Req());
    scrna_run run = req->run;
    run.precision = SCRNA_PREC_FP32;     // served requests run the sparse FP32 stepper
    rc = compile_network(&req->net, &req->kin, &run, nullptr, 0, &r->cn);
    if (rc) return rc;
    r->cn.seed = req->run.seed;
    r->cn.net_id = req->run.net_id;
    r->capture = cap != nullptr;
    if (cap) r->cap = *cap;
    r->readout = req->run.readout;
    r->t_submit = Clock::now();
    const size_t GC = (size_t)r->cn.G * (size_t)r->cn.cells;
    const size_t need = 2 * GC * 4 + GC + (r->capture ? GC * 4 : 0) + (size_t)r->cn.G * 40 + 4096;
    std::lock_guard<std::mutex> lk(S->mu);
    if (need > S->arena_bytes) {
        S->rejected++;
        return set_error(SCRNA_ERR_BUDGET, "request needs %zu bytes, serving arena holds %zu", need, S->arena_bytes);
    }
    r->id = S->next_id++;
    *ticket = r->id;
    S->pending.push_back(r->id);
    S->reqs[r->id] = std::move(r);
    S->cv_work.notify_one();
    return SCRNA_OK;
}

int scrna_wait(scrna_server* S, int64_t ticket, int64_t timeout_ms, scrna_output* out) {
    if (!S) return set_error(SCRNA_ERR_INVALID, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return set_error(SCRNA_ERR_INVALID, "unknown ticket %lld", (long long)ticket);
    Req* r = it->second.get();
    auto ready = [&] {
        return (r->st == St::Done || r->st == St::Failed || r->st == St::Cancelled) && !r->in_worker;
    };
    if (timeout_ms < 0) S->cv_done.wait(lk, ready);
    else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return set_error(SCRNA_ERR_TIMEOUT, "request %lld still running", (long long)ticket);
    // Take the finished request out of the table and copy it out unlocked,
    // so that the worker is not held up by the conversion.
    std::unique_ptr<Req> own = std::move(it->second);
    S->reqs.erase(it);
    lk.unlock();
    r = own.get();
    int rc = SCRNA_OK;
    if (r->st == St::Cancelled) rc = set_error(SCRNA_ERR_CANCELLED, "request %lld was cancelled", (long long)ticket);
    else if (r->st == St::Failed) rc = set_error(r->rc, "%s", r->err.c_str());
    else if (out) {
        const size_t C = (size_t)r->cn.cells, G = (size_t)r->cn.G;
        const int64_t ld = out->ld > 0 ? out->ld : (int64_t)C;
        if (ld < (int64_t)C) {
            rc = set_error(SCRNA_ERR_INVALID, "output leading dimension too small");
        } else {
            if (out->y) {
                const size_t ncol = C ? r->y.size() / C : 0;
                for (size_t col = 0; col < ncol; ++col) {
                    const float* src = r->y.data() + col * C;
                    double* dst = out->y + col * ld;
                    for (size_t c = 0; c < C; ++c) dst[c] = (double)src[c];
                }
            }
            if (out->counts && r->capture)
                for (size_t g = 0; g < G; ++g)
                    std::memcpy(out->counts + g * ld, r->counts.data() + g * C, C * sizeof(int32_t));
            if (out->moments) std::memcpy(out->moments, r->moments.data(), G * 5 * sizeof(double));
        }
    }
    return rc;
}

int scrna_cancel(scrna_server* S, int64_t ticket) {
    if (!S) return set_error(SCRNA_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return set_error(SCRNA_ERR_INVALID, "unknown ticket %lld", (long long)ticket);
    Req& r = *it->second;
    if (r.st == St::Pending || r.st == St::Active) {
        if (r.st == St::Pending)
            S->pending.erase(std::remove(S->pending.begin(), S->pending.end(), ticket), S->pending.end());
        r.st = St::Cancelled;
        r.t_done = Clock::now();
        S->cancelled++;
        S->cv_work.notify_one();
        S->cv_done.notify_all();
    }
    return SCRNA_OK;
}

int scrna_server_stats(scrna_server* S, scrna_latency* out) {
    if (!S || !out) return set_error(SCRNA_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    std::vector<double> v = S->latencies_ms;
    std::sort(v.begin(), v.end());
    auto q = [&](double f) -> double {
        if (v.empty()) return 0.0;
        const size_t i = std::min(v.size() - 1, (size_t)std::ceil(f * (double)v.size()) - (f > 0 ? 1 : 0));
        return v[i];
    };
    double sum = 0;
    for (double x : v) sum += x;
    out->p50_ms = q(0.50);
    out->p99_ms = q(0.99);
    out->mean_ms = v.empty() ? 0.0 : sum / (double)v.size();
    out->max_ms = v.empty() ? 0.0 : v.back();
    out->cell_gene_steps = S->cgs_done;
    out->busy_s = S->busy_s;
    out->completed = S->completed;
    out->rejected = S->rejected;
    out->cancelled = S->cancelled;
    return SCRNA_OK;
}

int scrna_server_destroy(scrna_server* S) {
    if (!S) return SCRNA_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    (void)hipSetDevice(S->ctx->device);
    destroy_streams(S);
    ctx_free(S->ctx, S->d_tiles, S->tiles_cap * sizeof(Tile));
    ctx_free(S->ctx, S->d_sched, S->sched_cap * sizeof(int));
    ctx_free(S->ctx, S->d_mjobs, S->jobs_cap * sizeof(MomentJob));
    ctx_free(S->ctx, S->d_cjobs, S->jobs_cap * sizeof(CaptureJob));
    ctx_free(S->ctx, S->arena, S->arena_bytes);
    ctx_free(S->ctx, S->blob_arena, S->blob_words * 4);
    delete S;
    return SCRNA_OK;
}

}  // extern "C"
