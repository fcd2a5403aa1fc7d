// Tile-shape autotuner for the sparse stepper.
//
// The stepper maps one cell per lane and CH genes per lane (NW = ceil(G/CH)
// waves per workgroup).  Small CH means more waves (better latency hiding and
// load balance), large CH means fewer barriers and more register reuse; the
// best choice depends on G and on the device.  The tuner times every legal CH
// on a synthetic network per G bucket, keeps the fastest per bucket in the
// context, and caches the table as JSON keyed by the device architecture.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "runtime_internal.h"

namespace scrna {

int ctx_tile_for(scrna_ctx* ctx, int G) {
    std::lock_guard<std::mutex> lk(ctx->tile_mu);
    for (const TileRule& r : ctx->tiles)
        if (G >= r.G_lo && G <= r.G_hi) return r.ch;
    return ctx->force_ch;
}

namespace {

// Deterministic synthetic network: ~2 additive in-edges per gene, one AND and
// one OR gate, telegraph switching on every gene.
struct SynthNet {
    int G;
    std::vector<double> basals, A, kon, koff, and_s, or_s;
    std::vector<int32_t> and_t, and_off, and_src, or_t, or_off, or_src;
    scrna_network net{};
};

void synth(int G, SynthNet& s) {
    uint64_t x = 0x9E3779B97F4A7C15ull ^ (uint64_t)G;
    auto rnd = [&]() {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        return (double)(x >> 11) * (1.0 / 9007199254740992.0);
    };
    s.G = G;
    s.basals.resize(G);
    s.A.assign((size_t)G * G, 0.0);
    s.kon.resize(G);
    s.koff.resize(G);
    for (int i = 0; i < G; ++i) {
        s.basals[i] = 0.5 + 2.5 * rnd();
        s.kon[i] = 0.1 + 0.9 * rnd();
        s.koff[i] = 0.5 + 4.5 * rnd();
        for (int e = 0; e < 2 && G > 1; ++e) {
            int j = (int)(rnd() * G) % G;
            if (j == i) j = (j + 1) % G;
            s.A[i + (size_t)j * G] = (rnd() < 0.5 ? -1.0 : 1.0) * (1.0 + 3.0 * rnd());
        }
    }
    s.net.G = G;
    s.net.basals = s.basals.data();
    s.net.A = s.A.data();
    if (G >= 3) {
        s.and_t = {G - 1}; s.and_off = {0, 2}; s.and_src = {0, 1}; s.and_s = {3.0};
        s.or_t = {G - 2}; s.or_off = {0, 2}; s.or_src = {0, 1}; s.or_s = {2.5};
        s.net.n_coop = 1; s.net.coop_target = s.and_t.data(); s.net.coop_off = s.and_off.data();
        s.net.coop_src = s.and_src.data(); s.net.coop_strength = s.and_s.data();
        s.net.n_redun = 1; s.net.redun_target = s.or_t.data(); s.net.redun_off = s.or_off.data();
        s.net.redun_src = s.or_src.data(); s.net.redun_strength = s.or_s.data();
    }
    s.net.k_on = s.kon.data();
    s.net.k_off = s.koff.data();
}

// phase > 0: the plan runs population dynamics with a division check every
// `phase` steps, so the stepper is launched `phase` steps at a time.
int time_shape(scrna_ctx* ctx, SynthNet& s, int ch, int64_t cells, int steps, int phase, double* ms) {
    scrna_kinetics kin{5.0, 10.0, 2.0, 1.0, 0.2, 0.0};
    scrna_run run{};
    run.dt = 0.01;
    run.T = steps * 0.01;
    run.cell_num = cells;
    run.readout = SCRNA_READOUT_PROTEIN;
    run.seed = 12345;
    scrna_population pop{};
    if (phase > 0) {
        pop.enabled = 1;
        pop.growth_rate = 0.05;
        pop.V_div = 2.0;
        pop.V_init_lo = 0.8;
        pop.V_init_hi = 1.2;
        pop.div_check_interval = phase;
        pop.partition = 1;
        pop.replace_mode = 0;
    }
    scrna_plan* P = nullptr;
    int rc = plan_build(ctx, 1, &s.net, &kin, &run, phase > 0 ? &pop : nullptr, nullptr, ch, &P);
    if (rc) return rc;
    double best = 1e30;
    // The fastest of four timed launches after a warm-up one: shapes whose
    // times differ by a few percent are then told apart run after run.
    for (int rep = 0; rep < 5 && rc == SCRNA_OK; ++rep) {
        rc = plan_launch(P, ctx->stream);
        if (rc) break;
        if (hipEventSynchronize(P->ev_stop) != hipSuccess) { rc = SCRNA_ERR_DEVICE; break; }
        float f = 0.0f;
        (void)hipEventElapsedTime(&f, P->ev_start, P->ev_stop);
        if (rep > 0) best = std::min(best, (double)f);   // first launch warms up
    }
    (void)hipStreamSynchronize(ctx->stream);
    plan_release(P);
    delete P;
    *ms = best;
    return rc;
}

bool load_cache(scrna_ctx* ctx, const char* path, const std::string& arch, int max_G) {
    std::ifstream f(path);
    if (!f) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string t = ss.str();
    if (t.find("\"arch\": \"" + arch + "\"") == std::string::npos) return false;
    const size_t a = t.find("\"table\"");
    if (a == std::string::npos) return false;
    std::vector<TileRule> rules;
    int lo, hi, ch, covered = 0;
    const char* p = t.c_str() + a;
    while ((p = std::strchr(p, '[')) != nullptr) {
        ++p;
        if (std::sscanf(p, "%d , %d , %d", &lo, &hi, &ch) == 3 || std::sscanf(p, "%d,%d,%d", &lo, &hi, &ch) == 3) {
            if (ch == 2 || ch == 4 || ch == 8 || ch == 16) {
                rules.push_back({lo, hi, ch});
                covered = std::max(covered, hi);
            }
        }
    }
    if (rules.empty() || covered < max_G) return false;
    std::lock_guard<std::mutex> lk(ctx->tile_mu);
    ctx->tiles = rules;
    return true;
}

void save_cache(scrna_ctx* ctx, const char* path, const std::string& arch) {
    std::ofstream f(path);
    if (!f) return;
    f << "{\"arch\": \"" << arch << "\", \"cus\": " << ctx->num_cu << ", \"table\": [";
    std::lock_guard<std::mutex> lk(ctx->tile_mu);
    for (size_t i = 0; i < ctx->tiles.size(); ++i)
        f << (i ? ", " : "") << "[" << ctx->tiles[i].G_lo << ", " << ctx->tiles[i].G_hi << ", " << ctx->tiles[i].ch
          << "]";
    f << "]}\n";
}

}  // namespace
}  // namespace scrna

using namespace scrna;

extern "C" {

int scrna_ctx_set_tile(scrna_ctx* ctx, int32_t G_lo, int32_t G_hi, int32_t ch) {
    if (!ctx) return set_error(SCRNA_ERR_INVALID, "null context");
    if (ch != 0 && ch != 2 && ch != 4 && ch != 8 && ch != 16)
        return set_error(SCRNA_ERR_INVALID, "genes per lane must be 2, 4, 8 or 16 (0 clears)");
    if (G_lo < 1 || G_hi < G_lo) return set_error(SCRNA_ERR_INVALID, "invalid G range");
    if (ch && (G_hi + ch - 1) / ch > 16)
        return set_error(SCRNA_ERR_INVALID, "%d genes per lane needs more than 16 waves at G=%d", ch, G_hi);
    std::lock_guard<std::mutex> lk(ctx->tile_mu);
    auto& t = ctx->tiles;
    t.erase(std::remove_if(t.begin(), t.end(),
                           [&](const TileRule& r) { return r.G_lo >= G_lo && r.G_hi <= G_hi; }),
            t.end());
    if (ch) t.insert(t.begin(), TileRule{G_lo, G_hi, ch});
    return SCRNA_OK;
}

int scrna_autotune(scrna_ctx* ctx, int32_t max_G, const char* cache_path, int32_t* n_buckets) {
    return scrna_autotune_phases(ctx, max_G, 0, cache_path, n_buckets);
}

int scrna_autotune_phases(scrna_ctx* ctx, int32_t max_G, int32_t phase_steps, const char* cache_path,
                          int32_t* n_buckets) {
    if (!ctx) return set_error(SCRNA_ERR_INVALID, "null context");
    if (max_G < 1 || max_G > kMaxGenes) return set_error(SCRNA_ERR_INVALID, "max_G must be in [1, %d]", kMaxGenes);
    if (phase_steps < 0) return set_error(SCRNA_ERR_INVALID, "phase_steps must be >= 0");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess)
        return set_error(SCRNA_ERR_DEVICE, "hipGetDeviceProperties failed");
    // The cache key holds the launch length the table was tuned for.
    const std::string arch = std::string(prop.gcnArchName).substr(0, 6) + "/" + std::to_string(ctx->num_cu) +
                             (phase_steps > 0 ? "/phase" + std::to_string(phase_steps) : std::string());
    if (cache_path && load_cache(ctx, cache_path, arch, max_G)) {
        if (n_buckets) *n_buckets = (int32_t)ctx->tiles.size();
        return SCRNA_OK;
    }
    static const int buckets[] = {4, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256};
    // Enough cells for eight 64-cell tiles per CU, so that every shape is timed
    // at the occupancy it reaches in a full launch rather than on a half-empty
    // device (where the shape with the most waves per tile always wins).
    const int64_t cells = std::max<int64_t>(16384, (int64_t)ctx->num_cu * 64 * 8);
    std::vector<TileRule> rules;
    int prev = 0;
    for (int Gb : buckets) {
        if (prev >= max_G) break;
        SynthNet s;
        synth(Gb, s);
        int best_ch = 0;
        double best = 1e30;
        // The default shape (4 genes per lane where 16 waves allow it) keeps
        // the bucket unless another is faster by more than kTuneMargin: shapes
        // within a few percent of each other otherwise trade places from run
        // to run, and with them the tables.
        constexpr double kTuneMargin = 0.03;
        const int def_ch = (Gb + 3) / 4 <= 16 ? 4 : (Gb + 7) / 8 <= 16 ? 8 : 16;
        double def_ms = 1e30;
        for (int ch : {2, 4, 8, 16}) {
            if ((Gb + ch - 1) / ch > 16) continue;
            double ms = 0;
            // Whole 128-step launches, or eight launches of phase_steps steps.
            const int steps = phase_steps > 0 ? 8 * phase_steps : 128;
            int rc = time_shape(ctx, s, ch, cells, steps, phase_steps, &ms);
            if (rc) return rc;
            if (ch == def_ch) def_ms = ms;
            if (ms < best) { best = ms; best_ch = ch; }
        }
        if (def_ms < 1e30 && best > (1.0 - kTuneMargin) * def_ms) best_ch = def_ch;
        rules.push_back({prev + 1, Gb, best_ch});
        prev = Gb;
    }
    {
        std::lock_guard<std::mutex> lk(ctx->tile_mu);
        ctx->tiles = rules;
    }
    if (cache_path) save_cache(ctx, cache_path, arch);
    if (n_buckets) *n_buckets = (int32_t)rules.size();
    return SCRNA_OK;
}

}  // extern "C"
