// libscrna host runtime: C ABI (include/scrna.h), network compilation into
// device blobs, HBM-budgeted plans, launches and downloads.
//
// Reference behaviour mirrored here:
//   n_steps = ceil(T/dt) in double                  cle.jl:56, ext:69
//   A_eff = A .* beta (rows), zero diagonal, split   network.jl:206-219
//   mu_eff = mu + dilution (no population)           cle.jl:60-61
//   intrinsic decay only with population             cle.jl:166-167
//   init ceil(steady state)                          cle.jl:52-53, kinetics.jl:65-70
//   telegraph init / switching probabilities         tauleap_binomial.jl:64-76, 110-121
//   readout :mrna / :protein / :both                 cle.jl:123-129
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/scrna.h"
#include "blob.h"
#include "cle_kernels.h"
#include "runtime_internal.h"

namespace scrna {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess)                                                                    \
            return set_error(SCRNA_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(_e));  \
    } while (0)

static inline uint32_t f2w(float f) {
    uint32_t w;
    std::memcpy(&w, &f, 4);
    return w;
}

// float -> bfloat16 bits, round to nearest even (finite inputs).
static inline uint16_t f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
static inline float bf2f(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// u = u32 * 2^-32 in [0,1); "u < q" <=> u32 < ceil(q * 2^32).
static uint64_t prob_threshold(double q) {
    if (!(q > 0.0)) return 0;
    if (q >= 1.0) return 4294967296ull;
    double t = std::ceil(q * 4294967296.0);
    if (t > 4294967296.0) t = 4294967296.0;
    return (uint64_t)t;
}

int64_t steps_for(double T, double dt) { return (int64_t)std::ceil(T / dt); }

// ---------------------------------------------------------------------------
// Context

int ctx_alloc(scrna_ctx* ctx, void** p, size_t bytes) {
    if (bytes == 0) {
        *p = nullptr;
        return SCRNA_OK;
    }
    // Reserve the bytes against the budget first (compare-exchange, so that
    // concurrent callers cannot both pass the check), then allocate.
    uint64_t used = ctx->used.load();
    do {
        if (used + bytes > ctx->budget)
            return set_error(SCRNA_ERR_BUDGET, "HBM budget exceeded: need %zu more bytes, %zu of %zu in use",
                             bytes, (size_t)used, (size_t)ctx->budget);
    } while (!ctx->used.compare_exchange_weak(used, used + bytes));
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        ctx->used -= bytes;
        *p = nullptr;
        return set_error(SCRNA_ERR_DEVICE, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return SCRNA_OK;
}

void ctx_free(scrna_ctx* ctx, void* p, size_t bytes) {
    if (!p) return;
    (void)hipSetDevice(ctx->device);
    (void)hipFree(p);
    ctx->used -= bytes;
}

// ---------------------------------------------------------------------------
// Network compilation

int validate_inputs(const scrna_network* net, const scrna_kinetics* kin, const scrna_run* run,
                    const scrna_population* pop, const scrna_capture* cap) {
    if (!net || !kin || !run) return set_error(SCRNA_ERR_INVALID, "null network, kinetics or run");
    const int G = net->G;
    if (G <= 0) return set_error(SCRNA_ERR_INVALID, "G must be positive (got %d)", G);
    if (G > kMaxGenes) {
        // Past the fused steppers' limit the CLE runs on the wide path (one library GEMM
        // and one fused elementwise kernel per step: FP32, BF16 or BF16x3 products, like
        // ext:93-94 for any G) or, when asked for, on the reference-structured FP32 path.
        // Neither holds population dynamics, the tau-leap or the sparse VALU stepper.
        if (G > kMaxGenesUnfused || run->algorithm != SCRNA_ALG_CLE || run->precision == SCRNA_PREC_FP32_VALU ||
            (pop && pop->enabled))
            return set_error(SCRNA_ERR_UNSUPPORTED,
                             "G=%d exceeds the fused steppers' maximum %d: it runs as a CLE without population "
                             "dynamics on the wide path (precision fp32, fp32_mfma, bf16 or bf16x3) or the "
                             "reference-structured path (fp32_unfused), G <= %d",
                             G, kMaxGenes, kMaxGenesUnfused);
    }
    if (!net->basals || !net->A) return set_error(SCRNA_ERR_INVALID, "basals and A are required");
    for (int i = 0; i < G; ++i) {
        if (!std::isfinite(net->basals[i])) return set_error(SCRNA_ERR_INVALID, "basal %d is not finite", i);
        if (net->A[i + (size_t)i * G] != 0.0)   // network.jl:68
            return set_error(SCRNA_ERR_INVALID, "interaction diagonal must be zero (A[%d,%d] != 0)", i, i);
    }
    for (size_t k = 0; k < (size_t)G * G; ++k)
        if (!std::isfinite(net->A[k])) return set_error(SCRNA_ERR_INVALID, "interaction entry %zu is not finite", k);
    auto check_gates = [&](int n, const int32_t* tgt, const int32_t* off, const int32_t* src, const double* str,
                           const char* what) -> int {
        if (n < 0) return set_error(SCRNA_ERR_INVALID, "negative %s gate count", what);
        if (n == 0) return SCRNA_OK;
        if (!tgt || !off || !src || !str) return set_error(SCRNA_ERR_INVALID, "%s gate arrays missing", what);
        if (off[0] != 0) return set_error(SCRNA_ERR_INVALID, "%s gate offsets must start at 0", what);
        for (int e = 0; e < n; ++e) {
            if (tgt[e] < 0 || tgt[e] >= G) return set_error(SCRNA_ERR_INVALID, "%s gate %d target out of range", what, e);
            if (off[e + 1] < off[e] || off[e + 1] - off[e] > 0xffff)
                return set_error(SCRNA_ERR_INVALID, "%s gate %d has invalid source offsets", what, e);
            for (int k = off[e]; k < off[e + 1]; ++k)
                if (src[k] < 0 || src[k] >= G) return set_error(SCRNA_ERR_INVALID, "%s gate %d source out of range", what, e);
            if (!std::isfinite(str[e])) return set_error(SCRNA_ERR_INVALID, "%s gate %d strength not finite", what, e);
        }
        return SCRNA_OK;
    };
    int rc = check_gates(net->n_coop, net->coop_target, net->coop_off, net->coop_src, net->coop_strength, "AND");
    if (rc) return rc;
    rc = check_gates(net->n_redun, net->redun_target, net->redun_off, net->redun_src, net->redun_strength, "OR");
    if (rc) return rc;
    if ((net->k_on == nullptr) != (net->k_off == nullptr))
        return set_error(SCRNA_ERR_INVALID, "k_on and k_off must both be given or both be NULL");
    if (net->k_on)
        for (int g = 0; g < G; ++g)
            if (std::isnan(net->k_on[g]) || std::isnan(net->k_off[g]) || net->k_on[g] < 0 || net->k_off[g] < 0)
                return set_error(SCRNA_ERR_INVALID, "switching rates of gene %d must be >= 0 or Inf", g);
    if (!(kin->K_d > 0.0) || !(kin->n > 0.0) || !std::isfinite(kin->K_d) || !std::isfinite(kin->n))
        return set_error(SCRNA_ERR_INVALID, "K_d and n must be positive and finite");
    if (!std::isfinite(kin->k_t) || !std::isfinite(kin->mu_m) || !std::isfinite(kin->mu_p) ||
        !std::isfinite(kin->dilution) || kin->k_t < 0 || kin->mu_m < 0 || kin->mu_p < 0 || kin->dilution < 0)
        return set_error(SCRNA_ERR_INVALID, "kinetic rates must be finite and non-negative");
    if (!(run->dt > 0.0) || !std::isfinite(run->dt)) return set_error(SCRNA_ERR_INVALID, "dt must be positive");
    if (!(run->T >= 0.0) || !std::isfinite(run->T)) return set_error(SCRNA_ERR_INVALID, "T must be finite and >= 0");
    if (run->cell_num <= 0 || run->cell_num > (int64_t)INT32_MAX)
        return set_error(SCRNA_ERR_INVALID, "cell_num must be in [1, 2^31)");
    const int64_t n_steps = steps_for(run->T, run->dt);
    if (n_steps > (int64_t)INT32_MAX) return set_error(SCRNA_ERR_INVALID, "too many steps (%lld)", (long long)n_steps);
    if (run->readout < 0 || run->readout > 2) return set_error(SCRNA_ERR_INVALID, "readout must be 0, 1 or 2");
    if (run->init_mode < 0 || run->init_mode > 1) return set_error(SCRNA_ERR_INVALID, "init_mode must be 0 or 1");
    if (run->snapshot_every < 0) return set_error(SCRNA_ERR_INVALID, "snapshot_every must be >= 0");
    if (run->precision < 0 || run->precision > 5) return set_error(SCRNA_ERR_INVALID, "unknown precision %d", run->precision);
    if (run->algorithm != SCRNA_ALG_CLE && run->algorithm != SCRNA_ALG_BINOMIAL_TAU_LEAP)
        return set_error(SCRNA_ERR_INVALID, "unknown algorithm %d", run->algorithm);
    if (run->reserved0 != 0) return set_error(SCRNA_ERR_INVALID, "scrna_run.reserved0 must be 0");
    if (run->algorithm == SCRNA_ALG_BINOMIAL_TAU_LEAP) {
        if (run->precision != SCRNA_PREC_FP32 && run->precision != SCRNA_PREC_FP32_VALU)
            return set_error(SCRNA_ERR_UNSUPPORTED, "the tau-leap computes propensities in FP64; precision must be 0");
        if (run->injected_noise)
            return set_error(SCRNA_ERR_INVALID, "injected noise applies to the CLE only");
        if (pop && pop->enabled && pop->partition != 1)
            return set_error(SCRNA_ERR_UNSUPPORTED,
                             "the tau-leap keeps integer counts: population runs need binomial partitioning");
    }
    if (run->precision == SCRNA_PREC_FP32_UNFUSED) {
        bool tele = false;
        if (net->k_on)
            for (int g = 0; g < G; ++g) tele |= std::isfinite(net->k_on[g]) || std::isfinite(net->k_off[g]);
        if (tele || (pop && pop->enabled) || run->injected_noise)
            return set_error(SCRNA_ERR_UNSUPPORTED,
                             "the reference-structured FP32 path has no telegraph, population or injected noise");
    }
    if (run->net_id >= (1u << 28)) return set_error(SCRNA_ERR_INVALID, "net_id must be < 2^28");
    if ((int64_t)G * run->cell_num > ((int64_t)1 << 40)) return set_error(SCRNA_ERR_INVALID, "problem too large");
    if (pop && pop->enabled) {
        if (!(pop->growth_rate >= 0.0) || !(pop->V_div > 0.0) || !(pop->V_init_lo > 0.0) ||
            !(pop->V_init_hi >= pop->V_init_lo) || pop->div_check_interval <= 0)
            return set_error(SCRNA_ERR_INVALID, "invalid population configuration");
        if (pop->partition < 0 || pop->partition > 1 || pop->replace_mode < 0 || pop->replace_mode > 1)
            return set_error(SCRNA_ERR_INVALID, "invalid population partition / replace mode");
        if (!(pop->growth_noise >= 0.0) || !std::isfinite(pop->growth_noise))
            return set_error(SCRNA_ERR_INVALID, "growth_noise must be finite and >= 0");
    }
    if (cap && cap->enabled) {
        // CaptureModel asserts readout in (:mrna, :protein) (population.jl:63)
        if (cap->readout != SCRNA_READOUT_MRNA && cap->readout != SCRNA_READOUT_PROTEIN)
            return set_error(SCRNA_ERR_INVALID, "capture readout must be mrna or protein");
        if (!(cap->efficiency > 0.0) || !std::isfinite(cap->efficiency) || !(cap->efficiency_std >= 0.0) ||
            !std::isfinite(cap->efficiency_std))
            return set_error(SCRNA_ERR_INVALID, "capture efficiency must be > 0 and efficiency_std >= 0");
    }
    return SCRNA_OK;
}

// Default genes per lane when no autotuner rule applies: 4 (measured fastest
// on MI355X for a G=20 telegraph network: 42.4 ms vs 49.6 ms with 2 and 46.9 ms
// with 8 per 1e10 cell*gene*steps), larger where needed to keep NW <= 16 waves.
static void choose_shape(int G, int* CH, int* NW) {
    int ch = 4;
    if (G > 64) ch = 8;
    if (G > 128) ch = 16;
    *CH = ch;
    *NW = (G + ch - 1) / ch;
}

int compile_network(const scrna_network* net, const scrna_kinetics* kin, const scrna_run* run,
                    const scrna_population* pop, int force_ch, CompiledNet* out) {
    const int G = net->G;
    const bool population = pop && pop->enabled;
    out->G = G;
    out->cells = run->cell_num;
    out->n_steps = steps_for(run->T, run->dt);
    out->snap_every = run->snapshot_every;
    out->n_snap = run->snapshot_every > 0 ? out->n_steps / run->snapshot_every : 0;
    out->population = population;
    if (population) out->pop = *pop;
    int CH, NW;
    choose_shape(G, &CH, &NW);
    if (force_ch == 2 || force_ch == 4 || force_ch == 8 || force_ch == 16) {
        CH = force_ch;
        NW = (G + CH - 1) / CH;
        if (NW > 16) return set_error(SCRNA_ERR_INVALID, "tile override CH=%d gives %d waves > 16", CH, NW);
    }
    if (run->algorithm == SCRNA_ALG_BINOMIAL_TAU_LEAP) {
        // Tau-leap: the FP64 samplers hold the kernel at 2 waves per SIMD, 8 per
        // CU.  Workgroups of 4 waves let two share a CU, so one computes while
        // the other waits at its per-step barrier behind its slowest wave (the
        // samplers' loop counts vary by lane); more waves only where a wave would
        // exceed 32 genes.  G=20 benchmark: 233 ms at 4 waves, 237 at 2, 321 at 7.
        // Genes per wave is a run-time value.
        NW = std::min(tauleap_max_waves(), std::max(std::min(4, G), (G + 31) / 32));
        CH = (G + NW - 1) / NW;
        NW = (G + CH - 1) / CH;
    }
    out->CH = CH;
    out->NW = NW;
    // Dense MFMA path: the additive regulation becomes act * W^T + b.  Kernel
    // modes: 1 BF16, 2 BF16X3, 3 FP32.
    const bool dense = run->precision == SCRNA_PREC_BF16 || run->precision == SCRNA_PREC_BF16X3 ||
                       run->precision == SCRNA_PREC_FP32_MFMA;
    out->dense = dense;
    out->wide = dense && G > kMaxGenes;   // wide path: no tiles, no packed MFMA operands
    out->mode = run->precision == SCRNA_PREC_FP32_MFMA ? 3 : run->precision;
    out->Gp = (G + 31) / 32 * 32;
    std::vector<double> Wd, bias;
    const bool packed_w = dense && G <= kMaxGenes;   // the wide path uploads its own W (wide.hip)
    if (dense) {
        if (packed_w) Wd.assign((size_t)G * G, 0.0);
        bias.assign(G, 0.0);
        out->NW = out->Gp / (16 * dense_nb(out->mode));
    }

    // A_eff split (network.jl:206-219), as per-target CSR over source genes
    // (sparse path) or as the signed dense W = A_pos - A_neg plus
    // b = rowsum(A_neg) (dense path, using rep = 1 - act).
    std::vector<int> is_src(G, 0);
    std::vector<std::vector<std::pair<int, double>>> in_edges(G);
    for (int i = 0; i < G; ++i)
        for (int j = 0; j < G; ++j) {
            if (i == j) continue;
            const double aeff = net->A[i + (size_t)j * G] * net->basals[i];
            if (aeff == 0.0) continue;
            if (dense) {
                if (packed_w) Wd[(size_t)i * G + j] = aeff;
                if (aeff < 0.0) bias[i] += -aeff;
            } else {
                in_edges[i].push_back({j, aeff});
                is_src[j] = 1;
            }
        }
    struct GateRec {
        uint32_t type;
        std::vector<int> src;
        double strength_beta;
    };
    std::vector<std::vector<GateRec>> gates_of(G);
    for (int e = 0; e < net->n_coop; ++e) {
        GateRec r{GATE_AND, {}, net->coop_strength[e] * net->basals[net->coop_target[e]]};
        for (int k = net->coop_off[e]; k < net->coop_off[e + 1]; ++k) {
            r.src.push_back(net->coop_src[k]);
            is_src[net->coop_src[k]] = 1;
        }
        gates_of[net->coop_target[e]].push_back(r);
    }
    for (int e = 0; e < net->n_redun; ++e) {
        GateRec r{GATE_OR, {}, net->redun_strength[e] * net->basals[net->redun_target[e]]};
        for (int k = net->redun_off[e]; k < net->redun_off[e + 1]; ++k) {
            r.src.push_back(net->redun_src[k]);
            is_src[net->redun_src[k]] = 1;
        }
        gates_of[net->redun_target[e]].push_back(r);
    }
    std::vector<int32_t> slot_of(G, -1);
    int S = 0;
    for (int g = 0; g < G; ++g)
        if (is_src[g]) slot_of[g] = S++;
    out->S = S;
    const bool any_gates = net->n_coop > 0 || net->n_redun > 0;

    bool tele = false;
    if (net->k_on)
        for (int g = 0; g < G; ++g)
            if (std::isfinite(net->k_on[g]) || std::isfinite(net->k_off[g])) tele = true;   // network.jl:103

    std::vector<uint32_t>& b = out->blob;
    b.assign(H_WORDS, 0u);
    auto align16 = [&]() {
        while (b.size() % 16) b.push_back(0u);
    };
    const double mu_m = population ? kin->mu_m : kin->mu_m + kin->dilution;
    const double mu_p = population ? kin->mu_p : kin->mu_p + kin->dilution;
    uint32_t flags = 0;
    if (tele) flags |= F_TELE;
    if (kin->n == 2.0) flags |= F_HILL2;
    if (any_gates) flags |= F_GATES;
    if (run->injected_noise) flags |= F_INJECT;
    if (population) flags |= F_POP;
    b[H_G] = (uint32_t)G;
    b[H_S] = (uint32_t)S;
    b[H_FLAGS] = flags;
    b[H_NSTEPS] = (uint32_t)out->n_steps;
    b[H_SNAP_EVERY] = (uint32_t)run->snapshot_every;
    b[H_NSNAP] = (uint32_t)out->n_snap;
    b[H_NETID] = run->net_id;
    b[H_CELLS] = (uint32_t)run->cell_num;
    b[H_SEED_LO] = (uint32_t)run->seed;
    b[H_SEED_HI] = (uint32_t)(run->seed >> 32);
    b[H_DT] = f2w((float)run->dt);
    b[H_NEG2LN2DT] = f2w((float)(-2.0 * std::log(2.0) * run->dt));
    b[H_KT] = f2w((float)kin->k_t);
    b[H_MUM] = f2w((float)mu_m);
    b[H_MUP] = f2w((float)mu_p);
    b[H_INVK] = f2w((float)(1.0 / kin->K_d));
    b[H_HILLN] = f2w((float)kin->n);
    b[H_SQRTDT] = f2w((float)std::sqrt(run->dt));
    b[H_POPMODE] = population ? (uint32_t)((pop->partition == 1 ? 1u : 0u) | (pop->replace_mode == 1 ? 2u : 0u)) : 0u;
    b[H_DIVINT] = population ? (uint32_t)pop->div_check_interval : 0u;
    {
        const double growth = population ? std::exp(pop->growth_rate * run->dt) : 1.0;   // cle.jl:169
        const double vdiv = population ? pop->V_div : 0.0;
        uint64_t gb, vb;
        std::memcpy(&gb, &growth, 8);
        std::memcpy(&vb, &vdiv, 8);
        b[H_GROWTH_LO] = (uint32_t)gb;
        b[H_GROWTH_HI] = (uint32_t)(gb >> 32);
        b[H_VDIV_LO] = (uint32_t)vb;
        b[H_VDIV_HI] = (uint32_t)(vb >> 32);
    }
    b[H_CH] = (uint32_t)CH;
    b[H_NW] = (uint32_t)NW;

    b[H_OFF_BETA] = (uint32_t)b.size();
    for (int g = 0; g < G; ++g) b.push_back(f2w((float)net->basals[g]));
    align16();
    b[H_OFF_SLOT] = (uint32_t)b.size();
    for (int g = 0; g < G; ++g) b.push_back((uint32_t)slot_of[g]);
    align16();
    b[H_OFF_INOFF] = (uint32_t)b.size();
    {
        int e = 0;
        for (int g = 0; g < G; ++g) {
            b.push_back((uint32_t)e);
            e += (int)in_edges[g].size();
        }
        b.push_back((uint32_t)e);
        out->n_edges = e;
    }
    align16();
    b[H_OFF_EDGES] = (uint32_t)b.size();
    for (int g = 0; g < G; ++g)
        for (auto& ed : in_edges[g]) {
            const int kind = ed.second > 0.0 ? 0 : 1;   // A_pos reads act, A_neg reads rep
            b.push_back((uint32_t)(2 * slot_of[ed.first] + kind));
            b.push_back(f2w((float)std::fabs(ed.second)));
        }
    align16();
    b[H_OFF_GATEOFF] = (uint32_t)b.size();
    {
        uint32_t w = 0;
        for (int g = 0; g < G; ++g) {
            b.push_back(w);
            for (auto& r : gates_of[g]) w += 2 + (uint32_t)r.src.size();
        }
        b.push_back(w);
    }
    align16();
    b[H_OFF_GATES] = (uint32_t)b.size();
    std::vector<double> gate_w;   // FP64 strength*beta by gate word (tau-leap)
    for (int g = 0; g < G; ++g)
        for (auto& r : gates_of[g]) {
            gate_w.push_back(r.strength_beta);
            gate_w.push_back(0.0);
            for (size_t i = 0; i < r.src.size(); ++i) gate_w.push_back(0.0);
            b.push_back((r.type << 24) | (uint32_t)r.src.size());
            b.push_back(f2w((float)r.strength_beta));
            for (int s : r.src) b.push_back((uint32_t)slot_of[s]);
        }
    align16();
    b[H_OFF_TELE] = (uint32_t)b.size();
    std::vector<uint32_t> t18_on(G, 0u), t18_off(G, 0u), tflags(G, 0u);
    for (int g = 0; g < G; ++g) {
        uint32_t tf = 0;
        uint64_t thr_on = 0, thr_off = 0, thr_init = 0;
        if (net->k_on) {
            const double kon = net->k_on[g], koff = net->k_off[g];
            if (std::isfinite(kon)) {                                     // tauleap_binomial.jl:112
                tf |= 1u;
                thr_on = prob_threshold(1.0 - std::exp(-kon * run->dt));  // tauleap_binomial.jl:118
                thr_off = prob_threshold(1.0 - std::exp(-koff * run->dt));// tauleap_binomial.jl:113
            }
            if (std::isfinite(kon) && std::isfinite(koff)) {              // tauleap_binomial.jl:68-70
                tf |= 2u;
                thr_init = prob_threshold(kon / (kon + koff));
            }
        }
        b.push_back(tf);
        b.push_back(0u);
        b.push_back((uint32_t)thr_on);
        b.push_back((uint32_t)(thr_on >> 32));
        b.push_back((uint32_t)thr_off);
        b.push_back((uint32_t)(thr_off >> 32));
        b.push_back((uint32_t)thr_init);
        b.push_back((uint32_t)(thr_init >> 32));
        tflags[g] = tf;
        if (tf & 1u) {
            t18_on[g] = thr_on >> 32 ? 1u << 18 : (uint32_t)thr_on >> 14;
            t18_off[g] = thr_off >> 32 ? 1u << 18 : (uint32_t)thr_off >> 14;
        }
    }
    if (CH <= 4) {
        // Wave records (RecWord), the values the stepper's waves hold in registers.
        align16();
        b[H_OFF_WREC] = (uint32_t)b.size();
        const int stride = rec_stride(CH);
        const uint32_t zero_row = (uint32_t)(2 * S * 64 * 4);
        for (int w = 0; w < NW; ++w) {
            std::vector<uint32_t> r((size_t)stride, 0u);
            for (int k = 0; k < CH; ++k) {
                const int g = w * CH + k;
                uint32_t* rg = r.data() + (size_t)k * R_GENE_WORDS;
                for (int i = 0; i < kRecEdgeSlots; ++i) rg[R_EDGE + 2 * i] = zero_row;
                if (g >= G) continue;
                rg[R_TON] = t18_on[g];
                rg[R_TOFF] = t18_off[g];
                rg[R_BETA] = f2w((float)net->basals[g]);
                if (slot_of[g] >= 0) {
                    rg[R_HROW] = (uint32_t)(2 * slot_of[g] * 64 * 4);
                    r[R_GENE_WORDS * CH + RM_SOURCE] |= 1u << k;
                }
                for (int i = 0; i < kRecEdgeSlots && i < (int)in_edges[g].size(); ++i) {
                    const auto& ed = in_edges[g][i];
                    const int kind = ed.second > 0.0 ? 0 : 1;
                    rg[R_EDGE + 2 * i] = (uint32_t)((2 * slot_of[ed.first] + kind) * 64 * 4);
                    rg[R_EDGE + 2 * i + 1] = f2w((float)std::fabs(ed.second));
                }
                if ((int)in_edges[g].size() > kRecEdgeSlots) r[R_GENE_WORDS * CH + RM_EXTRA] |= 1u << k;
                if (!gates_of[g].empty()) r[R_GENE_WORDS * CH + RM_EXTRA] |= 1u << (16 + k);
                r[R_GENE_WORDS * CH + RM_SWITCH] |= (tflags[g] & 1u) << k;
            }
            b.insert(b.end(), r.begin(), r.end());
        }
    }
    align16();
    b[H_OFF_INIT] = (uint32_t)b.size();
    for (int g = 0; g < G; ++g) {
        const double mss = net->basals[g] / (kin->mu_m + kin->dilution);   // kinetics.jl:67
        const double pss = kin->k_t * mss / (kin->mu_p + kin->dilution);   // kinetics.jl:68
        const float m0 = run->init_mode == 0 ? (float)std::ceil(mss) : 0.0f;
        const float p0 = run->init_mode == 0 ? (float)std::ceil(pss) : 0.0f;
        if (m0 < 0.0f || p0 < 0.0f) b[H_FLAGS] |= F_NEG_INIT;
        b.push_back(f2w(m0));
        b.push_back(f2w(p0));
    }
    align16();
    b[H_OFF_SLOTGENE] = (uint32_t)b.size();
    for (int g = 0; g < G; ++g)
        if (slot_of[g] >= 0) b.push_back((uint32_t)g);
    align16();
    b[H_ALG] = (uint32_t)run->algorithm;
    out->algorithm = run->algorithm;
    if (run->algorithm == SCRNA_ALG_BINOMIAL_TAU_LEAP) {
        // FP64 section of the tau-leap: the values the CPU restatement uses,
        // computed with the same expressions (tauleap_binomial.jl:61-62, 145, 158).
        std::vector<double> d(TL_HEAD, 0.0);
        // with population dynamics decay is intrinsic only, dilution being emergent
        // (tauleap_binomial.jl:228, 242); otherwise mu + dilution (:144, 157)
        const double pdm = population ? 1.0 - std::exp(-kin->mu_m * run->dt)
                                      : 1.0 - std::exp(-(kin->mu_m + kin->dilution) * run->dt);
        const double pdp = population ? 1.0 - std::exp(-kin->mu_p * run->dt)
                                      : 1.0 - std::exp(-(kin->mu_p + kin->dilution) * run->dt);
        d[TL_KN] = std::pow(kin->K_d, kin->n);
        d[TL_HILLN] = kin->n;
        d[TL_KT] = kin->k_t;
        d[TL_DT] = run->dt;
        d[TL_PDM] = pdm;
        d[TL_LQM] = std::log(1.0 - pdm);
        d[TL_LPM] = std::log(pdm);
        d[TL_PDP] = pdp;
        d[TL_LQP] = std::log(1.0 - pdp);
        d[TL_LPP] = std::log(pdp);
        d[TL_GROWTH_RATE] = population ? pop->growth_rate : 0.0;
        d[TL_GROWTH_NOISE] = population ? pop->growth_noise : 0.0;
        for (int g = 0; g < G; ++g) d.push_back(net->basals[g]);
        for (int g = 0; g < G; ++g)
            for (auto& ed : in_edges[g]) d.push_back(std::fabs(ed.second));
        d.insert(d.end(), gate_w.begin(), gate_w.end());
        for (int g = 0; g < G; ++g) {
            const double mss = net->basals[g] / (kin->mu_m + kin->dilution);   // kinetics.jl:67
            const double pss = kin->k_t * mss / (kin->mu_p + kin->dilution);   // kinetics.jl:68
            d.push_back(run->init_mode == 0 ? std::nearbyint(mss) : 0.0);      // round half to even
            d.push_back(run->init_mode == 0 ? std::nearbyint(pss) : 0.0);
        }
        // Start values of the inversion samplers over a count n, with the
        // kernel's and the CPU restatement's expressions, so that the kernel
        // reads the values the host library's exp returns.
        for (int n = 0; n < kTlTable; ++n) d.push_back(std::exp(-(kin->k_t * (double)n * run->dt)));
        for (int n = 0; n < kTlTable; ++n) d.push_back(std::exp((double)n * (pdm > 0.5 ? d[TL_LPM] : d[TL_LQM])));
        for (int n = 0; n < kTlTable; ++n) d.push_back(std::exp((double)n * (pdp > 0.5 ? d[TL_LPP] : d[TL_LQP])));

        b[H_TL_NEDGES] = (uint32_t)out->n_edges;
        b[H_TL_NGATEW] = (uint32_t)gate_w.size();
        b[H_OFF_TL] = (uint32_t)b.size();   // 16-word aligned, so the doubles are 8-byte aligned
        for (double v : d) {
            uint64_t u;
            std::memcpy(&u, &v, 8);
            b.push_back((uint32_t)u);
            b.push_back((uint32_t)(u >> 32));
        }
        align16();
    }
    b[H_GP] = (uint32_t)out->Gp;
    b[H_OFF_BIAS] = (uint32_t)b.size();
    for (int g = 0; g < G; ++g) b.push_back(f2w(dense ? (float)bias[g] : 0.0f));
    align16();
    if (out->wide) {
        // the wide path uploads W itself, column-major (wide.hip)
    } else if (dense && out->mode == 3) {
        // FP32 W^T as v_mfma_f32_16x16x4_f32 B operands: [gene block][t][lane] f32, lane l
        // of block blk holding W[16 blk + (l&15)][(l>>4) Gp/4 + t] (the kernel's K order).
        const int Gp = out->Gp, NBLK = Gp / 16, KT = Gp / 4;
        out->wpack.assign((size_t)NBLK * KT * 64 * 2, 0);
        for (int blk = 0; blk < NBLK; ++blk)
            for (int t = 0; t < KT; ++t)
                for (int l = 0; l < 64; ++l) {
                    const int gi = 16 * blk + (l & 15), gj = (l >> 4) * KT + t;
                    const float wv = (gi < G && gj < G) ? (float)Wd[(size_t)gi * G + gj] : 0.0f;
                    std::memcpy(&out->wpack[(((size_t)blk * KT + t) * 64 + l) * 2], &wv, 4);
                }
    } else if (dense) {
        // W^T as MFMA B operands: [plane][gene block][K chunk][lane][8] bf16,
        // lane l of block blk / chunk kc holding W[16 blk + (l&15)][32 kc + 8 (l>>4) + e].
        const int Gp = out->Gp, NBLK = Gp / 16, KC = Gp / 32;
        const int planes = out->mode == SCRNA_PREC_BF16X3 ? 2 : 1;
        out->wpack.assign((size_t)planes * NBLK * KC * 64 * 8, 0);
        for (int blk = 0; blk < NBLK; ++blk)
            for (int kc = 0; kc < KC; ++kc)
                for (int l = 0; l < 64; ++l)
                    for (int e = 0; e < 8; ++e) {
                        const int gi = 16 * blk + (l & 15), gj = 32 * kc + 8 * (l >> 4) + e;
                        const float wv = (gi < G && gj < G) ? (float)Wd[(size_t)gi * G + gj] : 0.0f;
                        const size_t idx = (((size_t)blk * KC + kc) * 64 + l) * 8 + e;
                        const uint16_t hi = f2bf(wv);
                        out->wpack[idx] = hi;
                        if (planes == 2) out->wpack[(size_t)NBLK * KC * 64 * 8 + idx] = f2bf(wv - bf2f(hi));
                    }
    }
    return SCRNA_OK;
}

static void put_u64(std::vector<uint32_t>& b, int idx, uint64_t v) {
    b[idx] = (uint32_t)v;
    b[idx + 1] = (uint32_t)(v >> 32);
}

// ---------------------------------------------------------------------------
// Plans

void plan_release(scrna_plan* p) {
    if (!p) return;
    scrna_ctx* ctx = p->ctx;
    for (auto& a : p->allocs) ctx_free(ctx, a.first, a.second);
    p->allocs.clear();
    if (p->ev_start) (void)hipEventDestroy(p->ev_start);
    if (p->ev_stop) (void)hipEventDestroy(p->ev_stop);
    if (p->ev_done) (void)hipEventDestroy(p->ev_done);
    p->ev_start = p->ev_stop = p->ev_done = nullptr;
    for (int i = 0; i < scrna_plan::kSideStreams; ++i) {
        if (p->side[i]) {
            (void)hipStreamSynchronize(p->side[i]);
            (void)hipStreamDestroy(p->side[i]);
        }
        if (p->ev_join[i]) (void)hipEventDestroy(p->ev_join[i]);
        p->side[i] = nullptr;
        p->ev_join[i] = nullptr;
    }
    if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
    p->ev_fork = nullptr;
    for (auto& u : p->unfused) unfused_release(u);
    for (auto& w : p->wide) wide_release(w);
    if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->graph_exec = nullptr;
    p->graph = nullptr;
}

// Precision policy: SCRNA_PREC_FP32 runs the dense FP32 contraction on the
// matrix cores once the network is large and dense enough that streaming
// in-edges on the VALU costs more.  Measured on the G=256 network with 64
// additive in-edges per gene: sparse VALU stepper 670 ms, matrix-core kernel
// 185 ms; sparse networks (a few in-edges per gene) stay on the VALU, whose
// cost grows with the in-degree rather than with G.
static int effective_precision(const scrna_network& net, const scrna_run& run, bool population) {
    if (run.precision == SCRNA_PREC_FP32_VALU) return SCRNA_PREC_FP32;
    if (run.precision == SCRNA_PREC_FP32 && net.G > kMaxGenes) return SCRNA_PREC_FP32_MFMA;   // wide path
    if (run.precision != SCRNA_PREC_FP32 || run.algorithm != SCRNA_ALG_CLE || population || net.G < 128)
        return run.precision;
    int64_t nnz = 0;
    for (int i = 0; i < net.G; ++i)
        for (int j = 0; j < net.G; ++j)
            if (i != j && net.A[i + (size_t)j * net.G] * net.basals[i] != 0.0) ++nnz;
    return nnz >= 16 * (int64_t)net.G ? SCRNA_PREC_FP32_MFMA : SCRNA_PREC_FP32;
}

int plan_alloc(scrna_plan* p, void** ptr, size_t bytes) {
    if (p->arena) {
        Arena& a = *p->arena;
        const size_t off = (a.used + 255) & ~(size_t)255;
        if (off > a.size || bytes > a.size - off)
            return set_error(SCRNA_ERR_INVALID,
                             "run needs more than the %zu-byte fixed scratch of the allocation-free entry point",
                             a.size);
        *ptr = bytes ? a.base + off : nullptr;
        a.used = off + bytes;
        p->bytes += bytes;
        return SCRNA_OK;
    }
    int rc = ctx_alloc(p->ctx, ptr, bytes);
    if (rc) return rc;
    if (*ptr) p->allocs.push_back({*ptr, bytes});
    p->bytes += bytes;
    return SCRNA_OK;
}

int plan_build(scrna_ctx* ctx, int32_t n, const scrna_network* nets, const scrna_kinetics* kins,
               const scrna_run* runs, const scrna_population* pops, const scrna_capture* caps,
               int force_ch, scrna_plan** out, Arena* arena) {
    if (!ctx || !out || n <= 0 || !nets || !kins || !runs)
        return set_error(SCRNA_ERR_INVALID, "plan needs a context, n > 0 and network/kinetics/run arrays");
    for (int i = 0; i < n; ++i) {
        int rc = validate_inputs(&nets[i], &kins[i], &runs[i], pops ? &pops[i] : nullptr, caps ? &caps[i] : nullptr);
        if (rc) return set_error(rc, "network %d: %s", i, g_last_error.c_str());
    }
    std::unique_ptr<scrna_plan> P(new scrna_plan());
    P->ctx = ctx;
    P->arena = arena;
    if (arena) {
        // Arena plans run on one stream with no events, and hold no
        // population buffers or library handles of their own.
        for (int i = 0; i < n; ++i)
            if ((pops && pops[i].enabled) || nets[i].G > kMaxGenes ||
                effective_precision(nets[i], runs[i], false) == SCRNA_PREC_FP32_UNFUSED)
                return set_error(SCRNA_ERR_UNSUPPORTED,
                                 "population runs, the reference-structured FP32 path and the wide path (every "
                                 "network with G > %d) allocate their own buffers", kMaxGenes);
    }
    P->nets.resize(n);
    std::vector<uint32_t> blobs;
    size_t state_elems = 0, snap_elems = 0, prom_elems = 0, vol_elems = 0, inj_elems = 0, count_elems = 0,
           mom_elems = 0;
    for (int i = 0; i < n; ++i) {
        CompiledNet& cn = P->nets[i];
        const int ch = force_ch ? force_ch : ctx_tile_for(ctx, nets[i].G);   // autotuner table
        scrna_run run = runs[i];
        run.precision = effective_precision(nets[i], run, pops && pops[i].enabled);
        int rc = compile_network(&nets[i], &kins[i], &run, pops ? &pops[i] : nullptr, ch, &cn);
        if (rc) return rc;
        cn.readout = runs[i].readout;
        cn.precision = run.precision;
        cn.capture = caps && caps[i].enabled;
        if (cn.capture) cn.cap = caps[i];
        cn.seed = runs[i].seed;
        cn.net_id = runs[i].net_id;
        const size_t GC = (size_t)cn.G * (size_t)cn.cells;
        cn.off_state = state_elems;
        state_elems += 2 * GC;
        cn.off_snap = snap_elems;
        snap_elems += (size_t)cn.n_snap * 2 * GC;
        cn.off_prom = prom_elems;
        prom_elems += GC;
        cn.off_vol = vol_elems;
        if (cn.population) vol_elems += (size_t)cn.cells;
        cn.off_inj = inj_elems;
        if (runs[i].injected_noise) inj_elems += (size_t)cn.n_steps * 2 * GC;
        cn.off_counts = count_elems;
        if (cn.capture) count_elems += GC;
        cn.n_mom = cn.n_snap > 0 ? cn.n_snap : 1;
        cn.off_mom = mom_elems;
        mom_elems += (size_t)cn.n_mom * cn.G * 5;
        put_u64(cn.blob, H_OUT_FINAL_LO, cn.off_state);
        put_u64(cn.blob, H_OUT_SNAP_LO, cn.off_snap);
        put_u64(cn.blob, H_OUT_PROM_LO, cn.off_prom);
        put_u64(cn.blob, H_OUT_V_LO, cn.off_vol);
        put_u64(cn.blob, H_INJ_LO, cn.off_inj);
        cn.blob_off = blobs.size();
        blobs.insert(blobs.end(), cn.blob.begin(), cn.blob.end());
        if (blobs.size() > (size_t)INT32_MAX) return set_error(SCRNA_ERR_INVALID, "network descriptors too large");
        if (cn.population) P->has_population = true;
        if (cn.population && cn.dense)
            return set_error(SCRNA_ERR_UNSUPPORTED, "network %d: population dynamics run with FP32 regulation only", i);
    }
    if (P->has_population) {
        for (int i = 0; i < n; ++i)
            if (!P->nets[i].population)
                return set_error(SCRNA_ERR_UNSUPPORTED, "a plan mixes population and non-population networks");
    }

    // Launch groups: sparse networks sharing (CH, NW) in tiles of 64 cells;
    // dense networks sharing (mode, NW, telegraph, injected noise) in tiles of 16.
    struct GKey {
        int kind, a, NW, tele, inj;
        bool operator<(const GKey& o) const {
            return std::tie(kind, a, NW, tele, inj) < std::tie(o.kind, o.a, o.NW, o.tele, o.inj);
        }
    };
    std::map<GKey, std::vector<Tile>> groups;
    std::map<GKey, std::pair<int, int>> group_SG;   // max S, max Gp
    std::map<GKey, int64_t> group_steps;            // longest horizon
    std::map<GKey, bool> group_tele;                // some network has telegraph promoters
    std::map<GKey, bool> group_clamp;               // some network starts from negative counts
    size_t w_elems = 0;
    for (int i = 0; i < n; ++i) {
        CompiledNet& cn = P->nets[i];
        if (cn.precision == SCRNA_PREC_FP32_UNFUSED) {
            // Reference-structured baseline: no tiles, its own per-step launch sequence.
            UnfusedNet u;
            u.net = i;
            u.hill_n = (float)kins[i].n;
            u.Kn = (float)std::pow(kins[i].K_d, kins[i].n);          // K_d^n in double, then FP32 (ext:71)
            u.has_gates = nets[i].n_coop + nets[i].n_redun > 0;
            const int G = cn.G;
            u.host_A.assign(2 * (size_t)G * G, 0.0f);
            for (int a = 0; a < G; ++a)
                for (int b2 = 0; b2 < G; ++b2) {
                    if (a == b2) continue;
                    const double aeff = nets[i].A[a + (size_t)b2 * G] * nets[i].basals[a];   // network.jl:212
                    u.host_A[a + (size_t)b2 * G] = (float)std::max(aeff, 0.0);             // A_pos
                    u.host_A[(size_t)G * G + a + (size_t)b2 * G] = (float)std::max(-aeff, 0.0);   // A_neg
                }
            P->unfused.push_back(std::move(u));
            continue;
        }
        if (cn.wide) {
            // Wide path: W = A_eff (network.jl:212) column-major, the element (g, k) at g + k G.
            WideNet w;
            w.net = i;
            w.mode = cn.mode;
            const int G = cn.G;
            w.host_W.assign((size_t)G * G, 0.0f);
            for (int a = 0; a < G; ++a)
                for (int b2 = 0; b2 < G; ++b2)
                    if (a != b2) w.host_W[a + (size_t)b2 * G] = (float)(nets[i].A[a + (size_t)b2 * G] * nets[i].basals[a]);
            P->wide.push_back(std::move(w));
            continue;
        }
        const bool tele = (cn.blob[H_FLAGS] & F_TELE) != 0, inj = (cn.blob[H_FLAGS] & F_INJECT) != 0;
        GKey key = cn.dense ? GKey{1, cn.mode, cn.NW, tele ? 1 : 0, inj ? 1 : 0}
                            : GKey{cn.algorithm == SCRNA_ALG_BINOMIAL_TAU_LEAP ? 2 : 0, cn.CH, cn.NW, 0, 0};
        auto& v = groups[key];
        group_tele[key] = group_tele[key] || tele;
        group_clamp[key] = group_clamp[key] || (cn.blob[H_FLAGS] & F_NEG_INIT) != 0;
        auto& sg = group_SG[key];
        sg.first = std::max(sg.first, cn.S);
        sg.second = std::max(sg.second, cn.Gp);
        group_steps[key] = std::max(group_steps[key], cn.n_steps);
        const int per = cn.dense ? 16 : 64;
        for (int64_t c0 = 0; c0 < cn.cells; c0 += per) {
            Tile t;
            t.blob = (int32_t)cn.blob_off;
            t.cell0 = (int32_t)c0;
            t.ncell = (int32_t)std::min<int64_t>(per, cn.cells - c0);
            t.net = i;
            t.s0 = 0;
            t.s1 = INT32_MAX;
            t.pad0 = t.pad1 = 0;
            v.push_back(t);
        }
        if (cn.dense) {
            cn.off_w = w_elems / 8;
            w_elems += cn.wpack.size();
            uint32_t* hb = blobs.data() + cn.blob_off;
            hb[H_DENSE_LO] = (uint32_t)cn.off_w;
            hb[H_DENSE_HI] = (uint32_t)((uint64_t)cn.off_w >> 32);
        }
    }
    // Population plans fix each launch group's stream lane here (largest work
    // first, each to the least loaded lane) and store a lane's tiles together,
    // so that the lane can run the division of its own networks and go on to
    // its next phase without waiting for the other lanes (population_launch).
    std::vector<std::pair<int, GKey>> key_order;
    for (auto& kv : groups) key_order.push_back({0, kv.first});
    if (P->has_population && !arena) {
        auto gcost = [&](const GKey& k) { return (double)groups[k].size() * k.NW * std::max(1, k.a); };
        std::vector<size_t> idx(key_order.size());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(),
                         [&](size_t x, size_t y) { return gcost(key_order[x].second) > gcost(key_order[y].second); });
        const int L = (int)std::min<size_t>(1 + scrna_plan::kSideStreams, key_order.size());
        std::vector<double> load((size_t)std::max(1, L), 0.0);
        for (size_t i : idx) {
            const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            key_order[i].first = l;
            load[l] += gcost(key_order[i].second);
        }
        std::stable_sort(key_order.begin(), key_order.end(),
                         [](const std::pair<int, GKey>& x, const std::pair<int, GKey>& y) { return x.first < y.first; });
    }
    std::vector<Tile> all_tiles;
    for (auto& lk : key_order) {
        const auto& kv = *groups.find(lk.second);
        LaunchGroup lg;
        lg.kind = kv.first.kind;
        lg.NW = kv.first.NW;
        lg.lane = lk.first;
        lg.tile0 = (int)all_tiles.size();
        lg.n_tiles = (int)kv.second.size();
        const int S = std::max(1, group_SG[kv.first].first);
        if (lg.kind == 0 || lg.kind == 2) {
            const bool tl = lg.kind == 2;
            lg.CH = kv.first.a;
            lg.tele = group_tele[kv.first];
            lg.clamp = group_clamp[kv.first];
            const size_t two = tl ? tauleap_lds_bytes(S, 2, lg.NW, lg.CH) : stepper_lds_bytes(S, 2);
            lg.nbuf = two <= ctx->lds_limit_soft ? 2 : 1;
            lg.lds = tl ? tauleap_lds_bytes(S, lg.nbuf, lg.NW, lg.CH) : stepper_lds_bytes(S, lg.nbuf);
            const int per_cu = tl ? tauleap_blocks_per_cu(lg.NW, lg.lds)
                                  : stepper_blocks_per_cu(lg.CH, lg.NW, lg.lds);
            // Population plans advance in division phases that always run one
            // workgroup per tile (launch_steps with segmented = false).
            lg.nseg = P->has_population ? 1
                                        : choose_segments(lg.n_tiles, (long long)per_cu * ctx->num_cu,
                                                          group_steps[kv.first], tl ? 8 : 16);
            if (lg.nseg > 1) {
                lg.sched_off = P->sched_words;
                P->sched_words += sched_words(lg.n_tiles);
            }
        } else {
            lg.mode = kv.first.a;
            lg.tele = kv.first.tele != 0;
            lg.clamp = group_clamp[kv.first];
            lg.inject = kv.first.inj != 0;
            lg.nbuf = 2;
            lg.lds = dense_lds_bytes(group_SG[kv.first].second, S, lg.mode);
        }
        if (lg.lds + (lg.kind != 1 ? kStepperStaticLds : 0) > ctx->lds_max)
            return set_error(SCRNA_ERR_UNSUPPORTED, "%d source genes need %zu B of LDS per workgroup (max %zu)", S,
                             lg.lds, ctx->lds_max);
        all_tiles.insert(all_tiles.end(), kv.second.begin(), kv.second.end());
        P->groups.push_back(lg);
    }

    // Capture and moment jobs.
    std::vector<CaptureJob> cjobs;
    std::vector<MomentJob> mjobs;
    for (int i = 0; i < n; ++i) {
        const CompiledNet& cn = P->nets[i];
        const size_t GC = (size_t)cn.G * (size_t)cn.cells;
        if (cn.capture) {
            CaptureJob j{};
            j.src_off = cn.off_state + (cn.cap.readout == SCRNA_READOUT_PROTEIN ? GC : 0);
            j.dst_off = cn.off_counts;
            j.src_ld = cn.cells;
            j.dst_ld = cn.cells;
            j.cells = (int32_t)cn.cells;
            j.G = cn.G;
            j.mu_log = std::log(cn.cap.efficiency) - cn.cap.efficiency_std * cn.cap.efficiency_std / 2.0;
            j.sigma_log = cn.cap.efficiency_std;
            j.k0 = (uint32_t)cn.seed;
            j.k1 = (uint32_t)(cn.seed >> 32);
            j.net_id = cn.net_id;
            j.src_int = cn.algorithm == SCRNA_ALG_BINOMIAL_TAU_LEAP ? 1u : 0u;
            cjobs.push_back(j);
            P->max_cap_elems = std::max<int64_t>(P->max_cap_elems, (int64_t)GC);
        }
        for (int k = 0; k < cn.n_mom; ++k) {
            MomentJob j{};
            if (cn.n_snap > 0) {
                j.m_off = cn.off_snap + (size_t)(2 * k) * GC;
                j.p_off = cn.off_snap + (size_t)(2 * k + 1) * GC;
                j.m_off |= kSnapPoolBit;
                j.p_off |= kSnapPoolBit;
            } else {
                j.m_off = cn.off_state;
                j.p_off = cn.off_state + GC;
                j.src_int = cn.algorithm == SCRNA_ALG_BINOMIAL_TAU_LEAP ? 1 : 0;   // the final state pool
            }
            j.out_off = cn.off_mom + (size_t)k * cn.G * 5;
            j.cells = (int32_t)cn.cells;
            j.G = cn.G;
            mjobs.push_back(j);
            P->max_mom_G = std::max(P->max_mom_G, cn.G);
        }
    }
    // Moment jobs read either pool; split them by pool.
    for (auto& j : mjobs) {
        if (j.m_off & kSnapPoolBit) {
            j.m_off &= ~kSnapPoolBit;
            j.p_off &= ~kSnapPoolBit;
            P->mjobs_snap.push_back(j);
        } else {
            P->mjobs_state.push_back(j);
        }
    }

    // Device allocations.
    int rc;
    void* ptr;
#define PLAN_ALLOC(field, type, count)                                 \
    do {                                                               \
        rc = plan_alloc(P.get(), &ptr, sizeof(type) * (size_t)(count)); \
        if (rc) {                                                      \
            plan_release(P.get());                                     \
            return rc;                                                 \
        }                                                              \
        P->field = (type*)ptr;                                         \
    } while (0)
    PLAN_ALLOC(d_blobs, uint32_t, blobs.size());
    PLAN_ALLOC(d_tiles, Tile, all_tiles.size());
    P->n_tiles_total = all_tiles.size();
    PLAN_ALLOC(d_state, float, state_elems);
    PLAN_ALLOC(d_snap, float, snap_elems);
    PLAN_ALLOC(d_prom, uint8_t, prom_elems);
    PLAN_ALLOC(d_vol, double, vol_elems);
    PLAN_ALLOC(d_inj, float, inj_elems);
    PLAN_ALLOC(d_counts, int32_t, count_elems);
    PLAN_ALLOC(d_mom, double, mom_elems);
    PLAN_ALLOC(d_cjobs, CaptureJob, cjobs.size());
    PLAN_ALLOC(d_mjobs_snap, MomentJob, P->mjobs_snap.size());
    PLAN_ALLOC(d_mjobs_state, MomentJob, P->mjobs_state.size());
    PLAN_ALLOC(d_wpool, uint16_t, w_elems);
    PLAN_ALLOC(d_sched, int, P->sched_words);
#undef PLAN_ALLOC
    // Every launch clears its block with a kernel first; clearing it here as
    // well means no error word is ever read uninitialised.  (Arena plans read
    // only the blocks their own launch cleared.)
    if (P->d_sched && !arena && hipMemset(P->d_sched, 0, P->sched_words * sizeof(int)) != hipSuccess) {
        plan_release(P.get());
        return set_error(SCRNA_ERR_DEVICE, "hipMemset of the schedule block failed");
    }
    P->wait_ticks = 300000LL * ctx->wall_khz;   // 300 s
    P->n_cjobs = (int)cjobs.size();
    for (auto& u : P->unfused) {
        if ((rc = unfused_prepare(P.get(), u))) {
            plan_release(P.get());
            return rc;
        }
    }
    for (auto& w : P->wide) {
        if ((rc = wide_prepare(P.get(), w))) {
            plan_release(P.get());
            return rc;
        }
    }
    if (P->has_population && (rc = population_prepare(P.get(), state_elems, prom_elems, vol_elems))) {
        plan_release(P.get());
        return rc;
    }
    hipStream_t st = ctx->stream;
    auto up = [&](void* dst, const void* src, size_t bytes) -> int {
        if (bytes == 0) return SCRNA_OK;
        hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return set_error(SCRNA_ERR_DEVICE, "upload failed: %s", hipGetErrorString(e));
        return SCRNA_OK;
    };
    if ((rc = up(P->d_blobs, blobs.data(), blobs.size() * 4)) ||
        (rc = up(P->d_tiles, all_tiles.data(), all_tiles.size() * sizeof(Tile))) ||
        (rc = up(P->d_cjobs, cjobs.data(), cjobs.size() * sizeof(CaptureJob))) ||
        (rc = up(P->d_mjobs_snap, P->mjobs_snap.data(), P->mjobs_snap.size() * sizeof(MomentJob))) ||
        (rc = up(P->d_mjobs_state, P->mjobs_state.data(), P->mjobs_state.size() * sizeof(MomentJob)))) {
        plan_release(P.get());
        return rc;
    }
    for (int i = 0; i < n; ++i) {
        const CompiledNet& cn = P->nets[i];
        if (cn.dense && (rc = up(P->d_wpool + cn.off_w * 8, cn.wpack.data(), cn.wpack.size() * sizeof(uint16_t)))) {
            plan_release(P.get());
            return rc;
        }
    }
    for (int i = 0; i < n; ++i) {
        if (!runs[i].injected_noise) continue;
        const CompiledNet& cn = P->nets[i];
        if ((rc = up(P->d_inj + cn.off_inj, runs[i].injected_noise,
                     (size_t)cn.n_steps * 2 * cn.G * cn.cells * sizeof(float)))) {
            plan_release(P.get());
            return rc;
        }
    }
    hipError_t e = hipStreamSynchronize(st);   // host arrays may be freed after return
    if (e != hipSuccess) {
        plan_release(P.get());
        return set_error(SCRNA_ERR_DEVICE, "upload sync failed: %s", hipGetErrorString(e));
    }
    for (auto& cn : P->nets) P->cell_gene_steps += (double)cn.G * (double)cn.cells * (double)cn.n_steps;
    if (arena) {
        for (auto& g : P->groups) g.lane = 0;
        P->n_lanes = 1;
        *out = P.release();
        return SCRNA_OK;
    }
    if (hipEventCreate(&P->ev_start) != hipSuccess || hipEventCreate(&P->ev_stop) != hipSuccess ||
        hipEventCreateWithFlags(&P->ev_done, hipEventDisableTiming) != hipSuccess) {
        plan_release(P.get());
        return set_error(SCRNA_ERR_DEVICE, "hipEventCreate failed");
    }
    // Stream lanes for the launch groups: largest work first, each to the
    // least loaded lane, so that small groups fill the device beside the
    // large ones instead of running after them with a mostly idle tail.
    {
        std::vector<int> order(P->groups.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
        auto cost = [&](int i) {
            const LaunchGroup& g = P->groups[i];
            return (double)g.n_tiles * (double)g.NW * (double)std::max(1, g.CH);
        };
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
        if (P->has_population) {
            // lanes were fixed per launch group when the groups were formed
            P->n_lanes = 1;
            for (const auto& g : P->groups) P->n_lanes = std::max(P->n_lanes, g.lane + 1);
            for (const auto& g : P->groups) {
                if (P->lane_nt[g.lane] == 0) P->lane_t0[g.lane] = g.tile0;
                P->lane_nt[g.lane] += g.n_tiles;
            }
        } else {
            P->n_lanes = (int)std::min<size_t>(1 + scrna_plan::kSideStreams, P->groups.size());
            std::vector<double> load((size_t)std::max(1, P->n_lanes), 0.0);
            for (int i : order) {
                const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                P->groups[i].lane = l;
                load[l] += cost(i);
            }
        }
        bool ok = true;
        for (int i = 0; ok && i + 1 < P->n_lanes; ++i)
            ok = hipStreamCreateWithFlags(&P->side[i], hipStreamNonBlocking) == hipSuccess &&
                 hipEventCreateWithFlags(&P->ev_join[i], hipEventDisableTiming) == hipSuccess;
        if (ok && P->n_lanes > 1) ok = hipEventCreateWithFlags(&P->ev_fork, hipEventDisableTiming) == hipSuccess;
        if (!ok) {
            plan_release(P.get());
            return set_error(SCRNA_ERR_DEVICE, "stream or event creation failed");
        }
    }
    *out = P.release();
    return SCRNA_OK;
}

// Segments per tile for a whole-horizon stepper launch.  With one workgroup
// per tile the device runs ceil(n_tiles / resident) rounds of a full horizon:
// the last round is often nearly empty, and even a single full round leaves
// SIMDs idle because a workgroup's waves load the four SIMDs unevenly and
// nothing refills them.  k segments per tile cut the work quantum to 1/k of
// the horizon and let the dispatcher refill freed slots: the makespan is
// about n_tiles/resident + 1/k full horizons.  Measured on the G=20 benchmark
// network (1536 resident workgroups): 1563 tiles took 33.8 ms with k = 8 and
// 33.2 ms with k = 16, while 1536 tiles in one unsegmented round took 37.1 ms.  So:
// segment whenever the tiles fill at least half the device, with the largest
// k <= max_k that keeps >= 64 steps per segment.  Below that, segments of a
// tile would only run one after another.  The tau-leap stepper (one 7-wave
// workgroup per CU) measured best at max_k = 8 (331 ms) rather than 16 (344 ms).
// Past kMaxSegmentedWaves rounds the dispatcher already refills slots as
// workgroups end and the last round is a small share, so segments only add
// their start-up and state hand-over: 1024 uneven networks (36k tiles) over
// 500 steps took 106.6 ms with 7 segments per tile and 88 ms with none.
constexpr long long kMaxSegmentedWaves = 4;

int choose_segments(int n_tiles, long long resident, int64_t steps, int max_k) {
    if (resident <= 0 || 2LL * n_tiles < resident) return 1;
    if ((n_tiles + resident - 1) / resident > kMaxSegmentedWaves) return 1;
    const int64_t k = std::min<int64_t>(max_k, steps / 64);
    return k < 1 ? 1 : (int)k;
}

// The dispatch a launch group makes, checked on the host before it is
// enqueued (or recorded into a graph): a workgroup of 64..1024 threads, a
// grid of 1..2^31-1 workgroups, dynamic plus static LDS within what the
// device grants a workgroup, and a stream lane that exists.  The plan builder
// only makes such groups; the check keeps a malformed dispatch packet from
// ever reaching the device, also from a graph that replays it many times.
static int check_dispatch(const scrna_plan* P, const LaunchGroup& lg, int nseg) {
    const int threads = 64 * lg.NW;
    const long long grid = (long long)lg.n_tiles * nseg;
    const size_t lds = lg.lds + (lg.kind != 1 ? kStepperStaticLds : 0);
    if (threads < 64 || threads > 1024 || grid < 1 || grid > INT32_MAX || lds > P->ctx->lds_max ||
        lg.lane < 0 || lg.lane >= std::max(P->n_lanes, 1) || lg.tile0 < 0 ||
        (size_t)lg.tile0 + (size_t)lg.n_tiles > P->n_tiles_total)
        return set_error(SCRNA_ERR_DEVICE,
                         "refusing a malformed dispatch: kind %d, %d threads, %lld workgroups, %zu B LDS (limit %zu), "
                         "lane %d of %d", lg.kind, threads, grid, lds, P->ctx->lds_max, lg.lane, P->n_lanes);
    return SCRNA_OK;
}

// Run steps [s_begin, s_end) of every network (clamped to each horizon).
// A tile resumes from the state pool exactly when its first step is > 0.
// `segmented` enables the segmented schedule of the FP32 and tau-leap stepper groups.
int launch_group_steps(scrna_plan* P, const LaunchGroup& lg, hipStream_t gs, int s_begin, int s_end,
                       bool segmented) {
    int rc = check_dispatch(P, lg, segmented && lg.kind != 1 ? lg.nseg : 1);
    if (rc) return rc;
    if (lg.kind == 1) {
        HIP_TRY(launch_dense(lg.mode, lg.tele, lg.inject, lg.clamp, lg.NW, lg.n_tiles, P->d_blobs, P->d_tiles + lg.tile0,
                             P->d_wpool, P->d_state, P->d_snap, P->d_prom, P->d_inj, s_begin, s_end, lg.lds, gs));
        return SCRNA_OK;
    }
    int* sched = nullptr;
    int nseg = 1;
    if (segmented && lg.nseg > 1) {
        sched = P->d_sched + lg.sched_off;
        nseg = lg.nseg;
        HIP_TRY(launch_sched_reset(sched, (int)sched_words(lg.n_tiles), gs));
    }
    if (lg.kind == 0)
        HIP_TRY(launch_stepper(lg.CH, lg.NW, lg.n_tiles, P->d_blobs, P->d_tiles + lg.tile0, P->d_state, P->d_snap,
                               P->d_prom, P->d_vol, P->d_inj, s_begin, s_end, lg.nbuf, lg.lds, sched, nseg,
                               P->wait_ticks, lg.tele, lg.clamp, gs));
    else
        HIP_TRY(launch_tauleap(lg.CH, lg.NW, lg.n_tiles, P->d_blobs, P->d_tiles + lg.tile0, P->d_state, P->d_snap,
                               P->d_prom, P->d_vol, s_begin, s_end, lg.nbuf, lg.lds, sched, nseg, P->wait_ticks,
                               gs));
    return SCRNA_OK;
}

int launch_steps(scrna_plan* P, hipStream_t st, int s_begin, int s_end, bool segmented) {
    for (const LaunchGroup& lg : P->groups) {
        int rc = check_dispatch(P, lg, segmented && lg.kind != 1 ? lg.nseg : 1);
        if (rc) return rc;
    }
    if (P->n_lanes > 1) {
        HIP_TRY(hipEventRecord(P->ev_fork, st));
        for (int i = 0; i + 1 < P->n_lanes; ++i) HIP_TRY(hipStreamWaitEvent(P->side[i], P->ev_fork, 0));
    }
    for (const LaunchGroup& lg : P->groups) {
        hipStream_t gs = lg.lane == 0 ? st : P->side[lg.lane - 1];
        int rc = launch_group_steps(P, lg, gs, s_begin, s_end, segmented);
        if (rc) return rc;
    }
    for (int i = 0; i + 1 < P->n_lanes; ++i) {
        HIP_TRY(hipEventRecord(P->ev_join[i], P->side[i]));
        HIP_TRY(hipStreamWaitEvent(st, P->ev_join[i], 0));
    }
    return SCRNA_OK;
}

// Every kernel of one simulation, in stream order (also the body of the graph).
int plan_enqueue(scrna_plan* P, hipStream_t st, bool events) {
    if (events) HIP_TRY(hipEventRecord(P->ev_start, st));
    if (P->has_population) {
        int rc = population_launch(P, st);
        if (rc) return rc;
    } else {
        int rc = launch_steps(P, st, 0, INT32_MAX, true);
        if (rc) return rc;
        for (auto& u : P->unfused)
            if ((rc = unfused_launch(P, u, st))) return rc;
        for (auto& w : P->wide)
            if ((rc = wide_launch(P, w, st))) return rc;
    }
    if (events) HIP_TRY(hipEventRecord(P->ev_stop, st));
    HIP_TRY(launch_moments((int)P->mjobs_snap.size(), P->max_mom_G, P->d_mjobs_snap, P->d_snap, P->d_mom, st));
    HIP_TRY(launch_moments((int)P->mjobs_state.size(), P->max_mom_G, P->d_mjobs_state, P->d_state, P->d_mom, st));
    HIP_TRY(launch_capture_f32(P->n_cjobs, P->max_cap_elems, P->d_cjobs, P->d_state, P->d_counts, st));
    return SCRNA_OK;
}

int plan_launch(scrna_plan* P, hipStream_t st) {
    if (!st) st = P->ctx->stream;
    HIP_TRY(hipSetDevice(P->ctx->device));
    if (P->graph_exec) {
        // Graph replay: the events bracket the whole graph (stepper kernels dominate it).
        HIP_TRY(hipEventRecord(P->ev_start, st));
        HIP_TRY(hipGraphLaunch(P->graph_exec, st));
        HIP_TRY(hipEventRecord(P->ev_stop, st));
    } else {
        int rc = plan_enqueue(P, st, true);
        if (rc) return rc;
    }
    HIP_TRY(hipEventRecord(P->ev_done, st));
    P->launched = true;
    return SCRNA_OK;
}

// After a finished launch: the error word of every segmented launch group.
int plan_check_schedule(scrna_plan* P) {
    for (const auto& lg : P->groups) {
        if (lg.kind == 1 || lg.nseg <= 1) continue;
        int err = 0;
        HIP_TRY(hipMemcpy(&err, P->d_sched + lg.sched_off + SCHED_ERROR, sizeof(int), hipMemcpyDeviceToHost));
        if (err) return set_error(SCRNA_ERR_DEVICE, "stepper segment waited past its time bound");
    }
    return SCRNA_OK;
}

// Capture the plan's kernel sequence once into a hipGraph; later launches
// replay it with a single hipGraphLaunch.  No memset nodes are captured.
int plan_make_graph(scrna_plan* P) {
    if (P->graph_exec) return SCRNA_OK;
    if (!P->unfused.empty() || !P->wide.empty())
        return set_error(SCRNA_ERR_UNSUPPORTED,
                         "the reference-structured and wide paths (library GEMMs) are not captured into graphs");
    HIP_TRY(hipSetDevice(P->ctx->device));
    hipStream_t cs = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    int rc = SCRNA_OK;
    if (e == hipSuccess) {
        rc = plan_enqueue(P, cs, false);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(cs, &g);
        if (rc == SCRNA_OK && e2 == hipSuccess) {
            e = hipGraphInstantiate(&P->graph_exec, g, nullptr, nullptr, 0);
            P->graph = g;
        } else if (rc == SCRNA_OK) {
            e = e2;
        }
    }
    (void)hipStreamDestroy(cs);
    if (rc) return rc;
    if (e != hipSuccess) {
        P->graph_exec = nullptr;
        return set_error(SCRNA_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
    }
    return SCRNA_OK;
}

int plan_download(scrna_plan* P, int32_t i, scrna_output* out) {
    if (i < 0 || i >= (int32_t)P->nets.size()) return set_error(SCRNA_ERR_INVALID, "network index %d out of range", i);
    if (!out) return set_error(SCRNA_ERR_INVALID, "null output");
    const CompiledNet& cn = P->nets[i];
    const size_t C = (size_t)cn.cells, G = (size_t)cn.G, GC = G * C;
    HIP_TRY(hipSetDevice(P->ctx->device));
    if (P->launched) HIP_TRY(hipEventSynchronize(P->ev_done));   // the copies below use the null stream
    if (out->y) {
        const int64_t ld = out->ld > 0 ? out->ld : (int64_t)C;
        if (ld < (int64_t)C) return set_error(SCRNA_ERR_INVALID, "output leading dimension %lld < cells", (long long)ld);
        std::vector<float> h(2 * GC);
        HIP_TRY(hipMemcpy(h.data(), P->d_state + cn.off_state, 2 * GC * sizeof(float), hipMemcpyDeviceToHost));
        const bool counts = cn.algorithm == SCRNA_ALG_BINOMIAL_TAU_LEAP;   // int32 counts in the pool
        // cle.jl:123-129 -- columns are genes; :both puts mRNA first.
        const int ncol = cn.readout == SCRNA_READOUT_BOTH ? 2 * (int)G : (int)G;
        for (int col = 0; col < ncol; ++col) {
            size_t src_row;
            if (cn.readout == SCRNA_READOUT_MRNA) src_row = col;
            else if (cn.readout == SCRNA_READOUT_PROTEIN) src_row = G + col;
            else src_row = col;   // both: rows 0..G-1 mRNA, G..2G-1 protein
            const float* s = h.data() + src_row * C;
            double* d = out->y + (size_t)col * ld;
            if (counts) {
                const int32_t* si = (const int32_t*)s;
                for (size_t c = 0; c < C; ++c) d[c] = (double)si[c];
            } else {
                for (size_t c = 0; c < C; ++c) d[c] = (double)s[c];
            }
        }
    }
    if (out->counts) {
        if (!cn.capture) return set_error(SCRNA_ERR_INVALID, "network %d was run without capture", i);
        const int64_t ld = out->ld > 0 ? out->ld : (int64_t)C;
        if (ld == (int64_t)C) {
            HIP_TRY(hipMemcpy(out->counts, P->d_counts + cn.off_counts, GC * sizeof(int32_t), hipMemcpyDeviceToHost));
        } else {
            for (size_t g = 0; g < G; ++g)
                HIP_TRY(hipMemcpy(out->counts + g * ld, P->d_counts + cn.off_counts + g * C, C * sizeof(int32_t),
                                  hipMemcpyDeviceToHost));
        }
    }
    if (out->snapshots && cn.n_snap > 0)
        HIP_TRY(hipMemcpy(out->snapshots, P->d_snap + cn.off_snap, (size_t)cn.n_snap * 2 * GC * sizeof(float),
                          hipMemcpyDeviceToHost));
    if (out->moments)
        HIP_TRY(hipMemcpy(out->moments, P->d_mom + cn.off_mom, (size_t)cn.n_mom * G * 5 * sizeof(double),
                          hipMemcpyDeviceToHost));
    if (out->volumes && cn.population) {
        HIP_TRY(hipMemcpy(out->volumes, P->d_vol + cn.off_vol, C * sizeof(double), hipMemcpyDeviceToHost));
    }
    return SCRNA_OK;
}

}  // namespace scrna

// ===========================================================================
// C ABI

using namespace scrna;

extern "C" {

int32_t scrna_abi_version(void) { return SCRNA_ABI_VERSION; }

const char* scrna_version(void) { return "libscrna 0.1.0 (gfx950)"; }

const char* scrna_last_error(void) { return g_last_error.c_str(); }

int scrna_ctx_create(int32_t device, uint64_t hbm_budget_bytes, scrna_ctx** out) {
    if (!out) return set_error(SCRNA_ERR_INVALID, "null output pointer");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
        return set_error(SCRNA_ERR_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= count) return set_error(SCRNA_ERR_INVALID, "device %d out of range [0,%d)", device, count);
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_error(SCRNA_ERR_DEVICE, "libscrna is built for gfx950, device %d is %s", device, prop.gcnArchName);
    std::unique_ptr<scrna_ctx> C(new scrna_ctx());
    C->device = device;
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    C->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)(0.9 * (double)free_b);
    int lds = 0;
    HIP_TRY(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    C->lds_max = (size_t)lds;
    C->lds_limit_soft = std::min<size_t>(C->lds_max, 64 * 1024);
    C->num_cu = prop.multiProcessorCount;
    int wall_khz = 0;
    if (hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && wall_khz > 0)
        C->wall_khz = wall_khz;
    if (C->lds_max > 65536 && (set_stepper_lds_limit(C->lds_max) != hipSuccess ||
                               set_dense_lds_limit(C->lds_max) != hipSuccess ||
                               set_tauleap_lds_limit(C->lds_max) != hipSuccess)) {
        (void)hipGetLastError();
        C->lds_max = 65536;   // large dynamic LDS not granted: stay within the default limit
    }
    C->lds_limit_soft = std::min(C->lds_limit_soft, C->lds_max);
    HIP_TRY(hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking));
    noalloc_scratch_init(device);   // this device's address of the allocation-free entry point's array
    *out = C.release();
    return SCRNA_OK;
}

int scrna_ctx_destroy(scrna_ctx* ctx) {
    if (!ctx) return SCRNA_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
    return SCRNA_OK;
}

int scrna_plan_create(scrna_ctx* ctx, int32_t n, const scrna_network* nets, const scrna_kinetics* kins,
                      const scrna_run* runs, const scrna_population* pops, const scrna_capture* captures,
                      scrna_plan** out) {
    if (!out) return set_error(SCRNA_ERR_INVALID, "null output pointer");
    *out = nullptr;
    if (!ctx) return set_error(SCRNA_ERR_INVALID, "null context");
    std::lock_guard<std::mutex> lk(ctx->mu);
    return plan_build(ctx, n, nets, kins, runs, pops, captures, 0, out);
}

int scrna_plan_launch(scrna_plan* plan, void* stream) {
    if (!plan) return set_error(SCRNA_ERR_INVALID, "null plan");
    return plan_launch(plan, (hipStream_t)stream);
}

int scrna_plan_sync(scrna_plan* plan) {
    if (!plan) return set_error(SCRNA_ERR_INVALID, "null plan");
    HIP_TRY(hipSetDevice(plan->ctx->device));
    // The launch stream's last event covers the side lanes too (they join it
    // before the moments and capture kernels): other work on the device is
    // not waited for.
    HIP_TRY(hipEventSynchronize(plan->ev_done));
    return plan_check_schedule(plan);
}

int scrna_plan_download(scrna_plan* plan, int32_t i, scrna_output* out) {
    if (!plan) return set_error(SCRNA_ERR_INVALID, "null plan");
    return plan_download(plan, i, out);
}

int scrna_plan_stepper_ms(scrna_plan* plan, double* ms, double* cgs) {
    if (!plan || !ms) return set_error(SCRNA_ERR_INVALID, "null argument");
    if (!plan->launched) return set_error(SCRNA_ERR_INVALID, "plan has not been launched");
    float f = 0.0f;
    HIP_TRY(hipEventSynchronize(plan->ev_stop));
    HIP_TRY(hipEventElapsedTime(&f, plan->ev_start, plan->ev_stop));
    *ms = (double)f;
    if (cgs) *cgs = plan->cell_gene_steps;
    return SCRNA_OK;
}

int scrna_plan_bytes(scrna_plan* plan, uint64_t* bytes) {
    if (!plan || !bytes) return set_error(SCRNA_ERR_INVALID, "null argument");
    *bytes = plan->bytes;
    return SCRNA_OK;
}

int scrna_plan_use_graph(scrna_plan* plan) {
    if (!plan) return set_error(SCRNA_ERR_INVALID, "null plan");
    std::lock_guard<std::mutex> lk(plan->ctx->mu);
    return plan_make_graph(plan);
}

int scrna_plan_destroy(scrna_plan* plan) {
    if (!plan) return SCRNA_OK;
    (void)hipSetDevice(plan->ctx->device);
    if (plan->launched && plan->ev_done) (void)hipEventSynchronize(plan->ev_done);
    plan_release(plan);
    delete plan;
    return SCRNA_OK;
}

int scrna_simulate_cle_batch(scrna_ctx* ctx, int32_t n, const scrna_network* nets, const scrna_kinetics* kins,
                             const scrna_run* runs, const scrna_population* pops, const scrna_capture* captures,
                             scrna_output* outs) {
    if (!ctx) return set_error(SCRNA_ERR_INVALID, "null context");
    if (!outs) return set_error(SCRNA_ERR_INVALID, "null outputs");
    scrna_plan* plan = nullptr;
    int rc = scrna_plan_create(ctx, n, nets, kins, runs, pops, captures, &plan);
    if (rc) return rc;
    rc = scrna_plan_launch(plan, nullptr);
    if (!rc) rc = scrna_plan_sync(plan);
    for (int i = 0; !rc && i < n; ++i) rc = scrna_plan_download(plan, i, &outs[i]);
    std::string err = rc ? g_last_error : std::string();
    scrna_plan_destroy(plan);
    if (rc) g_last_error = err;
    return rc;
}

int scrna_simulate_cle(scrna_ctx* ctx, const scrna_network* net, const scrna_kinetics* kin, const scrna_run* run,
                       const scrna_population* pop, const scrna_capture* capture, scrna_output* out) {
    return scrna_simulate_cle_batch(ctx, 1, net, kin, run, pop, capture, out);
}

int scrna_apply_capture(scrna_ctx* ctx, const double* X, int64_t cells, int32_t G, int64_t ld, double efficiency,
                        double efficiency_std, uint64_t seed, uint32_t net_id, int32_t* Y) {
    if (!ctx || !X || !Y) return set_error(SCRNA_ERR_INVALID, "null argument");
    if (cells <= 0 || G <= 0 || ld < cells || cells > INT32_MAX)
        return set_error(SCRNA_ERR_INVALID, "invalid shape cells=%lld G=%d ld=%lld", (long long)cells, G, (long long)ld);
    if (!(efficiency > 0.0) || !(efficiency_std >= 0.0) || !std::isfinite(efficiency) || !std::isfinite(efficiency_std))
        return set_error(SCRNA_ERR_INVALID, "capture efficiency must be > 0 and efficiency_std >= 0");
    if (net_id >= (1u << 28)) return set_error(SCRNA_ERR_INVALID, "net_id must be < 2^28");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const size_t GC = (size_t)G * (size_t)cells;
    void *dX = nullptr, *dY = nullptr, *dJ = nullptr;
    int rc = ctx_alloc(ctx, &dX, GC * sizeof(double));
    if (!rc) rc = ctx_alloc(ctx, &dY, GC * sizeof(int32_t));
    if (!rc) rc = ctx_alloc(ctx, &dJ, sizeof(CaptureJob));
    auto cleanup = [&]() {
        ctx_free(ctx, dX, GC * sizeof(double));
        ctx_free(ctx, dY, GC * sizeof(int32_t));
        ctx_free(ctx, dJ, sizeof(CaptureJob));
    };
    if (rc) {
        cleanup();
        return rc;
    }
    CaptureJob j{};
    j.src_off = 0;
    j.dst_off = 0;
    j.src_ld = cells;
    j.dst_ld = cells;
    j.cells = (int32_t)cells;
    j.G = G;
    j.mu_log = std::log(efficiency) - efficiency_std * efficiency_std / 2.0;   // capture.jl:30
    j.sigma_log = efficiency_std;
    j.k0 = (uint32_t)seed;
    j.k1 = (uint32_t)(seed >> 32);
    j.net_id = net_id;
    hipError_t e = hipSuccess;
    for (int g = 0; g < G && e == hipSuccess; ++g)
        e = hipMemcpy((double*)dX + (size_t)g * cells, X + (size_t)g * ld, cells * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dJ, &j, sizeof(j), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_capture_f64(1, (int64_t)GC, (CaptureJob*)dJ, (double*)dX, (int32_t*)dY, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    for (int g = 0; g < G && e == hipSuccess; ++g)
        e = hipMemcpy(Y + (size_t)g * ld, (int32_t*)dY + (size_t)g * cells, cells * sizeof(int32_t), hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) return set_error(SCRNA_ERR_DEVICE, "capture failed: %s", hipGetErrorString(e));
    return SCRNA_OK;
}

}  // extern "C"
