// Network prior sampler and theta constructor for bulk simulation-based
// inference batches (src/kinetics.jl:105-205, src/network.jl:233-256).
//
// The sampler follows sample_network's generative structure: LogNormal basals,
// a uniform number of active TFs drawn without replacement, a uniform number of
// targets per TF with strengths U([-10,-3] u [3,10]), then AND / OR pair edges
// whose (target, source) pairs avoid every pair already used, and for :mixed at
// least one non-additive edge.  Julia's MersenneTwister / StatsBase streams
// cannot be reproduced, so the draws come from Philox4x32-10 keyed by the seed
// with counter (network index, block, 0, purpose 9); each block gives two
// 52-bit uniforms consumed in a fixed order, which makes network `index` of a
// seed reproducible on its own (no shared state across networks).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <set>
#include <utility>
#include <vector>

#include "runtime_internal.h"

namespace scrna {

namespace {

constexpr uint32_t kPriorPurpose = 9;

void host_philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Sequential 52-bit uniforms in (0,1) of one network's stream.
class PriorStream {
  public:
    PriorStream(uint64_t seed, uint32_t index) : k0_((uint32_t)seed), k1_((uint32_t)(seed >> 32)), index_(index) {}
    double uniform() {
        if (pos_ >= 2) {
            u_[0] = index_;
            u_[1] = block_++;
            u_[2] = 0u;
            u_[3] = kPriorPurpose;
            host_philox(u_, k0_, k1_);
            pos_ = 0;
        }
        const uint64_t x = ((uint64_t)u_[2 * pos_] << 20) | (uint64_t)(u_[2 * pos_ + 1] >> 12);
        ++pos_;
        return ((double)x + 0.5) * 0x1p-52;
    }
    // Uniform integer in [lo, hi].
    int range(int lo, int hi) {
        const double u = uniform();
        int v = lo + (int)std::floor(u * (double)(hi - lo + 1));
        return std::min(v, hi);
    }
    // Standard normal (Box-Muller on two uniforms; the second of the pair is
    // kept for the next call).
    double normal() {
        if (have_spare_) {
            have_spare_ = false;
            return spare_;
        }
        const double u1 = uniform(), u2 = uniform();
        const double r = std::sqrt(-2.0 * std::log(u1));
        spare_ = r * std::sin(6.283185307179586 * u2);
        have_spare_ = true;
        return r * std::cos(6.283185307179586 * u2);
    }
    // Edge strength from U([-10,-3] u [3,10]) (kinetics.jl:140-144).
    double strength() {
        const bool neg = uniform() < 0.5;
        const double u = uniform();
        return neg ? -10.0 + 7.0 * u : 3.0 + 7.0 * u;
    }
    // k distinct values of `pool`, in selection order (partial Fisher-Yates).
    std::vector<int> choose(std::vector<int> pool, int k) {
        const int n = (int)pool.size();
        k = std::min(k, n);
        for (int i = 0; i < k; ++i) {
            const int j = i + (int)std::min<double>(std::floor(uniform() * (double)(n - i)), (double)(n - i - 1));
            std::swap(pool[i], pool[j]);
        }
        pool.resize(k);
        return pool;
    }

  private:
    uint32_t k0_, k1_, index_;
    uint32_t block_ = 0;
    uint32_t u_[4] = {0, 0, 0, 0};
    int pos_ = 2;
    bool have_spare_ = false;
    double spare_ = 0.0;
};

struct PairEdge {
    int target, s1, s2;
    double strength;
};

// _sample_pair_edge (kinetics.jl:183-205): two distinct sorted sources and a
// target outside them, both (target, source) pairs unused; 20 attempts.
bool pair_edge(int G, std::set<std::pair<int, int>>& used, PriorStream& rs, PairEdge* out) {
    if (G < 3) return false;
    std::vector<int> all(G);
    for (int g = 0; g < G; ++g) all[g] = g;
    for (int attempt = 0; attempt < 20; ++attempt) {
        std::vector<int> src = rs.choose(all, 2);
        std::sort(src.begin(), src.end());
        std::vector<int> targets;
        for (int g = 0; g < G; ++g)
            if (g != src[0] && g != src[1]) targets.push_back(g);
        const int t = targets[(size_t)rs.range(0, (int)targets.size() - 1)];
        const std::pair<int, int> p1{t, src[0]}, p2{t, src[1]};
        if (!used.count(p1) && !used.count(p2)) {
            used.insert(p1);
            used.insert(p2);
            *out = PairEdge{t, src[0], src[1], rs.strength()};
            return true;
        }
    }
    return false;
}

}  // namespace

}  // namespace scrna

using namespace scrna;

static void prior_defaults(const scrna_prior* in, int G, int r[8]) {
    // kinetics.jl:107-111
    const int d[8] = {G <= 5 ? 1 : G <= 10 ? 2 : 4, G <= 5 ? 2 : G <= 10 ? 3 : 5,
                      G <= 5 ? 1 : G <= 10 ? 2 : 3, G <= 5 ? 3 : G <= 10 ? 4 : 6,
                      G > 10 ? 1 : 0,             G > 10 ? 2 : 1,
                      G > 10 ? 1 : 0,             G > 10 ? 2 : 1};
    const int v[8] = {in->n_tf_lo, in->n_tf_hi, in->targets_lo, in->targets_hi,
                      in->coop_lo, in->coop_hi, in->redun_lo, in->redun_hi};
    for (int i = 0; i < 8; i += 2) {
        const bool given = v[i] >= 0 && v[i + 1] >= v[i];
        r[i] = given ? v[i] : d[i];
        r[i + 1] = given ? v[i + 1] : d[i + 1];
    }
}

int scrna_prior_max_gates(const scrna_prior* prior, int32_t* max_gates) {
    if (!prior || !max_gates) return set_error(SCRNA_ERR_INVALID, "null argument");
    if (prior->G < 1 || prior->G > kMaxGenes) return set_error(SCRNA_ERR_INVALID, "G must be in [1, %d]", kMaxGenes);
    int r[8];
    prior_defaults(prior, prior->G, r);
    *max_gates = r[5] + r[7] + 1;
    return SCRNA_OK;
}

int scrna_sample_network(const scrna_prior* prior, uint64_t seed, uint32_t index, double* basals, double* A,
                         int32_t* gate_type, int32_t* gate_target, int32_t* gate_src, double* gate_strength,
                         int32_t max_gates, int32_t* n_gates) {
    if (!prior || !basals || !A || !n_gates) return set_error(SCRNA_ERR_INVALID, "null argument");
    const int G = prior->G;
    if (G < 1 || G > kMaxGenes) return set_error(SCRNA_ERR_INVALID, "G must be in [1, %d]", kMaxGenes);
    if (prior->regulation < SCRNA_REG_MIXED || prior->regulation > SCRNA_REG_REDUNDANT)
        return set_error(SCRNA_ERR_INVALID, "unknown regulation %d", prior->regulation);
    int r[8];
    prior_defaults(prior, G, r);
    if (max_gates < r[5] + r[7] + 1 || (max_gates > 0 && (!gate_type || !gate_target || !gate_src || !gate_strength)))
        return set_error(SCRNA_ERR_INVALID, "gate buffers must hold scrna_prior_max_gates() = %d gates",
                         r[5] + r[7] + 1);
    PriorStream rs(seed, index);
    // Basals (kinetics.jl:115-119): exp(0.5 z - 1), z ~ N(0,1), or exp(-1).
    for (int g = 0; g < G; ++g) basals[g] = prior->infer_basals ? std::exp(rs.normal() * 0.5 - 1.0) : std::exp(-1.0);
    std::fill(A, A + (size_t)G * G, 0.0);
    // TF-level sparsity (kinetics.jl:121-147).
    std::vector<int> all(G);
    for (int g = 0; g < G; ++g) all[g] = g;
    const int n_tf = std::min(rs.range(r[0], r[1]), G);
    std::vector<int> tfs = rs.choose(all, n_tf);
    std::sort(tfs.begin(), tfs.end());
    std::set<std::pair<int, int>> used;
    for (int tf : tfs) {
        const int want = rs.range(r[2], r[3]);
        std::vector<int> possible;
        for (int g = 0; g < G; ++g)
            if (g != tf) possible.push_back(g);
        for (int t : rs.choose(possible, want)) {
            A[(size_t)t + (size_t)tf * G] = rs.strength();   // A[target, tf], column-major
            used.insert({t, tf});
        }
    }
    // AND / OR pair edges (kinetics.jl:150-176).
    int ng = 0;
    auto add = [&](int type, const PairEdge& e) {
        gate_type[ng] = type;
        gate_target[ng] = e.target;
        gate_src[2 * ng] = e.s1;
        gate_src[2 * ng + 1] = e.s2;
        gate_strength[ng] = e.strength;
        ++ng;
    };
    const int reg = prior->regulation;
    PairEdge e;
    if (reg == SCRNA_REG_MIXED || reg == SCRNA_REG_COOPERATIVE) {
        const int n = rs.range(r[4], r[5]);
        for (int i = 0; i < n; ++i)
            if (pair_edge(G, used, rs, &e)) add(0, e);
    }
    if (reg == SCRNA_REG_MIXED || reg == SCRNA_REG_REDUNDANT) {
        const int n = rs.range(r[6], r[7]);
        for (int i = 0; i < n; ++i)
            if (pair_edge(G, used, rs, &e)) add(1, e);
    }
    if (reg == SCRNA_REG_MIXED && ng == 0 && pair_edge(G, used, rs, &e)) add(1, e);   // kinetics.jl:169-175
    *n_gates = ng;
    return SCRNA_OK;
}

int scrna_network_from_theta(int32_t G, const double* theta, int64_t n_theta, double* basals, double* A) {
    if (!theta || !basals || !A) return set_error(SCRNA_ERR_INVALID, "null argument");
    if (G < 1 || G > kMaxGenes) return set_error(SCRNA_ERR_INVALID, "G must be in [1, %d]", kMaxGenes);
    const int64_t n_int = (int64_t)G * (G - 1);
    const double* a = nullptr;
    if (n_theta == n_int + G) {                                  // network.jl:238-240
        for (int g = 0; g < G; ++g) basals[g] = std::exp(theta[g]);
        a = theta + G;
    } else if (n_theta == n_int) {                               // network.jl:241-243
        for (int g = 0; g < G; ++g) basals[g] = std::exp(-1.0);
        a = theta;
    } else {
        return set_error(SCRNA_ERR_INVALID, "theta must have length %lld or %lld, got %lld", (long long)n_int,
                         (long long)(n_int + G), (long long)n_theta);
    }
    int64_t idx = 0;                                             // row-major off-diagonal (network.jl:248-256)
    for (int i = 0; i < G; ++i)
        for (int j = 0; j < G; ++j) A[(size_t)i + (size_t)j * G] = (i == j) ? 0.0 : a[idx++];
    return SCRNA_OK;
}
