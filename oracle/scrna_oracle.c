/*
 * CPU FP64 oracle for the GPU CLE path -- TEST INFRASTRUCTURE ONLY (the
 * checker; see scrna_oracle.h for the list of reference files it restates
 * and for who may load it).
 *
 * Every function below cites the reference lines it follows.  The random
 * stream layout (Philox counters) is this project's own and is documented
 * at orc_stream_* below; the GPU library draws the same stream.
 */
#include "scrna_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).                      */
/* ---------------------------------------------------------------------- */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * (uint64_t)c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += PHILOX_W0; k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Stream purposes (low 4 bits of counter word 3; the net id fills the rest). */
enum {
    ORC_P_NORMAL = 0,      /* ctr = (cell, step, gene>>1): two genes' (xi_m, xi_p); low 9 bits of
                              each word: the gene's CLE switching bits (orc_cle_flip) */
    ORC_P_SWITCH = 1,      /* ctr = (cell, step, gene>>2): one u32 per gene (tau-leap switching;
                              14 extension bits of the CLE switch) */
    ORC_P_PROM_INIT = 2,   /* ctr = (cell, 0, gene>>2) */
    ORC_P_CAPTURE_EFF = 3, /* ctr = (cell, 0, 0) */
    ORC_P_CAPTURE_BIN = 4, /* ctr = (cell, gene, block) */
    ORC_P_PARTITION = 5,   /* ctr = (mother, step, block), per (species,gene) sub-stream in word 3 bits */
    ORC_P_REPLACE = 6,     /* ctr = (mother, step, 0) */
    ORC_P_VINIT = 7,       /* ctr = (cell, 0, 0) */
    ORC_P_TAULEAP = 8,     /* ctr = (cell, step, (gene << 16) + block): one 52-bit uniform stream per
                              (cell, step, gene), read by transcription, mRNA decay, translation and
                              protein decay in that order */
    ORC_P_GROWTH = 9       /* ctr = (cell, step, 0): the volume growth noise of one cell and step */
};

static inline void orc_draw(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                            uint32_t out[4]) {
    uint32_t ctr[4] = {c0, c1, c2, c3};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr, key, out);
}

/* Box-Muller on two u32: radius uniform has 23+1 exact bits in (0,1), angle
 * 23 exact bits in revolutions [0,1).  Both are exactly representable in
 * float, so the GPU's float transform and this double one see identical
 * uniforms. */
void orc_normal_pair(uint32_t a, uint32_t b, double* z0, double* z1) {
    double u1 = ((double)(a >> 9) + 0.5) * (1.0 / 8388608.0);   /* 2^-23 */
    double th = (double)(b >> 9) * (1.0 / 8388608.0);           /* 2^-23 */
    double r = sqrt(-2.0 * log(u1));
    double ang = 2.0 * M_PI * th;
    *z0 = r * cos(ang);
    *z1 = r * sin(ang);
}

/* 52-bit uniform in (0,1) from two u32 (capture / population draws). */
static inline double orc_u52(uint32_t a, uint32_t b) {
    uint64_t x = ((uint64_t)a << 20) | (uint64_t)(b >> 12);
    return ((double)x + 0.5) * (1.0 / 4503599627370496.0);      /* 2^-52 */
}

/* ---------------------------------------------------------------------- */
/* Regulation                                                               */
/* ---------------------------------------------------------------------- */

/* Matrix form, cle.jl:79-106 with A_pos/A_neg from network.jl:206-219. */
void orc_regulation_matrix(const orc_network* net, const orc_kinetics* kin, const double* p,
                           double* reg) {
    const int G = net->G;
    const double Kn = pow(kin->K_d, kin->n);
    double act[4096], rep[4096];
    double* a = act; double* r = rep;
    double* heap = NULL;
    if (G > 4096) { heap = (double*)malloc(sizeof(double) * 2 * (size_t)G); a = heap; r = heap + G; }
    for (int j = 0; j < G; ++j) {
        double pn = pow(p[j], kin->n);        /* cle.jl:79 */
        double den = pn + Kn;                  /* cle.jl:80 */
        a[j] = pn / den;                       /* cle.jl:81 */
        r[j] = Kn / den;                       /* cle.jl:82 */
    }
    for (int i = 0; i < G; ++i) {
        double s_pos = 0.0, s_neg = 0.0;
        for (int j = 0; j < G; ++j) {
            if (i == j) continue;              /* network.jl:211 */
            double aeff = net->A[i + (size_t)j * G] * net->basals[i];   /* network.jl:212 */
            double apos = aeff > 0.0 ? aeff : 0.0;    /* network.jl:216 */
            double aneg = aeff < 0.0 ? -aeff : 0.0;   /* network.jl:217 */
            s_pos += apos * a[j];              /* cle.jl:84 */
            s_neg += aneg * r[j];              /* cle.jl:85 */
        }
        reg[i] = s_pos + s_neg;                /* cle.jl:86 */
    }
    for (int e = 0; e < net->n_coop; ++e) {   /* cle.jl:91-97 */
        double prod = 1.0;
        for (int k = net->coop_off[e]; k < net->coop_off[e + 1]; ++k) prod *= a[net->coop_src[k]];
        int t = net->coop_target[e];
        reg[t] += net->coop_strength[e] * net->basals[t] * prod;
    }
    for (int e = 0; e < net->n_redun; ++e) {  /* cle.jl:98-104 */
        double prod = 1.0;
        for (int k = net->redun_off[e]; k < net->redun_off[e + 1]; ++k) prod *= (1.0 - a[net->redun_src[k]]);
        int t = net->redun_target[e];
        reg[t] += net->redun_strength[e] * net->basals[t] * (1.0 - prod);
    }
    free(heap);
}

/* Loop form, network.jl:112-143 and 158-193. */
static double orc_hill_act(double p, double K, double n) {   /* network.jl:112-115 */
    double pn = pow(p, n);
    return pn / (pn + pow(K, n));
}
static double orc_hill_rep(double p, double K, double n) {   /* network.jl:122-125 */
    double pn = pow(p, n);
    return pow(K, n) / (pn + pow(K, n));
}
void orc_regulation_loop(const orc_network* net, const orc_kinetics* kin, const double* p,
                         double* reg) {
    const int G = net->G;
    for (int i = 0; i < G; ++i) {
        double acc = 0.0;
        for (int j = 0; j < G; ++j) {                          /* network.jl:163-170 */
            if (j == i) continue;
            double aij = net->A[i + (size_t)j * G];
            if (aij == 0.0) continue;
            double a = aij * net->basals[i];
            if (a > 0.0) acc += a * orc_hill_act(p[j], kin->K_d, kin->n);       /* network.jl:136-137 */
            else if (a < 0.0) acc += (-a) * orc_hill_rep(p[j], kin->K_d, kin->n);
        }
        for (int e = 0; e < net->n_coop; ++e) {                /* network.jl:173-180 */
            if (net->coop_target[e] != i) continue;
            double prod = 1.0;
            for (int k = net->coop_off[e]; k < net->coop_off[e + 1]; ++k)
                prod *= orc_hill_act(p[net->coop_src[k]], kin->K_d, kin->n);
            acc += net->coop_strength[e] * net->basals[i] * prod;
        }
        for (int e = 0; e < net->n_redun; ++e) {               /* network.jl:183-190 */
            if (net->redun_target[e] != i) continue;
            double prod = 1.0;
            for (int k = net->redun_off[e]; k < net->redun_off[e + 1]; ++k)
                prod *= (1.0 - orc_hill_act(p[net->redun_src[k]], kin->K_d, kin->n));
            acc += net->redun_strength[e] * net->basals[i] * (1.0 - prod);
        }
        reg[i] = acc;
    }
}

/* ---------------------------------------------------------------------- */
/* Telegraph helpers                                                        */
/* ---------------------------------------------------------------------- */

/* u = u32 * 2^-32 in [0,1); "u < q" <=> u32 < ceil(q * 2^32). */
static inline uint64_t orc_threshold(double q) {
    if (!(q > 0.0)) return 0;
    if (q >= 1.0) return 4294967296ull;
    double t = ceil(q * 4294967296.0);
    if (t > 4294967296.0) t = 4294967296.0;
    return (uint64_t)t;
}

typedef struct {
    int telegraph_any;
    uint8_t* switching;      /* isfinite(k_on)  (tauleap_binomial.jl:112) */
    uint8_t* random_init;    /* isfinite(k_on) && isfinite(k_off) (tauleap_binomial.jl:68) */
    uint64_t* thr_on;        /* OFF -> ON */
    uint64_t* thr_off;       /* ON -> OFF */
    uint64_t* thr_init;      /* init ON */
} orc_tele;

static void orc_tele_init(const orc_network* net, double dt, orc_tele* t) {
    int G = net->G;
    t->telegraph_any = 0;
    t->switching = (uint8_t*)calloc((size_t)G, 1);
    t->random_init = (uint8_t*)calloc((size_t)G, 1);
    t->thr_on = (uint64_t*)calloc((size_t)G, sizeof(uint64_t));
    t->thr_off = (uint64_t*)calloc((size_t)G, sizeof(uint64_t));
    t->thr_init = (uint64_t*)calloc((size_t)G, sizeof(uint64_t));
    if (!net->k_on || !net->k_off) return;
    for (int g = 0; g < G; ++g) {
        double kon = net->k_on[g], koff = net->k_off[g];
        if (isfinite(kon) || isfinite(koff)) t->telegraph_any = 1;   /* network.jl:103 */
        if (isfinite(kon)) {
            t->switching[g] = 1;
            t->thr_on[g] = orc_threshold(1.0 - exp(-kon * dt));      /* tauleap_binomial.jl:118 */
            t->thr_off[g] = orc_threshold(1.0 - exp(-koff * dt));    /* tauleap_binomial.jl:113 */
        }
        if (isfinite(kon) && isfinite(koff)) {
            t->random_init[g] = 1;
            t->thr_init[g] = orc_threshold(kon / (kon + koff));      /* tauleap_binomial.jl:69-70 */
        }
    }
}
static void orc_tele_free(orc_tele* t) {
    free(t->switching); free(t->random_init); free(t->thr_on); free(t->thr_off); free(t->thr_init);
}

static void orc_promoter_init(const orc_tele* t, int G, uint64_t seed, uint32_t c3base, uint32_t cell,
                              uint8_t* on) {
    uint32_t u[4];
    int have = -1;
    for (int g = 0; g < G; ++g) {
        on[g] = 1;                                                   /* constitutive: always ON */
        if (!t->telegraph_any || !t->random_init[g]) continue;
        if (have != (g >> 2)) { orc_draw(seed, cell, 0, (uint32_t)(g >> 2), c3base | ORC_P_PROM_INIT, u); have = g >> 2; }
        on[g] = ((uint64_t)u[g & 3] < t->thr_init[g]) ? 1 : 0;
    }
}

/* Tau-leap switching: one u32 of the switch block per gene. */
static void orc_promoter_switch(const orc_tele* t, int G, uint64_t seed, uint32_t c3base, uint32_t cell,
                                uint32_t step, uint8_t* on) {
    uint32_t u[4];
    int have = -1;
    for (int g = 0; g < G; ++g) {                                    /* tauleap_binomial.jl:110-121 */
        if (!t->switching[g]) continue;
        if (have != (g >> 2)) { orc_draw(seed, cell, step, (uint32_t)(g >> 2), c3base | ORC_P_SWITCH, u); have = g >> 2; }
        uint64_t x = (uint64_t)u[g & 3];
        if (on[g]) { if (x < t->thr_off[g]) on[g] = 0; }
        else       { if (x < t->thr_on[g])  on[g] = 1; }
    }
}

/* Promoter switch of gene g in the CLE steppers (tauleap_binomial.jl:110-121):
 * flip iff the 32-bit switching uniform X is below the threshold of the
 * current state.  X = (a & 511) << 23 | (b & 511) << 14 | e, where a and b are
 * the gene's radius and angle words of this step's normal block (the
 * Box-Muller transform reads only their top 23 bits) and e is the top 14 bits
 * of the gene's word of the switch block (cell, step, g >> 2), which is drawn
 * only when the top 18 bits of X leave the comparison open. */
int orc_cle_flip(uint32_t a, uint32_t b, uint64_t thr, uint64_t seed, uint32_t cell, uint32_t step, int32_t g,
                 uint32_t c3base) {
    const uint64_t x = ((uint64_t)(a & 511u) << 23) | ((uint64_t)(b & 511u) << 14);
    if (x + 16384u <= thr) return 1;
    if (x >= thr) return 0;
    uint32_t u[4];
    orc_draw(seed, cell, step, (uint32_t)(g >> 2), c3base | ORC_P_SWITCH, u);
    return (x | (uint64_t)(u[g & 3] >> 18)) < thr;
}

/* ---------------------------------------------------------------------- */
/* CLE, independent cells                                                   */
/* ---------------------------------------------------------------------- */

static int orc_check(const orc_network* net, const orc_kinetics* kin, const orc_run* run) {
    if (!net || !kin || !run) return 1;
    if (net->G <= 0 || !net->basals || !net->A) return 1;
    if (!(run->dt > 0.0) || run->n_steps < 0 || run->cell_num < 0) return 1;
    for (int i = 0; i < net->G; ++i)
        if (net->A[i + (size_t)i * net->G] != 0.0) return 1;          /* network.jl:68 */
    return 0;
}

int orc_cle_simulate(const orc_network* net, const orc_kinetics* kin, const orc_run* run,
                     double* m_out, double* p_out, double* snaps_m, double* snaps_p,
                     uint8_t* promoter_out) {
    if (orc_check(net, kin, run)) return 1;
    const int G = net->G;
    const int64_t lo = run->cell_lo, hi = run->cell_hi, nc = hi - lo;
    if (lo < 0 || hi > run->cell_num || nc < 0) return 1;
    const double dt = run->dt, sqrt_dt = sqrt(dt);                     /* cle.jl:55, 62 */
    const double k_t = kin->k_t;
    const double mu_m = kin->mu_m + kin->dilution;                      /* cle.jl:60 */
    const double mu_p = kin->mu_p + kin->dilution;                      /* cle.jl:61 */
    const uint32_t c3base = run->net_id << 4;
    const int64_t n_snap = run->snapshot_every > 0 ? run->n_steps / run->snapshot_every : 0;
    orc_tele tele;
    orc_tele_init(net, dt, &tele);

    /* steady state, kinetics.jl:65-70 */
    double* m0 = (double*)malloc(sizeof(double) * (size_t)G);
    double* p0 = (double*)malloc(sizeof(double) * (size_t)G);
    for (int g = 0; g < G; ++g) {
        double mss = net->basals[g] / (kin->mu_m + kin->dilution);
        double pss = kin->k_t * mss / (kin->mu_p + kin->dilution);
        m0[g] = run->init_mode == 0 ? ceil(mss) : 0.0;                  /* cle.jl:52-53 */
        p0[g] = run->init_mode == 0 ? ceil(pss) : 0.0;
    }

    int nthreads = run->n_threads > 0 ? run->n_threads : 1;
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
    for (int64_t c = lo; c < hi; ++c) {
        double* M = (double*)malloc(sizeof(double) * (size_t)G * 3);
        double* P = M + G;
        double* reg = M + 2 * G;
        uint8_t* on = (uint8_t*)malloc((size_t)G);
        memcpy(M, m0, sizeof(double) * (size_t)G);
        memcpy(P, p0, sizeof(double) * (size_t)G);
        orc_promoter_init(&tele, G, run->seed, c3base, (uint32_t)c, on);
        int64_t snap = 0;
        for (int64_t s = 0; s < run->n_steps; ++s) {
            orc_regulation_matrix(net, kin, P, reg);                    /* cle.jl:79-106 */
            uint32_t u[4] = {0, 0, 0, 0};
            for (int g = 0; g < G; ++g) {
                double xm, xp;
                const int o = (g & 1) * 2;
                if ((g & 1) == 0 && (!run->injected || tele.telegraph_any))
                    orc_draw(run->seed, (uint32_t)c, (uint32_t)s, (uint32_t)(g >> 1), c3base | ORC_P_NORMAL, u);
                /* switching first; the regulation above does not read promoter states */
                if (tele.telegraph_any && tele.switching[g] &&
                    orc_cle_flip(u[o], u[o + 1], on[g] ? tele.thr_off[g] : tele.thr_on[g], run->seed,
                                 (uint32_t)c, (uint32_t)s, g, c3base))
                    on[g] ^= 1;
                if (run->injected) {
                    size_t base = ((size_t)s * G + g) * 2;
                    xm = run->injected[(base + 0) * (size_t)run->cell_num + c];
                    xp = run->injected[(base + 1) * (size_t)run->cell_num + c];
                } else {
                    orc_normal_pair(u[o], u[o + 1], &xm, &xp);
                }
                /* telegraph gating of the whole transcription propensity
                 * (tauleap_binomial.jl:134-137; generate_figures.jl:124-128) */
                double tx = on[g] ? (reg[g] + net->basals[g]) : 0.0;
                double mm = M[g];
                double var_m = tx + mu_m * (mm > 0.0 ? mm : 0.0);        /* cle.jl:111 */
                if (var_m < 0.0) var_m = 0.0;
                double nm = xm * sqrt_dt * sqrt(var_m);
                double mn = mm + (tx - mu_m * mm) * dt + nm;             /* cle.jl:112 */
                M[g] = mn > 0.0 ? mn : 0.0;
                double pp = P[g];
                double mcl = M[g] > 0.0 ? M[g] : 0.0;
                double var_p = k_t * mcl + mu_p * (pp > 0.0 ? pp : 0.0); /* cle.jl:118 */
                if (var_p < 0.0) var_p = 0.0;
                double np_ = xp * sqrt_dt * sqrt(var_p);
                double pn = pp + (k_t * M[g] - mu_p * pp) * dt + np_;   /* cle.jl:119 */
                P[g] = pn > 0.0 ? pn : 0.0;
            }
            if (n_snap > 0 && ((s + 1) % run->snapshot_every) == 0 && snap < n_snap) {
                for (int g = 0; g < G; ++g) {
                    size_t idx = ((size_t)snap * G + g) * (size_t)nc + (size_t)(c - lo);
                    if (snaps_m) snaps_m[idx] = M[g];
                    if (snaps_p) snaps_p[idx] = P[g];
                }
                ++snap;
            }
        }
        for (int g = 0; g < G; ++g) {
            if (m_out) m_out[(size_t)g * nc + (size_t)(c - lo)] = M[g];
            if (p_out) p_out[(size_t)g * nc + (size_t)(c - lo)] = P[g];
            if (promoter_out) promoter_out[(size_t)g * nc + (size_t)(c - lo)] = on[g];
        }
        free(M); free(on);
    }
    free(m0); free(p0);
    orc_tele_free(&tele);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Cell-vectorised CLE pieces for the CPU baseline restatement of          */
/* cle.jl:34-130 (oracle/cpu_ref.py runs the step; between orc_ref_hill and */
/* orc_ref_update it forms the regulation with two OpenBLAS products, as   */
/* cle.jl:84-86).  Arrays are Julia's G x cells column-major state: element */
/* (g, c) at c*G + g.  Draws come from the same Philox stream as           */
/* orc_cle_simulate, so the restatement reproduces the oracle's runs.      */
/* ---------------------------------------------------------------------- */
void orc_ref_hill(const double* P, int64_t n, double hill_n, double Kn, double* act, double* rep) {
    for (int64_t i = 0; i < n; ++i) {
        const double pn = pow(P[i], hill_n);  /* cle.jl:79 */
        const double den = pn + Kn;           /* cle.jl:80 */
        act[i] = pn / den;                    /* cle.jl:81 */
        rep[i] = Kn / den;                    /* cle.jl:82 */
    }
}

void orc_ref_init(const orc_network* net, const orc_kinetics* kin, const orc_run* run, double* M, double* P,
                  uint8_t* on) {
    const int G = net->G;
    orc_tele tele;
    orc_tele_init(net, run->dt, &tele);
    for (int64_t c = 0; c < run->cell_num; ++c) {
        for (int g = 0; g < G; ++g) {
            const double mss = net->basals[g] / (kin->mu_m + kin->dilution);   /* kinetics.jl:65-70 */
            const double pss = kin->k_t * mss / (kin->mu_p + kin->dilution);
            M[c * G + g] = run->init_mode == 0 ? ceil(mss) : 0.0;             /* cle.jl:52-53 */
            P[c * G + g] = run->init_mode == 0 ? ceil(pss) : 0.0;
        }
        orc_promoter_init(&tele, G, run->seed, run->net_id << 4, (uint32_t)c, on + c * G);
    }
    orc_tele_free(&tele);
}

void orc_ref_update(const orc_network* net, const orc_kinetics* kin, const orc_run* run, int64_t step, double* M,
                    double* P, double* reg, const double* act, uint8_t* on, const double* noise_m,
                    const double* noise_p) {
    const int G = net->G;
    const int64_t C = run->cell_num;
    const double dt = run->dt, sqrt_dt = sqrt(dt);
    const double k_t = kin->k_t, mu_m = kin->mu_m + kin->dilution, mu_p = kin->mu_p + kin->dilution;
    const uint32_t c3base = run->net_id << 4;
    orc_tele tele;
    orc_tele_init(net, dt, &tele);
    for (int64_t c = 0; c < C; ++c) {
        double* R = reg + c * G;
        const double* a = act + c * G;
        for (int e = 0; e < net->n_coop; ++e) {            /* cle.jl:91-97 */
            double prod = 1.0;
            for (int k = net->coop_off[e]; k < net->coop_off[e + 1]; ++k) prod *= a[net->coop_src[k]];
            const int t = net->coop_target[e];
            R[t] += net->coop_strength[e] * net->basals[t] * prod;
        }
        for (int e = 0; e < net->n_redun; ++e) {           /* cle.jl:98-104 */
            double prod = 1.0;
            for (int k = net->redun_off[e]; k < net->redun_off[e + 1]; ++k) prod *= (1.0 - a[net->redun_src[k]]);
            const int t = net->redun_target[e];
            R[t] += net->redun_strength[e] * net->basals[t] * (1.0 - prod);
        }
    }
    for (int64_t c = 0; c < C; ++c) {
        uint32_t u[4] = {0, 0, 0, 0};
        for (int g = 0; g < G; ++g) {
            const int64_t i = c * G + g;
            const int o = (g & 1) * 2;
            double xm, xp;
            if ((g & 1) == 0 && ((!run->injected && !noise_m) || tele.telegraph_any))
                orc_draw(run->seed, (uint32_t)c, (uint32_t)step, (uint32_t)(g >> 1), c3base | ORC_P_NORMAL, u);
            if (tele.telegraph_any && tele.switching[g] &&
                orc_cle_flip(u[o], u[o + 1], on[i] ? tele.thr_off[g] : tele.thr_on[g], run->seed, (uint32_t)c,
                             (uint32_t)step, g, c3base))
                on[i] ^= 1;
            if (noise_m) {                             /* this step's noise matrices (cle.jl:109, 116) */
                xm = noise_m[i];
                xp = noise_p[i];
            } else if (run->injected) {
                const size_t base = ((size_t)step * G + g) * 2;
                xm = run->injected[(base + 0) * (size_t)C + c];
                xp = run->injected[(base + 1) * (size_t)C + c];
            } else {
                orc_normal_pair(u[o], u[o + 1], &xm, &xp);
            }
            const double tx = on[i] ? (reg[i] + net->basals[g]) : 0.0;
            const double mm = M[i];
            double var_m = tx + mu_m * (mm > 0.0 ? mm : 0.0);                   /* cle.jl:111 */
            if (var_m < 0.0) var_m = 0.0;
            const double mn = mm + (tx - mu_m * mm) * dt + xm * sqrt_dt * sqrt(var_m);   /* cle.jl:112 */
            M[i] = mn > 0.0 ? mn : 0.0;
            const double pp = P[i];
            double var_p = k_t * M[i] + mu_p * (pp > 0.0 ? pp : 0.0);           /* cle.jl:118 */
            if (var_p < 0.0) var_p = 0.0;
            const double pn = pp + (k_t * M[i] - mu_p * pp) * dt + xp * sqrt_dt * sqrt(var_p);   /* cle.jl:119 */
            P[i] = pn > 0.0 ? pn : 0.0;
        }
    }
    orc_tele_free(&tele);
}

/* ---------------------------------------------------------------------- */
/* Binomial sampler: inversion (small n*p) and BTRD (Hormann 1993).        */
/* Uniforms come from consecutive Philox blocks, two 52-bit uniforms each. */
/* ---------------------------------------------------------------------- */
typedef struct {
    uint64_t seed; uint32_t c0, c1, c3;
    uint32_t block; int pos; uint32_t u[4];
} orc_ustream;

static double orc_next_u(orc_ustream* s) {
    if (s->pos >= 2) { orc_draw(s->seed, s->c0, s->c1, s->block, s->c3, s->u); s->block++; s->pos = 0; }
    double v = orc_u52(s->u[2 * s->pos], s->u[2 * s->pos + 1]);
    s->pos++;
    return v;
}

/* Stirling-series correction log(k!) - [log(sqrt(2 pi)) + (k+1/2) log(k+1) - (k+1)]. */
static double orc_fc(int64_t k) {
    static const double tab[10] = {
        0.08106146679532726, 0.04134069595540929, 0.02767792568499834,
        0.02079067210376509, 0.01664469118982119, 0.01387612882307075,
        0.01189670994589177, 0.01041126526197209, 0.009255462182712733,
        0.008330563433362871};
    if (k < 10) return tab[k];
    double kp = (double)k + 1.0;
    double kp2 = kp * kp;
    return (1.0 / 12.0 - (1.0 / 360.0 - 1.0 / 1260.0 / kp2) / kp2) / kp;
}

/* lq: NAN, or log(1-p) and log(p) precomputed by the caller ({log1mp, logp});
 * with them the inversion start q^n is exp(n * log q) instead of pow(q, n). */
static int64_t orc_binomial_stream_lq(int64_t n, double p, const double* lq, orc_ustream* st) {
    if (n <= 0 || !(p > 0.0)) return 0;
    if (p >= 1.0) return n;
    int flip = p > 0.5;
    double pp = flip ? 1.0 - p : p;
    double q = 1.0 - pp;
    int64_t x;
    if ((double)n * pp < 10.0) {
        /* inversion from 0 upwards (BINV) */
        double s = pp / q, a = (double)(n + 1) * s;
        double r0 = lq ? exp((double)n * (flip ? lq[1] : lq[0])) : pow(q, (double)n);
        for (;;) {
            double u = orc_next_u(st), r = r0;
            int64_t k = 0;
            int ok = 1;
            while (u > r) {
                u -= r;
                ++k;
                if (k > n) { ok = 0; break; }
                r *= (a / (double)k - s);
            }
            if (ok) { x = k; break; }
        }
    } else {
        /* BTRD */
        double m = floor((double)(n + 1) * pp);
        double r = pp / q, nr = (double)(n + 1) * r, npq = (double)n * pp * q;
        double spq = sqrt(npq);
        double b = 1.15 + 2.53 * spq;
        double a = -0.0873 + 0.0248 * b + 0.01 * pp;
        double c = (double)n * pp + 0.5;
        double alpha = (2.83 + 5.1 / b) * spq;
        double vr = 0.92 - 4.2 / b;
        double urvr = 0.86 * vr;
        for (;;) {
            double v = orc_next_u(st), u;
            if (v <= urvr) {
                u = v / vr - 0.43;
                x = (int64_t)floor((2.0 * a / (0.5 - fabs(u)) + b) * u + c);
                break;
            }
            if (v >= vr) {
                u = orc_next_u(st) - 0.5;
            } else {
                u = v / vr - 0.93;
                u = (u < 0.0 ? -0.5 : 0.5) - u;
                v = orc_next_u(st) * vr;
            }
            double us = 0.5 - fabs(u);
            double kf = floor((2.0 * a / us + b) * u + c);
            if (kf < 0.0 || kf > (double)n) continue;
            int64_t k = (int64_t)kf;
            v = v * alpha / (a / (us * us) + b);
            int64_t mi = (int64_t)m;
            int64_t km = k > mi ? k - mi : mi - k;
            if (km <= 15) {
                double f = 1.0;
                if (mi < k) { for (int64_t i = mi + 1; i <= k; ++i) f *= (nr / (double)i - r); }
                else if (mi > k) { for (int64_t i = k + 1; i <= mi; ++i) v *= (nr / (double)i - r); }
                if (v <= f) { x = k; break; }
                continue;
            }
            double lv = log(v);
            double kmd = (double)km;
            double rho = (kmd / npq) * (((kmd / 3.0 + 0.625) * kmd + 1.0 / 6.0) / npq + 0.5);
            double t = -kmd * kmd / (2.0 * npq);
            if (lv < t - rho) { x = k; break; }
            if (lv > t + rho) continue;
            double nm = (double)(n - mi + 1);
            double h = (m + 0.5) * log((m + 1.0) / (r * nm)) + orc_fc(mi) + orc_fc(n - mi);
            double nk = (double)(n - k + 1);
            if (lv <= h + (double)(n + 1) * log(nm / nk) + (kf + 0.5) * log(nk * r / (kf + 1.0))
                          - orc_fc(k) - orc_fc(n - k)) { x = k; break; }
        }
    }
    return flip ? n - x : x;
}

static int64_t orc_binomial_stream(int64_t n, double p, orc_ustream* st) {
    return orc_binomial_stream_lq(n, p, NULL, st);
}

/* log(k!) from the Stirling series and its correction table (orc_fc). */
static double orc_logfact(int64_t k) {
    double kd = (double)k;
    return 0.91893853320467274178 + (kd + 0.5) * log(kd + 1.0) - (kd + 1.0) + orc_fc(k);
}

/* Poisson(lam): sequential inversion from 0 for lam < 10, otherwise PTRS,
 * the transformed rejection of W. Hormann, "The transformed rejection method
 * for generating Poisson random variables", Insurance: Mathematics and
 * Economics 12 (1993) 39-45.  lam <= 0 (or NaN) draws nothing and returns 0. */
static int64_t orc_poisson_stream(double lam, orc_ustream* st) {
    if (!(lam > 0.0)) return 0;
    if (lam < 10.0) {
        const double p0 = exp(-lam);
        for (;;) {
            double u = orc_next_u(st), p = p0;
            int64_t k = 0;
            while (u > p) {
                u -= p;
                ++k;
                if (k > 200) break;               /* u beyond the rounded total mass: redraw */
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
        double u = orc_next_u(st) - 0.5;
        double v = orc_next_u(st);
        double us = 0.5 - fabs(u);
        double kf = floor((2.0 * a / us + b) * u + lam + 0.43);
        if (us >= 0.07 && v <= vr) return (int64_t)kf;
        if (kf < 0.0 || (us < 0.013 && v > us)) continue;
        double w = v * alpha / (a / (us * us) + b);
        if (log(w) <= -lam + kf * loglam - orc_logfact((int64_t)kf)) return (int64_t)kf;
    }
}

int64_t orc_poisson(double lam, uint32_t ctr0, uint32_t ctr1, uint32_t c3, uint64_t seed) {
    orc_ustream st = {seed, ctr0, ctr1, c3, 0, 2, {0, 0, 0, 0}};
    return orc_poisson_stream(lam, &st);
}

int64_t orc_binomial(int64_t n, double p, uint32_t ctr0, uint32_t ctr1, uint32_t c3, uint64_t seed) {
    orc_ustream st = {seed, ctr0, ctr1, c3, 0, 2, {0, 0, 0, 0}};
    return orc_binomial_stream(n, p, &st);
}

/* Julia's round(Int, x): round half to even (capture.jl:38). */
static inline int64_t orc_round_even(double x) {
    double r = nearbyint(x);   /* default rounding mode: to nearest, ties to even */
    return (int64_t)r;
}

/* ---------------------------------------------------------------------- */
/* Capture, capture.jl:25-47                                                */
/* ---------------------------------------------------------------------- */
int orc_capture(const double* X, int64_t cells, int32_t G, int64_t ld, double efficiency,
                double efficiency_std, uint64_t seed, uint32_t net_id, int32_t* Y, double* p_cell) {
    if (!X || !Y || cells < 0 || G <= 0 || ld < cells) return 1;
    if (!(efficiency > 0.0) || efficiency_std < 0.0) return 1;
    const double mu_log = log(efficiency) - efficiency_std * efficiency_std / 2.0;   /* capture.jl:30 */
    const double sigma_log = efficiency_std;                                         /* capture.jl:31 */
    const uint32_t c3base = net_id << 4;
    for (int64_t c = 0; c < cells; ++c) {
        uint32_t u[4];
        orc_draw(seed, (uint32_t)c, 0, 0, c3base | ORC_P_CAPTURE_EFF, u);
        double U1 = orc_u52(u[0], u[1]), U2 = orc_u52(u[2], u[3]);
        double z = sqrt(-2.0 * log(U1)) * cos(2.0 * M_PI * U2);
        double pc = exp(mu_log + sigma_log * z);                                     /* capture.jl:36 */
        if (pc < 0.0) pc = 0.0;
        if (pc > 1.0) pc = 1.0;
        if (p_cell) p_cell[c] = pc;
        for (int g = 0; g < G; ++g) {
            double xv = X[(size_t)g * ld + c];
            int64_t n = orc_round_even(xv);                                          /* capture.jl:38 */
            if (n < 0) n = 0;
            int64_t y = 0;
            if (n != 0 && pc != 0.0)                                                 /* capture.jl:39-43 */
                y = orc_binomial(n, pc, (uint32_t)c, (uint32_t)g, c3base | ORC_P_CAPTURE_BIN, seed);
            Y[(size_t)g * ld + c] = (int32_t)y;
        }
    }
    return 0;
}

/* What one division check did, for the tests (any pointer may be NULL):
 * n_div dividers in ascending cell order, the mothers' molecules after
 * partitioning and the daughters' (cell-major [k*G + g]), the slot each
 * daughter replaced, and the ambiguity taint (below) per cell. */
typedef struct orc_div_record {
    int64_t n_div;
    int64_t* dividers;            /* [N] */
    double *mother_m, *mother_p;  /* [N*G] */
    double *daughter_m, *daughter_p;
    int64_t* replaced;            /* [N] */
} orc_div_record;

/* Julia's round half to even, and whether x lies within amb * max(1, |x|) of
 * a rounding boundary (a half-integer): a restatement in other arithmetic
 * (the GPU's FP32 state) may round such a value the other way. */
static inline int orc_round_ambiguous(double x, double amb) {
    if (!(amb > 0.0)) return 0;
    const double f = x - floor(x);
    const double scale = fabs(x) > 1.0 ? fabs(x) : 1.0;
    return fabs(f - 0.5) < amb * scale;
}

/* One division check (population.jl:157-205; cle.jl:240-272): every cell
 * with V > V_div divides, then each daughter replaces a cell (Moran).
 * Daughter buffers are caller-provided ([N] and [N*G]).  With amb > 0 and
 * taint != NULL, a divider whose binomial partition rounds a value within
 * amb of a half-integer is marked, and a replaced slot takes its daughter's
 * mark (the mother's, so marks follow the lineage). */
static void orc_division_check(const orc_population* pop, int G, int64_t N, uint64_t seed, uint32_t c3base,
                               int64_t s, double* M, double* P, double* V, uint8_t* on, int64_t* div_idx,
                               double* dM, double* dP, double* dV, uint8_t* dOn, double amb, uint8_t* taint,
                               orc_div_record* rec) {
    int64_t nd = 0;
    for (int64_t c = 0; c < N; ++c) if (V[c] > pop->V_div) div_idx[nd++] = c;   /* cle.jl:241 */
    uint8_t* dtaint = taint ? (uint8_t*)calloc((size_t)(nd > 0 ? nd : 1), 1) : NULL;
    /* Phase 1: partition mothers, buffer daughters (population.jl:166-192) */
    for (int64_t k = 0; k < nd; ++k) {
        int64_t c = div_idx[k];
        for (int g = 0; g < G; ++g) {
            for (int sp = 0; sp < 2; ++sp) {
                double* X = sp == 0 ? &M[c * G + g] : &P[c * G + g];
                double* D = sp == 0 ? &dM[k * G + g] : &dP[k * G + g];
                if (pop->partition == 0) {                      /* cle.jl:251-257 */
                    double h = *X / 2.0; *X = h; *D = h;
                } else {
                    if (taint && orc_round_ambiguous(*X, amb)) taint[c] = 1;
                    int64_t n = orc_round_even(*X);
                    if (n < 0) n = 0;
                    int64_t d = 0;
                    if (n > 0) {
                        uint32_t c3 = c3base | ORC_P_PARTITION;
                        /* sub-stream per (species, gene) in counter word 2 high bits */
                        orc_ustream st = {seed, (uint32_t)c, (uint32_t)s, c3,
                                          ((uint32_t)(g * 2 + sp)) << 16, 2, {0, 0, 0, 0}};
                        d = orc_binomial_stream(n, 0.5, &st);
                    }
                    *X = (double)(n - d); *D = (double)d;       /* population.jl:174-176 */
                }
            }
            dOn[k * G + g] = on[c * G + g];                     /* daughter inherits promoter state */
        }
        V[c] /= 2.0;                                            /* population.jl:190 */
        dV[k] = V[c];
        if (dtaint) dtaint[k] = taint[c];
        if (rec) {
            rec->dividers[k] = c;
            for (int g = 0; g < G; ++g) {
                rec->mother_m[k * G + g] = M[c * G + g];
                rec->mother_p[k * G + g] = P[c * G + g];
                rec->daughter_m[k * G + g] = dM[k * G + g];
                rec->daughter_p[k * G + g] = dP[k * G + g];
            }
        }
    }
    /* Phase 2: Moran replacement in divider order (population.jl:194-202) */
    for (int64_t k = 0; k < nd; ++k) {
        int64_t mother = div_idx[k];
        uint32_t u[4];
        orc_draw(seed, (uint32_t)mother, (uint32_t)s, 0, c3base | ORC_P_REPLACE, u);
        uint64_t x = ((uint64_t)u[0] << 32) | u[1];
        int64_t replaced;
        if (pop->replace_mode == 1 && N > 1) {                  /* cle.jl:265-268 */
            replaced = (int64_t)(((unsigned __int128)x * (uint64_t)(N - 1)) >> 64);
            if (replaced >= mother) replaced += 1;
        } else {
            replaced = (int64_t)(((unsigned __int128)x * (uint64_t)N) >> 64);
        }
        for (int g = 0; g < G; ++g) {
            M[replaced * G + g] = dM[k * G + g];
            P[replaced * G + g] = dP[k * G + g];
            on[replaced * G + g] = dOn[k * G + g];
        }
        V[replaced] = dV[k];
        if (taint) taint[replaced] = dtaint[k];
        if (rec) rec->replaced[k] = replaced;
    }
    if (rec) rec->n_div = nd;
    free(dtaint);
}

int orc_divide(const orc_population* pop, int32_t G, int64_t N, uint64_t seed, uint32_t net_id, int64_t step,
               double* m, double* p, double* V, int64_t* n_div, int64_t* dividers, double* mother_m,
               double* mother_p, double* daughter_m, double* daughter_p, int64_t* replaced) {
    if (!pop || G <= 0 || N <= 0 || !m || !p || !V) return 1;
    double* M = (double*)malloc(sizeof(double) * (size_t)G * (size_t)N);
    double* P = (double*)malloc(sizeof(double) * (size_t)G * (size_t)N);
    uint8_t* on = (uint8_t*)calloc((size_t)G * (size_t)N, 1);
    uint8_t* dOn = (uint8_t*)calloc((size_t)G * (size_t)N, 1);
    double* dM = (double*)malloc(sizeof(double) * (size_t)G * (size_t)N);
    double* dP = (double*)malloc(sizeof(double) * (size_t)G * (size_t)N);
    double* dV = (double*)malloc(sizeof(double) * (size_t)N);
    int64_t* div_idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)N);
    int64_t* rdiv = dividers ? dividers : (int64_t*)malloc(sizeof(int64_t) * (size_t)N);
    int64_t* rrep = replaced ? replaced : (int64_t*)malloc(sizeof(int64_t) * (size_t)N);
    double* scratch = (double*)malloc(sizeof(double) * (size_t)G * (size_t)N * 4);
    orc_div_record rec = {0, rdiv, mother_m ? mother_m : scratch, mother_p ? mother_p : scratch + (size_t)G * N,
                          daughter_m ? daughter_m : scratch + 2 * (size_t)G * N,
                          daughter_p ? daughter_p : scratch + 3 * (size_t)G * N, rrep};
    for (int64_t c = 0; c < N; ++c)
        for (int g = 0; g < G; ++g) {
            M[c * G + g] = m[(size_t)g * N + c];
            P[c * G + g] = p[(size_t)g * N + c];
        }
    orc_division_check(pop, G, N, seed, net_id << 4, step, M, P, V, on, div_idx, dM, dP, dV, dOn, 0.0, NULL, &rec);
    for (int64_t c = 0; c < N; ++c)
        for (int g = 0; g < G; ++g) {
            m[(size_t)g * N + c] = M[c * G + g];
            p[(size_t)g * N + c] = P[c * G + g];
        }
    if (n_div) *n_div = rec.n_div;
    if (!dividers) free(rdiv);
    if (!replaced) free(rrep);
    free(M); free(P); free(on); free(dOn); free(dM); free(dP); free(dV); free(div_idx); free(scratch);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Population CLE, cle.jl:143-284 + population.jl:157-205                  */
/* ---------------------------------------------------------------------- */
int orc_cle_population(const orc_network* net, const orc_kinetics* kin, const orc_run* run,
                       const orc_population* pop, double* m_out, double* p_out, double* V_out,
                       uint8_t* promoter_out) {
    return orc_cle_population_ex(net, kin, run, pop, 0.0, m_out, p_out, V_out, promoter_out, NULL);
}

int orc_cle_population_ex(const orc_network* net, const orc_kinetics* kin, const orc_run* run,
                          const orc_population* pop, double amb, double* m_out, double* p_out, double* V_out,
                          uint8_t* promoter_out, uint8_t* taint_out) {
    if (orc_check(net, kin, run) || !pop || pop->div_check_interval <= 0) return 1;
    const int G = net->G;
    const int64_t N = run->cell_num;
    if (N <= 0) return 1;
    const double dt = run->dt, sqrt_dt = sqrt(dt);
    const double k_t = kin->k_t, mu_m = kin->mu_m, mu_p = kin->mu_p;  /* cle.jl:166-167: intrinsic only */
    const double growth = exp(pop->growth_rate * dt);                 /* cle.jl:169 */
    const uint32_t c3base = run->net_id << 4;
    orc_tele tele;
    orc_tele_init(net, dt, &tele);

    double* M = (double*)malloc(sizeof(double) * (size_t)G * (size_t)N);
    double* P = (double*)malloc(sizeof(double) * (size_t)G * (size_t)N);
    double* V = (double*)malloc(sizeof(double) * (size_t)N);
    uint8_t* on = (uint8_t*)malloc((size_t)G * (size_t)N);
    double* reg = (double*)malloc(sizeof(double) * (size_t)G);
    double* pcol = (double*)malloc(sizeof(double) * (size_t)G);
    int64_t* div_idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)N);
    double* dM = (double*)malloc(sizeof(double) * (size_t)G * (size_t)N);
    double* dP = (double*)malloc(sizeof(double) * (size_t)G * (size_t)N);
    double* dV = (double*)malloc(sizeof(double) * (size_t)N);
    uint8_t* dOn = (uint8_t*)malloc((size_t)G * (size_t)N);
    uint8_t* taint = taint_out ? (uint8_t*)calloc((size_t)N, 1) : NULL;

    for (int64_t c = 0; c < N; ++c) {
        for (int g = 0; g < G; ++g) {
            /* steady_state includes kin.dilution (cle.jl:152, kinetics.jl:65-70) */
            double mss = net->basals[g] / (kin->mu_m + kin->dilution);
            double pss = kin->k_t * mss / (kin->mu_p + kin->dilution);
            M[c * G + g] = run->init_mode == 0 ? ceil(mss) : 0.0;       /* cle.jl:155-156 */
            P[c * G + g] = run->init_mode == 0 ? ceil(pss) : 0.0;
        }
        uint32_t u[4];
        orc_draw(run->seed, (uint32_t)c, 0, 0, c3base | ORC_P_VINIT, u);
        V[c] = pop->V_init_lo + (pop->V_init_hi - pop->V_init_lo) * orc_u52(u[0], u[1]);   /* cle.jl:159 */
        orc_promoter_init(&tele, G, run->seed, c3base, (uint32_t)c, on + c * G);
    }

    for (int64_t s = 0; s < run->n_steps; ++s) {
        for (int64_t c = 0; c < N; ++c) {
            double* Mc = M + c * G; double* Pc = P + c * G; uint8_t* oc = on + c * G;
            for (int g = 0; g < G; ++g) pcol[g] = Pc[g];
            orc_regulation_matrix(net, kin, pcol, reg);                 /* cle.jl:190-217 */
            uint32_t u[4] = {0, 0, 0, 0};
            for (int g = 0; g < G; ++g) {
                double xm, xp;
                const int o = (g & 1) * 2;
                if ((g & 1) == 0 && (!run->injected || tele.telegraph_any))
                    orc_draw(run->seed, (uint32_t)c, (uint32_t)s, (uint32_t)(g >> 1), c3base | ORC_P_NORMAL, u);
                if (tele.telegraph_any && tele.switching[g] &&
                    orc_cle_flip(u[o], u[o + 1], oc[g] ? tele.thr_off[g] : tele.thr_on[g], run->seed,
                                 (uint32_t)c, (uint32_t)s, g, c3base))
                    oc[g] ^= 1;
                if (run->injected) {
                    size_t base = ((size_t)s * G + g) * 2;
                    xm = run->injected[(base + 0) * (size_t)N + c];
                    xp = run->injected[(base + 1) * (size_t)N + c];
                } else {
                    orc_normal_pair(u[o], u[o + 1], &xm, &xp);
                }
                double tx = oc[g] ? (reg[g] + net->basals[g]) * V[c] : 0.0;   /* cle.jl:224-225 */
                double mm = Mc[g];
                double var_m = tx + mu_m * (mm > 0.0 ? mm : 0.0);
                if (var_m < 0.0) var_m = 0.0;
                double mn = mm + (tx - mu_m * mm) * dt + xm * sqrt_dt * sqrt(var_m);
                Mc[g] = mn > 0.0 ? mn : 0.0;
                double pp = Pc[g];
                double var_p = k_t * Mc[g] + mu_p * (pp > 0.0 ? pp : 0.0);   /* cle.jl:231-233 */
                if (var_p < 0.0) var_p = 0.0;
                double pn = pp + (k_t * Mc[g] - mu_p * pp) * dt + xp * sqrt_dt * sqrt(var_p);
                Pc[g] = pn > 0.0 ? pn : 0.0;
            }
            V[c] *= growth;                                             /* cle.jl:237 */
        }
        if (((s + 1) % pop->div_check_interval) == 0)                   /* cle.jl:240 */
            orc_division_check(pop, G, N, run->seed, c3base, s, M, P, V, on, div_idx, dM, dP, dV, dOn, amb, taint,
                               NULL);
    }
    for (int64_t c = 0; c < N; ++c) {
        for (int g = 0; g < G; ++g) {
            if (m_out) m_out[(size_t)g * N + c] = M[c * G + g];
            if (p_out) p_out[(size_t)g * N + c] = P[c * G + g];
            if (promoter_out) promoter_out[(size_t)g * N + c] = on[c * G + g];
        }
        if (V_out) V_out[c] = V[c];
        if (taint_out) taint_out[c] = taint[c];
    }
    free(M); free(P); free(V); free(on); free(reg); free(pcol); free(div_idx);
    free(dM); free(dP); free(dV); free(dOn); free(taint);
    orc_tele_free(&tele);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Moments (validate.jl:70-77 uses mean/var over cells per gene)            */
/* ---------------------------------------------------------------------- */
void orc_moments(const double* m, const double* p, int64_t cells, int32_t G, int64_t ld,
                 double* out) {
    for (int g = 0; g < G; ++g) {
        double s[5] = {0, 0, 0, 0, 0};
        for (int64_t c = 0; c < cells; ++c) {
            double a = m[(size_t)g * ld + c], b = p[(size_t)g * ld + c];
            s[0] += a; s[1] += a * a; s[2] += b; s[3] += b * b; s[4] += a * b;
        }
        for (int k = 0; k < 5; ++k) out[g * 5 + k] = s[k];
    }
}

/* ---------------------------------------------------------------------- */
/* Binomial tau-leap, independent cells                                     */
/* (src/algorithms/tauleap_binomial.jl:31-180, regulation_mode :transcription) */
/* ---------------------------------------------------------------------- */
int orc_tauleap_simulate(const orc_network* net, const orc_kinetics* kin, const orc_run* run,
                         double* m_out, double* p_out, double* snaps_m, double* snaps_p,
                         uint8_t* promoter_out) {
    if (orc_check(net, kin, run) || run->injected) return 1;
    const int G = net->G;
    const int64_t lo = run->cell_lo, hi = run->cell_hi, nc = hi - lo;
    if (lo < 0 || hi > run->cell_num || nc < 0) return 1;
    const double dt = run->dt;
    const uint32_t c3 = (run->net_id << 4) | ORC_P_TAULEAP;
    const int64_t n_snap = run->snapshot_every > 0 ? run->n_steps / run->snapshot_every : 0;
    /* decay probabilities with dilution (tauleap_binomial.jl:144, 157) and their logs */
    const double pdm = 1.0 - exp(-(kin->mu_m + kin->dilution) * dt);
    const double pdp = 1.0 - exp(-(kin->mu_p + kin->dilution) * dt);
    const double lqm[2] = {log(1.0 - pdm), log(pdm)};
    const double lqp[2] = {log(1.0 - pdp), log(pdp)};
    orc_tele tele;
    orc_tele_init(net, dt, &tele);

    /* init: round(Int, steady state), half to even (tauleap_binomial.jl:61-62) */
    int64_t* m0 = (int64_t*)malloc(sizeof(int64_t) * (size_t)G);
    int64_t* p0 = (int64_t*)malloc(sizeof(int64_t) * (size_t)G);
    for (int g = 0; g < G; ++g) {
        double mss = net->basals[g] / (kin->mu_m + kin->dilution);            /* kinetics.jl:65-70 */
        double pss = kin->k_t * mss / (kin->mu_p + kin->dilution);
        m0[g] = run->init_mode == 0 ? orc_round_even(mss) : 0;
        p0[g] = run->init_mode == 0 ? orc_round_even(pss) : 0;
    }

    int nthreads = run->n_threads > 0 ? run->n_threads : 1;
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
    for (int64_t c = lo; c < hi; ++c) {
        int64_t* M = (int64_t*)malloc(sizeof(int64_t) * (size_t)G * 2);
        int64_t* P = M + G;
        double* pf = (double*)malloc(sizeof(double) * (size_t)G * 2);
        double* reg = pf + G;
        uint8_t* on = (uint8_t*)malloc((size_t)G);
        memcpy(M, m0, sizeof(int64_t) * (size_t)G);
        memcpy(P, p0, sizeof(int64_t) * (size_t)G);
        orc_promoter_init(&tele, G, run->seed, run->net_id << 4, (uint32_t)c, on);   /* :64-76 */
        int64_t snap = 0;
        for (int64_t s = 0; s < run->n_steps; ++s) {
            for (int g = 0; g < G; ++g) pf[g] = (double)P[g];                     /* :81 */
            if (tele.telegraph_any)                                               /* :84-123 */
                orc_promoter_switch(&tele, G, run->seed, run->net_id << 4, (uint32_t)c, (uint32_t)s, on);
            orc_regulation_loop(net, kin, pf, reg);                              /* :128 */
            for (int g = 0; g < G; ++g) {
                orc_ustream st = {run->seed, (uint32_t)c, (uint32_t)s, c3, (uint32_t)g << 16, 2, {0, 0, 0, 0}};
                double prop = reg[g] + net->basals[g];                           /* :129 */
                if (!(prop > 0.0)) prop = 0.0;
                if (tele.telegraph_any && !on[g]) prop = 0.0;                    /* :134-137 */
                int64_t n_tx = orc_poisson_stream(prop * dt, &st);                /* :140 */
                int64_t n_md = M[g] > 0 ? orc_binomial_stream_lq(M[g], pdm, lqm, &st) : 0;   /* :143-149 */
                M[g] += n_tx - n_md;                                             /* :150 */
                int64_t mpos = M[g] > 0 ? M[g] : 0;
                int64_t n_tl = orc_poisson_stream(kin->k_t * (double)mpos * dt, &st);     /* :153 */
                int64_t n_pd = P[g] > 0 ? orc_binomial_stream_lq(P[g], pdp, lqp, &st) : 0;   /* :156-161 */
                P[g] += n_tl - n_pd;                                             /* :163 */
            }
            if (n_snap > 0 && ((s + 1) % run->snapshot_every) == 0 && snap < n_snap) {
                for (int g = 0; g < G; ++g) {
                    size_t idx = ((size_t)snap * G + g) * (size_t)nc + (size_t)(c - lo);
                    if (snaps_m) snaps_m[idx] = (double)M[g];
                    if (snaps_p) snaps_p[idx] = (double)P[g];
                }
                ++snap;
            }
        }
        for (int g = 0; g < G; ++g) {
            if (m_out) m_out[(size_t)g * nc + (size_t)(c - lo)] = (double)M[g];
            if (p_out) p_out[(size_t)g * nc + (size_t)(c - lo)] = (double)P[g];
            if (promoter_out) promoter_out[(size_t)g * nc + (size_t)(c - lo)] = on[g];
        }
        free(M); free(pf); free(on);
    }
    free(m0); free(p0);
    orc_tele_free(&tele);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Binomial tau-leap with population dynamics                              */
/* (tauleap_binomial.jl:195-261; population.jl:96-116, 128-139, 157-205)   */
/* ---------------------------------------------------------------------- */
int orc_tauleap_population(const orc_network* net, const orc_kinetics* kin, const orc_run* run,
                           const orc_population* pop, double* m_out, double* p_out, double* V_out,
                           uint8_t* promoter_out) {
    if (orc_check(net, kin, run) || !pop || pop->div_check_interval <= 0 || run->injected) return 1;
    if (pop->partition != 1) return 1;          /* counts stay integers only with binomial partitioning */
    const int G = net->G;
    const int64_t N = run->cell_num;
    if (N <= 0) return 1;
    const double dt = run->dt;
    const uint32_t c3base = run->net_id << 4, c3t = c3base | ORC_P_TAULEAP;
    /* intrinsic decay only (tauleap_binomial.jl:228, 242); growth factor (population.jl:136) */
    const double pdm = 1.0 - exp(-kin->mu_m * dt);
    const double pdp = 1.0 - exp(-kin->mu_p * dt);
    const double lqm[2] = {log(1.0 - pdm), log(pdm)};
    const double lqp[2] = {log(1.0 - pdp), log(pdp)};
    const double growth = exp(pop->growth_rate * dt);
    orc_tele tele;
    orc_tele_init(net, dt, &tele);

    double* M = (double*)malloc(sizeof(double) * (size_t)G * (size_t)N);
    double* P = (double*)malloc(sizeof(double) * (size_t)G * (size_t)N);
    double* V = (double*)malloc(sizeof(double) * (size_t)N);
    uint8_t* on = (uint8_t*)malloc((size_t)G * (size_t)N);
    double* reg = (double*)malloc(sizeof(double) * (size_t)G);
    double* pf = (double*)malloc(sizeof(double) * (size_t)G);
    int64_t* div_idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)N);
    double* dM = (double*)malloc(sizeof(double) * (size_t)G * (size_t)N);
    double* dP = (double*)malloc(sizeof(double) * (size_t)G * (size_t)N);
    double* dV = (double*)malloc(sizeof(double) * (size_t)N);
    uint8_t* dOn = (uint8_t*)malloc((size_t)G * (size_t)N);

    for (int64_t c = 0; c < N; ++c) {
        for (int g = 0; g < G; ++g) {
            double mss = net->basals[g] / (kin->mu_m + kin->dilution);         /* kinetics.jl:65-70 */
            double pss = kin->k_t * mss / (kin->mu_p + kin->dilution);
            M[c * G + g] = run->init_mode == 0 ? (double)orc_round_even(mss) : 0.0;   /* population.jl:104-105 */
            P[c * G + g] = run->init_mode == 0 ? (double)orc_round_even(pss) : 0.0;
        }
        uint32_t u[4];
        orc_draw(run->seed, (uint32_t)c, 0, 0, c3base | ORC_P_VINIT, u);
        V[c] = pop->V_init_lo + (pop->V_init_hi - pop->V_init_lo) * orc_u52(u[0], u[1]);   /* population.jl:108-109 */
        orc_promoter_init(&tele, G, run->seed, c3base, (uint32_t)c, on + c * G);
    }

    for (int64_t s = 0; s < run->n_steps; ++s) {
        for (int64_t c = 0; c < N; ++c) {
            double* Mc = M + c * G; double* Pc = P + c * G; uint8_t* oc = on + c * G;
            if (tele.telegraph_any)
                orc_promoter_switch(&tele, G, run->seed, c3base, (uint32_t)c, (uint32_t)s, oc);
            for (int g = 0; g < G; ++g) pf[g] = Pc[g];                          /* counts, :213-215 */
            orc_regulation_loop(net, kin, pf, reg);                             /* :218 */
            for (int g = 0; g < G; ++g) {
                orc_ustream st = {run->seed, (uint32_t)c, (uint32_t)s, c3t, (uint32_t)g << 16, 2, {0, 0, 0, 0}};
                double prop = reg[g] + net->basals[g];
                if (!(prop > 0.0)) prop = 0.0;
                prop = prop * V[c];                                             /* :221 */
                if (tele.telegraph_any && !oc[g]) prop = 0.0;
                int64_t m = (int64_t)Mc[g], p = (int64_t)Pc[g];
                int64_t n_tx = orc_poisson_stream(prop * dt, &st);               /* :224 */
                int64_t n_md = m > 0 ? orc_binomial_stream_lq(m, pdm, lqm, &st) : 0;   /* :227-233 */
                m += n_tx - n_md;
                int64_t mpos = m > 0 ? m : 0;
                int64_t n_tl = orc_poisson_stream(kin->k_t * (double)mpos * dt, &st);   /* :237-238 */
                int64_t n_pd = p > 0 ? orc_binomial_stream_lq(p, pdp, lqp, &st) : 0;   /* :241-247 */
                p += n_tl - n_pd;
                Mc[g] = (double)m;
                Pc[g] = (double)p;
            }
        }
        for (int64_t c = 0; c < N; ++c) {                                       /* :253 */
            if (pop->growth_noise > 0.0) {                                      /* population.jl:130-134 */
                uint32_t u[4];
                orc_draw(run->seed, (uint32_t)c, (uint32_t)s, 0, c3base | ORC_P_GROWTH, u);
                const double xi = sqrt(-2.0 * log(orc_u52(u[0], u[1]))) * cos(6.283185307179586 * orc_u52(u[2], u[3]));
                V[c] *= exp((pop->growth_rate + pop->growth_noise * xi) * dt);
            } else {
                V[c] *= growth;                                                 /* population.jl:136-137 */
            }
        }
        if (((s + 1) % pop->div_check_interval) == 0)                            /* :256-258 */
            orc_division_check(pop, G, N, run->seed, c3base, s, M, P, V, on, div_idx, dM, dP, dV, dOn, 0.0, NULL,
                               NULL);
    }
    for (int64_t c = 0; c < N; ++c) {
        for (int g = 0; g < G; ++g) {
            if (m_out) m_out[(size_t)g * N + c] = M[c * G + g];
            if (p_out) p_out[(size_t)g * N + c] = P[c * G + g];
            if (promoter_out) promoter_out[(size_t)g * N + c] = on[c * G + g];
        }
        if (V_out) V_out[c] = V[c];
    }
    free(M); free(P); free(V); free(on); free(reg); free(pf); free(div_idx);
    free(dM); free(dP); free(dV); free(dOn);
    orc_tele_free(&tele);
    return 0;
}
