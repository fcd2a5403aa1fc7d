/*
 * CPU FP64 oracle for the GPU Chemical Langevin Equation (CLE) path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libscrna.so) never links, loads or calls it.
 *
 * It restates, in plain C and double precision, the semantics of the
 * reference Julia package SyntheticscRNAseq.jl:
 *   - CLE step (src/algorithms/cle.jl:78-121), init (cle.jl:52-53), mu_eff
 *     (cle.jl:60-61), step count (cle.jl:56);
 *   - Hill functions and the A_eff split (src/network.jl:112-125, 206-219),
 *     the loop form of regulation (src/network.jl:158-193);
 *   - AND / OR gates (cle.jl:89-106);
 *   - telegraph promoters, which no reference CLE has: init and switching
 *     follow src/algorithms/tauleap_binomial.jl:64-76 and :110-121, the
 *     gating of the whole transcription propensity :134-137, and the CLE
 *     form scripts/generate_figures.jl:124-134;
 *   - population CLE (cle.jl:143-284) with the binomial partitioning and
 *     Moran replacement of src/population.jl:157-205;
 *   - LogNormal-Binomial capture (src/capture.jl:25-47);
 *   - moment statistics (src/validate.jl:70-77);
 *   - binomial tau-leap (src/algorithms/tauleap_binomial.jl:31-180, regulation
 *     mode :transcription) with exact Poisson (inversion / PTRS) and binomial
 *     (inversion / BTRD) draws.
 *
 * Random numbers: the reference draws from Julia's global or MersenneTwister
 * streams, which cannot be reproduced here.  This oracle and the GPU library
 * both draw from Philox4x32-10 (Salmon et al., SC'11) keyed by the 64-bit
 * run seed, with a counter made of (cell, step, index, net_id<<4 | purpose),
 * so that the oracle can replay the GPU's exact uniform stream.
 *
 * Index convention: genes and cells are 0-based here (Julia is 1-based).
 * Matrices are column-major like Julia's: state[g + 0] ... is not used;
 * instead every per-gene array over cells is stored gene-major with cells
 * fastest: x[g * ld + c], which is exactly Julia's column-major cells x G.
 */
#ifndef SCRNA_ORACLE_H
#define SCRNA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_network {
    int32_t G;
    const double* basals;        /* [G] */
    const double* A;             /* [G*G] column-major: A[i + j*G] = effect of protein j on gene i */
    int32_t n_coop;              /* AND gates */
    const int32_t* coop_target;  /* [n_coop] */
    const int32_t* coop_off;     /* [n_coop+1] offsets into coop_src */
    const int32_t* coop_src;
    const double* coop_strength; /* [n_coop] */
    int32_t n_redun;             /* OR gates */
    const int32_t* redun_target;
    const int32_t* redun_off;
    const int32_t* redun_src;
    const double* redun_strength;
    const double* k_on;          /* [G] or NULL; INFINITY = constitutive */
    const double* k_off;         /* [G] or NULL */
} orc_network;

typedef struct orc_kinetics {
    double k_t, K_d, n, mu_m, mu_p, dilution;
} orc_kinetics;

typedef struct orc_population {
    int32_t enabled;
    double growth_rate;
    double V_div;
    double V_init_lo, V_init_hi;
    int32_t div_check_interval;
    int32_t partition;           /* 0: halve continuously (cle.jl:249-261); 1: Binomial(round(x), 1/2) */
    int32_t replace_mode;        /* 0: uniform over all N (population.jl:194-202); 1: exclude mother (cle.jl:264-268) */
    double growth_noise;         /* population.jl:128-139 (binomial tau-leap; the CLE ignores it, cle.jl:169) */
} orc_population;

typedef struct orc_run {
    double dt;
    int64_t n_steps;
    int64_t cell_num;            /* population size (global cell count) */
    int64_t cell_lo, cell_hi;    /* simulate cells [cell_lo, cell_hi) only (no population) */
    uint64_t seed;
    uint32_t net_id;
    int32_t init_mode;           /* 0: ceil(steady state) (cle.jl:52-53); 1: zeros */
    int64_t snapshot_every;      /* 0: no snapshots */
    const float* injected;       /* optional noise [((s*G+g)*2+k)*cell_num + c] */
    int32_t n_threads;           /* OpenMP threads for the cell loop */
} orc_run;

/* Philox4x32-10: out = philox(ctr, key). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* The Box-Muller pair used by the CLE stepper for two u32 inputs. */
void orc_normal_pair(uint32_t a, uint32_t b, double* z0, double* z1);
/* CLE promoter switch decision of gene g from its normal words a, b (1 = flip). */
int orc_cle_flip(uint32_t a, uint32_t b, uint64_t thr, uint64_t seed, uint32_t cell, uint32_t step, int32_t g,
                 uint32_t c3base);

/* Regulation for one cell, matrix form (cle.jl:79-106) and loop form
 * (network.jl:158-193).  p: [G] proteins; reg: [G] output. */
void orc_regulation_matrix(const orc_network* net, const orc_kinetics* kin,
                           const double* p, double* reg);
void orc_regulation_loop(const orc_network* net, const orc_kinetics* kin,
                         const double* p, double* reg);

/* Independent-cell CLE.  m_out/p_out: [G * nc] with nc = cell_hi-cell_lo,
 * element (g, c) at g*nc + (c-cell_lo).  snaps_m/snaps_p (optional):
 * [n_snap * G * nc].  promoter_out (optional): [G * nc].
 * Returns 0 on success, nonzero on invalid input. */
int orc_cle_simulate(const orc_network* net, const orc_kinetics* kin, const orc_run* run,
                     double* m_out, double* p_out, double* snaps_m, double* snaps_p,
                     uint8_t* promoter_out);

/* Population CLE over all cell_num cells (steps outer).  V_out: [cell_num]. */
int orc_cle_population(const orc_network* net, const orc_kinetics* kin, const orc_run* run,
                       const orc_population* pop, double* m_out, double* p_out,
                       double* V_out, uint8_t* promoter_out);

/* The same with an ambiguity margin: taint_out[c] (optional, [cell_num])
 * is 1 for a cell one of whose values rounded at a binomial partition lay
 * within amb * max(1, |x|) of a half-integer, or that holds (by Moran
 * replacement) a daughter of such a cell.  An FP32 restatement whose state
 * stays within that margin of this oracle's partitions every other cell
 * identically. */
int orc_cle_population_ex(const orc_network* net, const orc_kinetics* kin, const orc_run* run,
                          const orc_population* pop, double amb, double* m_out, double* p_out,
                          double* V_out, uint8_t* promoter_out, uint8_t* taint_out);

/* One division check (population.jl:157-205) on a given state, for tests.
 * m, p: [G*N] gene-major (element (g, c) at g*N + c), updated in place; V:
 * [N] updated in place; promoter states are not modelled.  Outputs (each
 * optional): *n_div, dividers [N] (ascending), mothers' molecules after the
 * partition and daughters' ([n_div*G], divider-major), and the slot each
 * daughter replaced [N].  `step` is the 0-based step whose end the check
 * follows (it keys the partition and replacement draws). */
int orc_divide(const orc_population* pop, int32_t G, int64_t N, uint64_t seed, uint32_t net_id, int64_t step,
               double* m, double* p, double* V, int64_t* n_div, int64_t* dividers, double* mother_m,
               double* mother_p, double* daughter_m, double* daughter_p, int64_t* replaced);

/* Cell-vectorised CLE pieces for the CPU baseline restatement of cle.jl:34-130
 * (state G x cell_num column-major, element (g, c) at c*G + g; all
 * cell_num cells).  orc_ref_hill: act, rep of cle.jl:79-82.  orc_ref_init:
 * initial state and promoters.  orc_ref_update: given reg = A_pos*act +
 * A_neg*rep, the AND/OR corrections (cle.jl:89-106), telegraph switching and
 * the mRNA then protein update of step `step` (cle.jl:108-120), with the
 * draws of orc_cle_simulate, or with noise_m / noise_p (same layout as M,
 * both or neither) as this step's two noise matrices. */
void orc_ref_hill(const double* P, int64_t n, double hill_n, double Kn, double* act, double* rep);
void orc_ref_init(const orc_network* net, const orc_kinetics* kin, const orc_run* run, double* M, double* P,
                  uint8_t* on);
void orc_ref_update(const orc_network* net, const orc_kinetics* kin, const orc_run* run, int64_t step, double* M,
                    double* P, double* reg, const double* act, uint8_t* on, const double* noise_m,
                    const double* noise_p);

/* Binomial tau-leap of cells [cell_lo, cell_hi) (tauleap_binomial.jl:31-180):
 * same layouts as orc_cle_simulate; counts are returned as doubles.  init_mode
 * 0 rounds the steady state half to even (tauleap_binomial.jl:61-62).
 * injected must be NULL. */
int orc_tauleap_simulate(const orc_network* net, const orc_kinetics* kin, const orc_run* run,
                         double* m_out, double* p_out, double* snaps_m, double* snaps_p,
                         uint8_t* promoter_out);

/* Binomial tau-leap with Moran population dynamics over all cell_num cells
 * (tauleap_binomial.jl:195-261): Hill on counts, transcription scaled by the
 * cell volume, intrinsic decay only, binomial partitioning (pop->partition
 * must be 1) and Moran replacement.  V_out: [cell_num]. */
int orc_tauleap_population(const orc_network* net, const orc_kinetics* kin, const orc_run* run,
                           const orc_population* pop, double* m_out, double* p_out,
                           double* V_out, uint8_t* promoter_out);

/* Poisson(lam) drawn from the Philox stream with the given counter base
 * (blocks counted in ctr[2]).  Exposed for tests. */
int64_t orc_poisson(double lam, uint32_t ctr0, uint32_t ctr1, uint32_t c3, uint64_t seed);

/* Binomial(n, p) drawn from the Philox stream with the given counter base
 * (ctr0, ctr1, key, c3); blocks are counted in ctr[2].  Exposed for tests. */
int64_t orc_binomial(int64_t n, double p, uint32_t ctr0, uint32_t ctr1, uint32_t c3,
                     uint64_t seed);

/* LogNormal-Binomial capture (capture.jl:25-47).  X: [G*ld] element (g,c)
 * at g*ld + c; Y: same layout, int32.  Also returns per-cell efficiencies
 * when p_cell != NULL. */
int orc_capture(const double* X, int64_t cells, int32_t G, int64_t ld, double efficiency,
                double efficiency_std, uint64_t seed, uint32_t net_id, int32_t* Y,
                double* p_cell);

/* Per-gene sums over cells in double: out[g*5 + {0..4}] = sum m, sum m^2,
 * sum p, sum p^2, sum m*p. */
void orc_moments(const double* m, const double* p, int64_t cells, int32_t G, int64_t ld,
                 double* out);

#ifdef __cplusplus
}
#endif
#endif
