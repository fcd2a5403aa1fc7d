/*
 * libscrna -- MI355X (gfx950) Chemical Langevin Equation simulator for
 * synthetic scRNA-seq, C ABI.
 *
 * This is the drop-in boundary for the GPU CLE path of SyntheticscRNAseq.jl.
 * A Julia shim `ccall`s these entry points in place of the CUDA.jl package
 * extension ext/SyntheticscRNAseqCUDAExt.jl; every entry point below names
 * the reference method it replaces.
 *
 * Conventions
 *  - Plain C types only.  Every pointer is a host pointer unless the name
 *    says `_dev`.  Doubles in, doubles or int32 out, like the reference.
 *  - Gene and cell indices are 0-based (the Julia shim subtracts 1).
 *  - Matrices use Julia's column-major layout.  The interaction matrix is
 *    A[i + j*G] = effect of protein j on gene i (network.jl:43-45).  Result
 *    matrices are cells x G column-major, i.e. element (cell c, gene g) at
 *    out[g*ld + c], which is exactly `Matrix{Float64}(cell_num, G)` with
 *    ld = cell_num.
 *  - Every function returns SCRNA_OK (0) or a negative scrna_status;
 *    scrna_last_error() returns a thread-local message for the last failure.
 *  - Blocking calls never keep host pointers after they return.
 */
#ifndef SCRNA_H
#define SCRNA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCRNA_ABI_VERSION 3

typedef enum scrna_status {
    SCRNA_OK = 0,
    SCRNA_ERR_INVALID = -1,     /* invalid argument (Julia: ArgumentError / AssertionError) */
    SCRNA_ERR_BUDGET = -2,      /* request exceeds the context's HBM budget */
    SCRNA_ERR_DEVICE = -3,      /* HIP runtime / device failure */
    SCRNA_ERR_CANCELLED = -4,   /* serving request was cancelled */
    SCRNA_ERR_TIMEOUT = -5,     /* scrna_wait timed out */
    SCRNA_ERR_UNSUPPORTED = -6  /* combination not supported (Julia: error()) */
} scrna_status;

typedef enum scrna_readout {
    SCRNA_READOUT_MRNA = 0,     /* :mrna */
    SCRNA_READOUT_PROTEIN = 1,  /* :protein (reference default, cle.jl:37 / ext:52) */
    SCRNA_READOUT_BOTH = 2      /* :both -> cells x 2G, mRNA columns first (cle.jl:128) */
} scrna_readout;

typedef enum scrna_precision {
    SCRNA_PREC_FP32 = 0,        /* FP32 state and regulation; the library picks the form: the dense matrix-core
                                   contraction (SCRNA_PREC_FP32_MFMA) for a CLE network without population
                                   dynamics with G >= 128 and at least 16 additive in-edges per gene on
                                   average, or with G > 256; the sparse VALU stepper (SCRNA_PREC_FP32_VALU)
                                   otherwise.  Above 256 genes the BF16, BF16X3 and FP32_MFMA forms run as
                                   one library GEMM and one fused elementwise kernel per step (the wide
                                   path), without population dynamics */
    SCRNA_PREC_BF16 = 1,        /* dense regulation on BF16 MFMA, FP32 accumulate */
    SCRNA_PREC_BF16X3 = 2,      /* dense regulation as hi*hi + hi*lo + lo*hi BF16 MFMA */
    SCRNA_PREC_FP32_UNFUSED = 3,/* the reference GPU structure (ext:87-108): one kernel per
                                   broadcast, two SGEMMs per step, state in HBM -- a baseline */
    SCRNA_PREC_FP32_MFMA = 4,   /* dense regulation in FP32 on the matrix cores (FP32 products and
                                   accumulation), FP32 state */
    SCRNA_PREC_FP32_VALU = 5    /* sparse regulation in FP32 on the VALU (in-edges per gene), FP32 state */
} scrna_precision;

/* Simulation algorithm of a run (the reference dispatches on the algorithm
 * type: simulate(net, alg::CLE, ...) ext:48-52, simulate(net,
 * alg::BinomialTauLeap, ...) ext:48-117 / tauleap_binomial.jl:31-38). */
typedef enum scrna_algorithm {
    SCRNA_ALG_CLE = 0,                 /* Chemical Langevin Equation, Euler-Maruyama (cle.jl) */
    SCRNA_ALG_BINOMIAL_TAU_LEAP = 1    /* binomial tau-leap (tauleap_binomial.jl), exact Poisson and
                                          binomial draws, FP64 propensities, integer counts */
} scrna_algorithm;

/* GeneNetwork (network.jl:55-63) with CoopEdge / RedunEdge (network.jl:16-34).
 * Gate sources are listed in CSR form: gate e has sources
 * src[off[e] .. off[e+1]-1].  k_on / k_off may be NULL (all constitutive);
 * INFINITY marks a constitutive gene (network.jl:48-49). */
typedef struct scrna_network {
    int32_t G;
    const double* basals;         /* [G] */
    const double* A;              /* [G*G] column-major, diagonal must be 0 */
    int32_t n_coop;
    const int32_t* coop_target;   /* [n_coop] */
    const int32_t* coop_off;      /* [n_coop+1] */
    const int32_t* coop_src;
    const double* coop_strength;  /* [n_coop] */
    int32_t n_redun;
    const int32_t* redun_target;
    const int32_t* redun_off;
    const int32_t* redun_src;
    const double* redun_strength;
    const double* k_on;           /* [G] or NULL */
    const double* k_off;          /* [G] or NULL */
} scrna_network;

/* KineticParams (kinetics.jl:25-32). */
typedef struct scrna_kinetics {
    double k_t, K_d, n, mu_m, mu_p, dilution;
} scrna_kinetics;

/* PopulationConfig (population.jl:27-45) plus the two choices the three
 * reference implementations disagree on. */
typedef struct scrna_population {
    int32_t enabled;
    double growth_rate;
    double V_div;
    double V_init_lo, V_init_hi;
    int32_t div_check_interval;
    int32_t partition;            /* 0: halve (cle.jl:249-261, ext:494-505); 1: Binomial(round(x), 1/2) (population.jl:166-192) */
    int32_t replace_mode;         /* 0: uniform over 1:N (population.jl:196, ext:509); 1: exclude mother (cle.jl:265-268) */
    double growth_noise;          /* sigma of population.jl:128-139: with the binomial tau-leap each cell's volume
                                     grows by exp((growth_rate + sigma*xi)*dt) per step (tauleap_binomial.jl:253);
                                     the population CLE grows deterministically and ignores it (cle.jl:169) */
} scrna_population;

/* CaptureModel (population.jl:57-74). */
typedef struct scrna_capture {
    int32_t enabled;
    double efficiency;
    double efficiency_std;
    int32_t readout;              /* SCRNA_READOUT_MRNA or SCRNA_READOUT_PROTEIN */
} scrna_capture;

/* Run options: CLE(dt) (cle.jl:21-23), cell_num, T, readout, plus the
 * options the reference GPU path lacks (seed, init, snapshots, precision). */
typedef struct scrna_run {
    double dt;
    double T;                     /* n_steps = ceil(T/dt) in double (cle.jl:56, ext:69) */
    int64_t cell_num;
    int32_t readout;              /* scrna_readout */
    int32_t init_mode;            /* 0: steady state, ceil for the CLE (cle.jl:52-53), rounded half to
                                     even for the tau-leap (tauleap_binomial.jl:61-62); 1: zeros */
    uint64_t seed;                /* Philox key */
    uint32_t net_id;              /* stream id of this network inside its request */
    int32_t precision;            /* scrna_precision */
    int64_t snapshot_every;       /* steps between snapshots; 0 = none */
    const float* injected_noise;  /* optional, host: [((s*G+g)*2+k)*cell_num + c], k=0 mRNA, 1 protein */
    int32_t algorithm;            /* scrna_algorithm (0 = CLE) */
    int32_t reserved0;            /* must be 0 */
} scrna_run;

/* Outputs of one network (all optional except y).
 *  y:          readout, cells x G (x2G for BOTH) column-major, leading dim ld
 *  counts:     captured counts, cells x G column-major int32, leading dim ld
 *  snapshots:  [n_snap][2][G][cell_num] float (species 0 = mRNA, 1 = protein);
 *              tau-leap counts in snapshots are exact up to 2^24, y, counts and
 *              moments carry them exactly (int32 state) up to 2^31 - 1
 *  moments:    [n_snap_or_1][G][5] double sums over cells: m, m^2, p, p^2, m*p
 *              (final state when there are no snapshots)
 *  volumes:    [cell_num] final cell volumes (population runs) */
typedef struct scrna_output {
    double* y;
    int64_t ld;
    int32_t* counts;
    float* snapshots;
    double* moments;
    double* volumes;
} scrna_output;

typedef struct scrna_ctx scrna_ctx;
typedef struct scrna_plan scrna_plan;

/* ---- library -------------------------------------------------------- */
int32_t scrna_abi_version(void);
const char* scrna_version(void);
const char* scrna_last_error(void);

/* Context = one device + one HBM budget.  hbm_budget_bytes 0 = 90% of free. */
int scrna_ctx_create(int32_t device, uint64_t hbm_budget_bytes, scrna_ctx** out);
int scrna_ctx_destroy(scrna_ctx* ctx);

/* ---- blocking drop-ins ---------------------------------------------- */

/* Replaces SyntheticscRNAseq.simulate(net, ::CLE, kin, ::Val{:gpu};
 * cell_num, T, readout)  (ext/SyntheticscRNAseqCUDAExt.jl:48-117), and with
 * pop != NULL the population method (ext:409-529).  capture != NULL fuses
 * apply_capture (capture.jl:25-47) on the final state. */
int scrna_simulate_cle(scrna_ctx* ctx, const scrna_network* net, const scrna_kinetics* kin,
                       const scrna_run* run, const scrna_population* pop,
                       const scrna_capture* capture, scrna_output* out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation): scrna_simulate_cle for one network without population
 * dynamics (there is no population argument: population runs need daughter
 * and claim buffers), allocating no device memory.  Every device buffer of the run --
 * network descriptor, tiles, schedule block, job tables, state, promoter and
 * snapshot pools, counts, moments, injected noise -- is carved from one fixed
 * 128 MiB array per device (scrna_noalloc_scratch_bytes), which each device
 * the library runs on holds until the process ends.  A run that does not fit
 * returns SCRNA_ERR_INVALID (roughly cells x G x 9 bytes of state and
 * promoters, plus 4 per captured count and 8 per snapshot element).  Results
 * are bit-identical to scrna_simulate_cle.  All calls on a device share the
 * array, which keeps the last call's contents; calls that overlapped on one
 * device would race on it, so the library serialises them per device.
 * Outputs are host buffers, as for scrna_simulate_cle. */
int scrna_simulate_cle_noalloc(scrna_ctx* ctx, const scrna_network* net, const scrna_kinetics* kin,
                               const scrna_run* run, const scrna_capture* capture, scrna_output* out);
/* Size in bytes of the fixed per-device array of scrna_simulate_cle_noalloc. */
uint64_t scrna_noalloc_scratch_bytes(void);

/* Replaces SyntheticscRNAseq.simulate_gpu_batch(networks, ::CLE, kin;
 * cell_num, T, readout)  (ext:132-213).  Unlike the reference, networks may
 * differ in G, cell_num, kinetics and horizon: runs[i] belongs to nets[i]. */
int scrna_simulate_cle_batch(scrna_ctx* ctx, int32_t n, const scrna_network* nets,
                             const scrna_kinetics* kins, const scrna_run* runs,
                             const scrna_population* pops, const scrna_capture* captures,
                             scrna_output* outs);

/* Replaces apply_capture(X, CaptureModel; rng)  (capture.jl:25-47) on the
 * GPU.  X is cells x G column-major double (ld >= cells); Y int32 same
 * layout.  The random stream is keyed by (seed, net_id). */
int scrna_apply_capture(scrna_ctx* ctx, const double* X, int64_t cells, int32_t G, int64_t ld,
                        double efficiency, double efficiency_std, uint64_t seed, uint32_t net_id,
                        int32_t* Y);

/* ---- device-resident plans (benchmarks, graphs, serving) ------------ */

/* Compile and upload n networks, allocate their device outputs. */
int scrna_plan_create(scrna_ctx* ctx, int32_t n, const scrna_network* nets,
                      const scrna_kinetics* kins, const scrna_run* runs,
                      const scrna_population* pops, const scrna_capture* captures,
                      scrna_plan** out);
/* Enqueue the whole simulation on `stream` (a hipStream_t; NULL = the
 * context's stream).  Allocates nothing; may be captured into a hipGraph. */
int scrna_plan_launch(scrna_plan* plan, void* stream);
/* Wait for the plan's stream. */
int scrna_plan_sync(scrna_plan* plan);
/* Copy network i's results to host outputs. */
int scrna_plan_download(scrna_plan* plan, int32_t i, scrna_output* out);
/* Milliseconds of the last launch's stepper kernels (HIP events recorded
 * on the launch stream around them), and the cell*gene*steps it covered. */
int scrna_plan_stepper_ms(scrna_plan* plan, double* ms, double* cell_gene_steps);
/* Device bytes held by the plan. */
int scrna_plan_bytes(scrna_plan* plan, uint64_t* bytes);
/* Capture the plan's whole kernel sequence (all steps, division phases,
 * moments, capture) into one hipGraph; later scrna_plan_launch calls replay
 * it with a single hipGraphLaunch.  Replaces the reference's host-driven
 * per-step launch loop (ext:87-108, ext:184-205, ext:450-520). */
int scrna_plan_use_graph(scrna_plan* plan);
int scrna_plan_destroy(scrna_plan* plan);

/* ---- tile-shape autotuner ------------------------------------------- */

/* Time every legal genes-per-lane chunk (2, 4, 8, 16) of the FP32 stepper on
 * a synthetic network per G bucket up to max_G and keep the fastest per
 * bucket in the context; plans created afterwards use it.  With cache_path
 * non-NULL the table is read from that JSON file when it matches the device,
 * otherwise measured and written there.  No reference counterpart (the
 * reference GPU path has no tunable tiles). */
int scrna_autotune(scrna_ctx* ctx, int32_t max_G, const char* cache_path, int32_t* n_buckets);
/* The same, timing each shape the way population plans run it: launches of
 * phase_steps steps (their div_check_interval) with the division kernels
 * between them; phase_steps = 0 is scrna_autotune.  No reference counterpart. */
int scrna_autotune_phases(scrna_ctx* ctx, int32_t max_G, int32_t phase_steps, const char* cache_path,
                          int32_t* n_buckets);
/* Pin the chunk for networks with G in [G_lo, G_hi]; genes_per_lane 0 clears. */
int scrna_ctx_set_tile(scrna_ctx* ctx, int32_t G_lo, int32_t G_hi, int32_t genes_per_lane);

/* ---- serving: continuous batching of simulate() requests ------------ */

typedef struct scrna_server scrna_server;

typedef struct scrna_request {
    scrna_network net;
    scrna_kinetics kin;
    scrna_run run;
    scrna_capture capture;
} scrna_request;

typedef struct scrna_latency {
    double p50_ms, p99_ms, mean_ms, max_ms;
    double cell_gene_steps;        /* completed work */
    double busy_s;                 /* wall time with work on the device */
    int64_t completed, rejected, cancelled;
} scrna_latency;

/* Start the device-resident request loop.  max_slots bounds concurrent
 * requests; admission is by the context's HBM budget. */
int scrna_server_create(scrna_ctx* ctx, int32_t max_slots, scrna_server** out);
/* Copies the request (network arrays included); returns a ticket id. */
int scrna_submit(scrna_server* srv, const scrna_request* req, int64_t* ticket);
/* Block until the ticket finishes (timeout_ms < 0: forever), then copy its
 * results to out.  Returns SCRNA_ERR_TIMEOUT if still running. */
int scrna_wait(scrna_server* srv, int64_t ticket, int64_t timeout_ms, scrna_output* out);
int scrna_cancel(scrna_server* srv, int64_t ticket);
int scrna_server_stats(scrna_server* srv, scrna_latency* out);
int scrna_server_destroy(scrna_server* srv);

/* ---- Network prior and theta constructor (simulation-based inference) ----
 * Host-only calls (no device work): they fill scrna_network arrays for
 * scrna_simulate_cle_batch / scrna_plan_create. */

/* Edge types of sample_network's `regulation` keyword (kinetics.jl:83-84). */
typedef enum scrna_regulation {
    SCRNA_REG_MIXED = 0,        /* additive + AND + OR, at least one non-additive edge */
    SCRNA_REG_ADDITIVE = 1,
    SCRNA_REG_COOPERATIVE = 2,  /* additive + AND */
    SCRNA_REG_REDUNDANT = 3     /* additive + OR */
} scrna_regulation;

/* sample_network's keywords (kinetics.jl:105-112).  A range whose lo is < 0
 * (or hi < lo) takes the G-dependent default of kinetics.jl:107-111. */
typedef struct scrna_prior {
    int32_t G;
    int32_t regulation;                 /* scrna_regulation */
    int32_t n_tf_lo, n_tf_hi;           /* n_tf_range */
    int32_t targets_lo, targets_hi;     /* targets_per_tf_range */
    int32_t coop_lo, coop_hi;           /* n_coop_range */
    int32_t redun_lo, redun_hi;         /* n_redun_range */
    int32_t infer_basals;               /* 1: LogNormal basals exp(0.5 z - 1); 0: exp(-1) */
} scrna_prior;

/* Capacity (gates) the output buffers of scrna_sample_network need. */
int scrna_prior_max_gates(const scrna_prior* prior, int32_t* max_gates);
/* Replaces sample_network(G; rng, ...) (kinetics.jl:105-181).  Network `index`
 * of `seed` is a pure function of (seed, index): draws come from Philox4x32-10
 * with counter (index, block, 0, 9), so batches can be generated in any order
 * or in parallel.  Outputs: basals [G]; A [G*G] column-major (A[i + j*G] =
 * effect of protein j on gene i, diagonal 0); gates: type (0 AND = CoopEdge,
 * 1 OR = RedunEdge), target, 2 sources (ascending), strength; *n_gates set.
 * All k_on / k_off of the sampled network are Inf (constitutive). */
int scrna_sample_network(const scrna_prior* prior, uint64_t seed, uint32_t index, double* basals, double* A,
                         int32_t* gate_type, int32_t* gate_target, int32_t* gate_src, double* gate_strength,
                         int32_t max_gates, int32_t* n_gates);
/* Replaces GeneNetwork(theta, G) (network.jl:233-256): n_theta = G(G-1)
 * (basals exp(-1)) or G(G-1)+G (first G entries are log-basals); interactions
 * in row-major order of the off-diagonal entries A[i,j]. */
int scrna_network_from_theta(int32_t G, const double* theta, int64_t n_theta, double* basals, double* A);

/* ---- Count-matrix writer ----
 * Writes a cells x G int32 count matrix (column-major, element (c, g) at
 * counts[g*ld + c], e.g. scrna_output.counts) as a 10x Genomics v2 directory,
 * the format the reference's real-data comparison reads
 * (experiments/compare_real_scrnaseq.jl:62-96): <dir>/matrix.mtx
 * (MatrixMarket coordinate integer general, genes x cells, 1-based),
 * <dir>/genes.tsv ("SYNG000001<TAB>name", names default to gene1..geneG) and
 * <dir>/barcodes.tsv (16-nt barcodes of the cell index with "-1", after an
 * optional prefix).  Creates <dir> if needed; *nnz_out (nullable) receives
 * the number of nonzeros.  Host-only. */
int scrna_write_10x(const char* dir, const int32_t* counts, int64_t cells, int32_t G, int64_t ld,
                    const char* const* gene_names, const char* barcode_prefix, int64_t* nnz_out);

#ifdef __cplusplus
}
#endif
#endif
