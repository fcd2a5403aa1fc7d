// Packed per-network descriptor ("blob") read by the CLE kernels.
//
// One blob per network, 32-bit words, 16-word aligned inside a single device
// buffer.  Every workgroup serves one network, so each word is uniform across
// the wave and is fetched through the scalar cache.
//
// Header words (floats stored by bit pattern):
#pragma once
#include <stdint.h>

namespace scrna {

enum BlobHeader : int {
    H_G = 0,          // genes
    H_S,              // source slots (genes whose Hill value some edge reads)
    H_FLAGS,          // BlobFlags
    H_NSTEPS,         // total steps n = ceil(T/dt)
    H_SNAP_EVERY,     // steps between snapshots (0 = none)
    H_NSNAP,          // number of snapshots
    H_NETID,          // Philox stream id (counter word 3 = net_id<<4 | purpose)
    H_CELLS,          // cells of this network
    H_SEED_LO,
    H_SEED_HI,
    H_DT,             // f32 dt
    H_NEG2LN2DT,      // f32 -2 ln(2) dt  (Box-Muller radius, scaled by sqrt(dt))
    H_KT,             // f32 k_t
    H_MUM,            // f32 effective mRNA decay
    H_MUP,            // f32 effective protein decay
    H_INVK,           // f32 1/K_d
    H_HILLN,          // f32 Hill coefficient
    H_SQRTDT,         // f32 sqrt(dt) (injected-noise scaling)
    H_POPMODE,        // bit0: binomial partition (else halve), bit1: replacement excludes the mother
    H_DIVINT,         // steps between division checks
    H_OFF_BETA,       // [G] f32 basal rates
    H_OFF_SLOT,       // [G] i32 source slot of each gene, -1 if none
    H_OFF_INOFF,      // [G+1] i32 CSR offsets (edge index) of additive in-edges per target
    H_OFF_EDGES,      // [2*E] (lds_row = 2*slot + kind, f32 weight); kind 0 = act, 1 = rep
    H_OFF_GATEOFF,    // [G+1] i32 word offsets (relative to H_OFF_GATES) of each target's gates
    H_OFF_GATES,      // gate records: (type<<24 | nsrc), f32 strength*beta_t, nsrc slot words
    H_OFF_TELE,       // [G][8] telegraph words, see TeleWord
    H_OFF_INIT,       // [G][2] f32 initial m, p
    H_OUT_FINAL_LO,   // element offset of [2][G][cells] f32 final state in the state pool
    H_OUT_FINAL_HI,
    H_OUT_SNAP_LO,    // element offset of [nsnap][2][G][cells] f32 snapshots in the snapshot pool
    H_OUT_SNAP_HI,
    H_OUT_PROM_LO,    // element offset of [G][cells] u8 promoter state in the promoter pool
    H_OUT_PROM_HI,
    H_OUT_V_LO,       // element offset of [cells] f32 volumes in the volume pool
    H_OUT_V_HI,
    H_INJ_LO,         // element offset of injected noise [n][G][2][cells] f32 (F_INJECT)
    H_INJ_HI,
    H_CH,             // genes per lane the network was compiled for
    H_NW,             // waves per workgroup (NW*CH >= G)
    H_GROWTH_LO,      // f64 exp(growth_rate*dt) (1 without population)
    H_GROWTH_HI,
    H_VDIV_LO,        // f64 division volume
    H_VDIV_HI,
    H_GP,             // dense path: G rounded up to a multiple of 32
    H_OFF_BIAS,       // dense path: [G] f32 rowsum(A_neg)
    H_DENSE_LO,       // dense path: offset (bf16x8 units) of the packed W^T in the weight pool
    H_DENSE_HI,
    H_OFF_SLOTGENE,   // [S] i32 gene of each source slot
    H_ALG,            // scrna_algorithm the network was compiled for
    H_OFF_TL,         // tau-leap: FP64 section (TlWord head, then beta[G], edge weights[E],
                      // gate weights by gate word [W], initial counts [G][2])
    H_TL_NEDGES,      // tau-leap: E
    H_TL_NGATEW,      // tau-leap: W (gate record words)
    H_OFF_WREC,       // FP32 stepper with up to 4 genes per lane: [NW][rec_stride(CH)] wave records (RecWord)
    H_WORDS = 64
};

// Doubles at the head of the tau-leap FP64 section.
enum TlWord : int {
    TL_KN = 0,        // K_d^n
    TL_HILLN,         // n
    TL_KT,            // k_t
    TL_DT,            // dt
    TL_PDM,           // 1 - exp(-(mu_m + dilution) dt), tauleap_binomial.jl:144
    TL_LQM,           // log(1 - TL_PDM)
    TL_LPM,           // log(TL_PDM)
    TL_PDP,           // 1 - exp(-(mu_p + dilution) dt), tauleap_binomial.jl:157
    TL_LQP,           // log(1 - TL_PDP)
    TL_LPP,           // log(TL_PDP)
    TL_GROWTH_RATE,   // population growth rate lambda (volume noise: V *= exp((lambda + sigma xi) dt))
    TL_GROWTH_NOISE,  // its sigma (population.jl:128-139); 0 = deterministic growth by H_GROWTH
    TL_HEAD = 16
};

// After the initial counts the tau-leap FP64 section holds three tables over
// a count n < kTlTable: exp(-(k_t n dt)) (translation's Poisson start),
// exp(n log q_m) and exp(n log q_p) (the decay binomials' inversion start,
// q the complement of the decay probability or, above 1/2, the probability).
constexpr int kTlTable = 256;

enum BlobFlags : uint32_t {
    F_TELE = 1u,      // some gene has a finite switching rate
    F_HILL2 = 2u,     // Hill coefficient is exactly 2
    F_GATES = 4u,     // AND / OR gates present
    F_INJECT = 8u,    // read normals from the injected-noise buffer
    F_POP = 16u,      // population dynamics (volume-scaled transcription)
    F_NEG_INIT = 32u  // some initial count is negative (negative basal rates)
};

// Wave record of the FP32 stepper with up to 4 genes per lane: the values a
// wave keeps in registers for the whole horizon, laid out by the host so that
// a workgroup's start-up reads them with a few wide scalar loads.  Per gene
// (kRecGeneWords words): the switch thresholds in units of 2^14 (0 when the
// gene never switches), the byte offset of its Hill row pair in an LDS buffer
// (sources only), its basal rate and kRecEdgeSlots in-edges as (byte offset of
// the LDS row read, f32 weight), unused ones reading the buffer's zero row
// with weight 0.  Then the wave's masks: switching genes, sources, and genes
// with more in-edges (bits 0..15) or gates (bits 16..31).
constexpr int kRecEdgeSlots = 2;
enum RecWord : int {
    R_TON = 0,
    R_TOFF,
    R_HROW,
    R_BETA,
    R_EDGE,           // kRecEdgeSlots (row offset, weight) pairs
    R_GENE_WORDS = R_EDGE + 2 * kRecEdgeSlots
};
enum RecMask : int { RM_SWITCH = 0, RM_SOURCE, RM_EXTRA, RM_WORDS };
constexpr int rec_stride(int CH) { return ((R_GENE_WORDS * CH + RM_WORDS + 15) / 16) * 16; }

enum TeleWord : int {
    T_FLAGS = 0,      // bit0: switches (isfinite k_on), bit1: random init (both finite)
    T_ON_LO = 2,      // u64 threshold OFF->ON: switch when u32 < thr
    T_ON_HI = 3,
    T_OFF_LO = 4,     // u64 threshold ON->OFF
    T_OFF_HI = 5,
    T_INIT_LO = 6,    // u64 threshold for starting ON
    T_INIT_HI = 7
};

enum GateType : uint32_t { GATE_AND = 0u, GATE_OR = 1u };

// Tile of work: one workgroup = up to 64 cells (16 on the dense path) of one
// network, advanced over steps [s0, s1) intersected with the launch's range
// and the network's horizon.  A tile whose range starts after step 0 resumes
// from the state pool.
struct Tile {
    int32_t blob;        // word offset of the network's blob
    int32_t cell0;       // first cell
    int32_t ncell;       // cells in this tile
    int32_t net;         // network index inside the plan
    int32_t s0, s1;      // step range of this tile
    int32_t pad0, pad1;
};

}  // namespace scrna
