// Reference-structured FP32 CLE: the per-step launch sequence of the CUDA.jl
// extension (ext/SyntheticscRNAseqCUDAExt.jl:87-108) on MI355X -- one
// elementwise kernel per broadcast line, two SGEMMs (rocBLAS) for
// A_pos*act and A_neg*rep, normals drawn into buffers -- with every
// intermediate in HBM.  It exists as the "plain FP32 GPU CLE" baseline that
// the fused and MFMA steppers are measured against, and computes the same
// trajectories: same Philox stream, same update, AND/OR gates in one extra
// kernel.  Telegraph promoters and population dynamics are not offered here
// (the reference GPU path has neither).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <stdint.h>

#include <cstring>

#include "blob.h"
#include "device_rng.h"
#include "runtime_internal.h"

namespace scrna {

__device__ __forceinline__ float u_as_f(uint32_t w) { return __uint_as_float(w); }

// ext:88  P_n = P^n
__global__ void k_pow(const float* __restrict__ P, float* __restrict__ Pn, float n, int64_t N) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) Pn[i] = powf(P[i], n);
}
// ext:89  denom = P_n + K_n
__global__ void k_denom(const float* __restrict__ Pn, float* __restrict__ den, float Kn, int64_t N) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) den[i] = Pn[i] + Kn;
}
// ext:90  act = P_n / denom
__global__ void k_act(const float* __restrict__ Pn, const float* __restrict__ den, float* __restrict__ act,
                      int64_t N) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) act[i] = Pn[i] / den[i];
}
// ext:91  rep = K_n / denom
__global__ void k_rep(const float* __restrict__ den, float* __restrict__ rep, float Kn, int64_t N) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) rep[i] = Kn / den[i];
}
// ext:95  reg += temp
__global__ void k_add(float* __restrict__ reg, const float* __restrict__ tmp, int64_t N) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) reg[i] += tmp[i];
}
// cle.jl:89-106 (absent from the CUDA extension): one thread per cell walks the gate records.
__global__ void k_gates(const uint32_t* __restrict__ B, const float* __restrict__ act, float* __restrict__ reg, int G,
                        int C) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const int32_t* gate_off = (const int32_t*)(B + B[H_OFF_GATEOFF]);
    const uint32_t* gates = B + B[H_OFF_GATES];
    const int32_t* gene_of_slot = (const int32_t*)(B + B[H_OFF_SLOTGENE]);
    for (int g = 0; g < G; ++g) {
        for (int q = gate_off[g]; q < gate_off[g + 1];) {
            const uint32_t hd = gates[q];
            const uint32_t type = hd >> 24;
            const int ns = (int)(hd & 0xffffu);
            float prod = 1.0f;
            for (int i = 0; i < ns; ++i) {
                const float a = act[(size_t)gene_of_slot[gates[q + 2 + i]] * C + c];
                prod *= type == GATE_AND ? a : 1.0f - a;
            }
            reg[(size_t)g * C + c] += u_as_f(gates[q + 1]) * (type == GATE_AND ? prod : 1.0f - prod);
            q += 2 + ns;
        }
    }
}
// ext:97 / ext:103  randn!: normals of one species for every (gene, cell) at step s,
// drawn from the same Philox blocks as the fused steppers.
__global__ void k_randn(float* __restrict__ out, int G, int C, uint32_t step, uint32_t c3, uint32_t k0, uint32_t k1,
                        int species) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)G * C) return;
    const int g = (int)(i / C), c = (int)(i - (int64_t)g * C);
    const U4 u = philox4x32_10((uint32_t)c, step, (uint32_t)(g >> 1), c3, k0, k1);
    float z0, z1;
    normal_pair_scaled((g & 1) ? u.z : u.x, (g & 1) ? u.w : u.y, -1.3862943611198906f, z0, z1);   // -2 ln 2
    out[i] = species == 0 ? z0 : z1;
}
// ext:99-100  noise_m = noise*sqrt(dt)*sqrt(max(reg+beta+mu_m*max(M,0),0)); M = max(M + (reg+beta-mu_m*M)*dt + noise, 0)
__global__ void k_noise_m(float* __restrict__ nz, const float* __restrict__ reg, const float* __restrict__ M,
                          const uint32_t* __restrict__ beta_w, float mu_m, float sqrt_dt, int C, int64_t N) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float beta = u_as_f(beta_w[i / C]);
    nz[i] = nz[i] * sqrt_dt * sqrtf(fmaxf(reg[i] + beta + mu_m * fmaxf(M[i], 0.0f), 0.0f));
}
__global__ void k_update_m(float* __restrict__ M, const float* __restrict__ reg, const float* __restrict__ nz,
                           const uint32_t* __restrict__ beta_w, float mu_m, float dt, int C, int64_t N) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float beta = u_as_f(beta_w[i / C]);
    M[i] = fmaxf(M[i] + (reg[i] + beta - mu_m * M[i]) * dt + nz[i], 0.0f);
}
// ext:105-106
__global__ void k_noise_p(float* __restrict__ nz, const float* __restrict__ M, const float* __restrict__ P, float k_t,
                          float mu_p, float sqrt_dt, int64_t N) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    nz[i] = nz[i] * sqrt_dt * sqrtf(fmaxf(k_t * fmaxf(M[i], 0.0f) + mu_p * fmaxf(P[i], 0.0f), 0.0f));
}
__global__ void k_update_p(float* __restrict__ P, const float* __restrict__ M, const float* __restrict__ nz, float k_t,
                           float mu_p, float dt, int64_t N) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    P[i] = fmaxf(P[i] + (k_t * M[i] - mu_p * P[i]) * dt + nz[i], 0.0f);
}
// ext:63-66 initial state
__global__ void k_init(float* __restrict__ M, float* __restrict__ P, const uint32_t* __restrict__ initw, int C,
                       int64_t N) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int g = (int)(i / C);
    M[i] = u_as_f(initw[2 * g]);
    P[i] = u_as_f(initw[2 * g + 1]);
}

static inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

// One network: allocate work buffers, then the reference's per-step launch sequence.
int unfused_prepare(scrna_plan* P, UnfusedNet& u) {
    const CompiledNet& cn = P->nets[u.net];
    const size_t GC = (size_t)cn.G * (size_t)cn.cells;
    void* p = nullptr;
    int rc;
    for (float** b : {&u.Pn, &u.den, &u.act, &u.rep, &u.reg, &u.tmp, &u.nz}) {
        if ((rc = plan_alloc(P, &p, GC * sizeof(float)))) return rc;
        *b = (float*)p;
    }
    if ((rc = plan_alloc(P, &p, 2 * (size_t)cn.G * cn.G * sizeof(float)))) return rc;
    u.Apos = (float*)p;
    u.Aneg = u.Apos + (size_t)cn.G * cn.G;
    if (hipMemcpy(u.Apos, u.host_A.data(), u.host_A.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return set_error(SCRNA_ERR_DEVICE, "upload of A_pos/A_neg failed");
    if (rocblas_create_handle((rocblas_handle*)&u.blas) != rocblas_status_success)
        return set_error(SCRNA_ERR_DEVICE, "rocblas_create_handle failed");
    return SCRNA_OK;
}

void unfused_release(UnfusedNet& u) {
    if (u.blas) rocblas_destroy_handle((rocblas_handle)u.blas);
    u.blas = nullptr;
}

#define HIP_TRY_U(expr)                                                                          \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess)                                                                    \
            return set_error(SCRNA_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(_e));  \
    } while (0)

int unfused_launch(scrna_plan* P, UnfusedNet& u, hipStream_t st) {
    const CompiledNet& cn = P->nets[u.net];
    const uint32_t* hb = cn.blob.data();
    const uint32_t* B = P->d_blobs + cn.blob_off;
    const int G = cn.G, C = (int)cn.cells;
    const int64_t N = (int64_t)G * C;
    float* M = P->d_state + cn.off_state;
    float* Pp = M + N;
    const uint32_t* beta_w = B + hb[H_OFF_BETA];
    auto hf = [](uint32_t w) {
        float f;
        std::memcpy(&f, &w, 4);
        return f;
    };
    const float dt = hf(hb[H_DT]), sqrt_dt = hf(hb[H_SQRTDT]);
    const float k_t = hf(hb[H_KT]), mu_m = hf(hb[H_MUM]), mu_p = hf(hb[H_MUP]);
    const float n = u.hill_n, Kn = u.Kn;
    const uint32_t k0 = hb[H_SEED_LO], k1 = hb[H_SEED_HI], c3 = (hb[H_NETID] << 4) | kNormal;
    const dim3 gr = grid_for(N), bl(256);
    rocblas_handle h = (rocblas_handle)u.blas;
    if (rocblas_set_stream(h, st) != rocblas_status_success) return set_error(SCRNA_ERR_DEVICE, "rocblas_set_stream");
    const float one = 1.0f, zero = 0.0f;
    hipLaunchKernelGGL(k_init, gr, bl, 0, st, M, Pp, B + hb[H_OFF_INIT], C, N);
    for (int64_t s = 0; s < cn.n_steps; ++s) {
        hipLaunchKernelGGL(k_pow, gr, bl, 0, st, Pp, u.Pn, n, N);
        hipLaunchKernelGGL(k_denom, gr, bl, 0, st, u.Pn, u.den, Kn, N);
        hipLaunchKernelGGL(k_act, gr, bl, 0, st, u.Pn, u.den, u.act, N);
        hipLaunchKernelGGL(k_rep, gr, bl, 0, st, u.den, u.rep, Kn, N);
        // reg[C x G] = act[C x G] * A_pos^T, tmp = rep * A_neg^T  (A stored column-major G x G)
        if (rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_transpose, C, G, G, &one, u.act, C, u.Apos, G,
                          &zero, u.reg, C) != rocblas_status_success ||
            rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_transpose, C, G, G, &one, u.rep, C, u.Aneg, G,
                          &zero, u.tmp, C) != rocblas_status_success)
            return set_error(SCRNA_ERR_DEVICE, "rocblas_sgemm failed");
        hipLaunchKernelGGL(k_add, gr, bl, 0, st, u.reg, u.tmp, N);
        if (u.has_gates) hipLaunchKernelGGL(k_gates, grid_for(C), bl, 0, st, B, u.act, u.reg, G, C);
        hipLaunchKernelGGL(k_randn, gr, bl, 0, st, u.nz, G, C, (uint32_t)s, c3, k0, k1, 0);
        hipLaunchKernelGGL(k_noise_m, gr, bl, 0, st, u.nz, u.reg, M, beta_w, mu_m, sqrt_dt, C, N);
        hipLaunchKernelGGL(k_update_m, gr, bl, 0, st, M, u.reg, u.nz, beta_w, mu_m, dt, C, N);
        hipLaunchKernelGGL(k_randn, gr, bl, 0, st, u.nz, G, C, (uint32_t)s, c3, k0, k1, 1);
        hipLaunchKernelGGL(k_noise_p, gr, bl, 0, st, u.nz, M, Pp, k_t, mu_p, sqrt_dt, N);
        hipLaunchKernelGGL(k_update_p, gr, bl, 0, st, Pp, M, u.nz, k_t, mu_p, dt, N);
        if (cn.snap_every > 0 && ((s + 1) % cn.snap_every) == 0) {
            const int64_t snap = (s + 1) / cn.snap_every - 1;
            HIP_TRY_U(hipMemcpyAsync(P->d_snap + cn.off_snap + (size_t)snap * 2 * N, M, 2 * N * sizeof(float),
                                     hipMemcpyDeviceToDevice, st));
        }
    }
    HIP_TRY_U(hipGetLastError());
    return SCRNA_OK;
}

}  // namespace scrna
