// Allocation-free simulate(): the same work as scrna_simulate_cle for one
// network without population dynamics, with every device buffer (network
// descriptor, tile table, schedule block, job tables, state, promoter and
// snapshot pools, counts, moments) carved from one fixed file-scope device
// array instead of allocated.  For callers that cannot allocate device
// memory: code under HIP graph capture, tight serving loops, harnesses that
// forbid allocation.
//
// The array lives in every device's copy of the code object.  Host code
// never uses its name as an address: scrna_ctx_create looks its address up
// once per device with hipGetSymbolAddress (with that device current), and
// the entry point hands that address to the kernels.  All calls on a device
// share the array and it keeps the last call's contents; every buffer the
// kernels read before writing (the schedule block) is cleared by a kernel
// at the start of each launch.  Two calls that overlapped on one device would
// race on the array, so calls on a device are serialised by a lock.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "runtime_internal.h"

namespace scrna {

constexpr size_t kNoAllocScratchBytes = 128ull << 20;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
struct DeviceScratch {
    bool looked_up = false;
    void* ptr = nullptr;           // device address of `scratch` on this device
    std::string err;
    std::mutex run_mu;             // serialises the entry point's calls on this device
};
DeviceScratch g_scratch[kNoAllocMaxDevices];
std::mutex g_lookup_mu;
}  // namespace

// Called by scrna_ctx_create with `device` current.  Never fails the
// context: a failed lookup is reported by the entry point.
void noalloc_scratch_init(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_lookup_mu);
    DeviceScratch& d = g_scratch[device];
    if (d.looked_up) return;
    d.looked_up = true;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
        (void)hipGetLastError();
        d.err = std::string("hipGetSymbolAddress of the scratch array failed: ") + hipGetErrorString(e);
        return;
    }
    d.ptr = ptr;
}

}  // namespace scrna

using namespace scrna;

extern "C" int scrna_simulate_cle_noalloc(scrna_ctx* ctx, const scrna_network* net, const scrna_kinetics* kin,
                                          const scrna_run* run, const scrna_capture* capture, scrna_output* out) {
    if (!ctx || !net || !kin || !run || !out) return set_error(SCRNA_ERR_INVALID, "null argument");
    if (ctx->device < 0 || ctx->device >= kNoAllocMaxDevices)
        return set_error(SCRNA_ERR_UNSUPPORTED, "device %d has no scratch slot", ctx->device);
    DeviceScratch& d = g_scratch[ctx->device];
    {
        std::lock_guard<std::mutex> lk(g_lookup_mu);
        if (!d.ptr)
            return set_error(SCRNA_ERR_DEVICE, "%s",
                             d.err.empty() ? "scratch array not looked up for this device" : d.err.c_str());
    }
    std::lock_guard<std::mutex> run_lk(d.run_mu);
    const hipError_t se = hipSetDevice(ctx->device);
    if (se != hipSuccess) return set_error(SCRNA_ERR_DEVICE, "hipSetDevice failed: %s", hipGetErrorString(se));
    Arena arena;
    arena.base = (char*)d.ptr;
    arena.size = kNoAllocScratchBytes;
    scrna_plan* P = nullptr;
    int rc;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        rc = plan_build(ctx, 1, net, kin, run, nullptr, capture, 0, &P, &arena);
    }
    if (rc) return rc;
    rc = plan_enqueue(P, ctx->stream, false);
    if (!rc) {
        const hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = set_error(SCRNA_ERR_DEVICE, "simulation failed: %s", hipGetErrorString(e));
    }
    if (!rc) rc = plan_check_schedule(P);
    if (!rc) rc = plan_download(P, 0, out);
    plan_release(P);
    delete P;
    return rc;
}

extern "C" uint64_t scrna_noalloc_scratch_bytes(void) { return kNoAllocScratchBytes; }
