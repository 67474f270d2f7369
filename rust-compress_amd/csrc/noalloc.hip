// Allocation-free entry point: rcg_reconstruct_dev for callers that cannot allocate device memory (code under HIP
// graph capture, serving loops and harnesses that forbid hipMalloc). Its only scratch -- the IDCT's sample planes,
// which rcg_reconstruct_dev takes from the context's growing workspace -- lives in a static device array of
// RCG_NOALLOC_SCRATCH_BYTES. The loader gives every GPU its own copy of the array, at its own address, held until the
// process ends. Host code never uses the array's name as a device address: the address of each device's copy is
// looked up once, at context creation on that device, and only that pointer is handed to the kernels.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "rcg_host.h"

static __device__ __align__(256) uint8_t scratch[RCG_NOALLOC_SCRATCH_BYTES];

namespace rcg {
namespace {

constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device, once set up
std::mutex g_setup_mu;
// Calls on one GPU share its array, so they would race: they are serialised per device.
std::mutex g_device_mu[kMaxDevices];

}  // namespace

int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return RCG_OK;  // the entry point refuses such devices
    std::lock_guard<std::mutex> g(g_setup_mu);
    if (g_scratch[device]) return RCG_OK;
    void* p = nullptr;
    RCG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
        set_error("allocation-free scratch: no device address");
        return RCG_DEVICE;
    }
    g_scratch[device] = p;
    return RCG_OK;
}

}  // namespace rcg

using namespace rcg;

extern "C" int rcg_reconstruct_dev_noalloc(rcg_ctx* ctx, const rcg_coef_image* d_ci, uint8_t* d_pixels) {
    return rcg::guarded([&]() -> int {
        if (!d_pixels) return RCG_INVALID;
        uint32_t hmax, vmax;
        int rgb_direct;
        RCG_TRY(check_coef_image(d_ci, hmax, vmax, rgb_direct));
        const size_t need = sample_plane_bytes(d_ci);
        if (need > RCG_NOALLOC_SCRATCH_BYTES) {
            set_error("the image's sample planes take " + std::to_string(need) + " bytes, more than the " +
                      std::to_string(RCG_NOALLOC_SCRATCH_BYTES) + "-byte allocation-free scratch; use rcg_reconstruct_dev");
            return RCG_UNSUPPORTED;
        }
        const int16_t* d_co[4] = {nullptr, nullptr, nullptr, nullptr};
        for (uint32_t c = 0; c < d_ci->ncomp; c++) {
            if (reinterpret_cast<uintptr_t>(d_ci->comps[c].coefs) & 15) {
                set_error("coefficient buffers must be 16-byte aligned device pointers");
                return RCG_INVALID;
            }
            d_co[c] = d_ci->comps[c].coefs;
        }
        if (!ctx) return RCG_INVALID;
        if (ctx->device < 0 || ctx->device >= kMaxDevices || !g_scratch[ctx->device]) {
            set_error("allocation-free scratch is not set up on this device");
            return RCG_DEVICE;
        }
        std::lock_guard<std::mutex> ctx_lock(ctx->mu);
        std::lock_guard<std::mutex> dev_lock(g_device_mu[ctx->device]);
        RCG_HIP(hipSetDevice(ctx->device));
        uint8_t* base = static_cast<uint8_t*>(g_scratch[ctx->device]);
        uint8_t* planes[4] = {nullptr, nullptr, nullptr, nullptr};
        size_t off = 0;
        for (uint32_t c = 0; c < d_ci->ncomp; c++) {
            planes[c] = base + off;
            off += Arena::align(size_t(d_ci->comps[c].blocks_w) * d_ci->comps[c].blocks_h * 64);
        }
        // Profiling would create HIP events; this call records none.
        struct Suspend {
            rcg_ctx* c;
            explicit Suspend(rcg_ctx* x) : c(x) { c->prof_suspended = true; }
            ~Suspend() { c->prof_suspended = false; }
        } suspend(ctx);
        // The two-kernel form (IDCT into the scratch planes, then the colour stage), for every layout.
        RCG_TRY(reconstruct_on_device(ctx, d_ci, d_co, planes, d_pixels, hmax, vmax, rgb_direct, false));
        RCG_HIP(hipStreamSynchronize(ctx->stream));
        return RCG_OK;
    });
}
