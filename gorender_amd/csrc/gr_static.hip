// The draw that allocates nothing (gr_draw_device_noalloc): every per-draw
// buffer of the render path (batch header, triangle records, tile counters and
// lists, control words) is carved from one fixed device array of this file,
// whose address is resolved once per device when a context is created.
//
// The array is shared by every call on a GPU and keeps the last call's
// contents: the draw clears what must start at zero itself, and calls that
// overlap on one GPU (two contexts or streams at once) race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "gorender_hip.h"
#include "gr_error.h"
#include "gr_internal.h"

namespace {

// 32 MiB per GPU, held from the library's load on the device until the
// process ends.
constexpr size_t kScratchBytes = 32u << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices];  // device address of `scratch` per device

}  // namespace

namespace gr_internal {

// With `device` current: resolve the array's address on it (once per device).
gr_status static_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices)
        return fail(GR_ERR_INVALID_ARGUMENT, "device index beyond the no-allocation table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return GR_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(GR_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                       hipGetErrorString(e));
    g_scratch[device] = p;
    return GR_OK;
}

void* static_scratch(int device, size_t* bytes) {
    *bytes = kScratchBytes;
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_scratch[device];
}

}  // namespace gr_internal

extern "C" gr_status gr_draw_device_noalloc(gr_context* ctx, const gr_frame* frame,
                                            const gr_object* objects, int32_t num_objects,
                                            void* d_rgba, void* d_depth) {
    if (!ctx || !frame || num_objects < 0 || (num_objects > 0 && !objects))
        return gr_internal::fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    size_t bytes = 0;
    void* base = gr_internal::static_scratch(gr_internal::context_device(ctx), &bytes);
    if (!base)
        return gr_internal::fail(GR_ERR_DEVICE, "no-allocation scratch not resolved on this device");
    return gr_internal::enqueue_noalloc(ctx, frame, objects, num_objects, d_rgba, d_depth, base,
                                        bytes);
}
