// lh_frame_noalloc (include/lh/lh.h): the frame of lh_frame with its raster
// scratch in a file-scope device array instead of the caller's workspace, for
// callers that cannot allocate device memory.  No counterpart in the
// reference.
//
// The array's name is a host-side stand-in in host code, not a device
// address: the device address is looked up with hipGetSymbolAddress once per
// device (with that device current) and only that address is handed to the
// kernels, through the view's scratch field.  Nothing here allocates.
#include <mutex>

#include "capi_common.h"

namespace {

static __device__ __align__(256) unsigned char scratch[LH_STATIC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};

// Device address of this device's copy of `scratch`, looked up on first use.
lh_status device_scratch(void** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return LH_ERR_HIP;
    if (dev < 0 || dev >= kMaxDevices) return LH_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return LH_ERR_HIP;
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return LH_OK;
}

}  // namespace

extern "C" lh_status lh_frame_noalloc(const lh_scene* s, const lh_view* v, lh_stream stream) {
    if (!s || !v) return LH_ERR_INVALID;
    lh_view w = *v;
    if (lhs::scratch_layout(w).end > LH_STATIC_SCRATCH_BYTES) return LH_ERR_CAPACITY;
    void* p = nullptr;
    const lh_status st = device_scratch(&p);
    if (st != LH_OK) return st;
    w.scratch = p;
    w.scratch_bytes = LH_STATIC_SCRATCH_BYTES;
    if (!lhc::valid_frame(s, &w)) return LH_ERR_INVALID;
    return lhc::hip_status(lhk::record_frame(*s, w, lhc::stream_of(stream)));
}
