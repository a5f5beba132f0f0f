// gps_noalloc.hip — one frame of the render pass with no device allocation.
//
// gps_frame_render_noalloc is gps_pass_render of a pass created for that one
// frame (block and occlusion culling off), whose buffers — weights, valid
// masks, splat records, tile sums and offsets, compacted entries, chunk
// starts, counters and the framebuffer — live in `scratch`, a static device
// array that every device the library runs on holds until the process ends.
// Host code never passes `scratch` itself to a kernel (its name is a host-side
// stand-in there): each device's address is looked up once with
// hipGetSymbolAddress, with that device current, and used from then on.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/gps.h"
#include "gps_internal.h"

namespace gps {
gps_status render_in_workspace(gps_scene* scene, const gps_frame_params* fp, int64_t max_points, float* d_color,
                               float* d_depth, hipStream_t st, uint8_t* ws, uint64_t ws_bytes, int cus);
}

static __device__ __align__(256) unsigned char scratch[GPS_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
struct DeviceScratch {
    std::once_flag once;
    hipError_t err = hipSuccess;
    uint8_t* ptr = nullptr;  // this device's address of `scratch`
    int cus = 256;
};
DeviceScratch g_dev[kMaxDevices];

}  // namespace

extern "C" gps_status gps_frame_render_noalloc(gps_scene* scene, const gps_frame_params* frame, int64_t max_points,
                                               float* d_color, float* d_depth, void* stream) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return GPS_ERR_HIP;
    if (dev < 0 || dev >= kMaxDevices) return GPS_ERR_UNSUPPORTED;
    DeviceScratch& d = g_dev[dev];
    std::call_once(d.once, [&d, dev]() {
        void* p = nullptr;
        d.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (d.err == hipSuccess) d.ptr = static_cast<uint8_t*>(p);
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0) d.cus = n;
    });
    if (d.err != hipSuccess || !d.ptr) return GPS_ERR_HIP;
    return gps::render_in_workspace(scene, frame, max_points, d_color, d_depth, static_cast<hipStream_t>(stream),
                                    d.ptr, (uint64_t)GPS_NOALLOC_SCRATCH_BYTES, d.cus);
}
