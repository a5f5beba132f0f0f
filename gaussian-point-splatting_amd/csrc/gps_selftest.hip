// gps_selftest.hip — device-side evaluation of the shared scalar math and RNG
// for the host<->device bit-exactness checks (no counterpart in the
// reference; diagnostics of this implementation).
#include "gps_fb.h"
#include "gps_internal.h"
#include "gps_rng.h"

namespace gps {

__global__ void selftest_math_kernel(int kind, const float* __restrict__ in, float* __restrict__ out, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = in[i];
    float r = 0.0f;
    switch (kind) {
        case 0: r = det_expf(x); break;
        case 1: r = det_logf(x); break;
        case 2: r = det_sinf(x); break;
        case 3: r = det_cosf(x); break;
        case 4: r = det_sqrt(x); break;
        case 5: r = 1.0f / x; break;
        case 6: r = dilog(x); break;
        case 7: r = inv_dilog(x); break;
        case 8: r = (float)det_f2i_rn(x); break;
        case 9: r = box_muller(x, in[(i + 1) % n]).x; break;
        case 10: r = corrected_box_muller(x, in[(i + 1) % n], 0.7f).y; break;
    }
    out[i] = r;
}

__global__ void selftest_rng_kernel(uint32_t n, uint32_t frame, int draws, uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u32x2 s = make_seed(i, frame);
    uint32_t* o = out + (uint64_t)i * (2 + 2 * draws);
    o[0] = s.x;
    o[1] = s.y;
    for (int d = 0; d < draws; ++d) {
        pcg2d(s);
        o[2 + 2 * d] = s.x;
        o[3 + 2 * d] = s.y;
    }
}

__global__ void selftest_poisson_kernel(uint32_t n, const float* __restrict__ lambda, uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u32x2 s = make_seed(i, 12345u);
    out[i] = poisson(s, lambda[i]);
}

}  // namespace gps

extern "C" gps_status gps_selftest_math(int32_t kind, const float* d_in, float* d_out, uint64_t n, void* stream) {
    if (!d_in || !d_out || kind < 0 || kind > 10) return GPS_ERR_INVALID_ARG;
    if (n == 0) return GPS_OK;
    hipLaunchKernelGGL(gps::selftest_math_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, kind, d_in, d_out, n);
    return hipGetLastError() == hipSuccess ? GPS_OK : GPS_ERR_HIP;
}

extern "C" gps_status gps_selftest_rng(uint32_t n, uint32_t frame, int32_t draws, uint32_t* d_out, void* stream) {
    if (!d_out || draws < 0) return GPS_ERR_INVALID_ARG;
    if (n == 0) return GPS_OK;
    hipLaunchKernelGGL(gps::selftest_rng_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, frame,
                       draws, d_out);
    return hipGetLastError() == hipSuccess ? GPS_OK : GPS_ERR_HIP;
}

extern "C" gps_status gps_selftest_poisson(uint32_t n, const float* d_lambda, uint32_t* d_out, void* stream) {
    if (!d_lambda || !d_out) return GPS_ERR_INVALID_ARG;
    if (n == 0) return GPS_OK;
    hipLaunchKernelGGL(gps::selftest_poisson_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n,
                       d_lambda, d_out);
    return hipGetLastError() == hipSuccess ? GPS_OK : GPS_ERR_HIP;
}
