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
        case 11: r = fp8_e4m3_to_float_hw((uint32_t)(int)x & 0xFFu, 0); break;  // FP8 code x
        case 12: r = det_log_quotient(x); break;
        case 13: r = det_sqrt_normal(x); break;
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

// Plain streaming copy: the reference point for HBM byte counters.
__global__ __launch_bounds__(256) void selftest_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                            uint64_t n16) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Random 64-bit atomicMin over `words` framebuffer-sized words: the
// memory-side atomic rate the splat's framebuffer updates run against.  Word
// index = (hash * words) >> 32, always in [0, words).
__global__ __launch_bounds__(256) void selftest_atomics_kernel(unsigned long long* __restrict__ buf, uint64_t words,
                                                               uint32_t iters, uint32_t seed) {
    uint32_t x = (blockIdx.x * blockDim.x + threadIdx.x) ^ seed;
    for (uint32_t i = 0; i < iters; ++i) {
        x = hash32(x + 0x9E3779B9u);
        const uint64_t idx = ((uint64_t)x * words) >> 32;
        atomicMin(buf + idx, ((unsigned long long)hash32(x) << 32) | x);
    }
}

}  // namespace gps

extern "C" gps_status gps_selftest_atomics(uint64_t* d_buf, uint64_t words, uint32_t blocks, uint32_t iters,
                                           void* stream) {
    if (!d_buf || words == 0 || words > 0xFFFFFFFFull || blocks == 0) return GPS_ERR_INVALID_ARG;
    hipLaunchKernelGGL(gps::selftest_atomics_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       reinterpret_cast<unsigned long long*>(d_buf), words, iters, 0x5EEDu);
    return hipGetLastError() == hipSuccess ? GPS_OK : GPS_ERR_HIP;
}

extern "C" gps_status gps_selftest_copy(void* d_dst, const void* d_src, uint64_t bytes, void* stream) {
    if (!d_dst || !d_src || bytes % 16) return GPS_ERR_INVALID_ARG;
    const uint64_t n16 = bytes / 16;
    if (n16 == 0) return GPS_OK;
    uint64_t blocks = (n16 + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(gps::selftest_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       (const uint4*)d_src, (uint4*)d_dst, n16);
    return hipGetLastError() == hipSuccess ? GPS_OK : GPS_ERR_HIP;
}

extern "C" gps_status gps_selftest_math(int32_t kind, const float* d_in, float* d_out, uint64_t n, void* stream) {
    if (!d_in || !d_out || kind < 0 || kind > 13) return GPS_ERR_INVALID_ARG;
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
