// gps_morton.cpp — load-time Morton ordering of a scene, the reference's
// --sort-morton-order (src/core/gaussians/gaussian_loader.cu:363-426):
// positions normalized to their bounding box, 21 bits per axis
// (x at bit 3k, y at 3k+1, z at 3k+2), stable sort by the 63-bit code.
// Here the sort is a stable LSD radix sort (7 passes of 9 bits) so equal
// codes keep their input order, as thrust::sort_by_key does.
#include <cstring>
#include <vector>

#include "../../include/gps.h"

namespace {

uint64_t spread21(uint32_t v) {
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x001F00000000FFFFull;
    x = (x | (x << 16)) & 0x001F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

}  // namespace

extern "C" gps_status gps_morton_order(const float* means, uint64_t n, uint32_t* perm) {
    if ((!means || !perm) && n) return GPS_ERR_INVALID_ARG;
    if (n >= (1ull << 32)) return GPS_ERR_UNSUPPORTED;
    if (n == 0) return GPS_OK;
    float mn[3] = {3.40282347e38f, 3.40282347e38f, 3.40282347e38f};
    float mx[3] = {-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
    for (uint64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            const float v = means[3 * i + c];
            mn[c] = v < mn[c] ? v : mn[c];
            mx[c] = mx[c] < v ? v : mx[c];
        }
    std::vector<uint64_t> key(n), key2(n);
    std::vector<uint32_t> idx(n), idx2(n);
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t q[3];
        for (int c = 0; c < 3; ++c) {
            const float den = mx[c] - mn[c];
            const float rel = den != 0.0f ? (means[3 * i + c] - mn[c]) / den : 0.0f;
            q[c] = (uint32_t)(int)(rel * (float)((1 << 21) - 1));
        }
        key[i] = spread21(q[0]) | (spread21(q[1]) << 1) | (spread21(q[2]) << 2);
        idx[i] = (uint32_t)i;
    }
    const int bits = 9;
    std::vector<uint64_t> count(1u << bits);
    for (int shift = 0; shift < 63; shift += bits) {
        std::fill(count.begin(), count.end(), 0);
        for (uint64_t i = 0; i < n; ++i) ++count[(key[i] >> shift) & ((1u << bits) - 1)];
        uint64_t sum = 0;
        for (auto& c : count) {
            const uint64_t t = c;
            c = sum;
            sum += t;
        }
        for (uint64_t i = 0; i < n; ++i) {
            const uint64_t d = count[(key[i] >> shift) & ((1u << bits) - 1)]++;
            key2[d] = key[i];
            idx2[d] = idx[i];
        }
        key.swap(key2);
        idx.swap(idx2);
    }
    std::memcpy(perm, idx.data(), n * sizeof(uint32_t));
    return GPS_OK;
}
