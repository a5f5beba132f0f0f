// gps_synth.cpp — deterministic synthetic scenes for the benchmark
// configurations (uniform box, clustered).  splitmix64 streams with FP64
// Box-Muller normals; every Gaussian's attributes come from its own stream
// seeded by hash(seed, index), so any index range can be generated
// independently (chunked generation of 100M-Gaussian scenes) and the result
// does not depend on the chunking.
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/gps.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform double in [0, 1)
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    // standard normal, FP64 Box-Muller (one value per call)
    double normal() {
        double u1 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586476925286766559 * u2);
    }
};

uint64_t mix(uint64_t a, uint64_t b) {
    SplitMix64 m{a ^ (b * 0xD6E8FEB86659FD93ull)};
    m.next();
    return m.next();
}

}  // namespace

namespace {

void generate_range(const gps_synth_params* p, uint64_t first, const uint64_t* indices, uint64_t k0, uint64_t k1,
                    float* means, float* rotations, float* scales, float* opacities, float* shs) {
    const int ncoef = (p->sh_degree + 1) * (p->sh_degree + 1);
    const double h = p->half_extent;
    for (uint64_t k = k0; k < k1; ++k) {
        const uint64_t g = indices ? indices[k] : first + k;
        SplitMix64 r{mix(p->seed, g)};
        double pos[3];
        if (p->kind == 0) {
            for (int c = 0; c < 3; ++c) pos[c] = -h + 2.0 * h * r.uniform();
        } else {
            const uint64_t cl = r.next() % (uint64_t)p->clusters;
            SplitMix64 cr{mix(p->seed ^ 0xC1A57E5ull, cl)};
            for (int c = 0; c < 3; ++c) pos[c] = -h + 2.0 * h * cr.uniform() + p->cluster_sigma * r.normal();
        }
        for (int c = 0; c < 3; ++c) means[3 * k + c] = (float)(pos[c] + p->offset[c]);
        if (!rotations) continue;  // positions only
        double q[4], n2 = 0.0;
        for (int c = 0; c < 4; ++c) {
            q[c] = r.normal();
            n2 += q[c] * q[c];
        }
        const double inv = 1.0 / std::sqrt(n2 > 0.0 ? n2 : 1.0);
        for (int c = 0; c < 4; ++c) rotations[4 * k + c] = (float)(q[c] * inv);
        for (int c = 0; c < 3; ++c) scales[3 * k + c] = (float)std::exp(p->log_scale_mu + p->log_scale_sigma * r.normal());
        opacities[k] = (float)(1.0 / (1.0 + std::exp(-p->opacity_sigma * r.normal())));
        float* sh = shs + (uint64_t)3 * ncoef * k;
        for (int c = 0; c < 3; ++c) sh[c] = (float)(p->dc_sigma * r.normal());
        for (int i = 3; i < 3 * ncoef; ++i) sh[i] = (float)(p->rest_sigma * r.normal());
    }
}

}  // namespace

static gps_status generate(const gps_synth_params* p, uint64_t first, const uint64_t* indices, uint64_t count,
                           float* means, float* rotations, float* scales, float* opacities, float* shs) {
    if (!p || !means) return GPS_ERR_INVALID_ARG;
    const bool means_only = !rotations && !scales && !opacities && !shs;
    if (!means_only && (!rotations || !scales || !opacities || !shs)) return GPS_ERR_INVALID_ARG;
    if (p->sh_degree < 0 || p->sh_degree > 3 || (p->kind == 1 && p->clusters <= 0)) return GPS_ERR_INVALID_ARG;
    if (!indices && first + count > p->num_gaussians) return GPS_ERR_INVALID_ARG;
    if (indices)
        for (uint64_t k = 0; k < count; ++k)
            if (indices[k] >= p->num_gaussians) return GPS_ERR_INVALID_ARG;
    unsigned nt = std::thread::hardware_concurrency();
    nt = nt == 0 ? 1 : (nt > 32 ? 32 : nt);
    if (count < 65536) nt = 1;
    std::vector<std::thread> ts;
    const uint64_t chunk = (count + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const uint64_t k0 = (uint64_t)t * chunk;
        if (k0 >= count) break;
        const uint64_t k1 = k0 + chunk < count ? k0 + chunk : count;
        ts.emplace_back(generate_range, p, first, indices, k0, k1, means, rotations, scales, opacities, shs);
    }
    for (auto& t : ts) t.join();
    return GPS_OK;
}

extern "C" gps_status gps_synth_generate(const gps_synth_params* p, uint64_t first, uint64_t count, float* means,
                                         float* rotations, float* scales, float* opacities, float* shs) {
    return generate(p, first, nullptr, count, means, rotations, scales, opacities, shs);
}

extern "C" gps_status gps_synth_generate_indices(const gps_synth_params* p, const uint64_t* indices, uint64_t count,
                                                 float* means, float* rotations, float* scales, float* opacities,
                                                 float* shs) {
    if (!indices && count) return GPS_ERR_INVALID_ARG;
    return generate(p, 0, indices, count, means, rotations, scales, opacities, shs);
}
