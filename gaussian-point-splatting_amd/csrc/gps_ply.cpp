// gps_ply.cpp — 3D Gaussian Splatting PLY files, the scene format the
// reference loads and writes (src/core/gaussians/gaussian_loader.cu:105-427,
// 455-562; src/core/gaussians/ply_comments.h).
//
// Reading follows the reference's field rules: properties are located by
// name in the header (binary little-endian records, unknown properties
// skipped by their type's width), opacity = clamp(sigmoid(x), 0, 1), scale =
// exp(x), rotation = normalize(rot_0..3) kept as (w, x, y, z), SH DC from
// f_dc_0..2 and the 45 f_rest values de-interleaved from channel-major
// (R 0..14, G 15..29, B 30..44) to coefficient-major RGB — all zero unless
// f_rest_44 is declared.  Fields must be float32 (the reference reads every
// field as 4 float bytes; other types are rejected here instead).
// --sort-morton-order reorders the loaded Gaussians (gps_morton_order)
// unless a "bvh_sorted_order" comment says the file is already ordered.  Writing emits the reference's full layout (x y z,
// zero normals, f_dc, f_rest, opacity logit, log scale, rot_0..3).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gps.h"

namespace {

constexpr size_t kMissing = ~size_t(0);

struct Fields {
    size_t x = kMissing, y = kMissing, z = kMissing;
    size_t dc[3] = {kMissing, kMissing, kMissing};
    std::vector<size_t> rest;
    size_t opacity = kMissing;
    size_t scale[3] = {kMissing, kMissing, kMissing};
    size_t rot[4] = {kMissing, kMissing, kMissing, kMissing};
    size_t record = 0;
    uint64_t count = 0;
    bool sorted = false;
    bool complete() const {
        if (x == kMissing || y == kMissing || z == kMissing || opacity == kMissing) return false;
        for (size_t o : dc)
            if (o == kMissing) return false;
        for (size_t o : scale)
            if (o == kMissing) return false;
        for (size_t o : rot)
            if (o == kMissing) return false;
        return true;
    }
    // The reference reads f_rest only when indices up to 44 were declared
    // (a missing index in between reads as 0).
    bool full_rest() const { return rest.size() >= 45; }
};

size_t type_width(const std::string& t) {
    if (t == "double" || t == "float64" || t == "int64" || t == "uint64") return 8;
    if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
    if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
    return 4;
}

// Parses the header; the stream is left at the first record.
bool parse_header(std::istream& in, Fields& f) {
    std::string line;
    size_t off = 0;
    bool first = true;
    while (std::getline(in, line)) {
        const size_t b = line.find_first_not_of(" \t\r\n");
        line = b == std::string::npos ? std::string() : line.substr(b);
        while (!line.empty() && (line.back() == '\r' || line.back() == ' ')) line.pop_back();
        if (first) {
            first = false;
            if (line != "ply") return false;
            continue;
        }
        if (line.rfind("format", 0) == 0) {
            if (line.find("binary_little_endian") == std::string::npos) return false;
        } else if (line.rfind("comment", 0) == 0) {
            if (line.size() > 8 && line.substr(8) == "bvh_sorted_order") f.sorted = true;
        } else if (line.rfind("element vertex", 0) == 0) {
            std::istringstream ss(line);
            std::string a, b2;
            ss >> a >> b2 >> f.count;
        } else if (line.rfind("property", 0) == 0) {
            std::istringstream ss(line);
            std::string kw, type, name;
            ss >> kw >> type >> name;
            if (name.empty()) continue;
            const bool is_float = type == "float" || type == "float32";
            auto take = [&](size_t& slot) {
                if (is_float) slot = off;
            };
            if (name == "x") take(f.x);
            else if (name == "y") take(f.y);
            else if (name == "z") take(f.z);
            else if (name == "opacity") take(f.opacity);
            else if (name.size() == 6 && name.rfind("f_dc_", 0) == 0 && name[5] >= '0' && name[5] <= '2')
                take(f.dc[name[5] - '0']);
            else if (name.size() == 7 && name.rfind("scale_", 0) == 0 && name[6] >= '0' && name[6] <= '2')
                take(f.scale[name[6] - '0']);
            else if (name.size() == 5 && name.rfind("rot_", 0) == 0 && name[4] >= '0' && name[4] <= '3')
                take(f.rot[name[4] - '0']);
            else if (name.rfind("f_rest_", 0) == 0) {
                const int i = std::atoi(name.c_str() + 7);
                if (i >= 0 && i < 4096) {
                    if ((size_t)i >= f.rest.size()) f.rest.resize(i + 1, kMissing);
                    take(f.rest[i]);
                }
            }
            off += type_width(type);
        } else if (line == "end_header") {
            f.record = off;
            return true;
        }
    }
    return false;
}

inline float getf(const uint8_t* r, size_t off) {
    if (off == kMissing) return 0.0f;
    float v;
    std::memcpy(&v, r + off, 4);
    return v;
}

void convert(const Fields& f, const uint8_t* buf, uint64_t lo, uint64_t hi, float* means, float* rots, float* scales,
             float* ops, float* shs) {
    const bool rest = f.full_rest();
    for (uint64_t i = lo; i < hi; ++i) {
        const uint8_t* r = buf + (i - lo) * f.record;
        means[3 * i] = getf(r, f.x);
        means[3 * i + 1] = getf(r, f.y);
        means[3 * i + 2] = getf(r, f.z);
        const float s = 1.0f / (1.0f + std::exp(-getf(r, f.opacity)));
        ops[i] = s < 0.0f ? 0.0f : (s > 1.0f ? 1.0f : s);
        for (int c = 0; c < 3; ++c) scales[3 * i + c] = std::exp(getf(r, f.scale[c]));
        float q[4];
        for (int c = 0; c < 4; ++c) q[c] = getf(r, f.rot[c]);
        // glm::normalize(v) = v * inversesqrt(dot(v, v))
        const float d = (q[0] * q[0] + q[1] * q[1]) + (q[2] * q[2] + q[3] * q[3]);
        const float inv = 1.0f / std::sqrt(d);
        for (int c = 0; c < 4; ++c) rots[4 * i + c] = q[c] * inv;
        float* sh = shs + 48 * i;
        for (int c = 0; c < 3; ++c) sh[c] = getf(r, f.dc[c]);
        for (int j = 1; j < 16; ++j) {
            sh[3 * j] = rest ? getf(r, f.rest[j - 1]) : 0.0f;
            sh[3 * j + 1] = rest ? getf(r, f.rest[j - 1 + 15]) : 0.0f;
            sh[3 * j + 2] = rest ? getf(r, f.rest[j - 1 + 30]) : 0.0f;
        }
    }
}

}  // namespace

extern "C" {

gps_status gps_ply_info_read(const char* path, gps_ply_info* out) {
    if (!path || !out) return GPS_ERR_INVALID_ARG;
    std::ifstream in(path, std::ios::binary);
    if (!in) return GPS_ERR_IO;
    Fields f;
    if (!parse_header(in, f) || !f.complete()) return GPS_ERR_IO;
    std::memset(out, 0, sizeof(*out));
    out->num_gaussians = f.count;
    out->record_bytes = f.record;
    out->has_rest = f.full_rest() ? 1 : 0;
    out->bvh_sorted = f.sorted ? 1 : 0;
    return GPS_OK;
}

gps_status gps_ply_load(const char* path, int32_t sort_morton_order, uint64_t capacity, float* means,
                        float* rotations, float* scales, float* opacities, float* shs, uint64_t* loaded) {
    if (!path || !means || !rotations || !scales || !opacities || !shs || !loaded) return GPS_ERR_INVALID_ARG;
    *loaded = 0;
    std::ifstream in(path, std::ios::binary);
    if (!in) return GPS_ERR_IO;
    Fields f;
    if (!parse_header(in, f) || !f.complete() || f.count == 0) return GPS_ERR_IO;
    if (f.count > capacity) return GPS_ERR_INVALID_ARG;
    if (f.count >= (1ull << 32)) return GPS_ERR_UNSUPPORTED;
    // Read and convert in blocks of 1M records, each converted on all cores.
    const uint64_t block = 1u << 20;
    std::vector<uint8_t> buf(block * f.record);
    uint64_t done = 0;
    unsigned nt = std::thread::hardware_concurrency();
    nt = nt == 0 ? 1 : (nt > 32 ? 32 : nt);
    while (done < f.count) {
        const uint64_t n = std::min<uint64_t>(block, f.count - done);
        in.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)(n * f.record));
        const uint64_t got = (uint64_t)in.gcount() / f.record;  // a short file keeps its whole records
        std::vector<std::thread> ts;
        const uint64_t per = (got + nt - 1) / nt;
        for (unsigned t = 0; t < nt && per; ++t) {
            const uint64_t lo = done + t * per, hi = std::min<uint64_t>(done + got, lo + per);
            if (lo >= hi) break;
            ts.emplace_back([&, lo, hi] {
                convert(f, buf.data() + (lo - done) * f.record, lo, hi, means, rotations, scales, opacities, shs);
            });
        }
        for (auto& t : ts) t.join();
        done += got;
        if (got < n) break;
    }
    *loaded = done;
    if (sort_morton_order && !f.sorted && done > 1) {
        std::vector<uint32_t> perm(done);
        gps_status st = gps_morton_order(means, done, perm.data());
        if (st != GPS_OK) return st;
        auto permute = [&](float* a, int w) {
            std::vector<float> tmp((size_t)w * done);
            for (uint64_t i = 0; i < done; ++i) std::memcpy(&tmp[(size_t)w * i], a + (size_t)w * perm[i], 4 * w);
            std::memcpy(a, tmp.data(), tmp.size() * 4);
        };
        permute(means, 3);
        permute(rotations, 4);
        permute(scales, 3);
        permute(opacities, 1);
        permute(shs, 48);
    }
    return done == f.count ? GPS_OK : GPS_ERR_IO;
}

gps_status gps_ply_write(const char* path, const gps_scene_desc* d, int32_t mark_sorted) {
    if (!path || !d) return GPS_ERR_INVALID_ARG;
    if (d->sh_degree < 0 || d->sh_degree > 3) return GPS_ERR_INVALID_ARG;
    if (d->num_gaussians && (!d->means || !d->rotations || !d->scales || !d->opacities || !d->shs))
        return GPS_ERR_INVALID_ARG;
    std::ofstream out(path, std::ios::binary);
    if (!out) return GPS_ERR_IO;
    std::ostringstream h;
    h << "ply\nformat binary_little_endian 1.0\n";
    if (mark_sorted) h << "comment bvh_sorted_order\n";
    h << "element vertex " << d->num_gaussians << "\n"
      << "property float x\nproperty float y\nproperty float z\n"
      << "property float nx\nproperty float ny\nproperty float nz\n"
      << "property float f_dc_0\nproperty float f_dc_1\nproperty float f_dc_2\n";
    for (int i = 0; i < 45; ++i) h << "property float f_rest_" << i << "\n";
    h << "property float opacity\nproperty float scale_0\nproperty float scale_1\nproperty float scale_2\n"
      << "property float rot_0\nproperty float rot_1\nproperty float rot_2\nproperty float rot_3\nend_header\n";
    const std::string hs = h.str();
    out.write(hs.data(), (std::streamsize)hs.size());
    const int ncoef = (d->sh_degree + 1) * (d->sh_degree + 1);
    std::vector<float> rec(62);
    for (uint64_t g = 0; g < d->num_gaussians; ++g) {
        std::fill(rec.begin(), rec.end(), 0.0f);
        for (int c = 0; c < 3; ++c) rec[c] = d->means[3 * g + c];
        const float* sh = d->shs + (uint64_t)3 * ncoef * g;
        for (int c = 0; c < 3; ++c) rec[6 + c] = sh[c];
        for (int j = 1; j < ncoef; ++j)
            for (int c = 0; c < 3; ++c) rec[9 + (j - 1) + 15 * c] = sh[3 * j + c];
        const float o = d->opacities[g];
        rec[54] = std::log(o / (1.0f - o));
        for (int c = 0; c < 3; ++c) rec[55 + c] = std::log(d->scales[3 * g + c]);
        for (int c = 0; c < 4; ++c) rec[58 + c] = d->rotations[4 * g + c];
        out.write(reinterpret_cast<const char*>(rec.data()), 62 * 4);
    }
    return out ? GPS_OK : GPS_ERR_IO;
}

}  // extern "C"
