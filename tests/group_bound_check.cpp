// Soundness check of the group occlusion test on the CPU.
//
// The preprocess skips a group of 64 Gaussians when group_occluded() says
// every member is rejected by the per-Gaussian Hi-Z test.  This program runs
// the renderer's own code (gps_hiz.h, gps_project.h; the same source the GPU
// kernels compile) on random cameras, max pyramids and groups, and checks
// that claim member by member: whenever the group test fires, each member
// must be rejected by project_gaussian, the screen-bounds test or
// hiz_occluded, in the preprocess's order.  Occluders are placed just in
// front of a group's nearest member, so many groups sit at the margins of the
// test.  Prints one summary line; exits 1 on a violation.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "gps.h"
#include "gps_hiz.h"

namespace gps {
void make_frame_consts(const gps_frame_params* fp, int spp, FrameConsts* out);
}
using namespace gps;

namespace {

struct Pyramid {
    std::vector<float> mip;
    HiZView h{};
};

// The renderer's chain layout: levels = ceil(log2(longest side)) + 1, level l
// has ceil(size / 2^(levels - 1 - l)) texels per side; each texel the maximum
// of its 2x2 children.
void build_pyramid(int W, int H, const std::vector<float>& depth, Pyramid& p) {
    HiZView& h = p.h;
    h.width = W;
    h.height = H;
    int e = 0;
    while ((1 << e) < std::max(W, H)) ++e;
    h.levels = e + 1;
    uint32_t total = 0;
    for (int l = 0; l < h.levels; ++l) {
        const int s = h.levels - 1 - l;
        h.lw[l] = std::max(1, (W + (1 << s) - 1) >> s);
        h.lh[l] = std::max(1, (H + (1 << s) - 1) >> s);
        h.off[l] = total;
        total += (uint32_t)(h.lw[l] * h.lh[l]);
    }
    p.mip.assign(total, 0.0f);
    const int fl = h.levels - 1;
    std::copy(depth.begin(), depth.end(), p.mip.begin() + h.off[fl]);
    for (int l = fl - 1; l >= 0; --l)
        for (int y = 0; y < h.lh[l]; ++y)
            for (int x = 0; x < h.lw[l]; ++x) {
                float v = -INFINITY;
                for (int dy = 0; dy < 2; ++dy)
                    for (int dx = 0; dx < 2; ++dx) {
                        const int cx = 2 * x + dx, cy = 2 * y + dy;
                        if (cx < h.lw[l + 1] && cy < h.lh[l + 1])
                            v = std::max(v, p.mip[h.off[l + 1] + cy * h.lw[l + 1] + cx]);
                    }
                p.mip[h.off[l] + y * h.lw[l] + x] = v;
            }
    h.mip = p.mip.data();
}

struct Member {
    float a[10];  // centre xyz, quaternion wxyz, scale xyz
};

// The preprocess's per-Gaussian steps up to the Hi-Z test (the opacity test
// and the small-eigenvalue cull are left out: both only reject more).
bool survives(const FrameConsts& f, const HiZView& h, const Member& m) {
    Projected pr;
    if (!project_gaussian(f, m.a[0], m.a[1], m.a[2], m.a[3], m.a[4], m.a[5], m.a[6], m.a[7], m.a[8], m.a[9], pr))
        return false;
    if (pr.minx >= f.up_w || pr.miny >= f.up_h || pr.maxx < 0 || pr.maxy < 0) return false;
    return !hiz_occluded(h, pr.pz, div_ss(f, pr.minx), div_ss(f, pr.miny), div_ss(f, pr.maxx), div_ss(f, pr.maxy));
}

}  // namespace

int main(int argc, char** argv) {
    const int trials = argc > 1 ? atoi(argv[1]) : 4000;
    const unsigned seed = argc > 2 ? (unsigned)atoi(argv[2]) : 1u;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    auto uni = [&](double a, double b) { return a + (b - a) * U(rng); };
    auto logu = [&](double a, double b) { return std::exp(uni(std::log(a), std::log(b))); };
    std::normal_distribution<double> N01(0.0, 1.0);

    long groups_culled = 0, members_checked = 0, violations = 0, all_rejected = 0, tight = 0;
    for (int t = 0; t < trials; ++t) {
        // camera: random orientation and position, glm::perspective projection
        double q[4] = {N01(rng), N01(rng), N01(rng), N01(rng)};
        const double qn = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        for (double& c : q) c /= qn;
        const double w = q[0], x = q[1], y = q[2], z = q[3];
        const double R[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)},
                                {2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)},
                                {2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}};
        const double cam[3] = {uni(-500, 500), uni(-500, 500), uni(-500, 500)};
        gps_frame_params fp{};
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) fp.view[4 * j + i] = (float)R[i][j];
            fp.view[12 + i] = (float)(-(R[i][0] * cam[0] + R[i][1] * cam[1] + R[i][2] * cam[2]));
        }
        fp.view[15] = 1.0f;
        const int sses[4] = {1, 2, 3, 4};
        const int ss = sses[t % 4];
        fp.width = (int)uni(24, 720);
        fp.height = (int)uni(16, 480);
        fp.supersampling_factor = ss;
        fp.fovy_degrees = (float)uni(20.0, 110.0);
        fp.near_plane = (float)logu(0.01, 1.0);
        fp.far_plane = (float)uni(500.0, 5000.0);
        const double th = std::tan(fp.fovy_degrees * 0.017453292519943295 * 0.5);
        const double aspect = (double)fp.width / fp.height;
        fp.proj[0] = (float)(1.0 / (aspect * th));
        fp.proj[5] = (float)(1.0 / th);
        fp.proj[10] = -(fp.far_plane + fp.near_plane) / (fp.far_plane - fp.near_plane);
        fp.proj[11] = -1.0f;
        fp.proj[14] = -(2.0f * fp.far_plane * fp.near_plane) / (fp.far_plane - fp.near_plane);
        fp.enable_occlusion_culling = 1;
        FrameConsts f;
        make_frame_consts(&fp, 1, &f);

        // group: 64 members around a point in (or near) the view
        const double dist = logu(0.3, 800.0);
        const double ndx = uni(-1.4, 1.4) * th * aspect, ndy = uni(-1.4, 1.4) * th;
        double c0[3];
        for (int i = 0; i < 3; ++i)  // camera axes are the rows of R; it looks down -z
            c0[i] = cam[i] + dist * (ndx * R[0][i] + ndy * R[1][i] - R[2][i]);
        const double rho = dist * logu(1e-4, 0.5);
        const double smax = dist * logu(1e-5, 0.05);
        std::vector<Member> g(64);
        for (Member& m : g) {
            for (int i = 0; i < 3; ++i) m.a[i] = (float)(c0[i] + rho * uni(-1, 1));
            double r[4] = {N01(rng), N01(rng), N01(rng), N01(rng)};
            const double rn = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]) / uni(0.97, 1.03);
            for (int i = 0; i < 4; ++i) m.a[3 + i] = (float)(r[i] / rn);
            for (int i = 0; i < 3; ++i) m.a[7 + i] = (float)(smax * logu(1e-3, 1.0));
        }
        GroupBox bx{};
        bx.ok = 1;
        for (int i = 0; i < 3; ++i) {
            bx.lo[i] = INFINITY;
            bx.hi[i] = -INFINITY;
        }
        for (const Member& m : g) {
            float lam, sig[3];
            if (!group_member_bounds(m.a, &lam, sig)) bx.ok = 0;
            bx.lam = std::max(bx.lam, lam);
            for (int i = 0; i < 3; ++i) {
                bx.lo[i] = std::min(bx.lo[i], m.a[i]);
                bx.hi[i] = std::max(bx.hi[i], m.a[i]);
                bx.sig[i] = std::max(bx.sig[i], sig[i]);
            }
        }

        // depth image: far background, random occluders, and one occluder just
        // in front of the group's nearest member over (part of) its footprint
        const int W = fp.width, H = fp.height;
        std::vector<float> depth((size_t)W * H, 1.0f);
        auto fill = [&](int x0, int y0, int x1, int y1, float d) {
            for (int yy = std::max(0, y0); yy <= std::min(H - 1, y1); ++yy)
                for (int xx = std::max(0, x0); xx <= std::min(W - 1, x1); ++xx)
                    depth[(size_t)yy * W + xx] = std::min(depth[(size_t)yy * W + xx], d);
        };
        for (int k = 0; k < 4; ++k) {
            const int x0 = (int)uni(0, W), y0 = (int)uni(0, H);
            fill(x0, y0, x0 + (int)uni(1, W), y0 + (int)uni(1, H), (float)uni(-1.0, 1.0));
        }
        float zmin = INFINITY;
        int px0 = W, py0 = H, px1 = -1, py1 = -1;
        for (const Member& m : g) {
            Projected pr;
            if (!project_gaussian(f, m.a[0], m.a[1], m.a[2], m.a[3], m.a[4], m.a[5], m.a[6], m.a[7], m.a[8], m.a[9], pr))
                continue;
            zmin = std::min(zmin, pr.pz);
            px0 = std::min(px0, div_ss(f, pr.minx));
            py0 = std::min(py0, div_ss(f, pr.miny));
            px1 = std::max(px1, div_ss(f, pr.maxx));
            py1 = std::max(py1, div_ss(f, pr.maxy));
        }
        if (px1 >= 0 && zmin <= 1.0f) {
            const double gap = U(rng) < 0.3 ? 0.0 : logu(1e-7, 1e-2);
            const int mx = (int)((px1 - px0 + 1) * uni(-0.2, 1.0)), my = (int)((py1 - py0 + 1) * uni(-0.2, 1.0));
            fill(px0 - mx, py0 - my, px1 + mx, py1 + my, (float)(zmin - gap));
        }
        Pyramid pyr;
        build_pyramid(W, H, depth, pyr);

        bool every = true;
        for (const Member& m : g) every = every && !survives(f, pyr.h, m);
        all_rejected += every ? 1 : 0;
        if (!group_occluded(bx, f, pyr.h)) continue;
        ++groups_culled;
        tight += (zmin - 1e-4f <= 1.0f && px1 >= 0) ? 1 : 0;
        for (const Member& m : g) {
            ++members_checked;
            if (survives(f, pyr.h, m)) {
                ++violations;
                if (violations <= 5)
                    printf("violation: trial %d ss %d %dx%d member (%g %g %g) scale (%g %g %g)\n", t, ss, W, H, m.a[0],
                           m.a[1], m.a[2], m.a[7], m.a[8], m.a[9]);
            }
        }
    }
    printf("trials %d groups_culled %ld groups_all_rejected %ld members_checked %ld violations %ld\n", trials,
           groups_culled, all_rejected, members_checked, violations);
    return violations ? 1 : 0;
}
