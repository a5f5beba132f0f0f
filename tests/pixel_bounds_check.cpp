// Compares det_pixel_in (gps_detmath.h, the source the GPU splat compiles)
// with the bounds test on det_f2i_rz it replaces, on the CPU: every
// integral float in [-70000, 70000], NaN, infinities, values around and
// beyond the int range, against bounds 0 <= lo < hi <= 65536 (tile edges,
// zero-width and full-width).  Prints the number of disagreements.
#include <cmath>
#include <cstdio>
#include <vector>

#include "gps_detmath.h"

using namespace gps;

int main() {
    std::vector<float> vals;
    for (int v = -70000; v <= 70000; ++v) vals.push_back((float)v);
    const float special[] = {NAN, -NAN, INFINITY, -INFINITY, 2147483520.0f, 2147483648.0f, -2147483648.0f,
                             -2147483904.0f, 4294967296.0f, 1e30f, -1e30f, 65535.0f, 65536.0f, 65537.0f, -0.0f};
    for (float v : special) vals.push_back(v);
    const int bounds[][2] = {{0, 16}, {0, 65536}, {16, 32}, {4080, 65536}, {65520, 65536}, {0, 0}, {32, 32},
                             {1024, 2048}, {0, 1}, {65535, 65536}};
    long bad = 0, checked = 0;
    for (const auto& b : bounds) {
        for (float v : vals) {
            const int32_t x = det_f2i_rz(v);
            const bool ref = !(x < b[0] || x >= b[1]);
            int32_t p = -1;
            const bool got = det_pixel_in(v, b[0], b[1], &p);
            ++checked;
            if (got != ref || (got && p != x)) {
                if (bad < 5) std::printf("mismatch v %a bounds [%d, %d) ref %d got %d p %d x %d\n", v, b[0], b[1], ref, got, p, x);
                ++bad;
            }
        }
    }
    std::printf("checked %ld mismatches %ld\n", checked, bad);
    return 0;
}
