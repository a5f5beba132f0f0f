// Compares the banded alpha test (gps_rng.h alpha_test_banded, the source the
// GPU splat compiles) with the exact one (alpha_test) on the CPU.
//
// Usage: alpha_band_check <trials> <seed> [band]
// Opacities log-uniform in [1/512, 1] plus 1/255 and 1; exponents uniform in
// [-30, 0], within +-0.02 and +-1e-5 of the threshold, and runs of
// consecutive floats across the band edges and the threshold.  Prints the
// number of disagreements (must be 0 at the default band) and how often the
// exponential was still needed.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "gps_rng.h"

using namespace gps;

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const long trials = std::atol(argv[1]);
    std::mt19937_64 rng((uint64_t)std::atoll(argv[2]));
    const float band = argc > 3 ? (float)std::atof(argv[3]) : GPS_ALPHA_BAND;
    std::uniform_real_distribution<double> u01(0.0, 1.0);
    long bad = 0, checked = 0, in_band = 0;
    auto check = [&](float op, float t, float sigma) {
        ++checked;
        if (!(sigma < t - band) && !(sigma > t + band && !(sigma > 0.0f))) ++in_band;
        if (alpha_test_banded(op, t, sigma, band) != alpha_test(op, sigma)) {
            if (bad < 5) std::printf("mismatch op %a t %a sigma %a\n", op, t, sigma);
            ++bad;
        }
    };
    for (long i = 0; i < trials; ++i) {
        float op = (float)std::exp(std::log(1.0 / 512.0) * u01(rng));
        if (i % 97 == 0) op = 1.0f / 255.0f;
        if (i % 89 == 0) op = 1.0f;
        const float t = alpha_threshold(op);
        check(op, t, (float)(-30.0 * u01(rng)));
        check(op, t, t + (float)(0.04 * u01(rng) - 0.02));
        check(op, t, t + (float)(2e-5 * u01(rng) - 1e-5));
        if (i % 64 == 0) {
            for (float edge : {t - band, t, t + band}) {
                float x = edge;
                for (int k = 0; k < 64; ++k) x = std::nextafter(x, -INFINITY);
                for (int k = 0; k < 128; ++k, x = std::nextafter(x, INFINITY)) check(op, t, x);
            }
        }
    }
    std::printf("checked %ld mismatches %ld in_band %ld\n", checked, bad, in_band);
    return 0;
}
