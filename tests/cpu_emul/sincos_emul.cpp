// Host build of sincos_phase (shg_focused_gaussian_beams_amd/csrc/sincos_phase.h): prints
// the largest error, in units of the last place of the result's magnitude class
// (2^-24 for |v| in [0.5, 1), absolute 2^-25 below) and in absolute terms, against
// sin/cos of the same float argument in long double, over seeded arguments up to |x| < xmax.
//   usage: sincos_emul <xmax> <count> <seed>
#define SHG_HOST_EMULATION 1
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../../shg_focused_gaussian_beams_amd/csrc/sincos_phase.h"

int main(int argc, char **argv) {
    const double xmax = argc > 1 ? std::atof(argv[1]) : 100.0;
    const long count = argc > 2 ? std::atol(argv[2]) : 1000000;
    std::mt19937_64 g(argc > 3 ? std::atoll(argv[3]) : 1);
    std::uniform_real_distribution<double> u(-xmax, xmax);
    double max_abs = 0.0, max_ulp = 0.0;
    auto check = [&](float x) {
        float s, c;
        shg::sincos_phase(x, s, c);
        const long double xs = (long double)x;
        const long double rs = sinl(xs), rc = cosl(xs);
        for (int k = 0; k < 2; k++) {
            const long double ref = k ? rc : rs, got = k ? c : s;
            const double err = (double)fabsl(got - ref);
            const double mag = (double)fabsl(ref);
            const double ulp = std::ldexp(1.0, std::max(std::ilogb(std::max(mag, 1e-30)) - 23, -25));
            if (err > max_abs) max_abs = err;
            if (err / ulp > max_ulp) max_ulp = err / ulp;
        }
    };
    for (long i = 0; i < count; i++) check((float)u(g));
    // arguments near multiples of pi/4, where the reduction and the quadrant switch
    for (long k = -4000; k <= 4000; k++)
        for (int d = -3; d <= 3; d++) check(std::nextafter((float)(k * 0.78539816339744830962), d > 0 ? 1e9f : -1e9f) + d * 1e-7f * (float)(k % 7));
    std::printf("%.6e %.6e\n", max_abs, max_ulp);
    return 0;
}
