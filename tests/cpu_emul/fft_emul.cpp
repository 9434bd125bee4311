// Host build of the device FFT in shg_focused_gaussian_beams_amd/csrc/fft_lds.h.
// Each GPU thread of one transform group becomes a std::thread; the workgroup barrier
// becomes a std::barrier.  LDS is a bounds-checked array.  Reads complex input from
// stdin-free arguments: generates a seeded field, runs rows or columns mode, prints
// the transformed values for comparison with numpy.
//   usage: fft_emul <N> <sign +1|-1> <mode row|col> <seed> [E elements per thread, 0 = default plan]
//          [one: transform the first field only (fft1fe); the second must come back unchanged]
#include <barrier>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <thread>
#include <vector>

struct float2 { float x, y; };
inline float2 make_float2(float x, float y) { return {x, y}; }
#define __device__
#define __host__
#define __forceinline__ inline
#define SHG_HOST_EMULATION 1

// Bounds-checked LDS pointer.
struct Lds {
    float2 *base; long lo, hi;   // valid [lo, hi) relative to base
    float2 &operator[](long i) const {
        if (i < lo || i >= hi) { std::fprintf(stderr, "LDS index %ld out of [%ld,%ld)\n", i, lo, hi); std::abort(); }
        return base[i];
    }
};
#include "../../shg_focused_gaussian_beams_amd/csrc/fft_lds.h"

int main(int argc, char **argv) {
    if (argc < 5) return 2;
    const int N = std::atoi(argv[1]), sign = std::atoi(argv[2]);
    const bool col = std::string(argv[3]) == "col";
    const unsigned seed = std::atoi(argv[4]);
    srand(seed);
    const int CW = col ? (N == 512 ? 4 : 8) : 1;
    // data: CW transforms of length N for two fields
    std::vector<float2> A((size_t)N * CW), B((size_t)N * CW);
    for (auto &v : A) v = {(float)rand() / RAND_MAX - 0.5f, (float)rand() / RAND_MAX - 0.5f};
    for (auto &v : B) v = {(float)rand() / RAND_MAX - 0.5f, (float)rand() / RAND_MAX - 0.5f};
    std::vector<float2> tw(N);
    for (int j = 0; j < N; j++) tw[j] = {(float)std::cos(-2 * M_PI * j / N), (float)std::sin(-2 * M_PI * j / N)};
    const int L = N + N / 16;
    std::vector<float2> lds_a((size_t)L * CW), lds_b((size_t)L * CW);
    std::vector<float2> outA(A.size()), outB(B.size());
    int E = argc > 5 ? std::atoi(argv[5]) : 0;
    const bool one = argc > 6 && std::string(argv[6]) == "one";
    if (E == 0) E = (N == 64 || N == 512) ? 8 : 16;
    if (E != 4 && E != 8 && E != 16) return 2;
    const int T = N / E;
    std::barrier bar(T * CW);
    auto worker = [&](int tid) {
        const int c = tid % CW, t = tid / CW;
        Lds la{lds_a.data() + c, -c, (long)L * CW - c}, lb{lds_b.data() + c, -c, (long)L * CW - c};
        auto sync = [&] { bar.arrive_and_wait(); };
        auto run = [&](auto NN, auto EE) {
            constexpr int n = decltype(NN)::value;
            constexpr int e = decltype(EE)::value;
            float2 a[e], b[e];
            for (int j = 0; j < e; j++) { a[j] = A[(size_t)(t + j * T) * CW + c]; b[j] = B[(size_t)(t + j * T) * CW + c]; }
            if (one) {
                if (sign > 0) shg::fft1fe<n, e, +1>(a, la, t, CW, tw.data(), sync);
                else shg::fft1fe<n, e, -1>(a, la, t, CW, tw.data(), sync);
            } else {
                if (sign > 0) shg::fft2fe<n, e, +1>(a, b, la, lb, t, CW, tw.data(), sync);
                else shg::fft2fe<n, e, -1>(a, b, la, lb, t, CW, tw.data(), sync);
            }
            for (int j = 0; j < e; j++) { outA[(size_t)(t + j * T) * CW + c] = a[j]; outB[(size_t)(t + j * T) * CW + c] = b[j]; }
        };
        auto runE = [&](auto NN) {
            if (E == 4) run(NN, std::integral_constant<int, 4>{});
            else if (E == 8) run(NN, std::integral_constant<int, 8>{});
            else run(NN, std::integral_constant<int, 16>{});
        };
        switch (N) {
            case 64: runE(std::integral_constant<int, 64>{}); break;
            case 128: runE(std::integral_constant<int, 128>{}); break;
            case 256: runE(std::integral_constant<int, 256>{}); break;
            case 512: runE(std::integral_constant<int, 512>{}); break;
        }
    };
    std::vector<std::thread> th;
    for (int i = 0; i < T * CW; i++) th.emplace_back(worker, i);
    for (auto &x : th) x.join();
    // print: input then output, index-major (element i, column c)
    for (size_t i = 0; i < A.size(); i++)
        std::printf("%.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g\n", A[i].x, A[i].y, B[i].x, B[i].y, outA[i].x, outA[i].y,
                    outB[i].x, outB[i].y);
    return 0;
}
