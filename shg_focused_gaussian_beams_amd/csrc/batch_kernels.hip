// batch_kernels.hip — initial fields of a batch on the device: the reference's Gaussian
// pump (setPump, src/headers/Efields.h:74-104) per instance plus seeded complex Gaussian
// noise, SH = 0.  The per-instance constants (Ap0/M, w0^2 M, grid) are evaluated on the
// host in the reference's float arithmetic; the exponent, expf, cosf and sinf run here.
#include <hip/hip_runtime.h>

#include "batch_kernels.h"

namespace shg {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Two independent N(0,1) from a counter: Box-Muller in double, uniform from splitmix64.
__device__ __forceinline__ double2 gauss2(uint64_t seed, uint64_t counter) {
    const uint64_t a = mix64(seed + 0x9E3779B97F4A7C15ull * (2 * counter + 1));
    const uint64_t b = mix64(seed + 0x9E3779B97F4A7C15ull * (2 * counter + 2));
    const double u1 = ((double)(a >> 11) + 1.0) * (1.0 / 9007199254740993.0);
    const double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
    const double r = sqrt(-2.0 * log(u1));
    double s, c;
    sincospi(2.0 * u2, &s, &c);
    return make_double2(r * c, r * s);
}

__global__ __launch_bounds__(256) void k_pump_init(const PumpArgs *__restrict__ args) {
    const PumpArgs a = args[blockIdx.y];
    const int ncell = a.n * a.n;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ncell) return;
    const int ix = i % a.n, iy = i / a.n;
    const float xx = ((float)(unsigned)ix - a.ux) * a.dx, yy = ((float)(unsigned)iy - a.uy) * a.dy;
    const float e = -(xx * xx) - (yy * yy);
    const float er = (e * a.qr) / a.qden, ei = (-e * a.qi) / a.qden;
    const float mag = expf(er);
    float sn, cs;
    sincosf(ei, &sn, &cs);
    const float gr = mag * cs, gi = mag * sn;
    float pr = a.ar * gr - a.ai * gi, pi = a.ar * gi + a.ai * gr;
    if (a.sigma > 0.0f) {
        const double2 g = gauss2(a.seed, (uint64_t)i);
        pr += (float)(a.sigma * g.x);
        pi += (float)(a.sigma * g.y);
    }
    a.U[i] = make_float4(pr, pi, 0.0f, 0.0f);
}

hipError_t launch_pump_init(const PumpArgs *d_args, int count, int max_cells, hipStream_t st) {
    if (count <= 0 || count > 65535 || max_cells <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_pump_init, dim3((max_cells + 255) / 256, count), dim3(256), 0, st, d_args);
    return hipGetLastError();
}

}  // namespace shg
