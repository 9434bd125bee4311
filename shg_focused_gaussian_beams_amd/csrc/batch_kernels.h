// batch_kernels.h — device initialisation of a batch's z = 0 fields.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace shg {

struct PumpArgs {
    float4 *U;              // instance field pair, n*n cells
    float ar, ai;           // Ap0 / M
    float qr, qi, qden;     // w0^2 M and |w0^2 M|^2
    float dx, dy, ux, uy;   // grid spacing and centre (0.5*n)
    float sigma;            // noise standard deviation per component (0: none)
    int n;
    uint64_t seed;
};

hipError_t launch_pump_init(const PumpArgs *d_args, int count, int max_cells, hipStream_t st);

}  // namespace shg
