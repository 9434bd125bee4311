// ssfm_kernels.h — launch interface of the split-step kernels (device code in
// ssfm_kernels.hip).  Shapes: n in {64, 128, 256, 512}; U holds n*n float4 cells;
// tw holds n float2 twiddles W_n^j = exp(-2 pi i j/n); G holds 4n float2 (Gx_p, Gy_p,
// Gx_s, Gy_s).  The launchers check n and nothing else: callers size the buffers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace shg {

// Cell-independent constants of the nonlinear step (src/headers/Solver.h:65-83).
struct NlConst {
    float lossP;   // 0.5 * alpha_p
    float lossS;   // 0.5 * alpha_s
    float tpa;     // 0.5 * beta_s * 0.5 * c * eps0 * n_s
    float h6;      // RK4 step / 6
};

enum RowFlags : int { ROW_FWD = 1, ROW_NPREV = 2, ROW_NCUR = 4, ROW_BWD = 8 };

struct RowArgs {
    float4 cprev;   // (c1, c2) of the step whose second RK4 this pass applies
    float4 ccur;    // (c1, c2) of the step whose first RK4 this pass applies
    NlConst nl;
    int flags;
};

hipError_t launch_row_pass(int n, float4 *U, const float2 *tw, const RowArgs &a, hipStream_t st);
hipError_t launch_col_pass(int n, float4 *U, const float2 *tw, const float2 *G, hipStream_t st);
hipError_t launch_power(const float4 *U, int64_t ncell, double2 *part, int nparts, double2 *out, hipStream_t st);

}  // namespace shg
