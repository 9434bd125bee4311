// ssfm_kernels.h — launch interface of the plain split-step kernels (device code in
// ssfm_kernels.hip).  Shapes: n in {64, 128, 256, 512}; an instance's U holds n*n float4
// cells; tw holds n float2 twiddles W_n^j = exp(-2 pi i j/n); G holds 4n float2 (Gx_p,
// Gy_p, Gx_s, Gy_s); coef holds one float4 (c1, c2) per z-step.  The launchers check n and
// the grid shape; callers size the buffers.
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

// One solver instance as the plain passes see it.  A single context uses a table of one;
// a batch uses one table per transverse size, sorted by step count (longest first) so the
// instances still marching at step s are always a prefix.
struct InstDesc {
    float4 *U;              // field pair, n*n cells
    const float2 *G;        // separable propagator, 4n
    const float4 *coef;     // (c1, c2) per step
    NlConst nl;
    int steps;              // z-steps of this instance's march
    int pad[3];
};

// Row-pass modes: which parts of the fused chain run (see ssfm_kernels.hip).
enum RowMode : int {
    ROW_START = 0,   // RK4_s, R_bwd                     (first pass of a march)
    ROW_MID = 1,     // R_fwd, RK4_s, RK4_{s+1}, R_bwd   (between steps s and s+1)
    ROW_END = 2,     // R_fwd, RK4_s                     (last pass)
    ROW_AUTO = 3     // ROW_MID if s + 1 < steps of the instance, else ROW_END
};

// `variant` selects a thread layout (pass_cfg.h; 0 = default, rows 0..4, columns 0..9;
// row 4 and columns 5..7 are the split layouts through LDS, columns 8..9 the field-split
// register layouts).
constexpr int kRowVariantCount = 5, kColVariantCount = 10;
// A set `snap` (one instance only) receives the real-space state after step `step` (the
// fields after its second RK4) from a MID or END pass: the checkpoints of the adjoint.
hipError_t launch_row_pass(int n, const InstDesc *d, int ninst, const float2 *tw, int step, int mode, hipStream_t st,
                           int variant = 0, float4 *snap = nullptr);
hipError_t launch_col_pass(int n, const InstDesc *d, int ninst, const float2 *tw, hipStream_t st, int variant = 0);
hipError_t launch_power(const float4 *U, int64_t ncell, double2 *part, int nparts, double2 *out, hipStream_t st);
// out = sum of part[0 .. nparts) (componentwise, FP64; one block).
hipError_t launch_fold2(const double2 *part, int nparts, double2 *out, hipStream_t st);
// Per-instance FP64 sums of |FH|^2 and |SH|^2: out[i] for table entry i (one block each).
hipError_t launch_power_batch(const InstDesc *d, const int *ncell, int ninst, double2 *out, hipStream_t st);

}  // namespace shg
