// adjoint_kernels.h — launch interface of the adjoint z-march (adjoint_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "ssfm_kernels.h"

namespace shg {

// Tape of one checkpoint segment starting at step s0 (arrays of n*n float4 per step):
// a = the segment-start state (before the first RK4 of s0, i.e. its checkpoint),
// a2[k] = the state before the second RK4 of step s0+k, spec[k] = the 2D spectrum the
// column pass of step s0+k multiplies by H.  A set `ckpt` receives the segment-start
// state from the START row pass.
struct TapeArgs {
    float4 *a, *a2, *spec, *ckpt;
    int s0;
    int pad;
};

enum AdjMode : int { ADJ_START = 0, ADJ_MID = 1, ADJ_END = 2 };

hipError_t launch_row_tape(int n, const InstDesc *d, const float2 *tw, int step, int mode, const TapeArgs &tp,
                           hipStream_t st);
hipError_t launch_col_tape(int n, const InstDesc *d, const float2 *tw, int step, const TapeArgs &tp, hipStream_t st);
// sens = false: lambda only (the field gradient); g1, g2 and gH are not written.
hipError_t launch_adj_row(int n, const InstDesc *d, const float2 *tw, int step, int mode, const TapeArgs &tp,
                          double *g1, double *g2, hipStream_t st, bool sens = true);
hipError_t launch_adj_col(int n, const InstDesc *d, const float2 *tw, int step, const TapeArgs &tp, const float2 *qxy,
                          float a_p, float a_s, double *gH, hipStream_t st, bool sens = true);
hipError_t launch_adj_seed(const float4 *U, float4 *L, int64_t n, float scale, hipStream_t st);
// d eta / d A0 = lambda_FH + direct * A0_FH with direct = -2 cp J / P_in and
// J = cs * red->y / P_in (red: the FP64 output power sums of the forward), written to
// grad (float2 per cell); part[b] = block sums of Re(conj(g) pdf), Re(conj(g) pdn) (FP64,
// with g in FP64) for the focus and pump-index sensitivities.  nparts blocks of 256.
hipError_t launch_field_grad(const float4 *lam, const float4 *U0, const double2 *red, double cs, double cp, double pin,
                             const double2 *pdf, const double2 *pdn, float2 *grad, double2 *part, int nparts,
                             int64_t n, hipStream_t st);
hipError_t launch_copy_cells(float4 *dst, const float4 *src, int64_t n, hipStream_t st);
// Per-step sums of the gradient partials: out[s*6 + 0..4] from g1 + g2, out[s*6 + 5] from gH.
hipError_t launch_adj_reduce(const double *g1, const double *g2, int rparts, const double *gH, int cparts, int steps,
                             double *out, hipStream_t st);
// Partial records per step that the adjoint row pass (5 doubles each, one per row) and
// column pass (1 double per workgroup) write.
int adj_row_parts(int n);
int adj_col_parts(int n);

}  // namespace shg
