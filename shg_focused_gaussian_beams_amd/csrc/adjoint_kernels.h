// adjoint_kernels.h — launch interface of the adjoint z-march (adjoint_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "ssfm_kernels.h"

namespace shg {

// Tape of one checkpoint segment starting at step s0 (arrays of n*n float4 per step):
// a[k] = state before the first RK4 of step s0+k, a2[k] = before its second RK4,
// spec[k] = 2D spectrum the column pass of step s0+k multiplies by H.  `ckpt` receives
// the segment-start state when the tape is off (checkpoint-only forward).
struct TapeArgs {
    float4 *a, *a2, *spec, *ckpt;
    int s0;
    int pad;
};

enum AdjMode : int { ADJ_START = 0, ADJ_MID = 1, ADJ_END = 2 };

hipError_t launch_row_tape(int n, const InstDesc *d, const float2 *tw, int step, int mode, const TapeArgs &tp,
                           hipStream_t st);
hipError_t launch_col_tape(int n, const InstDesc *d, const float2 *tw, int step, const TapeArgs &tp, hipStream_t st);
hipError_t launch_adj_row(int n, const InstDesc *d, const float2 *tw, int step, int mode, const TapeArgs &tp,
                          double *g1, double *g2, hipStream_t st);
hipError_t launch_adj_col(int n, const InstDesc *d, const float2 *tw, int step, const TapeArgs &tp, const float2 *qxy,
                          float a_p, float a_s, double *gH, hipStream_t st);
hipError_t launch_adj_seed(const float4 *U, float4 *L, int64_t n, float scale, hipStream_t st);
// Partial records per step that the adjoint row pass (5 doubles each, one per wavefront)
// and column pass (1 double per workgroup) write.
int adj_row_parts(int n);
int adj_col_parts(int n);

}  // namespace shg
