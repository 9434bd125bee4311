// thermal_adjoint_kernels.h — launch interface of the adjoint of the THERMAL z-march with
// the temperature held fixed (thermal_adjoint_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "ssfm_kernels.h"
#include "thermal_kernels.h"

namespace shg {

// One adjoint row pass of a checkpoint segment starting at step s0 (modes AdjMode):
//   ADJ_START at s: RK4^T of the second RK4 of step s, R_bwd;
//   ADJ_MID   at s: R_fwd, RK4^T of the first RK4 of step s (at A_s, recomputed from the
//                   tape), RK4^T of the second RK4 of step s-1, R_bwd;
//   ADJ_END   at s0: R_fwd, RK4^T of the first RK4 of step s0 (at the checkpoint).
// Per heat voxel it accumulates Ap = dJ/dn_p and As = dJ/dn_s; per cell it carries G, the
// suffix sum over the later steps of dJ/dDKint, and Dp, the part of the current step's
// dJ/dDKint found by the previous pass.
struct ThermAdjArgs {
    ThermConst k;
    ThermStep cur, prev;          // steps s and s-1
    const float2 *NT;             // {n_p, n_s} per heat voxel (launch_index_table)
    const float4 *a0;             // segment-start fields (the checkpoint)
    const float4 *a2;             // tape: fields before the second RK4 of step s0 + k
    const float *dk;              // tape: DKint of step s0 + k
    double *Ap, *As;              // per heat voxel
    double *G, *Dp;               // per cell
    int s0;
    int pad;
};

// `d` is a one-entry instance table whose U is the adjoint field.  sens = false: lambda
// only (the field gradient); G, Dp, Ap and As are neither read nor written.
hipError_t launch_therm_adj_row(int n, const InstDesc *d, const float2 *tw, int mode, const ThermAdjArgs &a,
                                hipStream_t st, bool sens = true);
// p[0 .. n) = 0 (a kernel, so it may sit in a captured graph).
hipError_t launch_fill_zero(double *p, int64_t n, hipStream_t st);
// Ap[v] <- Ap[v] dn_p/dT(T[v]) + As[v] dn_s/dT(T[v]) (FP64 indices, central difference
// 1e-3 K) over n3 voxels, and the block sums of the result into part[0 .. nparts).
hipError_t launch_therm_dtmap(double *Ap, const double *As, const float *T, int64_t n3, int crystal, double lp,
                              double ls, double *part, int nparts, hipStream_t st);

}  // namespace shg
