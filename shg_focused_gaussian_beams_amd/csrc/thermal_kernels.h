// thermal_kernels.h — launch interface of the THERMAL-mode kernels (thermal_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "ssfm_kernels.h"

namespace shg {

// Weights of one Jacobi sweep of heatEquationTopOpen (src/headers/Tfield.h:118-125),
// evaluated on the host with the reference's float/double typing.
struct JacobiConst {
    float dxx, dyy, dzz, L2, kth;                 // inner points
    float ax, bx, fdenx, bixT;                    // x faces
    float ay, by, fdeny, biyT;                    // top face
    float az, bz, fdenz, bizT;                    // z faces
    float av, bv, cv, ah, bh, ch, ad, bd, cd;     // edges
    float AV, BV, CV, GV, tinfBiD;                // vertices
    double V, H, D, sv, sh, sd;                   // double parts of the edge formulas
    float rdxx, rdyy, rdzz, rkth;                 // RN(1/x) of the divisors above (division by a known
    float rfdenx, rfdeny, rfdenz;                 //   divisor in three operations, bit-identical)
    int bc;                                       // SHG_BC_*: which voxels a sweep writes
};

// Per-cell physics of the THERMAL nonlinear step (src/headers/Solver.h:44-63) and setQ
// (src/headers/Tfield.h:277-290).
struct ThermConst {
    float pi, dq, lp, ls, inv_period;             // dk = 2 pi (2 n_p/lp - n_s/ls + 1/Lambda)
    float rlp, rls;                               // RN(1/lp), RN(1/ls)
    float lossP, lossS, h6;                       // 0.5 alpha_p, 0.5 alpha_s, RK4 step / 6
    double tpa_base;                              // beta_s * 0.5 * eps0 * c   (times n_s(T))
    double q_p, q_s, q_i;                         // alpha_p*0.5*c*eps0, alpha_s*0.5*c*eps0, 0.5*c*eps0
    float beta_s;
    int crystal;
};

// Step-dependent scalars of one thermal RK4 application at field step s: the phase is
// DKint_s * zfac[k] / (s + 1) with zfac[k] = dz*inc_k + (s+1)*dz, inc in {0, .25, .5}.
struct ThermStep {
    double zfac[3];
    double zr[3];                                 // zfac / (s + 1)
    double sp1;                                   // s + 1
    int slice;                                    // heat slice of step s
    int step;                                     // s
};

enum ThermRowFlags : int {
    TROW_FWD = 1, TROW_NPREV = 2, TROW_QPREV = 4, TROW_SNAP = 8, TROW_NCUR = 16, TROW_BWD = 32, TROW_FIRST = 64,
    TROW_TAPE = 128, TROW_DKKEEP = 256
};

// With TROW_TAPE the pass also records, for the adjoint, the fields before the second RK4
// of step prev.step into tape_a2[prev.step - tape_s0] and DKint of step cur.step into
// tape_dk[cur.step - tape_s0] (n*n entries per step).
// The DKint array may lag: it holds the prefix of an earlier step w, and dk_pending =
// prev.step - w further additions of dk(T(slice(prev))) (every step in (w, prev.step]
// maps to that slice) give DKint_{prev}.  With TROW_DKKEEP the pass does not store
// DKint_cur (the next pass derives it the same way); stored or derived, the float
// additions are the same, in the same order.
constexpr int kMaxDkPending = 8;                  // a stored prefix lags at most this many steps
struct ThermRowArgs {
    ThermStep prev, cur;
    ThermConst k;
    int flags;
    int dk_pending;
    int tape_s0;
    float4 *tape_a2;
    float *tape_dk;
    // batched THERMAL instances (blockIdx.y = instance): the per-instance strides of the
    // field-plane arrays (U, Uout, DKint: cells) and of the heat-grid arrays (NT, Q:
    // voxels), and a per-instance flag (null: every instance) that lets instances which
    // no longer march skip the pass
    size_t inst_cells, inst_voxels;
    const int *active;
};

hipError_t launch_jacobi(float *Tf, const float *Ti, const float *Q, int nx, int ny, int hz, const JacobiConst &c,
                         double *part, int nparts, hipStream_t st);
// Two sweeps in one pass (z-marching LDS tiles, temporal blocking), bit-exact with two
// launch_jacobi sweeps; with `part` the second sweep's squared updates, one FP64 partial
// per block (plan.blocks of them).  Variants: 3 tile shapes x 3 wave counts; a plan
// (jacobi2_plan, which queries the current device) fixes the tiles and z-chunks for
// `count` instances run by one launch.
constexpr int kJacobi2VariantCount = 9;
struct Jacobi2Plan { int shape, tiles_x, zc, nchunks, blocks; };
int jacobi2_plan(int nx, int ny, int hz, int variant, Jacobi2Plan *plan, int count = 1);   // 0 on success
hipError_t launch_jacobi2(float *Tf, const float *Ti, const float *Q, int nx, int ny, int hz, const JacobiConst &c,
                          double *part, int nparts, hipStream_t st, const Jacobi2Plan &plan);
// The same two sweeps for instances ids[0 .. nids) of a batch (instance i at
// T + i nx ny hz, likewise Tf and Q; plan made for nids instances), without the residual.
hipError_t launch_jacobi2_batch(float *Tf, const float *Ti, const float *Q, int nx, int ny, int hz, int nids,
                                const int *ids, const JacobiConst &c, hipStream_t st, const Jacobi2Plan &plan);
hipError_t launch_sum_partials(const double *part, int nparts, double *out, hipStream_t st);
// Batched THERMAL instances of one heat grid (instance i at T + i nx ny hz, likewise Tf,
// Q): one single sweep of instances ids[0 .. nids); with `part`, the residual partials of
// each (jacobi_blocks of them, in the order of ids, summed as a single instance's), summed
// per entry of ids into out[j] by launch_sum_partials_batch.
hipError_t launch_jacobi_batch(float *Tf, const float *Ti, const float *Q, int nx, int ny, int hz, int nids,
                               const int *ids, const JacobiConst &c, double *part, hipStream_t st);
hipError_t launch_sum_partials_batch(const double *part, int per_instance, int count, double *out, hipStream_t st);
// Per heat voxel tab = {n_p(T), n_s(T)} with the reference's typing (T is frozen during
// one march, so the row pass reads them instead of evaluating the FP64 Sellmeier fits per
// cell and step).
hipError_t launch_index_table(const float *T, float2 *tab, int64_t n, const ThermConst &k, hipStream_t st);
// Row-pass layouts per size (the tuner times all of them): variants 0-3 are <E, W> row
// layouts, 4-5 wide layouts (one row per workgroup, one cell per thread or two), 6 the
// split layout (the two fields transformed by two waves side by side).
constexpr int kThermRowVariantCount = 7;
hipError_t launch_thermal_row_pass(int n, float4 *U, float4 *Uout, const float2 *tw, const float2 *NT, float *Q,
                                   float *DKint, const ThermRowArgs &a, hipStream_t st, int variant = 0,
                                   int count = 1);
// Planes per thread of the single sweep on a heat grid (1, 2, 4 or 8), and the grid size
// the single-sweep launcher uses (number of residual partials it writes).
int jacobi_planes(int nx, int ny, int hz);
int jacobi_blocks(int nx, int ny, int hz);

}  // namespace shg
