// thermal_kernels.hip — THERMAL mode on gfx950: the steady-state heat-equation Jacobi
// sweep, and the row pass whose nonlinear step reads the temperature, accumulates the
// phase mismatch and writes the heat source.
//
// Arithmetic that the reference evaluates in float (the Jacobi weights and sums, the
// refractive index and the phase mismatch) is evaluated here with FP contraction off
// and the same float/double typing, so a sweep and a dk(T) match the reference's
// float results bit for bit; the mismatch dk is a difference of O(4) terms whose
// rounding is amplified about a thousandfold.
#include <hip/hip_runtime.h>

#include "fft_lds.h"
#include "nl_step.h"
#include "thermal_cell.h"
#include "pass_cfg.h"
#include "thermal_kernels.h"
#include "../../include/shg.h"

namespace shg {

// ---------------------------------------------------------------------------
// One Jacobi sweep Tf <- J(Ti, Q) of heatEquationTopOpen (src/headers/Tfield.h:113-221):
// thread per voxel; the y = 0 plane (Dirichlet ovens) is never written.  With
// c.bc == SHG_BC_OVEN_SURROUNDED, kernelHeatEquationOvenSurrounded (:224-258): the four
// lateral faces x = 0, nx-1, y = 0, ny-1 are Dirichlet ovens and are never written; the
// voxels it updates (inner points and the z = 0, z = L faces) use the same formulas as
// the top-open sweep's inner points and z faces.  With RESID the FP64 sum of
// (Tf - Ti)^2 over the block goes to part[block].
template <bool RESID>
__global__ __launch_bounds__(256) void k_jacobi(float *__restrict__ Tf, const float *__restrict__ Ti,
                                                const float *__restrict__ Q, int nx, int ny, int hz, JacobiConst c,
                                                double *__restrict__ part) {
#pragma clang fp contract(off)
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    double r = 0.0;
    const bool oven = c.bc == SHG_BC_OVEN_SURROUNDED;
    if (x < nx && y >= 1 && y < ny && (!oven || (x >= 1 && x < nx - 1 && y < ny - 1))) {
        const size_t sx = 1, sy = (size_t)nx, sz = (size_t)nx * ny;
        const size_t i = (size_t)z * sz + (size_t)y * sy + x;
        const bool xi = x > 0 && x < nx - 1, zi = z > 0 && z < hz - 1, yi = y < ny - 1;
        const size_t ixn = x == 0 ? i + sx : i - sx;       // the one x-neighbour of an x face
        const size_t izn = z == 0 ? i + sz : i - sz;
        float out;
        if (yi) {
            if (xi && zi) {
                out = c.L2 * ((Ti[i + sx] + Ti[i - sx]) / c.dxx + (Ti[i + sy] + Ti[i - sy]) / c.dyy +
                              (Ti[i + sz] + Ti[i - sz]) / c.dzz + Q[i] / c.kth);
            } else if (!xi && zi) {
                out = (Ti[ixn] + c.ax * (Ti[i + sy] + Ti[i - sy]) + c.bx * (Ti[i + sz] + Ti[i - sz]) + c.bixT) / c.fdenx;
            } else if (xi && !zi) {
                out = (Ti[izn] + c.az * (Ti[i + sx] + Ti[i - sx]) + c.bz * (Ti[i + sy] + Ti[i - sy]) + c.bizT) / c.fdenz;
            } else {
                const float f = c.av * Ti[ixn] + c.bv * Ti[izn] + c.cv * (Ti[i + sy] + Ti[i - sy]);
                out = (float)(((double)f + c.sv) * c.V);
            }
        } else {
            if (xi && zi) {
                out = (Ti[i - sy] + c.ay * (Ti[i + sx] + Ti[i - sx]) + c.by * (Ti[i + sz] + Ti[i - sz]) + c.biyT) / c.fdeny;
            } else if (xi && !zi) {
                const float f = c.ah * Ti[i - sy] + c.bh * Ti[izn] + c.ch * (Ti[i - sx] + Ti[i + sx]);
                out = (float)(((double)f + c.sh) * c.H);
            } else if (!xi && zi) {
                const float f = c.ad * Ti[i - sy] + c.bd * Ti[ixn] + c.cd * (Ti[i + sz] + Ti[i - sz]);
                out = (float)(((double)f + c.sd) * c.D);
            } else {
                out = (c.AV * Ti[ixn] + c.BV * Ti[i - sy] + c.CV * Ti[izn] + c.tinfBiD) * c.GV;
            }
        }
        Tf[i] = out;
        if (RESID) {
            const double d = (double)(out - Ti[i]);
            r = d * d;
        }
    }
    if (RESID) {
        for (int off = 32; off > 0; off >>= 1) r += __shfl_down(r, off, 64);
        __shared__ double w[4];
        if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = r;
        __syncthreads();
        if (threadIdx.x == 0)
            part[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = w[0] + w[1] + w[2] + w[3];
    }
}

__global__ __launch_bounds__(256) void k_sum_partials(const double *__restrict__ part, int n, double *out) {
    __shared__ double s[256];
    double r = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) r += part[i];
    s[threadIdx.x] = r;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

// Refractive indices of every heat voxel, {n_p(T), n_s(T)}, and the coupling constants,
// phase mismatch and TPA coefficient derived from them: the temperature is frozen during
// a march, so the FP64 Sellmeier fits and the float divisions run once per voxel instead
// of twice per cell per row pass (the values are the same floats).
__global__ __launch_bounds__(256) void k_index_table(const float *__restrict__ T, float2 *__restrict__ tab,
                                                     float4 *__restrict__ ct, int64_t n, ThermConst k) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float t = T[i];
    const float2 nn = make_float2(index_dev(k.crystal, k.lp, t), index_dev(k.crystal, k.ls, t));
    const CellPhys c = cell_phys(k, nn);
    tab[i] = nn;
    ct[i] = make_float4(c.kp, c.ks, c.dk, c.tpa);
}

// Row pass of the THERMAL z-march (flags in ThermRowFlags):
//   FWD   R_fwd (end of the previous step's diffraction)
//   NPREV second RK4 of step s-1 with T(slice(s-1)) and DKint_{s-1}
//   QPREV heat source of slice(s-1) from the fields after step s-1
//   SNAP  copy of the fields after step s-1 to Uout (the reference's Pump_out state)
//   NCUR  DKint_s = DKint_{s-1} + dk(T(slice(s))), then the first RK4 of step s
//   BWD   R_bwd (start of step s's diffraction)
// DKint holds the running float prefix sum of dk over the executed steps: the reference
// recomputes the same float sum from z' = 0 twice per step (src/headers/PhaseMatching.h:
// 25-38, Solver.h:198-199); the sequential float additions are identical.
//
// One cell of the pass between the transforms (idx: cell index in the plane), from its
// already loaded phase accumulator and index-table entries.
__device__ __forceinline__ void therm_cell_pre(float2 &P, float2 &S, size_t idx, size_t n2, size_t sprev, float dkprev,
                                               float2 nnp, float4 cfp, float2 nnc, float4 cfc,
                                               float4 *__restrict__ Uout, float *__restrict__ Q,
                                               float *__restrict__ DKint, const ThermRowArgs &a) {
    const ThermConst &k = a.k;
    if (a.flags & (TROW_NPREV | TROW_QPREV)) {
        const CellPhys cp = cell_from_table(nnp, cfp);
        if ((a.flags & (TROW_TAPE | TROW_NPREV)) == (TROW_TAPE | TROW_NPREV))
            a.tape_a2[(size_t)(a.prev.step - a.tape_s0) * n2 + idx] = make_float4(P.x, P.y, S.x, S.y);
        if (a.flags & TROW_NPREV) rk4_thermal(P, S, cp, dkprev, a.prev, k);
        if (a.flags & TROW_QPREV) Q[sprev + idx] = heat_source(k, cp, P, S);
    }
    if (a.flags & TROW_SNAP) Uout[idx] = make_float4(P.x, P.y, S.x, S.y);
    if (a.flags & TROW_NCUR) {
        const CellPhys cc = cell_from_table(nnc, cfc);
        const float dk = dkprev + cc.dk;
        DKint[idx] = dk;
        if (a.flags & TROW_TAPE) a.tape_dk[(size_t)(a.cur.step - a.tape_s0) * n2 + idx] = dk;
        rk4_thermal(P, S, cc, dk, a.cur, k);
    }
}

// The same, loading the cell's inputs itself.
__device__ __forceinline__ void therm_cell(float2 &P, float2 &S, size_t idx, size_t n2, size_t sprev, size_t scur,
                                           float4 *__restrict__ Uout, const float2 *__restrict__ NT,
                                           float *__restrict__ Q, float *__restrict__ DKint, const ThermRowArgs &a) {
    const float dkprev = (a.flags & TROW_FIRST) ? 0.0f : DKint[idx];
    const bool pv = a.flags & (TROW_NPREV | TROW_QPREV), cu = a.flags & TROW_NCUR;
    const float2 nnp = pv ? NT[sprev + idx] : make_float2(0.f, 0.f);
    const float2 nnc = cu ? NT[scur + idx] : make_float2(0.f, 0.f);
    const float4 cfp = pv ? a.ct[sprev + idx] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 cfc = cu ? a.ct[scur + idx] : make_float4(0.f, 0.f, 0.f, 0.f);
    therm_cell_pre(P, S, idx, n2, sprev, dkprev, nnp, cfp, nnc, cfc, Uout, Q, DKint, a);
}

// Layout <N, E, W> as in pass_cfg.h.  Between the two transforms the thread's E cells
// are parked in the row's LDS slots and stepped by a rolled loop: the per-cell work (two
// RK4 with three sincos each, the heat source) is too large to unroll E times, and a
// rolled loop over register arrays would move them to scratch memory.
template <int N, int E, int W>
__global__ __launch_bounds__(64 * W) void k_therm_row_pass(float4 *__restrict__ U, float4 *__restrict__ Uout,
                                                           const float2 *__restrict__ twg,
                                                           const float2 *__restrict__ NT, float *__restrict__ Q,
                                                           float *__restrict__ DKint, ThermRowArgs a) {
    using C = RowCfgT<N, E, W>;
    constexpr int TT = C::T;
    __shared__ float2 lds[2][C::ROWS][C::L];
    __shared__ float2 tw[N];
    const int lane = threadIdx.x;
    const int t = lane % TT, rr = lane / TT;
    const int row = blockIdx.x * C::ROWS + rr;
    for (int j = lane; j < N; j += C::THREADS) tw[j] = twg[j];
    __syncthreads();
    const size_t n2 = (size_t)N * N;
    const size_t base = (size_t)row * N;
    float2 fa[E], fb[E];
#pragma unroll
    for (int j = 0; j < E; j++) {
        const float4 v = U[base + t + j * TT];
        fa[j] = make_float2(v.x, v.y);
        fb[j] = make_float2(v.z, v.w);
    }
    auto sync = [] { __syncthreads(); };
    float2 *la = lds[0][rr], *lb = lds[1][rr];
    if (a.flags & TROW_FWD) fft2fe<N, E, -1>(fa, fb, la, lb, t, 1, tw, sync);
    __syncthreads();                          // the transform's last LDS reads are done
#pragma unroll
    for (int j = 0; j < E; j++) {
        la[lds_pad(t + j * TT)] = fa[j];
        lb[lds_pad(t + j * TT)] = fb[j];
    }
    // each thread reads back only the slots it wrote: no barrier
    const size_t sprev = (size_t)a.prev.slice * n2, scur = (size_t)a.cur.slice * n2;
#pragma unroll 1
    for (int j = 0; j < E; j++) {
        const int o = lds_pad(t + j * TT);
        float2 P = la[o], S = lb[o];
        therm_cell(P, S, base + t + j * TT, n2, sprev, scur, Uout, NT, Q, DKint, a);
        la[o] = P;
        lb[o] = S;
    }
#pragma unroll
    for (int j = 0; j < E; j++) {
        fa[j] = la[lds_pad(t + j * TT)];
        fb[j] = lb[lds_pad(t + j * TT)];
    }
    if (a.flags & TROW_BWD) fft2fe<N, E, +1>(fa, fb, la, lb, t, 1, tw, sync);
#pragma unroll
    for (int j = 0; j < E; j++) U[base + t + j * TT] = make_float4(fa[j].x, fa[j].y, fb[j].x, fb[j].y);
}

// Wide layout <N, E, W>: one row per workgroup of W waves.  The first T = N/E lanes (one
// wavefront or less) run the two transforms; between them every thread of the group
// steps N/(64 W) cells of the row, so the per-cell work spreads over W times as many
// wavefronts as the transforms need.  Each thread loads its cells' phase accumulator and
// index-table entries before the first transform, which hides their latency behind it.
template <int N, int E, int W>
__global__ __launch_bounds__(64 * W) void k_therm_row_wide(float4 *__restrict__ U, float4 *__restrict__ Uout,
                                                           const float2 *__restrict__ twg,
                                                           const float2 *__restrict__ NT, float *__restrict__ Q,
                                                           float *__restrict__ DKint, ThermRowArgs a) {
    constexpr int TT = N / E, THREADS = 64 * W, L = LdsLen<N>::value, CPT = N / THREADS;
    static_assert(TT <= 64 && N % THREADS == 0, "the transforms run in one wavefront");
    __shared__ float2 lds[2][L];
    __shared__ float2 tw[N];
    const int tid = threadIdx.x;
    const size_t n2 = (size_t)N * N, base = (size_t)blockIdx.x * N;
    const size_t sprev = (size_t)a.prev.slice * n2, scur = (size_t)a.cur.slice * n2;
    float dkp[CPT];
    float2 nnp[CPT], nnc[CPT];
    float4 cfp[CPT], cfc[CPT];
    const bool pv = a.flags & (TROW_NPREV | TROW_QPREV), cu = a.flags & TROW_NCUR;
    const bool same = pv && a.prev.slice == a.cur.slice;      // consecutive steps share a heat slice
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const size_t idx = base + tid + q * THREADS;
        dkp[q] = (a.flags & TROW_FIRST) ? 0.0f : DKint[idx];
        nnp[q] = pv ? NT[sprev + idx] : make_float2(0.f, 0.f);
        cfp[q] = pv ? a.ct[sprev + idx] : make_float4(0.f, 0.f, 0.f, 0.f);
        nnc[q] = !cu ? make_float2(0.f, 0.f) : same ? nnp[q] : NT[scur + idx];
        cfc[q] = !cu ? make_float4(0.f, 0.f, 0.f, 0.f) : same ? cfp[q] : a.ct[scur + idx];
    }
    for (int j = tid; j < N; j += THREADS) tw[j] = twg[j];
    __syncthreads();
    auto sync = [] { wave_sync(); };
    if (tid < TT) {
        float2 fa[E], fb[E];
#pragma unroll
        for (int j = 0; j < E; j++) {
            const float4 v = U[base + tid + j * TT];
            fa[j] = make_float2(v.x, v.y);
            fb[j] = make_float2(v.z, v.w);
        }
        if (a.flags & TROW_FWD) fft2fe<N, E, -1>(fa, fb, lds[0], lds[1], tid, 1, tw, sync);
        wave_sync();
#pragma unroll
        for (int j = 0; j < E; j++) {
            lds[0][lds_pad(tid + j * TT)] = fa[j];
            lds[1][lds_pad(tid + j * TT)] = fb[j];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int cc = tid + q * THREADS, o = lds_pad(cc);
        float2 P = lds[0][o], S = lds[1][o];
        therm_cell_pre(P, S, base + cc, n2, sprev, dkp[q], nnp[q], cfp[q], nnc[q], cfc[q], Uout, Q, DKint, a);
        lds[0][o] = P;
        lds[1][o] = S;
    }
    __syncthreads();
    if (tid < TT) {
        float2 fa[E], fb[E];
#pragma unroll
        for (int j = 0; j < E; j++) {
            fa[j] = lds[0][lds_pad(tid + j * TT)];
            fb[j] = lds[1][lds_pad(tid + j * TT)];
        }
        if (a.flags & TROW_BWD) fft2fe<N, E, +1>(fa, fb, lds[0], lds[1], tid, 1, tw, sync);
#pragma unroll
        for (int j = 0; j < E; j++) U[base + tid + j * TT] = make_float4(fa[j].x, fa[j].y, fb[j].x, fb[j].y);
    }
}

// Split layout <N, E, W>, W >= 2: one row per workgroup; the row is loaded one cell per
// thread into LDS, waves 0 and 1 transform FH and SH side by side, every thread steps its
// cells, and the row is stored one cell per thread.
template <int N, int E, int W>
__global__ __launch_bounds__(64 * W) void k_therm_row_split(float4 *__restrict__ U, float4 *__restrict__ Uout,
                                                            const float2 *__restrict__ twg,
                                                            const float2 *__restrict__ NT, float *__restrict__ Q,
                                                            float *__restrict__ DKint, ThermRowArgs a) {
    constexpr int T = N / E, THREADS = 64 * W, L = LdsLen<N>::value, CPT = (N + THREADS - 1) / THREADS;
    static_assert(W >= 2 && T <= 64, "one wavefront per field transform");
    __shared__ float2 lds[2][L];
    __shared__ float2 tw[N];
    const int tid = threadIdx.x, f = tid >> 6, lane = tid & 63;
    const size_t n2 = (size_t)N * N, base = (size_t)blockIdx.x * N;
    const size_t sprev = (size_t)a.prev.slice * n2, scur = (size_t)a.cur.slice * n2;
    float dkp[CPT];
    float2 nnp[CPT], nnc[CPT];
    float4 cfp[CPT], cfc[CPT];
    const bool pv = a.flags & (TROW_NPREV | TROW_QPREV), cu = a.flags & TROW_NCUR;
    const bool same = pv && a.prev.slice == a.cur.slice;      // consecutive steps share a heat slice
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int c = tid + q * THREADS;
        const size_t idx = base + (c < N ? c : 0);
        dkp[q] = (a.flags & TROW_FIRST) ? 0.0f : DKint[idx];
        nnp[q] = pv ? NT[sprev + idx] : make_float2(0.f, 0.f);
        cfp[q] = pv ? a.ct[sprev + idx] : make_float4(0.f, 0.f, 0.f, 0.f);
        nnc[q] = !cu ? make_float2(0.f, 0.f) : same ? nnp[q] : NT[scur + idx];
        cfc[q] = !cu ? make_float4(0.f, 0.f, 0.f, 0.f) : same ? cfp[q] : a.ct[scur + idx];
    }
    for (int j = tid; j < N; j += THREADS) tw[j] = twg[j];
    for (int c = tid; c < N; c += THREADS) {
        const float4 v = U[base + c];
        lds[0][lds_pad(c)] = make_float2(v.x, v.y);
        lds[1][lds_pad(c)] = make_float2(v.z, v.w);
    }
    __syncthreads();
    if ((a.flags & TROW_FWD) && f < 2 && lane < T) fft_lds_field<N, E, -1>(lds[f], lane, tw);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int c = tid + q * THREADS;
        if (c < N) {
            float2 P = lds[0][lds_pad(c)], S = lds[1][lds_pad(c)];
            therm_cell_pre(P, S, base + c, n2, sprev, dkp[q], nnp[q], cfp[q], nnc[q], cfc[q], Uout, Q, DKint, a);
            lds[0][lds_pad(c)] = P;
            lds[1][lds_pad(c)] = S;
        }
    }
    __syncthreads();
    if ((a.flags & TROW_BWD) && f < 2 && lane < T) fft_lds_field<N, E, +1>(lds[f], lane, tw);
    __syncthreads();
    for (int c = tid; c < N; c += THREADS) {
        const float2 P = lds[0][lds_pad(c)], S = lds[1][lds_pad(c)];
        U[base + c] = make_float4(P.x, P.y, S.x, S.y);
    }
}

// ---------------------------------------------------------------------------
int jacobi_blocks(int nx, int ny, int hz) { return (nx / 64) * ((ny + 3) / 4) * hz; }

hipError_t launch_jacobi(float *Tf, const float *Ti, const float *Q, int nx, int ny, int hz, const JacobiConst &c,
                         double *part, int nparts, hipStream_t st) {
    if (nx % 64 != 0 || hz < 2 || ny < 2) return hipErrorInvalidValue;
    const dim3 grid(nx / 64, (ny + 3) / 4, hz);
    if (part) {
        if (nparts < jacobi_blocks(nx, ny, hz)) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_jacobi<true>, grid, dim3(256), 0, st, Tf, Ti, Q, nx, ny, hz, c, part);
    } else {
        hipLaunchKernelGGL(k_jacobi<false>, grid, dim3(256), 0, st, Tf, Ti, Q, nx, ny, hz, c, part);
    }
    return hipGetLastError();
}

hipError_t launch_sum_partials(const double *part, int nparts, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, part, nparts, out);
    return hipGetLastError();
}

hipError_t launch_index_table(const float *T, float2 *tab, float4 *ct, int64_t n, const ThermConst &k,
                              hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_index_table, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, T, tab, ct, n, k);
    return hipGetLastError();
}

template <int N, int E, int W>
static hipError_t therm_row_launch(float4 *U, float4 *Uout, const float2 *tw, const float2 *NT, float *Q,
                                   float *DKint, const ThermRowArgs &a, hipStream_t st) {
    using C = RowCfgT<N, E, W>;
    hipLaunchKernelGGL((k_therm_row_pass<N, E, W>), dim3(N / C::ROWS), dim3(C::THREADS), 0, st, U, Uout, tw, NT, Q,
                       DKint, a);
    return hipGetLastError();
}

template <int N, int E, int W>
static hipError_t therm_wide_launch(float4 *U, float4 *Uout, const float2 *tw, const float2 *NT, float *Q,
                                    float *DKint, const ThermRowArgs &a, hipStream_t st) {
    hipLaunchKernelGGL((k_therm_row_wide<N, E, W>), dim3(N), dim3(64 * W), 0, st, U, Uout, tw, NT, Q, DKint, a);
    return hipGetLastError();
}

template <int N, int E, int W>
static hipError_t therm_split_launch(float4 *U, float4 *Uout, const float2 *tw, const float2 *NT, float *Q,
                                     float *DKint, const ThermRowArgs &a, hipStream_t st) {
    hipLaunchKernelGGL((k_therm_row_split<N, E, W>), dim3(N), dim3(64 * W), 0, st, U, Uout, tw, NT, Q, DKint, a);
    return hipGetLastError();
}

#define THRV(N, E, W) return therm_row_launch<N, E, W>(U, Uout, tw, NT, Q, DKint, a, st)
#define THRS(N, E, W) return therm_split_launch<N, E, W>(U, Uout, tw, NT, Q, DKint, a, st)
#define THRW(N, E, W) return therm_wide_launch<N, E, W>(U, Uout, tw, NT, Q, DKint, a, st)

hipError_t launch_thermal_row_pass(int n, float4 *U, float4 *Uout, const float2 *tw, const float2 *NT, float *Q,
                                   float *DKint, const ThermRowArgs &a, hipStream_t st, int variant) {
    if (variant < 0 || variant >= kThermRowVariantCount) return hipErrorInvalidValue;
    switch (n * 16 + variant) {
        case 64 * 16 + 0: THRV(64, 8, 1);
        case 64 * 16 + 1: THRV(64, 4, 1);
        case 64 * 16 + 2: THRV(64, 16, 1);
        case 64 * 16 + 3: THRV(64, 8, 2);
        case 64 * 16 + 4: THRW(64, 4, 1);
        case 64 * 16 + 5: THRW(64, 8, 1);
        case 64 * 16 + 6: THRS(64, 4, 2);
        case 128 * 16 + 0: THRV(128, 16, 1);
        case 128 * 16 + 1: THRV(128, 8, 1);
        case 128 * 16 + 2: THRV(128, 4, 1);
        case 128 * 16 + 3: THRV(128, 4, 2);
        case 128 * 16 + 4: THRW(128, 4, 2);
        case 128 * 16 + 5: THRW(128, 4, 1);
        case 128 * 16 + 6: THRS(128, 4, 2);
        case 256 * 16 + 0: THRV(256, 16, 1);
        case 256 * 16 + 1: THRV(256, 8, 1);
        case 256 * 16 + 2: THRV(256, 4, 1);
        case 256 * 16 + 3: THRV(256, 4, 4);
        case 256 * 16 + 4: THRW(256, 4, 4);
        case 256 * 16 + 5: THRW(256, 4, 2);
        case 256 * 16 + 6: THRS(256, 4, 4);
        case 512 * 16 + 0: THRV(512, 8, 1);
        case 512 * 16 + 1: THRV(512, 4, 2);
        case 512 * 16 + 2: THRV(512, 16, 1);
        case 512 * 16 + 3: THRV(512, 8, 2);
        case 512 * 16 + 4: THRW(512, 8, 8);
        case 512 * 16 + 5: THRW(512, 8, 4);
        case 512 * 16 + 6: THRS(512, 8, 8);
        default: return hipErrorInvalidValue;
    }
}
#undef THRV
#undef THRW
#undef THRS

}  // namespace shg
