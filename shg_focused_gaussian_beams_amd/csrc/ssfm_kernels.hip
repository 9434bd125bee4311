// ssfm_kernels.hip — the split-step z-march on gfx950.
//
// State layout in HBM: one float4 per transverse cell, {FH.re, FH.im, SH.re, SH.im},
// row-major with x fastest (the reference's IDX, src/headers/Common.h:27-29), so one
// 16-byte load gives both fields of a cell.
//
// One z-step of the reference (Solver::SSFM, src/headers/Solver.h:236-242) is
//     RK4_s ; F_bwd ; x H ; F_fwd ; /(NX*NY) ; RK4_s
// with the 2D FFTs split as row transforms R and column transforms C.  Since H is
// diagonal in the 2D spectrum, the march becomes two kernels per step:
//     row pass  :  R_fwd (end of step s-1) -> RK4_{s-1} -> RK4_s -> R_bwd
//     col pass  :  C_bwd -> x H/(NX*NY) -> C_fwd
// so every step reads and writes the field pair twice (once per pass), and all FFT,
// propagator, scaling and RK4 work happens in registers/LDS in between.
#include <hip/hip_runtime.h>

#include "fft_lds.h"
#include "nl_step.h"
#include "pass_cfg.h"
#include "ssfm_kernels.h"

namespace shg {

__device__ __forceinline__ void block_sync() { __syncthreads(); }

__device__ __forceinline__ void rk4(float2 &P, float2 &S, float4 c, const NlConst &k) {
    rk4_cell(P, S, c, c, c, k.lossP, k.lossS, k.tpa, k.h6);
}

// Row pass over rows [blockIdx.x*ROWS, +ROWS) of instance blockIdx.y (see RowMode);
// layout <N, E, W> as in pass_cfg.h.
template <int N, int E, int W>
__global__ __launch_bounds__(64 * W) void k_row_pass(const InstDesc *__restrict__ descs, const float2 *__restrict__ twg,
                                                     int step, int mode, float4 *__restrict__ snap) {
    using C = RowCfgT<N, E, W>;
    constexpr int T = C::T;
    __shared__ float2 lds[2][C::ROWS][C::L];
    __shared__ float2 tw[N];
    const InstDesc d = descs[blockIdx.y];
    if (mode == ROW_AUTO) mode = step + 1 < d.steps ? ROW_MID : ROW_END;
    const int lane = threadIdx.x;
    const int t = lane % T, rr = lane / T;
    const int row = blockIdx.x * C::ROWS + rr;
    const float4 c0 = gp(d.coef)[step];
    const float4 c1 = gp(d.coef)[mode == ROW_MID ? step + 1 : step];
    // the row's loads go out before the twiddles are staged, so that the two latencies
    // overlap instead of adding up across the barrier
    const auto Ur = gp(d.U) + (size_t)row * N;
    float2 fa[E], fb[E];
#pragma unroll
    for (int j = 0; j < E; j++) {
        const float4 v = Ur[t + j * T];
        fa[j] = make_float2(v.x, v.y);
        fb[j] = make_float2(v.z, v.w);
    }
    for (int j = lane; j < N; j += C::THREADS) tw[j] = twg[j];
    __syncthreads();
    auto sync = [] { block_sync(); };
    if (mode != ROW_START) {
        fft2fe<N, E, -1>(fa, fb, lds[0][rr], lds[1][rr], t, 1, tw, sync);
#pragma unroll
        for (int j = 0; j < E; j++) rk4(fa[j], fb[j], c0, d.nl);
        if (snap)
#pragma unroll
            for (int j = 0; j < E; j++)
                gp(snap)[(size_t)row * N + t + j * T] = make_float4(fa[j].x, fa[j].y, fb[j].x, fb[j].y);
    }
    if (mode != ROW_END) {
#pragma unroll
        for (int j = 0; j < E; j++) rk4(fa[j], fb[j], c1, d.nl);
        fft2fe<N, E, +1>(fa, fb, lds[0][rr], lds[1][rr], t, 1, tw, sync);
    }
#pragma unroll
    for (int j = 0; j < E; j++) Ur[t + j * T] = make_float4(fa[j].x, fa[j].y, fb[j].x, fb[j].y);
}

// Column pass over columns [blockIdx.x*CW, +CW) of instance blockIdx.y: C_bwd, x H/(N*N),
// C_fwd.  G holds the separable propagator: Gx_p[N], Gy_p[N], Gx_s[N], Gy_s[N] (Gy
// carries the 1/N^2 of FFTscale).  Layout <N, E, CW> as in pass_cfg.h.
template <int N, int E, int CW>
__global__ __launch_bounds__(CW *(N / E)) void k_col_pass(const InstDesc *__restrict__ descs,
                                                          const float2 *__restrict__ twg) {
    using C = ColCfgT<N, E, CW>;
    constexpr int T = C::T;
    __shared__ float2 lds[2][C::L * CW];
    __shared__ float2 tw[N];
    const InstDesc d = descs[blockIdx.y];
    const auto U = gp(d.U);
    const auto G = gp(d.G);
    const int tid = threadIdx.x;
    const int c = tid % CW, t = tid / CW;
    const int x = xcd_group(blockIdx.x, gridDim.x) * CW + c;
    // the column's loads go out before the twiddles are staged, so that the two latencies
    // overlap instead of adding up across the barrier; the propagator factors, loaded
    // before the transform, hide theirs behind it
    float2 fa[E], fb[E];
#pragma unroll
    for (int j = 0; j < E; j++) {
        const float4 v = U[(size_t)(t + j * T) * N + x];
        fa[j] = make_float2(v.x, v.y);
        fb[j] = make_float2(v.z, v.w);
    }
    for (int j = tid; j < N; j += C::THREADS) tw[j] = twg[j];
    const float2 gxp = G[x], gxs = G[2 * N + x];
    float2 gyp[E], gys[E];
#pragma unroll
    for (int j = 0; j < E; j++) {
        gyp[j] = G[N + t + j * T];
        gys[j] = G[3 * N + t + j * T];
    }
    __syncthreads();
    auto sync = [] { block_sync(); };
    fft2fe<N, E, +1>(fa, fb, &lds[0][c], &lds[1][c], t, CW, tw, sync);
#pragma unroll
    for (int j = 0; j < E; j++) {
        fa[j] = c_mul(c_mul(gxp, gyp[j]), fa[j]);
        fb[j] = c_mul(c_mul(gxs, gys[j]), fb[j]);
    }
    fft2fe<N, E, -1>(fa, fb, &lds[0][c], &lds[1][c], t, CW, tw, sync);
#pragma unroll
    for (int j = 0; j < E; j++)
        U[(size_t)(t + j * T) * N + x] = make_float4(fa[j].x, fa[j].y, fb[j].x, fb[j].y);
}

// Column pass with the two fields on separate waves, straight from registers: the
// diffraction does not couple FH and SH, so each field of CW columns is owned by its own
// CW*T threads (field 0 the first half of the workgroup, field 1 the second), each
// holding E elements of one column; the halves run their transforms side by side (half
// the dependent chain of k_col_pass per thread, twice the waves) and meet only at the
// barriers.  Same operations as k_col_pass, so the same result bit for bit.
template <int N, int E, int CW>
__global__ __launch_bounds__(2 * CW *(N / E)) void k_col_fsplit(const InstDesc *__restrict__ descs,
                                                                const float2 *__restrict__ twg) {
    using C = ColCfgT<N, E, CW>;
    constexpr int T = C::T, HALF = C::THREADS;
    static_assert(HALF % 64 == 0, "each field on whole waves");
    __shared__ float2 lds[2][C::L * CW];
    __shared__ float2 tw[N];
    const InstDesc d = descs[blockIdx.y];
    const auto U = gp(reinterpret_cast<float2 *>(d.U));     // cell i, field f at 2 i + f
    const auto G = gp(d.G);
    const int tid = threadIdx.x, f = tid / HALF, r = tid - f * HALF;
    const int c = r % CW, t = r / CW;
    const int x = xcd_group(blockIdx.x, gridDim.x) * CW + c;
    // the column's loads go out before the twiddles are staged (the latencies overlap
    // instead of adding up across the barrier)
    float2 a[E], gy[E];
#pragma unroll
    for (int j = 0; j < E; j++) a[j] = U[2 * ((size_t)(t + j * T) * N + x) + f];
    for (int j = tid; j < N; j += 2 * HALF) tw[j] = twg[j];
    const float2 gx = G[2 * f * N + x];                     // Gx of the field, then its Gy
#pragma unroll
    for (int j = 0; j < E; j++) gy[j] = G[(2 * f + 1) * N + t + j * T];
    __syncthreads();
    auto sync = [] { block_sync(); };
    fft1fe<N, E, +1>(a, &lds[f][c], t, CW, tw, sync);
#pragma unroll
    for (int j = 0; j < E; j++) a[j] = c_mul(c_mul(gx, gy[j]), a[j]);
    fft1fe<N, E, -1>(a, &lds[f][c], t, CW, tw, sync);
#pragma unroll
    for (int j = 0; j < E; j++) U[2 * ((size_t)(t + j * T) * N + x) + f] = a[j];
}

// ---------------------------------------------------------------------------
// Split layouts: the workgroup loads its cells one per thread (coalesced, spread over all
// its waves) into LDS in natural order; each 1D transform of one field of one line runs
// on its own wavefront (so the two fields' transforms run side by side instead of one
// wave doing both); the per-cell work again runs one cell per thread.  Lower latency per
// pass for fields too small to fill the GPU in the register layouts.

// Row pass, one row per workgroup of W >= 2 waves: waves 0 and 1 transform FH and SH.
template <int N, int E, int W>
__global__ __launch_bounds__(64 * W) void k_row_split(const InstDesc *__restrict__ descs,
                                                      const float2 *__restrict__ twg, int step, int mode,
                                                      float4 *__restrict__ snap) {
    constexpr int T = N / E, THREADS = 64 * W, L = LdsLen<N>::value;
    static_assert(W >= 2 && T <= 64, "one wavefront per field transform");
    __shared__ float2 lds[2][L];
    __shared__ float2 tw[N];
    const InstDesc d = descs[blockIdx.y];
    if (mode == ROW_AUTO) mode = step + 1 < d.steps ? ROW_MID : ROW_END;
    const int tid = threadIdx.x, f = tid >> 6, lane = tid & 63;
    const auto Ur = gp(d.U) + (size_t)blockIdx.x * N;
    const float4 c0 = gp(d.coef)[step];
    const float4 c1 = gp(d.coef)[mode == ROW_MID ? step + 1 : step];
    for (int j = tid; j < N; j += THREADS) tw[j] = twg[j];
    for (int c = tid; c < N; c += THREADS) {
        const float4 v = Ur[c];
        lds[0][lds_pad(c)] = make_float2(v.x, v.y);
        lds[1][lds_pad(c)] = make_float2(v.z, v.w);
    }
    __syncthreads();
    if (mode != ROW_START && f < 2 && lane < T) fft_lds_field<N, E, -1>(lds[f], lane, tw);
    __syncthreads();
    for (int c = tid; c < N; c += THREADS) {
        float2 P = lds[0][lds_pad(c)], S = lds[1][lds_pad(c)];
        if (mode != ROW_START) {
            rk4(P, S, c0, d.nl);
            if (snap) gp(snap)[(size_t)blockIdx.x * N + c] = make_float4(P.x, P.y, S.x, S.y);
        }
        if (mode != ROW_END) rk4(P, S, c1, d.nl);
        lds[0][lds_pad(c)] = P;
        lds[1][lds_pad(c)] = S;
    }
    __syncthreads();
    if (mode != ROW_END && f < 2 && lane < T) fft_lds_field<N, E, +1>(lds[f], lane, tw);
    __syncthreads();
    for (int c = tid; c < N; c += THREADS) {
        const float2 P = lds[0][lds_pad(c)], S = lds[1][lds_pad(c)];
        Ur[c] = make_float4(P.x, P.y, S.x, S.y);
    }
}

// Column pass over CW adjacent columns per workgroup of 2*CW waves: loads and stores move
// CW*16-byte row segments; wave k transforms field k / CW of column k % CW.
template <int N, int E, int CW>
__global__ __launch_bounds__(128 * CW) void k_col_split(const InstDesc *__restrict__ descs,
                                                        const float2 *__restrict__ twg) {
    constexpr int T = N / E, THREADS = 128 * CW, L = LdsLen<N>::value;
    static_assert(T <= 64, "one wavefront per field transform");
    __shared__ float2 lds[2 * CW][L];
    __shared__ float2 tw[N];
    constexpr int CPT = (CW * N + THREADS - 1) / THREADS;   // cells per thread outside the transforms
    const InstDesc d = descs[blockIdx.y];
    const auto U = gp(d.U);
    const auto G = gp(d.G);
    const int tid = threadIdx.x, k = tid >> 6, lane = tid & 63;
    const int x0 = xcd_group(blockIdx.x, gridDim.x) * CW;
    for (int j = tid; j < N; j += THREADS) tw[j] = twg[j];
    for (int i = tid; i < CW * N; i += THREADS) {
        const int y = i / CW, c = i % CW;
        const float4 v = U[(size_t)y * N + x0 + c];
        lds[c][lds_pad(y)] = make_float2(v.x, v.y);
        lds[CW + c][lds_pad(y)] = make_float2(v.z, v.w);
    }
    // the propagator factors of the thread's cells, loaded before the transform
    float2 gxp[CPT], gyp[CPT], gxs[CPT], gys[CPT];
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int i = tid + q * THREADS, ii = i < CW * N ? i : 0;
        const int y = ii / CW, x = x0 + ii % CW;
        gxp[q] = G[x];
        gyp[q] = G[N + y];
        gxs[q] = G[2 * N + x];
        gys[q] = G[3 * N + y];
    }
    __syncthreads();
    if (lane < T) fft_lds_field<N, E, +1>(lds[k], lane, tw);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int i = tid + q * THREADS;
        if (i < CW * N) {
            const int y = i / CW, c = i % CW;
            lds[c][lds_pad(y)] = c_mul(c_mul(gxp[q], gyp[q]), lds[c][lds_pad(y)]);
            lds[CW + c][lds_pad(y)] = c_mul(c_mul(gxs[q], gys[q]), lds[CW + c][lds_pad(y)]);
        }
    }
    __syncthreads();
    if (lane < T) fft_lds_field<N, E, -1>(lds[k], lane, tw);
    __syncthreads();
    for (int i = tid; i < CW * N; i += THREADS) {
        const int y = i / CW, c = i % CW;
        const float2 P = lds[c][lds_pad(y)], S = lds[CW + c][lds_pad(y)];
        U[(size_t)y * N + x0 + c] = make_float4(P.x, P.y, S.x, S.y);
    }
}

// FP64 sums of |FH|^2 and |SH|^2: per-block partials, then one block folds them.
// Deterministic (fixed reduction tree).
__global__ __launch_bounds__(256) void k_power_partial(const float4 *__restrict__ U, int64_t n, double2 *part) {
    double sp = 0.0, ss = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float4 v = U[i];
        sp += (double)v.x * v.x + (double)v.y * v.y;
        ss += (double)v.z * v.z + (double)v.w * v.w;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sp += __shfl_down(sp, off, 64);
        ss += __shfl_down(ss, off, 64);
    }
    __shared__ double2 w[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) w[wv] = make_double2(sp, ss);
    __syncthreads();
    if (threadIdx.x == 0) {
        double2 r = w[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); k++) { r.x += w[k].x; r.y += w[k].y; }
        part[blockIdx.x] = r;
    }
}

__global__ __launch_bounds__(256) void k_power_final(const double2 *__restrict__ part, int nparts, double2 *out) {
    __shared__ double2 s[256];
    double2 r = make_double2(0.0, 0.0);
    for (int i = threadIdx.x; i < nparts; i += 256) { r.x += part[i].x; r.y += part[i].y; }
    s[threadIdx.x] = r;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) { s[threadIdx.x].x += s[threadIdx.x + w].x; s[threadIdx.x].y += s[threadIdx.x + w].y; }
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

// One block per table entry: FP64 sums of |FH|^2 and |SH|^2 of that instance.
__global__ __launch_bounds__(256) void k_power_batch(const InstDesc *__restrict__ descs, const int *__restrict__ ncell,
                                                     double2 *out) {
    const auto U = gp(descs[blockIdx.x].U);
    const int n = ncell[blockIdx.x];
    double sp = 0.0, ss = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const float4 v = U[i];
        sp += (double)v.x * v.x + (double)v.y * v.y;
        ss += (double)v.z * v.z + (double)v.w * v.w;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sp += __shfl_down(sp, off, 64);
        ss += __shfl_down(ss, off, 64);
    }
    __shared__ double2 w[4];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = make_double2(sp, ss);
    __syncthreads();
    if (threadIdx.x == 0)
        out[blockIdx.x] = make_double2(w[0].x + w[1].x + w[2].x + w[3].x, w[0].y + w[1].y + w[2].y + w[3].y);
}

// ---------------------------------------------------------------------------
// Launchers.  Variant tables (index 0 = default): rows {E, waves}, columns {E, CW}.

template <int N, int E, int W>
static hipError_t row_launch(const InstDesc *d, int ninst, const float2 *tw, int step, int mode, hipStream_t st,
                             float4 *snap) {
    using C = RowCfgT<N, E, W>;
    hipLaunchKernelGGL((k_row_pass<N, E, W>), dim3(N / C::ROWS, ninst), dim3(C::THREADS), 0, st, d, tw, step, mode,
                       snap);
    return hipGetLastError();
}
template <int N, int E, int CW>
static hipError_t col_launch(const InstDesc *d, int ninst, const float2 *tw, hipStream_t st) {
    using C = ColCfgT<N, E, CW>;
    hipLaunchKernelGGL((k_col_pass<N, E, CW>), dim3(N / CW, ninst), dim3(C::THREADS), 0, st, d, tw);
    return hipGetLastError();
}

template <int N, int E, int W>
static hipError_t row_split_launch(const InstDesc *d, int ninst, const float2 *tw, int step, int mode, hipStream_t st,
                                   float4 *snap) {
    hipLaunchKernelGGL((k_row_split<N, E, W>), dim3(N, ninst), dim3(64 * W), 0, st, d, tw, step, mode, snap);
    return hipGetLastError();
}
template <int N, int E, int CW>
static hipError_t col_split_launch(const InstDesc *d, int ninst, const float2 *tw, hipStream_t st) {
    hipLaunchKernelGGL((k_col_split<N, E, CW>), dim3(N / CW, ninst), dim3(128 * CW), 0, st, d, tw);
    return hipGetLastError();
}

template <int N, int E, int CW>
static hipError_t col_fsplit_launch(const InstDesc *d, int ninst, const float2 *tw, hipStream_t st) {
    using C = ColCfgT<N, E, CW>;
    hipLaunchKernelGGL((k_col_fsplit<N, E, CW>), dim3(N / CW, ninst), dim3(2 * C::THREADS), 0, st, d, tw);
    return hipGetLastError();
}

#define ROWV(N, E, W) return row_launch<N, E, W>(d, ninst, tw, step, mode, st, snap)
#define COLV(N, E, CW) return col_launch<N, E, CW>(d, ninst, tw, st)
#define ROWS(N, E, W) return row_split_launch<N, E, W>(d, ninst, tw, step, mode, st, snap)
#define COLS(N, E, CW) return col_split_launch<N, E, CW>(d, ninst, tw, st)
#define COLF(N, E, CW) return col_fsplit_launch<N, E, CW>(d, ninst, tw, st)

hipError_t launch_row_pass(int n, const InstDesc *d, int ninst, const float2 *tw, int step, int mode, hipStream_t st,
                           int variant, float4 *snap) {
    if (ninst <= 0 || ninst > 65535 || step < 0 || mode < ROW_START || mode > ROW_AUTO) return hipErrorInvalidValue;
    if (snap && ninst != 1) return hipErrorInvalidValue;
    switch (n * 16 + variant) {
        case 64 * 16 + 0: ROWV(64, 8, 1);
        case 64 * 16 + 1: ROWV(64, 4, 1);
        case 64 * 16 + 2: ROWV(64, 8, 4);
        case 64 * 16 + 3: ROWV(64, 16, 1);
        case 64 * 16 + 4: ROWS(64, 4, 2);
        case 128 * 16 + 0: ROWV(128, 16, 1);
        case 128 * 16 + 1: ROWV(128, 8, 1);
        case 128 * 16 + 2: ROWV(128, 4, 1);
        case 128 * 16 + 3: ROWV(128, 16, 2);
        case 128 * 16 + 4: ROWS(128, 4, 2);
        case 256 * 16 + 0: ROWV(256, 16, 1);
        case 256 * 16 + 1: ROWV(256, 8, 1);
        case 256 * 16 + 2: ROWV(256, 4, 1);
        case 256 * 16 + 3: ROWV(256, 16, 2);
        case 256 * 16 + 4: ROWS(256, 4, 4);
        case 512 * 16 + 0: ROWV(512, 8, 1);
        case 512 * 16 + 1: ROWV(512, 16, 1);
        case 512 * 16 + 2: ROWV(512, 4, 2);
        case 512 * 16 + 3: ROWV(512, 8, 4);
        case 512 * 16 + 4: ROWS(512, 8, 8);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_col_pass(int n, const InstDesc *d, int ninst, const float2 *tw, hipStream_t st, int variant) {
    if (ninst <= 0 || ninst > 65535) return hipErrorInvalidValue;
    switch (n * 16 + variant) {
        case 64 * 16 + 0: COLV(64, 8, 8);
        case 64 * 16 + 1: COLV(64, 4, 8);
        case 64 * 16 + 2: COLV(64, 8, 16);
        case 64 * 16 + 3: COLV(64, 4, 4);
        case 64 * 16 + 4: COLV(64, 4, 16);
        case 64 * 16 + 5: COLS(64, 4, 1);
        case 64 * 16 + 6: COLS(64, 4, 2);
        case 64 * 16 + 7: COLS(64, 4, 4);
        case 64 * 16 + 8: COLF(64, 4, 4);
        case 64 * 16 + 9: COLF(64, 4, 8);
        case 128 * 16 + 0: COLV(128, 16, 8);
        case 128 * 16 + 1: COLV(128, 8, 8);
        case 128 * 16 + 2: COLV(128, 4, 8);
        case 128 * 16 + 3: COLV(128, 4, 2);
        case 128 * 16 + 4: COLV(128, 4, 4);
        case 128 * 16 + 5: COLS(128, 4, 1);
        case 128 * 16 + 6: COLS(128, 4, 2);
        case 128 * 16 + 7: COLS(128, 4, 4);
        case 128 * 16 + 8: COLF(128, 4, 2);
        case 128 * 16 + 9: COLF(128, 4, 4);
        case 256 * 16 + 0: COLV(256, 16, 8);
        case 256 * 16 + 1: COLV(256, 8, 8);
        case 256 * 16 + 2: COLV(256, 4, 8);
        case 256 * 16 + 3: COLV(256, 4, 1);
        case 256 * 16 + 4: COLV(256, 4, 2);
        case 256 * 16 + 5: COLS(256, 4, 1);
        case 256 * 16 + 6: COLS(256, 4, 2);
        case 256 * 16 + 7: COLS(256, 4, 4);
        case 256 * 16 + 8: COLF(256, 4, 1);
        case 256 * 16 + 9: COLF(256, 4, 2);
        case 512 * 16 + 0: COLV(512, 8, 4);
        case 512 * 16 + 1: COLV(512, 16, 4);
        case 512 * 16 + 2: COLV(512, 4, 4);
        case 512 * 16 + 3: COLV(512, 8, 1);
        case 512 * 16 + 4: COLV(512, 8, 2);
        case 512 * 16 + 5: COLS(512, 8, 1);
        case 512 * 16 + 6: COLS(512, 8, 2);
        case 512 * 16 + 7: COLS(512, 8, 4);
        case 512 * 16 + 8: COLF(512, 8, 1);
        case 512 * 16 + 9: COLF(512, 8, 2);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_power(const float4 *U, int64_t ncell, double2 *part, int nparts, double2 *out, hipStream_t st) {
    hipLaunchKernelGGL(k_power_partial, dim3(nparts), dim3(256), 0, st, U, ncell, part);
    hipLaunchKernelGGL(k_power_final, dim3(1), dim3(256), 0, st, part, nparts, out);
    return hipGetLastError();
}

hipError_t launch_fold2(const double2 *part, int nparts, double2 *out, hipStream_t st) {
    if (nparts <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_power_final, dim3(1), dim3(256), 0, st, part, nparts, out);
    return hipGetLastError();
}

hipError_t launch_power_batch(const InstDesc *d, const int *ncell, int ninst, double2 *out, hipStream_t st) {
    if (ninst <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_power_batch, dim3(ninst), dim3(256), 0, st, d, ncell, out);
    return hipGetLastError();
}

}  // namespace shg
