// adjoint_kernels.hip — discrete adjoint of the plain z-march on gfx950.
//
// Forward (one checkpoint segment, steps [s0, s1)) with a tape:
//   row pass writes the real-space state before each step's second RK4 (A''_s); column
//   pass writes the 2D spectrum it multiplies by H.  The state before a step's first RK4,
//   A_s = RK4_{s-1}(A''_{s-1}), is recomputed by the adjoint (the segment's first one is
//   its checkpoint).
// Adjoint (reverse step order), with lambda the gradient dJ = Re sum conj(lambda) dA:
//   row pass  : R_fwd -> RK4_s^T at A_s -> RK4_{s-1}^T at A''_{s-1} -> R_bwd
//   col pass  : C_bwd -> x conj(H)/N^2 -> C_fwd   (L^H = (1/N^2) F_fwd conj(H) F_bwd)
// RK4^T recomputes the forward stages from the taped state and runs them backwards; the
// per-cell maps are not holomorphic (conj(P), |S|^2), so each stage uses the real adjoint
// zbar = conj(dg/dz) gbar + (dg/dconj z) conj(gbar).  FP64 partial sums per block give the
// sensitivities of J to the step coefficients (c1, c2, tpa) and to the propagator (for
// the temperature gradient).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "adjoint_kernels.h"
#include "fft_lds.h"
#include "adjoint_cell.h"
#include "nl_step.h"
#include "pass_cfg.h"

namespace shg {

// Sum of the five accumulators over the workgroup (wave shuffles, then LDS across the
// waves); thread 0 writes them to `out`.  All threads of the group must call it.
template <int WAVES>
__device__ __forceinline__ void block_store(AdjAcc a, double *out) {
    __shared__ double red[WAVES][5];
    for (int off = 32; off > 0; off >>= 1) {
        a.c1r += __shfl_down(a.c1r, off, 64);
        a.c1i += __shfl_down(a.c1i, off, 64);
        a.c2r += __shfl_down(a.c2r, off, 64);
        a.c2i += __shfl_down(a.c2i, off, 64);
        a.tpa += __shfl_down(a.tpa, off, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[w][0] = a.c1r; red[w][1] = a.c1i; red[w][2] = a.c2r; red[w][3] = a.c2i; red[w][4] = a.tpa;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        double s = 0.0;
        for (int k = 0; k < WAVES; k++) s += red[k][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------
// Forward row pass with tape (single instance, modes as RowMode without AUTO).
// START at s0: ckpt = input (if ckpt is set), then RK4_s0, R_bwd.
// MID at s   : R_fwd, tape_a2[s-s0] = A''_s, RK4_s, RK4_{s+1}, R_bwd.
// END at s   : R_fwd, tape_a2[s-s0] = A''_s, RK4_s  (real-space state of the segment end).
template <int N>
__global__ __launch_bounds__(64 * AdjWide<N>::W) void k_row_tape(const InstDesc *__restrict__ descs,
                                                                const float2 *__restrict__ twg, int step, int mode,
                                                                TapeArgs tp) {
    constexpr int E = AdjWide<N>::E, W = AdjWide<N>::W, T = N / E, THREADS = 64 * W, L = LdsLen<N>::value;
    static_assert(W >= 2 && T <= 64, "one wavefront per field transform");
    __shared__ float2 lds[2][L];
    __shared__ float2 tw[N];
    const InstDesc d = descs[0];
    const int tid = threadIdx.x, f = tid >> 6, lane = tid & 63;
    const size_t n2 = (size_t)N * N, base = (size_t)blockIdx.x * N;
    const int ls = step - tp.s0;
    const float4 c0 = gp(d.coef)[step];
    const float4 c1 = gp(d.coef)[mode == ROW_MID ? step + 1 : step];
    for (int j = tid; j < N; j += THREADS) tw[j] = twg[j];
    for (int c = tid; c < N; c += THREADS) {
        const float4 v = gp(d.U)[base + c];
        lds[0][lds_pad(c)] = make_float2(v.x, v.y);
        lds[1][lds_pad(c)] = make_float2(v.z, v.w);
    }
    __syncthreads();
    if (mode != ROW_START && f < 2 && lane < T) fft_lds_field<N, E, -1>(lds[f], lane, tw);
    __syncthreads();
    for (int c = tid; c < N; c += THREADS) {
        float2 P = lds[0][lds_pad(c)], S = lds[1][lds_pad(c)];
        if (mode == ROW_START) {
            if (tp.ckpt) tp.ckpt[base + c] = make_float4(P.x, P.y, S.x, S.y);
        } else {
            if (tp.a2) tp.a2[(size_t)ls * n2 + base + c] = make_float4(P.x, P.y, S.x, S.y);
            rk4_cell(P, S, c0, c0, c0, d.nl.lossP, d.nl.lossS, d.nl.tpa, d.nl.h6);
        }
        if (mode != ROW_END) rk4_cell(P, S, c1, c1, c1, d.nl.lossP, d.nl.lossS, d.nl.tpa, d.nl.h6);
        lds[0][lds_pad(c)] = P;
        lds[1][lds_pad(c)] = S;
    }
    __syncthreads();
    if (mode != ROW_END && f < 2 && lane < T) fft_lds_field<N, E, +1>(lds[f], lane, tw);
    __syncthreads();
    for (int c = tid; c < N; c += THREADS) {
        const float2 P = lds[0][lds_pad(c)], S = lds[1][lds_pad(c)];
        gp(d.U)[base + c] = make_float4(P.x, P.y, S.x, S.y);
    }
}

// Columns x0 .. x0+CW-1 of U into LDS (field k/CW of column k%CW at lds[k]).
template <int N, int CW, int THREADS>
__device__ __forceinline__ void cols_to_lds(gptr<const float4> U, int x0, float2 (*lds)[LdsLen<N>::value]) {
    for (int i = threadIdx.x; i < CW * N; i += THREADS) {
        const int y = i / CW, c = i % CW;
        const float4 v = U[(size_t)y * N + x0 + c];
        lds[c][lds_pad(y)] = make_float2(v.x, v.y);
        lds[CW + c][lds_pad(y)] = make_float2(v.z, v.w);
    }
}
template <int N, int CW, int THREADS>
__device__ __forceinline__ void lds_to_cols(gptr<float4> U, int x0, float2 (*lds)[LdsLen<N>::value]) {
    for (int i = threadIdx.x; i < CW * N; i += THREADS) {
        const int y = i / CW, c = i % CW;
        const float2 P = lds[c][lds_pad(y)], S = lds[CW + c][lds_pad(y)];
        U[(size_t)y * N + x0 + c] = make_float4(P.x, P.y, S.x, S.y);
    }
}

// Forward column pass with tape: tape_spec[s-s0] = the 2D spectrum before x H.
template <int N>
__global__ __launch_bounds__(128 * AdjCol<N>::CW) void k_col_tape(const InstDesc *__restrict__ descs,
                                                                 const float2 *__restrict__ twg, int step, TapeArgs tp) {
    constexpr int E = AdjCol<N>::E, CW = AdjCol<N>::CW, T = N / E, THREADS = 128 * CW, L = LdsLen<N>::value;
    static_assert(T <= 64, "one wavefront per field transform");
    __shared__ float2 lds[2 * CW][L];
    __shared__ float2 tw[N];
    constexpr int CPT = (CW * N + THREADS - 1) / THREADS;   // cells per thread (guarded: 64 x 2 < 256)
    const InstDesc d = descs[0];
    const auto G = gp(d.G);
    const int tid = threadIdx.x, k = tid >> 6, lane = tid & 63, x0 = xcd_group(blockIdx.x, gridDim.x) * CW;
    for (int j = tid; j < N; j += THREADS) tw[j] = twg[j];
    cols_to_lds<N, CW, THREADS>(gp(d.U), x0, lds);
    // the propagator factors of the thread's cells, loaded before the transform
    float2 gxp[CPT], gyp[CPT], gxs[CPT], gys[CPT];
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int i = tid + q * THREADS, ii = i < CW * N ? i : 0, y = ii / CW, x = x0 + ii % CW;
        gxp[q] = G[x];
        gyp[q] = G[N + y];
        gxs[q] = G[2 * N + x];
        gys[q] = G[3 * N + y];
    }
    __syncthreads();
    if (lane < T) fft_lds_field<N, E, +1>(lds[k], lane, tw);
    __syncthreads();
    const size_t soff = (size_t)(step - tp.s0) * N * N;
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int i = tid + q * THREADS, y = i / CW, c = i % CW, x = x0 + c;
        if (i >= CW * N) break;
        const float2 P = lds[c][lds_pad(y)], S = lds[CW + c][lds_pad(y)];
        if (tp.spec) gp(tp.spec)[soff + (size_t)y * N + x] = make_float4(P.x, P.y, S.x, S.y);
        lds[c][lds_pad(y)] = c_mul(c_mul(gxp[q], gyp[q]), P);
        lds[CW + c][lds_pad(y)] = c_mul(c_mul(gxs[q], gys[q]), S);
    }
    __syncthreads();
    if (lane < T) fft_lds_field<N, E, -1>(lds[k], lane, tw);
    __syncthreads();
    lds_to_cols<N, CW, THREADS>(gp(d.U), x0, lds);
}

// Field-split register forms of k_col_tape and k_adj_col (N >= 128): the column pass
// does not couple the fields, so each field of the CW columns gets its own CW*T threads
// (field 0 the first half of the workgroup), E elements of one column per thread in
// registers, and the halves transform side by side; no staging of the cells through LDS.
template <int N> struct AdjColF {
    static constexpr int E = AdjCol<N>::E, CW = AdjCol<N>::CW, T = N / E, HALF = CW * T, L = LdsLen<N>::value;
    static_assert(HALF % 64 == 0, "each field on whole waves");
};

template <int N>
__global__ __launch_bounds__(2 * AdjColF<N>::HALF) void k_col_tape_f(const InstDesc *__restrict__ descs,
                                                                    const float2 *__restrict__ twg, int step,
                                                                    TapeArgs tp) {
    using C = AdjColF<N>;
    constexpr int E = C::E, CW = C::CW, T = C::T, HALF = C::HALF;
    __shared__ float2 lds[2][C::L * CW];
    __shared__ float2 tw[N];
    const InstDesc d = descs[0];
    const auto U = gp(reinterpret_cast<float2 *>(d.U));    // cell i, field f at 2 i + f
    const auto G = gp(d.G);
    const int tid = threadIdx.x, f = tid / HALF, r = tid - f * HALF, c = r % CW, t = r / CW;
    const int x = xcd_group(blockIdx.x, gridDim.x) * CW + c;
    // the column's loads go out before the twiddles are staged (the latencies overlap
    // instead of adding up across the barrier)
    float2 a[E], gy[E];
#pragma unroll
    for (int j = 0; j < E; j++) a[j] = U[2 * ((size_t)(t + j * T) * N + x) + f];
    for (int j = tid; j < N; j += 2 * HALF) tw[j] = twg[j];
    const float2 gx = G[2 * f * N + x];
#pragma unroll
    for (int j = 0; j < E; j++) gy[j] = G[(2 * f + 1) * N + t + j * T];
    __syncthreads();
    auto sync = [] { __syncthreads(); };
    fft1fe<N, E, +1>(a, &lds[f][c], t, CW, tw, sync);
    if (tp.spec) {
        const auto spec = gp(reinterpret_cast<float2 *>(tp.spec)) + 2 * (size_t)(step - tp.s0) * N * N;
#pragma unroll
        for (int j = 0; j < E; j++) spec[2 * ((size_t)(t + j * T) * N + x) + f] = a[j];
    }
#pragma unroll
    for (int j = 0; j < E; j++) a[j] = c_mul(c_mul(gx, gy[j]), a[j]);
    fft1fe<N, E, -1>(a, &lds[f][c], t, CW, tw, sync);
#pragma unroll
    for (int j = 0; j < E; j++) U[2 * ((size_t)(t + j * T) * N + x) + f] = a[j];
}

template <int N, bool SENS>
__global__ __launch_bounds__(2 * AdjColF<N>::HALF) void k_adj_col_f(const InstDesc *__restrict__ descs,
                                                                   const float2 *__restrict__ twg, int step,
                                                                   TapeArgs tp, const float2 *__restrict__ qxy,
                                                                   float a_p, float a_s, double *__restrict__ gH) {
    using C = AdjColF<N>;
    constexpr int E = C::E, CW = C::CW, T = C::T, HALF = C::HALF;
    __shared__ float2 lds[2][C::L * CW];
    __shared__ float2 tw[N];
    __shared__ double red[2 * HALF / 64];
    const InstDesc d = descs[0];
    const auto U = gp(reinterpret_cast<float2 *>(d.U));
    const auto G = gp(d.G);
    const int tid = threadIdx.x, f = tid / HALF, r = tid - f * HALF, c = r % CW, t = r / CW;
    const int x = xcd_group(blockIdx.x, gridDim.x) * CW + c;
    // the column's loads go out before the twiddles are staged (the latencies overlap
    // instead of adding up across the barrier)
    float2 a[E], gy[E], xs[E];
    float q[E];
#pragma unroll
    for (int j = 0; j < E; j++) a[j] = U[2 * ((size_t)(t + j * T) * N + x) + f];
    for (int j = tid; j < N; j += 2 * HALF) tw[j] = twg[j];
    __syncthreads();
    const float2 gx = G[2 * f * N + x];
#pragma unroll
    for (int j = 0; j < E; j++) gy[j] = G[(2 * f + 1) * N + t + j * T];
    if constexpr (SENS) {
        const auto spec = gp(reinterpret_cast<const float2 *>(tp.spec)) + 2 * (size_t)(step - tp.s0) * N * N;
        const float qx = qxy[x].x;
#pragma unroll
        for (int j = 0; j < E; j++) {
            xs[j] = spec[2 * ((size_t)(t + j * T) * N + x) + f];
            q[j] = qx + qxy[t + j * T].y;
        }
    }
    auto sync = [] { __syncthreads(); };
    fft1fe<N, E, +1>(a, &lds[f][c], t, CW, tw, sync);
    const float af = f ? a_s : a_p;
    double gsum = 0.0;
#pragma unroll
    for (int j = 0; j < E; j++) {
        const float2 h = c_mul(gx, gy[j]);
        if constexpr (SENS) {
            // Re(conj(l) * h * xs * i a q) = -a q Im(conj(l) h xs)
            const float2 u = c_mul(c_mul(cj(a[j]), h), xs[j]);
            gsum += -(double)(af * q[j]) * (double)u.y;
        }
        a[j] = c_mul(cj(h), a[j]);
    }
    fft1fe<N, E, -1>(a, &lds[f][c], t, CW, tw, sync);
#pragma unroll
    for (int j = 0; j < E; j++) U[2 * ((size_t)(t + j * T) * N + x) + f] = a[j];
    if constexpr (SENS) {
        for (int off = 32; off > 0; off >>= 1) gsum += __shfl_down(gsum, off, 64);
        if ((tid & 63) == 0) red[tid >> 6] = gsum;
        __syncthreads();
        if (tid == 0) {
            double sum = 0.0;
            for (int w = 0; w < 2 * HALF / 64; w++) sum += red[w];
            gH[(size_t)step * gridDim.x + blockIdx.x] = sum;
        }
    }
}

// ---------------------------------------------------------------------------
// Adjoint row pass on lambda (descs[0].U is the adjoint state; coef/nl are the forward's).
// ADJ_START at s: RK4_s^T at A''_s, R_bwd.
// ADJ_MID   at s: R_fwd, RK4_s^T at A_s = RK4_{s-1}(A''_{s-1}), RK4_{s-1}^T at A''_{s-1}, R_bwd.
// ADJ_END   at s: R_fwd, RK4_s^T at A_s = tp.a (the segment's checkpoint); leaves the
//                 real-space lambda of the segment start.
// g1[(s*N + row)*5] gets the row's sensitivities of the first RK4 of step s (at A_s),
// g2[...] those of the second (at A''_s).  SENS = false propagates lambda only (the
// field gradient): the sensitivity sums are then never formed.
template <int N, bool SENS>
__global__ __launch_bounds__(64 * AdjWide<N>::W) void k_adj_row(const InstDesc *__restrict__ descs,
                                                               const float2 *__restrict__ twg, int step, int mode,
                                                               TapeArgs tp, double *__restrict__ g1,
                                                               double *__restrict__ g2) {
    constexpr int E = AdjWide<N>::E, W = AdjWide<N>::W, T = N / E, THREADS = 64 * W, L = LdsLen<N>::value;
    static_assert(W >= 2 && T <= 64 && THREADS >= N, "one wavefront per field transform; one cell per thread");
    __shared__ float2 lds[2][L];
    __shared__ float2 tw[N];
    const InstDesc d = descs[0];
    const int tid = threadIdx.x, f = tid >> 6, lane = tid & 63;
    const bool cell = tid < N;
    const size_t n2 = (size_t)N * N, base = (size_t)blockIdx.x * N, idx = base + (cell ? tid : 0);
    const int ls = step - tp.s0;
    // the cell's taped forward state, loaded before the transform hides its latency:
    // START: A''_s;  MID: A''_{s-1};  END: A_s (the checkpoint)
    const float4 tape = mode == ADJ_END ? tp.a[idx] : tp.a2[(size_t)(mode == ADJ_START ? ls : ls - 1) * n2 + idx];
    const float4 c_s = gp(d.coef)[step];
    const float4 c_p = gp(d.coef)[step > 0 ? step - 1 : 0];
    for (int j = tid; j < N; j += THREADS) tw[j] = twg[j];
    if (cell) {
        const float4 v = gp(d.U)[idx];
        lds[0][lds_pad(tid)] = make_float2(v.x, v.y);
        lds[1][lds_pad(tid)] = make_float2(v.z, v.w);
    }
    __syncthreads();
    if (mode != ADJ_START && f < 2 && lane < T) fft_lds_field<N, E, -1>(lds[f], lane, tw);
    __syncthreads();
    AdjAcc acc1, acc2;
    acc_zero(acc1);
    acc_zero(acc2);
    if (cell) {
        const int o = lds_pad(tid);
        float2 P = lds[0][o], S = lds[1][o];
        if (mode == ADJ_START) {
            rk4_adj(f2(tape.x, tape.y), f2(tape.z, tape.w), P, S, c_s, d.nl, acc2);
        } else if (mode == ADJ_END) {
            rk4_adj(f2(tape.x, tape.y), f2(tape.z, tape.w), P, S, c_s, d.nl, acc1);
        } else {
            float2 Ap = f2(tape.x, tape.y), As = f2(tape.z, tape.w);
            Rk4Stages kept;
            rk4_keep(Ap, As, c_p, c_p, c_p, d.nl.lossP, d.nl.lossS, d.nl.tpa, d.nl.h6, kept);   // A_s
            rk4_adj(Ap, As, P, S, c_s, d.nl, acc1);
            rk4_adj_kept(f2(tape.x, tape.y), f2(tape.z, tape.w), kept, P, S, c_p, d.nl, acc2);
        }
        lds[0][o] = P;
        lds[1][o] = S;
    }
    if constexpr (SENS) {
        if (mode != ADJ_START) block_store<W>(acc1, g1 + ((size_t)step * N + blockIdx.x) * 5);
        if (mode == ADJ_START) block_store<W>(acc2, g2 + ((size_t)step * N + blockIdx.x) * 5);
        else if (mode == ADJ_MID) {
            __syncthreads();                  // block_store's LDS slots are reused
            block_store<W>(acc2, g2 + ((size_t)(step - 1) * N + blockIdx.x) * 5);
        }
    }
    __syncthreads();
    if (mode != ADJ_END && f < 2 && lane < T) fft_lds_field<N, E, +1>(lds[f], lane, tw);
    __syncthreads();
    if (cell) {
        const float2 P = lds[0][lds_pad(tid)], S = lds[1][lds_pad(tid)];
        gp(d.U)[idx] = make_float4(P.x, P.y, S.x, S.y);
    }
}

// Adjoint column pass: C_bwd, x conj(Gx Gy), C_fwd.  gH[step*nblk + blk] receives
// sum Re(conj(lambda_hat) (Gx Gy) x_hat * i (a_p, a_s) (qx + qy)) over the block's cells,
// the sensitivity of J to a temperature change of the propagator (d log H/dT = i a q);
// not with SENS = false.
template <int N, bool SENS>
__global__ __launch_bounds__(128 * AdjCol<N>::CW) void k_adj_col(const InstDesc *__restrict__ descs,
                                                                const float2 *__restrict__ twg, int step, TapeArgs tp,
                                                                const float2 *__restrict__ qxy, float a_p, float a_s,
                                                                double *__restrict__ gH) {
    constexpr int E = AdjCol<N>::E, CW = AdjCol<N>::CW, T = N / E, THREADS = 128 * CW, L = LdsLen<N>::value;
    static_assert(T <= 64, "one wavefront per field transform");
    __shared__ float2 lds[2 * CW][L];
    __shared__ float2 tw[N];
    __shared__ double red[THREADS / 64];
    constexpr int CPT = (CW * N + THREADS - 1) / THREADS;   // cells per thread (guarded: 64 x 2 < 256)
    const InstDesc d = descs[0];
    const auto G = gp(d.G);
    const int tid = threadIdx.x, k = tid >> 6, lane = tid & 63, x0 = xcd_group(blockIdx.x, gridDim.x) * CW;
    for (int j = tid; j < N; j += THREADS) tw[j] = twg[j];
    cols_to_lds<N, CW, THREADS>(gp(d.U), x0, lds);
    // the thread's cells' propagator factors, frequencies and taped spectra, loaded
    // before the transform
    const size_t soff = (size_t)(step - tp.s0) * N * N;
    float2 gxp[CPT], gyp[CPT], gxs[CPT], gys[CPT];
    float qc[CPT];
    float4 xsp[CPT];
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int i = tid + q * THREADS, ii = i < CW * N ? i : 0, y = ii / CW, x = x0 + ii % CW;
        gxp[q] = G[x];
        gyp[q] = G[N + y];
        gxs[q] = G[2 * N + x];
        gys[q] = G[3 * N + y];
        qc[q] = qxy[x].x + qxy[y].y;
        xsp[q] = gp(tp.spec)[soff + (size_t)y * N + x];
    }
    __syncthreads();
    if (lane < T) fft_lds_field<N, E, +1>(lds[k], lane, tw);
    __syncthreads();
    double gsum = 0.0;
#pragma unroll
    for (int qq = 0; qq < CPT; qq++) {
        const int i = tid + qq * THREADS, y = i / CW, c = i % CW;
        if (i >= CW * N) break;
        const float q = qc[qq];
        const float2 hp = c_mul(gxp[qq], gyp[qq]), hs = c_mul(gxs[qq], gys[qq]);
        const float4 xs = xsp[qq];
        const float2 lp = lds[c][lds_pad(y)], lsg = lds[CW + c][lds_pad(y)];
        // Re(conj(l) * h * xs * i a q) = -a q Im(conj(l) h xs)
        const float2 tp_ = c_mul(c_mul(cj(lp), hp), f2(xs.x, xs.y));
        const float2 ts_ = c_mul(c_mul(cj(lsg), hs), f2(xs.z, xs.w));
        gsum += -(double)(a_p * q) * (double)tp_.y - (double)(a_s * q) * (double)ts_.y;
        lds[c][lds_pad(y)] = c_mul(cj(hp), lp);
        lds[CW + c][lds_pad(y)] = c_mul(cj(hs), lsg);
    }
    __syncthreads();
    if (lane < T) fft_lds_field<N, E, -1>(lds[k], lane, tw);
    __syncthreads();
    lds_to_cols<N, CW, THREADS>(gp(d.U), x0, lds);
    if constexpr (SENS) {
        for (int off = 32; off > 0; off >>= 1) gsum += __shfl_down(gsum, off, 64);
        if ((tid & 63) == 0) red[tid >> 6] = gsum;
        __syncthreads();
        if (tid == 0) {
            double sum = 0.0;
            for (int w = 0; w < THREADS / 64; w++) sum += red[w];
            gH[(size_t)step * gridDim.x + blockIdx.x] = sum;
        }
    }
}

// lambda at the output plane: dJ/dA_out with J = P_SH / P_in, P_SH = cs sum |S|^2:
// lambda_P = 0, lambda_S = 2 cs / P_in * S.
__global__ __launch_bounds__(256) void k_adj_seed(const float4 *__restrict__ U, float4 *__restrict__ L, int64_t n,
                                                  float scale) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 v = U[i];
    L[i] = make_float4(0.0f, 0.0f, scale * v.z, scale * v.w);
}

__global__ __launch_bounds__(256) void k_field_grad(const float4 *__restrict__ lam, const float4 *__restrict__ U0,
                                                    const double2 *__restrict__ red, double cs, double cp, double pin,
                                                    const double2 *__restrict__ pdf, const double2 *__restrict__ pdn,
                                                    float2 *__restrict__ grad, double2 *__restrict__ part, int64_t n) {
    const double J = cs * red->y / pin, direct = -J / pin * 2.0 * cp;
    double sf = 0.0, sn = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float4 l = lam[i], u = U0[i];
        const double gr = (double)l.x + direct * (double)u.x, gi = (double)l.y + direct * (double)u.y;
        grad[i] = make_float2((float)gr, (float)gi);
        const double2 f = pdf[i], m = pdn[i];
        sf += gr * f.x + gi * f.y;
        sn += gr * m.x + gi * m.y;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sf += __shfl_down(sf, off, 64);
        sn += __shfl_down(sn, off, 64);
    }
    __shared__ double2 w[4];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = make_double2(sf, sn);
    __syncthreads();
    if (threadIdx.x == 0)
        part[blockIdx.x] = make_double2(w[0].x + w[1].x + w[2].x + w[3].x, w[0].y + w[1].y + w[2].y + w[3].y);
}

// dst[i] = src[i] (field copies inside captured graphs: kernels, not copy nodes).
__global__ __launch_bounds__(256) void k_copy_cells(float4 *__restrict__ dst, const float4 *__restrict__ src, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}

// Per step s (one workgroup each): out[s*6 + q] = sum over the row-pass partials of
// g1 + g2 (q < 5: c1.re, c1.im, c2.re, c2.im, tpa) and out[s*6 + 5] = sum of the
// column-pass partials gH, in a fixed order.
__global__ __launch_bounds__(64) void k_adj_reduce(const double *__restrict__ g1, const double *__restrict__ g2,
                                                   int rparts, const double *__restrict__ gH, int cparts,
                                                   double *__restrict__ out) {
    const int s = blockIdx.x, lane = threadIdx.x;
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = lane; b < rparts; b += 64) {
        const size_t o = ((size_t)s * rparts + b) * 5;
#pragma unroll
        for (int q = 0; q < 5; q++) v[q] += g1[o + q] + g2[o + q];
    }
    for (int b = lane; b < cparts; b += 64) v[5] += gH[(size_t)s * cparts + b];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        double x = v[q];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) out[(size_t)s * 6 + q] = x;
    }
}

// ---------------------------------------------------------------------------
template <int N> static int col_blocks() { return N / AdjCol<N>::CW; }

#define SHG_DISPATCH(n, CALL)                \
    switch (n) {                             \
        case 64: { constexpr int NN = 64; CALL; } break;   \
        case 128: { constexpr int NN = 128; CALL; } break; \
        case 256: { constexpr int NN = 256; CALL; } break; \
        case 512: { constexpr int NN = 512; CALL; } break; \
        default: return hipErrorInvalidValue;            \
    }

int adj_row_parts(int n) { return n == 64 || n == 128 || n == 256 || n == 512 ? n : 0; }
int adj_col_parts(int n) {
    switch (n) {
        case 64: return col_blocks<64>();
        case 128: return col_blocks<128>();
        case 256: return col_blocks<256>();
        case 512: return col_blocks<512>();
        default: return 0;
    }
}

hipError_t launch_row_tape(int n, const InstDesc *d, const float2 *tw, int step, int mode, const TapeArgs &tp,
                           hipStream_t st) {
    if (mode != ROW_START && mode != ROW_MID && mode != ROW_END) return hipErrorInvalidValue;
    SHG_DISPATCH(n, hipLaunchKernelGGL(k_row_tape<NN>, dim3(NN), dim3(64 * AdjWide<NN>::W), 0, st, d, tw, step, mode,
                                       tp));
    return hipGetLastError();
}

// The column kernels of the adjoint: the field-split register forms from 128 on (their
// grid is the same N / CW blocks, so the gH partials keep their count), the LDS forms at 64.
#define SHG_DISPATCH_F(n, CALL)                             \
    switch (n) {                                            \
        case 128: { constexpr int NN = 128; CALL; } break; \
        case 256: { constexpr int NN = 256; CALL; } break; \
        case 512: { constexpr int NN = 512; CALL; } break; \
        default: return hipErrorInvalidValue;            \
    }

hipError_t launch_col_tape(int n, const InstDesc *d, const float2 *tw, int step, const TapeArgs &tp, hipStream_t st) {
    if (n == 64) {
        hipLaunchKernelGGL(k_col_tape<64>, dim3(col_blocks<64>()), dim3(128 * AdjCol<64>::CW), 0, st, d, tw, step, tp);
        return hipGetLastError();
    }
    SHG_DISPATCH_F(n, hipLaunchKernelGGL(k_col_tape_f<NN>, dim3(col_blocks<NN>()), dim3(2 * AdjColF<NN>::HALF), 0, st,
                                       d, tw, step, tp));
    return hipGetLastError();
}

hipError_t launch_adj_row(int n, const InstDesc *d, const float2 *tw, int step, int mode, const TapeArgs &tp,
                          double *g1, double *g2, hipStream_t st, bool sens) {
    if (mode != ADJ_START && mode != ADJ_MID && mode != ADJ_END) return hipErrorInvalidValue;
    if (sens) {
        SHG_DISPATCH(n, hipLaunchKernelGGL((k_adj_row<NN, true>), dim3(NN), dim3(64 * AdjWide<NN>::W), 0, st, d, tw,
                                           step, mode, tp, g1, g2));
    } else {
        SHG_DISPATCH(n, hipLaunchKernelGGL((k_adj_row<NN, false>), dim3(NN), dim3(64 * AdjWide<NN>::W), 0, st, d, tw,
                                           step, mode, tp, g1, g2));
    }
    return hipGetLastError();
}

hipError_t launch_adj_col(int n, const InstDesc *d, const float2 *tw, int step, const TapeArgs &tp, const float2 *qxy,
                          float a_p, float a_s, double *gH, hipStream_t st, bool sens) {
    if (n == 64) {
        if (sens)
            hipLaunchKernelGGL((k_adj_col<64, true>), dim3(col_blocks<64>()), dim3(128 * AdjCol<64>::CW), 0, st, d, tw,
                               step, tp, qxy, a_p, a_s, gH);
        else
            hipLaunchKernelGGL((k_adj_col<64, false>), dim3(col_blocks<64>()), dim3(128 * AdjCol<64>::CW), 0, st, d, tw,
                               step, tp, qxy, a_p, a_s, gH);
        return hipGetLastError();
    }
    if (sens) {
        SHG_DISPATCH_F(n, hipLaunchKernelGGL((k_adj_col_f<NN, true>), dim3(col_blocks<NN>()), dim3(2 * AdjColF<NN>::HALF),
                                           0, st, d, tw, step, tp, qxy, a_p, a_s, gH));
    } else {
        SHG_DISPATCH_F(n, hipLaunchKernelGGL((k_adj_col_f<NN, false>), dim3(col_blocks<NN>()), dim3(2 * AdjColF<NN>::HALF),
                                           0, st, d, tw, step, tp, qxy, a_p, a_s, gH));
    }
    return hipGetLastError();
}

hipError_t launch_adj_reduce(const double *g1, const double *g2, int rparts, const double *gH, int cparts, int steps,
                             double *out, hipStream_t st) {
    if (steps <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_reduce, dim3(steps), dim3(64), 0, st, g1, g2, rparts, gH, cparts, out);
    return hipGetLastError();
}

hipError_t launch_copy_cells(float4 *dst, const float4 *src, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(k_copy_cells, dim3((unsigned)blocks), dim3(256), 0, st, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_field_grad(const float4 *lam, const float4 *U0, const double2 *red, double cs, double cp, double pin,
                             const double2 *pdf, const double2 *pdn, float2 *grad, double2 *part, int nparts,
                             int64_t n, hipStream_t st) {
    if (n <= 0 || nparts <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_field_grad, dim3(nparts), dim3(256), 0, st, lam, U0, red, cs, cp, pin, pdf, pdn, grad, part, n);
    return hipGetLastError();
}

hipError_t launch_adj_seed(const float4 *U, float4 *L, int64_t n, float scale, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_seed, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, U, L, n, scale);
    return hipGetLastError();
}

}  // namespace shg
