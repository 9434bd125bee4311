// thermal_adjoint_kernels.hip — discrete adjoint of the THERMAL z-march (src/headers/
// Solver.h:44-63,193-233,245-276) with the temperature field held fixed.
//
// The THERMAL nonlinear step differs from the plain one per cell: n_p, n_s come from the
// cell's heat voxel, the phase phi_k = DKint_s * zfac_k / (s + 1) differs between the RK4
// stage groups (1 | 2-3 | 4), and DKint_s is the running sum of dk over the executed
// steps.  The transpose of each RK4 keeps the three stage groups' coefficient
// sensitivities apart and turns them, per cell, into
//   dJ/dkappa_p, dJ/dkappa_s, dJ/dtpa  -> dJ/dn_p, dJ/dn_s of the step's voxel,
//   dJ/dphi_k                          -> dJ/dDKint_s.
// Since DKint_s = sum_{s' <= s} dk(slice(s')), the dk of step s' sees G_s' = sum over
// s >= s' of dJ/dDKint_s: the adjoint runs steps last first, so G is a running sum per
// cell and every step adds G (d dk / d n) to its voxel.
#include <hip/hip_runtime.h>

#include "../../include/shg.h"
#include "adjoint_cell.h"
#include "adjoint_kernels.h"
#include "fft_lds.h"
#include "pass_cfg.h"
#include "physics.h"
#include "thermal_adjoint_kernels.h"
#include "thermal_cell.h"

namespace shg {

// The three stage-group phases of one THERMAL RK4 (rk4_thermal): sin and cos of
// phi_g = (float)(DKint * zr_g) (phase_sincos3) and the coefficients (c1, c2) = (i kp e^{-i phi}, i ks e^{+i phi}).
struct ThermPhase {
    float4 cf[3];
    float sn[3], cs[3];
};

__device__ __forceinline__ void therm_phase(const CellPhys &c, float dkint, const ThermStep &st, ThermPhase &ph) {
    phase_sincos3(dkint, st, ph.sn, ph.cs);
#pragma unroll
    for (int g = 0; g < 3; g++)
        ph.cf[g] = make_float4(c.kp * ph.sn[g], c.kp * ph.cs[g], -c.ks * ph.sn[g], c.ks * ph.cs[g]);
}

// RK4^T of one THERMAL RK4 application at forward input (P, S) with stage inputs `st`:
// updates the adjoint (lP, lS) and returns the cell's sensitivities to DKint (dDK) and,
// through the coupling constants and the TPA coefficient, to its indices (dnp, dns).
__device__ __forceinline__ void therm_rk4_adj(float2 P, float2 S, const Rk4Stages &stg, float2 &lP, float2 &lS,
                                              const CellPhys &c, const ThermPhase &ph, const ThermStep &st,
                                              const ThermConst &k, double &dDK, double &dnp, double &dns) {
    AdjAcc acc[3];
#pragma unroll
    for (int g = 0; g < 3; g++) acc_zero(acc[g]);
    rk4_adj_rev<false>(P, S, stg, lP, lS, ph.cf[0], ph.cf[1], ph.cf[2], k.lossP, k.lossS, c.tpa, k.h6, acc[0], acc[1],
                       acc[2]);
    double dkp = 0.0, dks = 0.0, dphi_dk = 0.0;
#pragma unroll
    for (int g = 0; g < 3; g++) {
        // c1 = i kp e^{-i phi}: dc1/dkp = (sin, cos), dc1/dphi = -i c1;
        // c2 = i ks e^{+i phi}: dc2/dks = (-sin, cos), dc2/dphi = +i c2.
        const float4 cf = ph.cf[g];
        dkp += acc[g].c1r * ph.sn[g] + acc[g].c1i * ph.cs[g];
        dks += -acc[g].c2r * ph.sn[g] + acc[g].c2i * ph.cs[g];
        const double dphi = acc[g].c1r * cf.y - acc[g].c1i * cf.x - acc[g].c2r * cf.w + acc[g].c2i * cf.z;
        dphi_dk += dphi * st.zr[g];
    }
    dDK += dphi_dk;
    // kp = pi dQ / (n_p lp), ks = pi dQ / (2 n_s ls), tpa = tpa_base n_s / 2
    dnp += dkp * (-(double)c.kp / (double)c.np);
    dns += dks * (-(double)c.ks / (double)c.ns) + (acc[0].tpa + acc[1].tpa + acc[2].tpa) * 0.5 * k.tpa_base;
}

// The same, evaluating the phases and stages itself.
__device__ __forceinline__ void therm_rk4_adj(float2 P, float2 S, float2 &lP, float2 &lS, const CellPhys &c,
                                              float dkint, const ThermStep &st, const ThermConst &k, double &dDK,
                                              double &dnp, double &dns) {
    ThermPhase ph;
    therm_phase(c, dkint, st, ph);
    Rk4Stages stg;
    rk4_stages(P, S, ph.cf[0], ph.cf[1], k.lossP, k.lossS, c.tpa, stg);
    therm_rk4_adj(P, S, stg, lP, lS, c, ph, st, k, dDK, dnp, dns);
}

// SENS = false propagates lambda only (the field gradient): the sensitivities to DKint
// and to the indices (G, Dp, Ap, As) are neither formed nor stored.
template <int N, bool SENS>
__global__ __launch_bounds__(64 * AdjWide<N>::W) void k_therm_adj_row(const InstDesc *__restrict__ descs,
                                                                     const float2 *__restrict__ twg, int mode,
                                                                     ThermAdjArgs a) {
    constexpr int E = AdjWide<N>::E, W = AdjWide<N>::W, T = N / E, THREADS = 64 * W, L = LdsLen<N>::value;
    static_assert(W >= 2 && T <= 64 && THREADS >= N, "one wavefront per field transform; one cell per thread");
    __shared__ float2 lds[2][L];
    __shared__ float2 tw[N];
    const auto U = gp(descs[0].U);
    const int tid = threadIdx.x, f = tid >> 6, lane = tid & 63;
    const bool cell = tid < N;
    const size_t n2 = (size_t)N * N, base = (size_t)blockIdx.x * N, idx = base + (cell ? tid : 0);
    const int ls = a.cur.step - a.s0;
    // The row is staged first, then the cell's inputs are loaded, from addresses chosen
    // per mode (uniform) with no per-lane branches or copies of loaded values: they are
    // then in flight behind the transform (a copy of a loaded value, or a wait for the row
    // issued after them, would make the wave wait for them before the row's loads).
    for (int j = tid; j < N; j += THREADS) tw[j] = twg[j];
    if (cell) {
        const float4 v = U[idx];
        lds[0][lds_pad(tid)] = make_float2(v.x, v.y);
        lds[1][lds_pad(tid)] = make_float2(v.z, v.w);
    }
    const size_t lp1 = (size_t)(ls > 0 ? ls - 1 : 0);
    const float4 *Ab = mode == ADJ_START ? a.a2 + (size_t)ls * n2 : mode == ADJ_END ? a.a0 : a.a2 + lp1 * n2;
    const float4 A = Ab[idx];
    const float2 nn_c = a.NT[(size_t)a.cur.slice * n2 + idx];
    const float dk_c = a.dk[(size_t)ls * n2 + idx];
    const float dk_p = a.dk[lp1 * n2 + idx];                  // used in ADJ_MID only
    const float2 nn_p = mode == ADJ_MID ? a.NT[(size_t)a.prev.slice * n2 + idx] : make_float2(0.f, 0.f);
    double G = SENS ? a.G[idx] : 0.0, Dp = SENS ? a.Dp[idx] : 0.0;
    __syncthreads();
    if (mode != ADJ_START && f < 2 && lane < T) fft_lds_field<N, E, -1>(lds[f], lane, tw);
    __syncthreads();
    if (cell) {
        const ThermConst &k = a.k;
        // d dk / d n_p = 4 pi / lp, d dk / d n_s = -2 pi / ls (src/headers/PhaseMatching.h:47)
        const double wp = 4.0 * (double)k.pi / (double)k.lp, ws = -2.0 * (double)k.pi / (double)k.ls;
        const int o = lds_pad(tid);
        float2 P = lds[0][o], S = lds[1][o];
        double *Apc = a.Ap + (size_t)a.cur.slice * n2 + idx, *Asc = a.As + (size_t)a.cur.slice * n2 + idx;
        const CellPhys cc = cell_phys(k, nn_c);
        if (mode == ADJ_START) {
            double dDK = 0.0, dnp = 0.0, dns = 0.0;
            therm_rk4_adj(f2(A.x, A.y), f2(A.z, A.w), P, S, cc, dk_c, a.cur, k, dDK, dnp, dns);
            if constexpr (SENS) {
                *Apc += dnp;
                *Asc += dns;
            }
            Dp = dDK;
        } else {
            float2 Ap_ = f2(A.x, A.y), As_ = f2(A.z, A.w);
            CellPhys cp{};
            ThermPhase php;
            Rk4Stages kept;
            if (mode == ADJ_MID) {
                // A_s from A''_{s-1} (rk4_thermal), keeping the phases and stages for RK4_{s-1}^T
                cp = cell_phys(k, nn_p);
                therm_phase(cp, dk_p, a.prev, php);
                rk4_keep(Ap_, As_, php.cf[0], php.cf[1], php.cf[2], k.lossP, k.lossS, cp.tpa, k.h6, kept);
            }
            double dDK = 0.0, dnp = 0.0, dns = 0.0;
            therm_rk4_adj(Ap_, As_, P, S, cc, dk_c, a.cur, k, dDK, dnp, dns);
            G += Dp + dDK;                                          // dJ/dDKint_s complete
            if constexpr (SENS) {
                *Apc += dnp + G * wp;
                *Asc += dns + G * ws;
            }
            Dp = 0.0;
            if (mode == ADJ_MID) {
                double dDK2 = 0.0, dnp2 = 0.0, dns2 = 0.0;
                therm_rk4_adj(f2(A.x, A.y), f2(A.z, A.w), kept, P, S, cp, php, a.prev, k, dDK2, dnp2, dns2);
                if constexpr (SENS) {
                    a.Ap[(size_t)a.prev.slice * n2 + idx] += dnp2;
                    a.As[(size_t)a.prev.slice * n2 + idx] += dns2;
                }
                Dp = dDK2;
            }
        }
        lds[0][o] = P;
        lds[1][o] = S;
        if constexpr (SENS) {
            a.G[idx] = G;
            a.Dp[idx] = Dp;
        }
    }
    __syncthreads();
    if (mode != ADJ_END && f < 2 && lane < T) fft_lds_field<N, E, +1>(lds[f], lane, tw);
    __syncthreads();
    if (cell) {
        const float2 P = lds[0][lds_pad(tid)], S = lds[1][lds_pad(tid)];
        U[idx] = make_float4(P.x, P.y, S.x, S.y);
    }
}

__global__ __launch_bounds__(256) void k_fill_zero(double *__restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = 0.0;
}

__global__ __launch_bounds__(256) void k_therm_dtmap(double *__restrict__ Ap, const double *__restrict__ As,
                                                     const float *__restrict__ T, int64_t n3, int crystal, double lp,
                                                     double ls, double *__restrict__ part) {
    const double h = 1e-3;
    double sum = 0.0;
    for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < n3; v += (int64_t)gridDim.x * 256) {
        const double t = (double)T[v];
        const double dnp = (phys::index_d(crystal, lp, t + h) - phys::index_d(crystal, lp, t - h)) / (2 * h);
        const double dns = (phys::index_d(crystal, ls, t + h) - phys::index_d(crystal, ls, t - h)) / (2 * h);
        const double m = Ap[v] * dnp + As[v] * dns;
        Ap[v] = m;
        sum += m;
    }
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off, 64);
    __shared__ double w[4];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = w[0] + w[1] + w[2] + w[3];
}

// ---------------------------------------------------------------------------
template <int N>
static void therm_adj_row_launch(const InstDesc *d, const float2 *tw, int mode, const ThermAdjArgs &a, hipStream_t st,
                                 bool sens) {
    if (sens)
        hipLaunchKernelGGL((k_therm_adj_row<N, true>), dim3(N), dim3(64 * AdjWide<N>::W), 0, st, d, tw, mode, a);
    else
        hipLaunchKernelGGL((k_therm_adj_row<N, false>), dim3(N), dim3(64 * AdjWide<N>::W), 0, st, d, tw, mode, a);
}

hipError_t launch_therm_adj_row(int n, const InstDesc *d, const float2 *tw, int mode, const ThermAdjArgs &a,
                                hipStream_t st, bool sens) {
    if (mode != ADJ_START && mode != ADJ_MID && mode != ADJ_END) return hipErrorInvalidValue;
    switch (n) {
        case 64: therm_adj_row_launch<64>(d, tw, mode, a, st, sens); break;
        case 128: therm_adj_row_launch<128>(d, tw, mode, a, st, sens); break;
        case 256: therm_adj_row_launch<256>(d, tw, mode, a, st, sens); break;
        case 512: therm_adj_row_launch<512>(d, tw, mode, a, st, sens); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fill_zero(double *p, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_therm_dtmap(double *Ap, const double *As, const float *T, int64_t n3, int crystal, double lp,
                              double ls, double *part, int nparts, hipStream_t st) {
    if (n3 <= 0 || nparts <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_therm_dtmap, dim3(nparts), dim3(256), 0, st, Ap, As, T, n3, crystal, lp, ls, part);
    return hipGetLastError();
}

}  // namespace shg
