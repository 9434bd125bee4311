// thermal_cell.h — per-cell physics of the THERMAL nonlinear step, shared by the THERMAL
// row pass and its adjoint: the refractive index in the reference's typing, the cell's
// coupling constants and phase mismatch, the phase-shifted RK4 and the heat source.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/shg.h"
#include "nl_step.h"
#include "thermal_kernels.h"

namespace shg {

// Refractive index with the reference's typing (src/headers/Crystal.h:55-71).
__device__ __forceinline__ float index_dev(int crystal, float L, float T) {
#pragma clang fp contract(off)
    if (crystal == SHG_CRYSTAL_PPLN) {
        const double t = (double)T, l2 = (double)L * (double)L;
        const double f = (t - 24.5) * (t + 570.82);
        const double p = 0.2020 + 6.113e-8 * f;
        return (float)sqrt(5.756 + 2.860e-6 * f + (0.0983 + 4.700e-8 * f) / (l2 - p * p) +
                           (189.32 + 1.516e-4 * f) / (l2 - 12.52 * 12.52) - 1.32e-2 * l2);
    }
    const float A = 4.502483f, B = 0.007294f, C = 0.185087f, D = -0.02357f;
    const float E = 0.073423f, F = 0.199595f, G = 0.001f, H = 7.99724f;
    const double tk = (double)T + 273.15;
    const float b = (float)(3.483933e-8 * (tk * tk));
    const float c = (float)(1.607839e-8 * (tk * tk));
    const double l2 = (double)L * (double)L;
    const float bb = B + b, cc = C + c;
    const double s = (double)A + (double)bb / (l2 - (double)cc * (double)cc) + (double)E / (l2 - (double)F * (double)F) +
                     (double)G / (l2 - (double)H * (double)H) + (double)D * l2;
    return (float)sqrt(s);
}

struct CellPhys {
    float np, ns, kp, ks, dk, tpa;
};

// The coupling magnitudes, the phase mismatch (src/headers/PhaseMatching.h:47) and the
// TPA coefficient from the cell's indices, with the reference's float operations.
__device__ __forceinline__ CellPhys cell_phys(const ThermConst &k, float2 nn) {
#pragma clang fp contract(off)
    CellPhys c;
    c.np = nn.x;
    c.ns = nn.y;
    c.kp = ((1.0f * k.pi) * k.dq) / (c.np * k.lp);
    c.ks = ((0.5f * k.pi) * k.dq) / (c.ns * k.ls);
    c.dk = (2.0f * k.pi) * (((2.0f * c.np) / k.lp - c.ns / k.ls) + k.inv_period);
    c.tpa = (float)(0.5 * (k.tpa_base * (double)c.ns));
    return c;
}

// The same values read from the per-march table: nn = {n_p, n_s}, cf = {kp, ks, dk, tpa}.
__device__ __forceinline__ CellPhys cell_from_table(float2 nn, float4 cf) {
    CellPhys c;
    c.np = nn.x;
    c.ns = nn.y;
    c.kp = cf.x;
    c.ks = cf.y;
    c.dk = cf.z;
    c.tpa = cf.w;
    return c;
}

// (c1, c2) = (i kp e^{-i phi}, i ks e^{+i phi}), phi = (float)(DKint * zfac / (s+1)) with
// zr = zfac / (s+1) precomputed in FP64 (the float phase is the same but for rare
// one-ulp double roundings).
__device__ __forceinline__ float4 phase_coef(const CellPhys &c, float dkint, double zr) {
    const float phi = (float)((double)dkint * zr);
    float sn, cs;
    sincosf(phi, &sn, &cs);
    return make_float4(c.kp * sn, c.kp * cs, -c.ks * sn, c.ks * cs);
}

// THERMAL solverRK4 (src/headers/Solver.h:197-213): stage increments 0, .25, .25, .5.
__device__ __forceinline__ void rk4_thermal(float2 &P, float2 &S, const CellPhys &c, float dkint, const ThermStep &st,
                                            const ThermConst &k) {
    const float4 ca = phase_coef(c, dkint, st.zr[0]);
    const float4 cb = phase_coef(c, dkint, st.zr[1]);
    const float4 cc = phase_coef(c, dkint, st.zr[2]);
    rk4_cell(P, S, ca, cb, cc, k.lossP, k.lossS, c.tpa, k.h6);
}

// setQ (src/headers/Tfield.h:277-290): Q = ap Ip + as Is + bs Is^2 with I_j = c eps0 n_j |A_j|^2 / 2.
__device__ __forceinline__ float heat_source(const ThermConst &k, const CellPhys &c, float2 P, float2 S) {
#pragma clang fp contract(off)
    const float p2 = P.x * P.x + P.y * P.y, s2 = S.x * S.x + S.y * S.y;
    const double ip = k.q_p * (double)c.np * (double)p2;
    const double is = k.q_s * (double)c.ns * (double)s2;
    const float i2 = (float)(k.q_i * (double)c.ns * (double)s2);
    const float tpa = k.beta_s * (i2 * i2);
    return (float)(ip + is + (double)tpa);
}

}  // namespace shg
