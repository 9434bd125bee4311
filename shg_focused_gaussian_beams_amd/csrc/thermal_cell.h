// thermal_cell.h — per-cell physics of the THERMAL nonlinear step, shared by the THERMAL
// row pass and its adjoint: the refractive index in the reference's typing, the cell's
// coupling constants and phase mismatch, the phase-shifted RK4 and the heat source.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/shg.h"
#include "nl_step.h"
#include "sincos_phase.h"
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

// a / d for a divisor d known in advance, with r = RN(1/d): q = RN(a r) is within an ulp
// of a/d, the FMA gives the exact remainder a - q d, and one more FMA step returns the
// correctly rounded quotient (Markstein's theorem; barring overflow and underflow, far
// from the values here).  Three operations instead of the IEEE division's sequence,
// bit-identical to it.
__device__ __forceinline__ float div_const(float a, float d, float r) {
    const float q = a * r;
    const float e = __builtin_fmaf(-q, d, a);
    return __builtin_fmaf(e, r, q);
}

struct CellPhys {
    float np, ns, kp, ks, dk, tpa;
};

// The coupling magnitudes, the phase mismatch (src/headers/PhaseMatching.h:47) and the
// TPA coefficient from the cell's indices, with the reference's float operations (the
// divisions by the wavelengths through their reciprocals, bit-identical).  The row pass
// evaluates them per cell from the march's {n_p, n_s} table.
__device__ __forceinline__ CellPhys cell_phys(const ThermConst &k, float2 nn) {
#pragma clang fp contract(off)
    CellPhys c;
    c.np = nn.x;
    c.ns = nn.y;
    c.kp = ((1.0f * k.pi) * k.dq) / (c.np * k.lp);
    c.ks = ((0.5f * k.pi) * k.dq) / (c.ns * k.ls);
    c.dk = (2.0f * k.pi) * ((div_const(2.0f * c.np, k.lp, k.rlp) - div_const(c.ns, k.ls, k.rls)) + k.inv_period);
    c.tpa = (float)(0.5 * (k.tpa_base * (double)c.ns));
    return c;
}

// The coefficients of a THERMAL RK4 stage group are (c1, c2) = (i kp e^{-i phi},
// i ks e^{+i phi}) with phi = (float)(DKint * zfac / (s+1)); zr = zfac / (s+1) is
// precomputed in FP64 (the float phase is the same but for rare one-ulp double roundings).
// sin and cos of the three stage-group phases phi_g = (float)(DKint * zr_g), g = 0, 1, 2,
// each by sincos_phase (within 2 ulp of the correctly rounded values, branch-free).
__device__ __forceinline__ void phase_sincos3(float dkint, const ThermStep &st, float sn[3], float cs[3]) {
#pragma unroll
    for (int g = 0; g < 3; g++) sincos_phase((float)((double)dkint * st.zr[g]), sn[g], cs[g]);
}

// THERMAL solverRK4 (src/headers/Solver.h:197-213): stage increments 0, .25, .25, .5.
__device__ __forceinline__ void rk4_thermal(float2 &P, float2 &S, const CellPhys &c, float dkint, const ThermStep &st,
                                            const ThermConst &k) {
    float sn[3], cs[3];
    phase_sincos3(dkint, st, sn, cs);
    const float4 ca = make_float4(c.kp * sn[0], c.kp * cs[0], -c.ks * sn[0], c.ks * cs[0]);
    const float4 cb = make_float4(c.kp * sn[1], c.kp * cs[1], -c.ks * sn[1], c.ks * cs[1]);
    const float4 cc = make_float4(c.kp * sn[2], c.kp * cs[2], -c.ks * sn[2], c.ks * cs[2]);
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
