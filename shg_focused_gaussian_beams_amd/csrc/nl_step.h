// nl_step.h — the nonlinear (coupled-wave) RK4 step at one cell, shared by the plain
// and THERMAL row passes.
#pragma once
#include <hip/hip_runtime.h>

#include "ssfm_kernels.h"

namespace shg {

// Right-hand side of dAdz (src/headers/Solver.h:44-83) with the phase folded into
//   c = (c1, c2),  c1 = i kappa_p e^{-i phi},  c2 = i kappa_s e^{+i phi}:
//   dP = c1 * S conj(P) - lossP * P
//   dS = c2 * P^2 - (lossS + tpa |S|^2) * S
__device__ __forceinline__ void nl_rhs(float2 P, float2 S, float4 c, float lossP, float lossS, float tpa, float2 &dP,
                                       float2 &dS) {
    const float ur = S.x * P.x + S.y * P.y, ui = S.y * P.x - S.x * P.y;
    dP.x = c.x * ur - c.y * ui - lossP * P.x;
    dP.y = c.x * ui + c.y * ur - lossP * P.y;
    const float vr = P.x * P.x - P.y * P.y, vi = 2.0f * P.x * P.y;
    const float g = lossS + tpa * (S.x * S.x + S.y * S.y);
    dS.x = c.z * vr - c.w * vi - g * S.x;
    dS.y = c.z * vi + c.w * vr - g * S.y;
}

// solverRK4 (src/headers/Solver.h:193-233): stages A + {.5,.5,1} k (not scaled by the
// step), final A += h/6 (k1 + 2k2 + 2k3 + k4).  ca/cb/cc are the coefficients of stage 1,
// stages 2-3 and stage 4 (equal in the plain build, phase-shifted in THERMAL).
__device__ __forceinline__ void rk4_cell(float2 &P, float2 &S, float4 ca, float4 cb, float4 cc, float lossP,
                                         float lossS, float tpa, float h6) {
    float2 k1p, k1s, k2p, k2s, k3p, k3s, k4p, k4s;
    nl_rhs(P, S, ca, lossP, lossS, tpa, k1p, k1s);
    nl_rhs(make_float2(P.x + 0.5f * k1p.x, P.y + 0.5f * k1p.y), make_float2(S.x + 0.5f * k1s.x, S.y + 0.5f * k1s.y), cb,
           lossP, lossS, tpa, k2p, k2s);
    nl_rhs(make_float2(P.x + 0.5f * k2p.x, P.y + 0.5f * k2p.y), make_float2(S.x + 0.5f * k2s.x, S.y + 0.5f * k2s.y), cb,
           lossP, lossS, tpa, k3p, k3s);
    nl_rhs(make_float2(P.x + k3p.x, P.y + k3p.y), make_float2(S.x + k3s.x, S.y + k3s.y), cc, lossP, lossS, tpa, k4p, k4s);
    P.x += h6 * (k1p.x + 2.0f * k2p.x + 2.0f * k3p.x + k4p.x);
    P.y += h6 * (k1p.y + 2.0f * k2p.y + 2.0f * k3p.y + k4p.y);
    S.x += h6 * (k1s.x + 2.0f * k2s.x + 2.0f * k3s.x + k4s.x);
    S.y += h6 * (k1s.y + 2.0f * k2s.y + 2.0f * k3s.y + k4s.y);
}

}  // namespace shg
