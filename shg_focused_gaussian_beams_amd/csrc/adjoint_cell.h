// adjoint_cell.h — the transpose of the per-cell RK4 step, shared by the plain and
// THERMAL adjoint row passes.
//
// The per-cell maps are not holomorphic (conj(P), |S|^2), so each stage uses the real
// adjoint zbar = conj(dg/dz) gbar + (dg/dconj z) conj(gbar) with the gradient convention
// dJ = Re sum conj(lambda) dA.  Accumulators collect the sensitivities of J to the
// coefficients, dJ = Re(conj(c1b) dc1) + Re(conj(c2b) dc2) + tpab dtpa: in float over the
// four stages of one RK4 of one cell, in FP64 beyond.
#pragma once
#include <hip/hip_runtime.h>

#include "fft_lds.h"
#include "nl_step.h"

namespace shg {

__device__ __forceinline__ float2 cj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 f2(float x, float y) { return make_float2(x, y); }
__device__ __forceinline__ float2 add2(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 axpy2(float s, float2 a, float2 b) { return make_float2(b.x + s * a.x, b.y + s * a.y); }
__device__ __forceinline__ float2 scl2(float s, float2 a) { return make_float2(s * a.x, s * a.y); }

struct AdjAcc {
    double c1r, c1i, c2r, c2i, tpa;
};

__device__ __forceinline__ void acc_zero(AdjAcc &a) { a.c1r = a.c1i = a.c2r = a.c2i = a.tpa = 0.0; }

// Per-cell, per-RK4 sums of the coefficient sensitivities (four stage terms at most; the
// caller folds them into the FP64 AdjAcc once per RK4).
struct AdjAccF {
    float c1r, c1i, c2r, c2i, tpa;
};

__device__ __forceinline__ void accf_zero(AdjAccF &a) { a.c1r = a.c1i = a.c2r = a.c2i = a.tpa = 0.0f; }
__device__ __forceinline__ void acc_add(AdjAcc &a, const AdjAccF &b) {
    a.c1r += (double)b.c1r; a.c1i += (double)b.c1i; a.c2r += (double)b.c2r; a.c2i += (double)b.c2i;
    a.tpa += (double)b.tpa;
}

// Adjoint of the right-hand side f(P, S) (nl_rhs) at (P, S) applied to (gP, gS).
__device__ __forceinline__ void rhs_adj(float2 P, float2 S, float2 gP, float2 gS, float4 c, float lossP, float lossS,
                                        float tpa, float2 &Pb, float2 &Sb, AdjAccF &acc) {
    const float2 c1 = f2(c.x, c.y), c2 = f2(c.z, c.w);
    const float s2 = S.x * S.x + S.y * S.y;
    // Pb = -lossP gP + c1 S conj(gP) + 2 conj(c2) conj(P) gS
    Pb = add2(add2(scl2(-lossP, gP), c_mul(c_mul(c1, S), cj(gP))), scl2(2.0f, c_mul(c_mul(cj(c2), cj(P)), gS)));
    // Sb = conj(c1) P gP - (lossS + 2 tpa |S|^2) gS - tpa S^2 conj(gS)
    Sb = add2(add2(c_mul(c_mul(cj(c1), P), gP), scl2(-(lossS + 2.0f * tpa * s2), gS)),
              scl2(-tpa, c_mul(c_mul(S, S), cj(gS))));
    // coefficient sensitivities: c1b += conj(S) P gP, c2b += conj(P^2) gS, tpab += -|S|^2 Re(conj(gS) S)
    const float2 u = c_mul(cj(S), P);
    const float2 w = c_mul(cj(c_mul(P, P)), gS);
    const float2 v = c_mul(u, gP);
    acc.c1r += v.x; acc.c1i += v.y;
    acc.c2r += w.x; acc.c2i += w.y;
    acc.tpa += -s2 * (gS.x * S.x + gS.y * S.y);
}

// RK4^T (rk4_cell) at forward input (P, S) with stage coefficients ca (stage 1), cb
// (stages 2-3), cc (stage 4): (lP, lS) is the adjoint of the output on entry and of the
// input on exit.  The sensitivities of the three coefficients go to a1, a23, a4 (one
// accumulator may be passed for all three), summed in float over the stages of this RK4.
__device__ __forceinline__ void rk4_adj3(float2 P, float2 S, float2 &lP, float2 &lS, float4 ca, float4 cb, float4 cc,
                                         float lossP, float lossS, float tpa, float h6, AdjAcc &a1, AdjAcc &a23,
                                         AdjAcc &a4) {
    float2 k1p, k1s, k2p, k2s, k3p, k3s;
    nl_rhs(P, S, ca, lossP, lossS, tpa, k1p, k1s);
    const float2 y2p = axpy2(0.5f, k1p, P), y2s = axpy2(0.5f, k1s, S);
    nl_rhs(y2p, y2s, cb, lossP, lossS, tpa, k2p, k2s);
    const float2 y3p = axpy2(0.5f, k2p, P), y3s = axpy2(0.5f, k2s, S);
    nl_rhs(y3p, y3s, cb, lossP, lossS, tpa, k3p, k3s);
    const float2 y4p = add2(P, k3p), y4s = add2(S, k3s);
    float2 ybp = lP, ybs = lS;
    float2 kb1p = scl2(h6, lP), kb1s = scl2(h6, lS);
    float2 kb2p = scl2(2.0f * h6, lP), kb2s = scl2(2.0f * h6, lS);
    float2 kb3p = kb2p, kb3s = kb2s;
    const float2 kb4p = kb1p, kb4s = kb1s;
    float2 rp, rs;
    AdjAccF f1, f23, f4;
    accf_zero(f1);
    accf_zero(f23);
    accf_zero(f4);
    rhs_adj(y4p, y4s, kb4p, kb4s, cc, lossP, lossS, tpa, rp, rs, f4);
    ybp = add2(ybp, rp); ybs = add2(ybs, rs);
    kb3p = add2(kb3p, rp); kb3s = add2(kb3s, rs);
    rhs_adj(y3p, y3s, kb3p, kb3s, cb, lossP, lossS, tpa, rp, rs, f23);
    ybp = add2(ybp, rp); ybs = add2(ybs, rs);
    kb2p = axpy2(0.5f, rp, kb2p); kb2s = axpy2(0.5f, rs, kb2s);
    rhs_adj(y2p, y2s, kb2p, kb2s, cb, lossP, lossS, tpa, rp, rs, f23);
    ybp = add2(ybp, rp); ybs = add2(ybs, rs);
    kb1p = axpy2(0.5f, rp, kb1p); kb1s = axpy2(0.5f, rs, kb1s);
    rhs_adj(P, S, kb1p, kb1s, ca, lossP, lossS, tpa, rp, rs, f1);
    lP = add2(ybp, rp);
    lS = add2(ybs, rs);
    acc_add(a1, f1);
    acc_add(a23, f23);
    acc_add(a4, f4);
}

// Layouts of the tape and adjoint passes (split layouts, as the forward's): a row per
// workgroup of W >= 2 waves, loaded and stored one cell per thread through LDS, the two
// fields transformed by waves 0 and 1, the per-cell work (RK4^T with its recomputed
// stages and FP64 accumulators) on every thread; CW columns per workgroup of 2*CW waves.
template <int N> struct AdjWide;
template <> struct AdjWide<64>  { static constexpr int E = 4, W = 2; };
template <> struct AdjWide<128> { static constexpr int E = 4, W = 2; };
template <> struct AdjWide<256> { static constexpr int E = 4, W = 4; };
template <> struct AdjWide<512> { static constexpr int E = 8, W = 8; };
template <int N> struct AdjCol { static constexpr int E = N == 512 ? 8 : 4, CW = 2; };

// The plain build's RK4^T (one coefficient for all stages).
__device__ __forceinline__ void rk4_adj(float2 P, float2 S, float2 &lP, float2 &lS, float4 c, const NlConst &k,
                                        AdjAcc &acc) {
    rk4_adj3(P, S, lP, lS, c, c, c, k.lossP, k.lossS, k.tpa, k.h6, acc, acc, acc);
}

}  // namespace shg
