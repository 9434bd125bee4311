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

// The stage inputs of one RK4 (rk4_cell): y2 = A + k1/2, y3 = A + k2/2, y4 = A + k3.
struct Rk4Stages {
    float2 y2p, y2s, y3p, y3s, y4p, y4s;
};

// Stage inputs of rk4_cell at (P, S) (k4 and the update are not needed by the transpose).
__device__ __forceinline__ void rk4_stages(float2 P, float2 S, float4 ca, float4 cb, float lossP, float lossS,
                                           float tpa, Rk4Stages &st) {
    float2 k1p, k1s, k2p, k2s, k3p, k3s;
    nl_rhs(P, S, ca, lossP, lossS, tpa, k1p, k1s);
    st.y2p = make_float2(P.x + 0.5f * k1p.x, P.y + 0.5f * k1p.y);
    st.y2s = make_float2(S.x + 0.5f * k1s.x, S.y + 0.5f * k1s.y);
    nl_rhs(st.y2p, st.y2s, cb, lossP, lossS, tpa, k2p, k2s);
    st.y3p = make_float2(P.x + 0.5f * k2p.x, P.y + 0.5f * k2p.y);
    st.y3s = make_float2(S.x + 0.5f * k2s.x, S.y + 0.5f * k2s.y);
    nl_rhs(st.y3p, st.y3s, cb, lossP, lossS, tpa, k3p, k3s);
    st.y4p = make_float2(P.x + k3p.x, P.y + k3p.y);
    st.y4s = make_float2(S.x + k3s.x, S.y + k3s.y);
}

// rk4_cell (the same operations, so the same floats) keeping its stage inputs: (P, S) is
// advanced in place and `st` receives the stages, so the transpose of this RK4 does not
// evaluate them again.
__device__ __forceinline__ void rk4_keep(float2 &P, float2 &S, float4 ca, float4 cb, float4 cc, float lossP,
                                         float lossS, float tpa, float h6, Rk4Stages &st) {
    float2 k1p, k1s, k2p, k2s, k3p, k3s, k4p, k4s;
    nl_rhs(P, S, ca, lossP, lossS, tpa, k1p, k1s);
    st.y2p = make_float2(P.x + 0.5f * k1p.x, P.y + 0.5f * k1p.y);
    st.y2s = make_float2(S.x + 0.5f * k1s.x, S.y + 0.5f * k1s.y);
    nl_rhs(st.y2p, st.y2s, cb, lossP, lossS, tpa, k2p, k2s);
    st.y3p = make_float2(P.x + 0.5f * k2p.x, P.y + 0.5f * k2p.y);
    st.y3s = make_float2(S.x + 0.5f * k2s.x, S.y + 0.5f * k2s.y);
    nl_rhs(st.y3p, st.y3s, cb, lossP, lossS, tpa, k3p, k3s);
    st.y4p = make_float2(P.x + k3p.x, P.y + k3p.y);
    st.y4s = make_float2(S.x + k3s.x, S.y + k3s.y);
    nl_rhs(st.y4p, st.y4s, cc, lossP, lossS, tpa, k4p, k4s);
    P.x += h6 * (k1p.x + 2.0f * k2p.x + 2.0f * k3p.x + k4p.x);
    P.y += h6 * (k1p.y + 2.0f * k2p.y + 2.0f * k3p.y + k4p.y);
    S.x += h6 * (k1s.x + 2.0f * k2s.x + 2.0f * k3s.x + k4s.x);
    S.y += h6 * (k1s.y + 2.0f * k2s.y + 2.0f * k3s.y + k4s.y);
}

// RK4^T (rk4_cell) at forward input (P, S) with stage inputs `st` and stage coefficients
// ca (stage 1), cb (stages 2-3), cc (stage 4): (lP, lS) is the adjoint of the output on
// entry and of the input on exit.  The sensitivities of the three coefficients, summed
// in float over the stages of this RK4, go to a1, a23, a4; with ONE (one coefficient for
// all stages) they are summed together and go to a1 alone.
template <bool ONE>
__device__ __forceinline__ void rk4_adj_rev(float2 P, float2 S, const Rk4Stages &st, float2 &lP, float2 &lS,
                                            float4 ca, float4 cb, float4 cc, float lossP, float lossS, float tpa,
                                            float h6, AdjAcc &a1, AdjAcc &a23, AdjAcc &a4) {
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
    rhs_adj(st.y4p, st.y4s, kb4p, kb4s, cc, lossP, lossS, tpa, rp, rs, ONE ? f1 : f4);
    ybp = add2(ybp, rp); ybs = add2(ybs, rs);
    kb3p = add2(kb3p, rp); kb3s = add2(kb3s, rs);
    rhs_adj(st.y3p, st.y3s, kb3p, kb3s, cb, lossP, lossS, tpa, rp, rs, ONE ? f1 : f23);
    ybp = add2(ybp, rp); ybs = add2(ybs, rs);
    kb2p = axpy2(0.5f, rp, kb2p); kb2s = axpy2(0.5f, rs, kb2s);
    rhs_adj(st.y2p, st.y2s, kb2p, kb2s, cb, lossP, lossS, tpa, rp, rs, ONE ? f1 : f23);
    ybp = add2(ybp, rp); ybs = add2(ybs, rs);
    kb1p = axpy2(0.5f, rp, kb1p); kb1s = axpy2(0.5f, rs, kb1s);
    rhs_adj(P, S, kb1p, kb1s, ca, lossP, lossS, tpa, rp, rs, f1);
    lP = add2(ybp, rp);
    lS = add2(ybs, rs);
    acc_add(a1, f1);
    if (!ONE) {
        acc_add(a23, f23);
        acc_add(a4, f4);
    }
}

// RK4^T evaluating its own stages.
__device__ __forceinline__ void rk4_adj3(float2 P, float2 S, float2 &lP, float2 &lS, float4 ca, float4 cb, float4 cc,
                                         float lossP, float lossS, float tpa, float h6, AdjAcc &a1, AdjAcc &a23,
                                         AdjAcc &a4) {
    Rk4Stages st;
    rk4_stages(P, S, ca, cb, lossP, lossS, tpa, st);
    rk4_adj_rev<false>(P, S, st, lP, lS, ca, cb, cc, lossP, lossS, tpa, h6, a1, a23, a4);
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

// The plain build's RK4^T (one coefficient for all stages), evaluating its own stages or
// from stages kept by rk4_keep.
__device__ __forceinline__ void rk4_adj(float2 P, float2 S, float2 &lP, float2 &lS, float4 c, const NlConst &k,
                                        AdjAcc &acc) {
    Rk4Stages st;
    rk4_stages(P, S, c, c, k.lossP, k.lossS, k.tpa, st);
    rk4_adj_rev<true>(P, S, st, lP, lS, c, c, c, k.lossP, k.lossS, k.tpa, k.h6, acc, acc, acc);
}
__device__ __forceinline__ void rk4_adj_kept(float2 P, float2 S, const Rk4Stages &st, float2 &lP, float2 &lS, float4 c,
                                             const NlConst &k, AdjAcc &acc) {
    rk4_adj_rev<true>(P, S, st, lP, lS, c, c, c, k.lossP, k.lossS, k.tpa, k.h6, acc, acc, acc);
}

}  // namespace shg
