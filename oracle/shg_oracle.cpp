// shg_oracle.cpp — CPU restatement of the reference focused-beam SHG split-step solver.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the CPU baseline.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the
// library built from it, and only as the checker: nothing in the product calls it.
//
// It follows the reference algorithm (alfredos84/SHG_Focused_Gaussian_Beams) in two
// precisions, selected per call:
//   prec 32  "oracle32": the reference's arithmetic — real_t = float with the exact
//            float/double mix of its expressions (double literals, std::pow promotion,
//            non-template complex<float> x real operators that evaluate left to right
//            in float, src/headers/Operators.h:13-101), float libm cosf/sinf/expf/powf.
//   prec 64  "oracle64": every quantity in double.
// The FFT is this file's own radix-2 transform (FFTW is not available here); its
// sign and normalisation follow FFTW's documented definitions (FFTW_BACKWARD = e^{+i},
// unnormalised).  The reference's FFTW rounding is therefore not reproduced bit for bit.
//
// Every function cites the reference file:line it restates.  Pins (tests/test_oracle*.py):
//   n(L,T), dk, pump amplitude  — values recorded by an independent FP64/FP32 probe of
//                                  the same formulas, and the reference's own (Lambda, T)
//                                  operating point (src/SHG.sh:34, src/headers/Crystal.h:43);
//   fft2                        — numpy.fft (brute DFT definition, Parseval, round trip);
//   diffraction chain           — closed-form spectral multiply H^m and Gaussian beam q-law;
//   rk4 / cwes (plain)          — plane-wave geometric-sum KAT, conservation (alpha=beta=0);
//   jacobi                      — fixed point, x-mirror symmetry, maximum principle;
//   jacobi (surrounded ovens)   — fixed point, oven faces untouched, steady state equal to a
//                                  sparse solve of the physics discretisation;
//   walk-off                    — centroid shift -m dz tan(rho) of the diffraction chain;
//   setQ                        — closed form of its three terms at alpha/beta toggles;
//   powers                      — Gaussian closed form 2P/(eps0 c n).
//   thermal outer loop          — parity unpinned beyond its parts (no reference run possible).

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../include/shg.h"

namespace {

// ---------------------------------------------------------------------------
// libm dispatch: the reference calls cosf/sinf/expf/powf/sqrtf on floats.
inline float  m_cos(float a)  { return cosf(a); }
inline double m_cos(double a) { return cos(a); }
inline float  m_sin(float a)  { return sinf(a); }
inline double m_sin(double a) { return sin(a); }
inline float  m_exp(float a)  { return expf(a); }
inline double m_exp(double a) { return exp(a); }
inline float  m_pow2(float a) { return powf(a, 2.0f); }
inline double m_pow2(double a) { return pow(a, 2.0); }
inline float  m_sqrt(float a) { return sqrtf(a); }
inline double m_sqrt(double a) { return sqrt(a); }
inline float  m_tan(float a)  { return tanf(a); }
inline double m_tan(double a) { return tan(a); }

// Complex type whose operators evaluate exactly like src/headers/Operators.h:13-125.
template <class R> struct cx { R re, im; };
template <class R> inline cx<R> operator+(cx<R> a, cx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <class R> inline cx<R> operator-(cx<R> a, cx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <class R> inline cx<R> operator-(cx<R> a) { return {-a.re, -a.im}; }
template <class R> inline cx<R> operator*(cx<R> a, cx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};            // Operators.h:78-82
}
template <class R> inline cx<R> rmul(R a, cx<R> b) { return {a * b.re, a * b.im}; }   // Operators.h:64-75
template <class R> inline cx<R> rsub(R a, cx<R> b) { return {a - b.re, -b.im}; }      // Operators.h:43-47
template <class R> inline cx<R> rdivc(cx<R> b, R a) { return {b.re / a, b.im / a}; }  // Operators.h:85-89
template <class R> inline cx<R> rdiv(R a, cx<R> b) {                                  // Operators.h:92-97
    R den = b.re * b.re + b.im * b.im;
    return {(+a * b.re) / den, (-a * b.im) / den};
}
template <class R> inline cx<R> cexpi(R a) { return {m_cos(a), m_sin(a)}; }            // Operators.h:105-109
template <class R> inline cx<R> cexpc(cx<R> a) {                                      // Operators.h:113-117
    return {m_exp(a.re) * m_cos(a.im), m_exp(a.re) * m_sin(a.im)};
}
template <class R> inline cx<R> conj(cx<R> a) { return {+a.re, -a.im}; }               // Operators.h:121-125
template <class R> inline R abs2(cx<R> a) { return a.re * a.re + a.im * a.im; }       // Operators.h:138-143

// ---------------------------------------------------------------------------
// Refractive indices.
//
// MgO:sPPLT extraordinary index, Bruner et al. (src/headers/Crystal.h:55-71): the
// coefficients are float; b(T), c(T) are double products rounded to float; pow()
// promotes to double, so the sum and sqrt are double; the result is rounded to R.
template <class R> R n_spplt(R L, R T) {
    const R A = (R)4.502483, B = (R)0.007294, C = (R)0.185087, D = (R)-0.02357;
    const R E = (R)0.073423, F = (R)0.199595, G = (R)0.001, H = (R)7.99724;
    const R b = (R)(3.483933e-8 * std::pow((double)T + 273.15, 2));
    const R c = (R)(1.607839e-8 * std::pow((double)T + 273.15, 2));
    const double L2 = std::pow((double)L, 2);
    const R Bb = B + b, Cc = C + c;                               // float adds in the reference
    return (R)std::sqrt((double)A + (double)Bb / (L2 - std::pow((double)Cc, 2)) +
                        (double)E / (L2 - std::pow((double)F, 2)) +
                        (double)G / (L2 - std::pow((double)H, 2)) + (double)D * L2);
}

// MgO:PPLN extraordinary index, Gayer et al., Appl. Phys. B 91, 343 (2008), 5% MgO:CLN.
// Absent from the reference (claimed in its README.md:48-49 only); evaluated in double.
template <class R> R n_ppln(R L, R T) {
    const double a1 = 5.756, a2 = 0.0983, a3 = 0.2020, a4 = 189.32, a5 = 12.52, a6 = 1.32e-2;
    const double b1 = 2.860e-6, b2 = 4.700e-8, b3 = 6.113e-8, b4 = 1.516e-4;
    const double t = (double)T, l2 = (double)L * (double)L;
    const double f = (t - 24.5) * (t + 570.82);
    const double p = a3 + b3 * f;
    return (R)std::sqrt(a1 + b1 * f + (a2 + b2 * f) / (l2 - p * p) + (a4 + b4 * f) / (l2 - a5 * a5) - a6 * l2);
}

template <class R> R n_index(int crystal, R L, R T) {
    return crystal == SHG_CRYSTAL_PPLN ? n_ppln<R>(L, T) : n_spplt<R>(L, T);
}

// ---------------------------------------------------------------------------
// The reference's global constants (src/headers/Crystal.h:18-53, src/SHG.cpp:7-20)
// at run-time grid sizes.  Each `const real_t X = <double expr>` is rounded to R once.
template <class R> struct K {
    int NX, NY, NZ, crystal, bc;
    R PI, C, EPS0, lp, ls, LX, LY, LZ, dx, dy, dz, deff, dQ, Lambda;
    R alpha_p, alpha_s, beta_s, rhop, rhos, k_th, h_conv;
    explicit K(const shg_params &p) {
        NX = p.nx; NY = p.ny; NZ = p.nz; crystal = p.crystal; bc = p.thermal_bc;
        PI = (R)3.14159265358979323846;
        C = (R)(299792458 * 1E6 / 1E12);
        EPS0 = (R)(8.8541878128E-12 * 1E12 / 1E6);
        lp = (R)p.lambda_p;
        ls = (R)(0.5 * (double)lp);
        LX = (R)p.lx; LY = (R)p.ly; LZ = (R)p.lz;
        dx = LX / (R)(unsigned)(NX - 1);
        dy = LY / (R)(unsigned)(NY - 1);
        dz = LZ / (R)(unsigned)(NZ - 1);
        deff = (R)p.deff;
        dQ = (R)((2.0 / (double)PI) * (double)deff);
        Lambda = (R)p.qpm_period;
        alpha_p = (R)p.alpha_p; alpha_s = (R)p.alpha_s; beta_s = (R)p.beta_s;
        rhop = (R)p.rho_p; rhos = (R)p.rho_s;
        k_th = (R)p.k_th; h_conv = (R)p.h_conv;
    }
    R n(R L, R T) const { return n_index<R>(crystal, L, T); }
    size_t idx(int x, int y, int z) const { return (size_t)z * NX * NY + (size_t)y * NX + x; }
};

// Phase mismatch, float cancellation included (src/headers/PhaseMatching.h:63, :47).
template <class R> R dk_of(const K<R> &k, R T) {
    return (R)2 * k.PI * ((R)2 * k.n(k.lp, T) / k.lp - k.n(k.ls, T) / k.ls + (R)1 / (k.Lambda));
}

// setPump (src/headers/Efields.h:74-104): TEM00 at z = 0 focused at `focalpoint`.
template <class R> void set_pump(const K<R> &k, R Power, R focalpoint, R waistX, R Temp, cx<R> *P) {
    const cx<R> Im{(R)0, (R)1};
    const R uX = (R)(0.5 * k.NX);
    const R uY = (R)(0.5 * k.NY);
    const R wX2 = waistX * waistX;
    const R zRX = k.PI * k.n(k.lp, Temp) * wX2 / k.lp;
    const R Ap0 = m_sqrt((R)4 * Power / (k.EPS0 * k.C * k.PI * k.n(k.lp, Temp) * wX2));
    const R etaX = focalpoint / zRX;
    const cx<R> MX = rsub((R)1, rmul(etaX, Im));      // 1 - Im*etaX
    for (int ix = 0; ix < k.NX; ix++)
        for (int iy = 0; iy < k.NY; iy++) {
            const R ex = -m_pow2(((R)(unsigned)ix - uX) * k.dx) - m_pow2(((R)(unsigned)iy - uY) * k.dy);
            const cx<R> arg = rdiv(ex, rmul(wX2, MX));  // real / (wX2*MX)
            P[k.idx(ix, iy, 0)] = rdiv(Ap0, MX) * cexpc(arg);
        }
}

// beamPropagator + fftShift2D (src/headers/Efields.h:107-126, 32-71, 186-192): the
// diffraction operator for one dz, shifted so that DC sits at index 0.
template <class R> void set_propagators(const K<R> &k, R Temp, cx<R> *Hp, cx<R> *Hs) {
    const R kp = (R)2 * k.PI * k.n(k.lp, Temp) / k.lp;
    const R ks = (R)2 * k.PI * k.n(k.ls, Temp) / k.ls;
    const R dfX = (R)1 / k.dx / (R)(unsigned)k.NX;
    const R dfY = (R)1 / k.dy / (R)(unsigned)k.NY;
    const R uX = (R)(0.5 * k.NX), uY = (R)(0.5 * k.NY);
    const int NX = k.NX, NY = k.NY;
    std::vector<cx<R>> tp((size_t)NX * NY), ts((size_t)NX * NY);
    for (int iy = 0; iy < NY; iy++)
        for (int ix = 0; ix < NX; ix++) {
            const R fx = (R)(unsigned)ix - uX, fy = (R)(unsigned)iy - uY;
            const R q = dfX * dfX * m_pow2(fx) + dfY * dfY * m_pow2(fy);
            const R twoPI2 = (R)2 * m_pow2(k.PI);
            tp[k.idx(ix, iy, 0)] = cexpi(-k.dz * (twoPI2 / kp * q + (R)2 * k.PI * dfX * m_tan(k.rhop) * fx));
            ts[k.idx(ix, iy, 0)] = cexpi(-k.dz * (twoPI2 / ks * q + (R)2 * k.PI * dfX * m_tan(k.rhos) * fx));
        }
    // fftshift by (NX/2, NY/2): out[(i+c) mod N] = in[i]
    const int cx_ = NX / 2, cy_ = NY / 2;
    for (int iy = 0; iy < NY; iy++)
        for (int ix = 0; ix < NX; ix++) {
            const int ox = (ix + cx_) % NX, oy = (iy + cy_) % NY;
            Hp[k.idx(ox, oy, 0)] = tp[k.idx(ix, iy, 0)];
            Hs[k.idx(ox, oy, 0)] = ts[k.idx(ix, iy, 0)];
        }
}

// ---------------------------------------------------------------------------
// 2D FFT (stands in for fftwf_plan_dft_2d + fftwf_execute, src/headers/Solver.h:161-187).
// sign = +1: FFTW_BACKWARD (e^{+2 pi i jk/N}); sign = -1: FFTW_FORWARD.  Unnormalised.
// Iterative radix-2 decimation in time; twiddles evaluated in double, rounded to R.
template <class R> void fft1d(cx<R> *a, int n, int stride, int sign, const std::vector<cx<R>> &w) {
    // w[j] = exp(sign * 2 pi i j / n), j < n/2
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { cx<R> t = a[(size_t)i * stride]; a[(size_t)i * stride] = a[(size_t)j * stride]; a[(size_t)j * stride] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len)
            for (int j = 0; j < half; j++) {
                cx<R> u = a[(size_t)(i + j) * stride];
                cx<R> v = w[(size_t)j * step] * a[(size_t)(i + j + half) * stride];
                a[(size_t)(i + j) * stride] = u + v;
                a[(size_t)(i + j + half) * stride] = u - v;
            }
    }
}

template <class R> std::vector<cx<R>> twiddles(int n, int sign) {
    std::vector<cx<R>> w(n / 2 > 0 ? n / 2 : 1);
    for (int j = 0; j < n / 2; j++) {
        const double th = sign * 2.0 * M_PI * (double)j / (double)n;
        w[j] = {(R)cos(th), (R)sin(th)};
    }
    return w;
}

template <class R> void fft2d(cx<R> *a, int nx, int ny, int sign) {
    const auto wx = twiddles<R>(nx, sign), wy = twiddles<R>(ny, sign);
    for (int y = 0; y < ny; y++) fft1d(a + (size_t)y * nx, nx, 1, sign, wx);
    for (int x = 0; x < nx; x++) fft1d(a + x, ny, nx, sign, wy);
}

// ---------------------------------------------------------------------------
// Solver state (the reference's Efields/Solver scratch, src/headers/Efields.h:147-172,
// src/headers/Solver.h:123-139).
template <class R> struct State {
    K<R> k;
    std::vector<cx<R>> P, S, PQ, SQ, Hp, Hs, k1p, k2p, k3p, k4p, k1s, k2s, k3s, k4s, auxp, auxs;
    R Temp;
    // THERMAL
    std::vector<R> Ti, Tf, Q, DK, DKint;
    explicit State(const shg_params &p) : k(p) {
        const size_t n2 = (size_t)k.NX * k.NY;
        for (auto *v : {&P, &S, &PQ, &SQ, &Hp, &Hs, &k1p, &k2p, &k3p, &k4p, &k1s, &k2s, &k3s, &k4s, &auxp, &auxs})
            v->assign(n2, cx<R>{(R)0, (R)0});
        Temp = (R)p.temp;
    }
};

// When set, dAdz evaluates n() once per cell as the reference does (src/headers/Solver.h:75,77)
// instead of once per call: same values, the reference's cost (used for CPU-baseline timing).
int g_percell_index = 0;

// dAdz, plain build (src/headers/Solver.h:65-83).  z is the step-start z, unchanged
// across the four stages; `increment` is unused by the reference.
template <class R> void dAdz_plain(State<R> &st, cx<R> *dP, cx<R> *dS, const cx<R> *P, const cx<R> *S,
                                   R DK, R z) {
    const K<R> &k = st.k;
    const cx<R> Im{(R)0, (R)1};
    const R np = k.n(k.lp, st.Temp), ns = k.n(k.ls, st.Temp);
    const cx<R> cP = rdivc(rmul(k.dQ, rmul(k.PI, rmul((R)1.0, Im))), np * k.lp);   // Im*1.0*PI*dQ/(n*lp)
    const cx<R> cS = rdivc(rmul(k.dQ, rmul(k.PI, rmul((R)0.5, Im))), ns * k.ls);   // Im*0.5*PI*dQ/(n*ls)
    const cx<R> eP = cexpc(rmul(DK, rmul(z, -Im)));                                 // CpxExp(-Im*z*DK)
    const cx<R> eS = cexpi(z * DK);                                                 // CpxExp(+z*DK)
    const R lossP = (R)(0.5 * (double)k.alpha_p);
    const double tpa = (double)k.beta_s * 0.5 * (double)k.C * (double)k.EPS0 * (double)ns;
    const size_t n2 = (size_t)k.NX * k.NY;
    if (g_percell_index) {
        for (size_t i = 0; i < n2; i++) {
            const R npc = k.n(k.lp, st.Temp), nsc = k.n(k.ls, st.Temp);
            const cx<R> cPc = rdivc(rmul(k.dQ, rmul(k.PI, rmul((R)1.0, Im))), npc * k.lp);
            const cx<R> cSc = rdivc(rmul(k.dQ, rmul(k.PI, rmul((R)0.5, Im))), nsc * k.ls);
            dP[i] = cPc * S[i] * conj(P[i]) * eP - rmul(lossP, P[i]);
            const double tpac = (double)k.beta_s * 0.5 * (double)k.C * (double)k.EPS0 * (double)k.n(k.ls, st.Temp);
            const R lossS = (R)(0.5 * ((double)k.alpha_s + tpac * (double)abs2(S[i])));
            dS[i] = cSc * P[i] * P[i] * eS - rmul(lossS, S[i]);
        }
        return;
    }
    for (size_t i = 0; i < n2; i++) {
        dP[i] = cP * S[i] * conj(P[i]) * eP - rmul(lossP, P[i]);
        const R lossS = (R)(0.5 * ((double)k.alpha_s + tpa * (double)abs2(S[i])));
        dS[i] = cS * P[i] * P[i] * eS - rmul(lossS, S[i]);
    }
}

// dAdz, THERMAL build (src/headers/Solver.h:44-63): per-cell n(T) from slice `s` of the
// temperature, and the phase DKint * (dz*inc + (s+1)*dz)/(s+1) in double, passed as float.
template <class R> void dAdz_thermal(State<R> &st, cx<R> *dP, cx<R> *dS, const cx<R> *P, const cx<R> *S,
                                     R inc, int s, int hs) {
    const K<R> &k = st.k;
    const cx<R> Im{(R)0, (R)1};
    const size_t n2 = (size_t)k.NX * k.NY;
    const R lossP = (R)(0.5 * (double)k.alpha_p);
    const double zfac = (double)(R)(k.dz * inc) + (s + 1.0) * (double)k.dz;
    for (size_t i = 0; i < n2; i++) {
        const R T = st.Tf[(size_t)hs * n2 + i];
        const R np = k.n(k.lp, T), ns = k.n(k.ls, T);
        const R phi = (R)((double)st.DKint[i] * zfac / (s + 1.0));
        const cx<R> cP = rdivc(rmul(k.dQ, rmul(k.PI, rmul((R)1.0, Im))), np * k.lp);
        const cx<R> cS = rdivc(rmul(k.dQ, rmul(k.PI, rmul((R)0.5, Im))), ns * k.ls);
        dP[i] = cP * S[i] * conj(P[i]) * cexpi(-phi) - rmul(lossP, P[i]);
        const R lossS = (R)(0.5 * ((double)k.alpha_s + (double)k.beta_s * 0.5 * (double)k.EPS0 * (double)k.C *
                                                           (double)ns * (double)abs2(S[i])));
        dS[i] = cS * P[i] * P[i] * cexpi(phi) - rmul(lossS, S[i]);
    }
}

// linealCombination (src/headers/Solver.h:103-117) and sumRK4 (:86-100).
template <class R> void lin_comb(State<R> &st, std::vector<cx<R>> &ap, std::vector<cx<R>> &as,
                                 const std::vector<cx<R>> &kp, const std::vector<cx<R>> &ks, R c) {
    for (size_t i = 0; i < ap.size(); i++) {
        ap[i] = st.P[i] + rmul(c, kp[i]);
        as[i] = st.S[i] + rmul(c, ks[i]);
    }
}
template <class R> void sum_rk4(State<R> &st, R h) {
    for (size_t i = 0; i < st.P.size(); i++) {
        st.P[i] = st.P[i] + rdivc(rmul(h, st.k1p[i] + rmul((R)2, st.k2p[i]) + rmul((R)2, st.k3p[i]) + st.k4p[i]), (R)6);
        st.S[i] = st.S[i] + rdivc(rmul(h, st.k1s[i] + rmul((R)2, st.k2s[i]) + rmul((R)2, st.k3s[i]) + st.k4s[i]), (R)6);
    }
}

// solverRK4 (src/headers/Solver.h:193-233).  Stage combinations use crk4 in {.5,.5,1}
// unscaled by dz; only the final sum carries h (the reference passes dz).
// thermal: s = field step, hs = heat slice of that step.
template <class R> void solver_rk4(State<R> &st, R z, R DK, R h, bool thermal, int s, int hs) {
    auto f = [&](std::vector<cx<R>> &kp, std::vector<cx<R>> &ks, const std::vector<cx<R>> &P,
                 const std::vector<cx<R>> &S, R inc) {
        if (thermal) dAdz_thermal(st, kp.data(), ks.data(), P.data(), S.data(), inc, s, hs);
        else dAdz_plain(st, kp.data(), ks.data(), P.data(), S.data(), DK, z);
    };
    if (thermal) {   // setDKInt0 + IntegrateDK: DKint = sum_{z'<=s} DK[z'] in float order
        const size_t n2 = (size_t)st.k.NX * st.k.NY;
        for (size_t i = 0; i < n2; i++) st.DKint[i] = (R)0;
        for (int zz = 0; zz <= s; zz++)
            for (size_t i = 0; i < n2; i++) st.DKint[i] += st.DK[(size_t)zz * n2 + i];
    }
    f(st.k1p, st.k1s, st.P, st.S, (R)0.5 * (R)0.0);
    lin_comb(st, st.auxp, st.auxs, st.k1p, st.k1s, (R)0.5);
    f(st.k2p, st.k2s, st.auxp, st.auxs, (R)0.5 * (R)0.5);
    lin_comb(st, st.auxp, st.auxs, st.k2p, st.k2s, (R)0.5);
    f(st.k3p, st.k3s, st.auxp, st.auxs, (R)0.5 * (R)0.5);
    lin_comb(st, st.auxp, st.auxs, st.k3p, st.k3s, (R)1.0);
    f(st.k4p, st.k4s, st.auxp, st.auxs, (R)0.5 * (R)1.0);
    sum_rk4(st, h);
}

// Solver::diffraction (src/headers/Solver.h:156-190): BACKWARD FFT, multiply by the
// shifted propagator, FORWARD FFT, divide by NX*NY (FFTscale, src/headers/Efields.h:15-30).
template <class R> void diffraction(State<R> &st) {
    const int NX = st.k.NX, NY = st.k.NY;
    const size_t n2 = (size_t)NX * NY;
    st.PQ = st.P; fft2d(st.PQ.data(), NX, NY, +1);
    st.SQ = st.S; fft2d(st.SQ.data(), NX, NY, +1);
    for (size_t i = 0; i < n2; i++) { st.PQ[i] = st.Hp[i] * st.PQ[i]; st.SQ[i] = st.Hs[i] * st.SQ[i]; }
    fft2d(st.PQ.data(), NX, NY, -1);
    fft2d(st.SQ.data(), NX, NY, -1);
    const R size = (R)(unsigned)(NX * NY);
    for (size_t i = 0; i < n2; i++) { st.P[i] = rdivc(st.PQ[i], size); st.S[i] = rdivc(st.SQ[i], size); }
}

// Solver::SSFM (src/headers/Solver.h:236-242); scheme STRANG halves the RK4 length.
template <class R> void ssfm(State<R> &st, R z, R DK, int scheme, bool thermal, int s, int hs) {
    const R h = scheme == SHG_SCHEME_STRANG ? st.k.dz * (R)0.5 : st.k.dz;
    solver_rk4(st, z, DK, h, thermal, s, hs);
    diffraction(st);
    solver_rk4(st, z, DK, h, thermal, s, hs);
}

// setQ (src/headers/Tfield.h:277-290): heat source of slice hs from the fields after step s.
template <class R> void set_q(State<R> &st, int hs) {
    const K<R> &k = st.k;
    const size_t n2 = (size_t)k.NX * k.NY;
    for (size_t i = 0; i < n2; i++) {
        const R T = st.Tf[(size_t)hs * n2 + i];
        const R np = k.n(k.lp, T), ns = k.n(k.ls, T);
        const double Ip = (double)k.alpha_p * 0.5 * (double)k.C * (double)k.EPS0 * (double)np * (double)abs2(st.P[i]);
        const double Is = (double)k.alpha_s * 0.5 * (double)k.C * (double)k.EPS0 * (double)ns * (double)abs2(st.S[i]);
        const R i2arg = (R)(0.5 * (double)k.C * (double)k.EPS0 * (double)ns * (double)abs2(st.S[i]));
        const R tpa = k.beta_s * m_pow2(i2arg);
        st.Q[(size_t)hs * n2 + i] = (R)(Ip + Is + (double)tpa);
    }
}

// kernelHeatEquationOvenSurrounded (src/headers/Tfield.h:224-258): one Jacobi sweep with
// the ovens on the four lateral faces (never written); inner points, and the z = 0 and
// z = L faces (convective, Biot number h dz / k) inside the lateral faces.
template <class R> void jacobi_sweep_oven(const K<R> &k, int NZ, R dz, R *Tf, const R *Ti, R T_inf, const R *Q) {
    const int NX = k.NX, NY = k.NY;
    const R dx = k.dx, dy = k.dy;
    auto I = [&](int x, int y, int z) { return (size_t)z * NX * NY + (size_t)y * NX + x; };
    const R Biz = k.h_conv * dz / k.k_th;
    const R az = dz * dz / ((R)2 * dx * dx), bz = dz * dz / ((R)2 * dy * dy);
    const R L2 = (R)1 / ((R)2 / (dx * dx) + (R)2 / (dy * dy) + (R)2 / (dz * dz));
    for (int z = 0; z < NZ; z++)
        for (int y = 1; y < NY - 1; y++)
            for (int x = 1; x < NX - 1; x++) {
                if (z > 0 && z < NZ - 1) {
                    Tf[I(x, y, z)] = L2 * ((Ti[I(x + 1, y, z)] + Ti[I(x - 1, y, z)]) / (dx * dx) +
                                           (Ti[I(x, y + 1, z)] + Ti[I(x, y - 1, z)]) / (dy * dy) +
                                           (Ti[I(x, y, z + 1)] + Ti[I(x, y, z - 1)]) / (dz * dz) +
                                           Q[I(x, y, z)] / k.k_th);
                } else {
                    const int zn = z == 0 ? z + 1 : z - 1;     // back (z = 0) or front (z = L)
                    Tf[I(x, y, z)] = (Ti[I(x, y, zn)] + az * (Ti[I(x + 1, y, z)] + Ti[I(x - 1, y, z)]) +
                                      bz * (Ti[I(x, y + 1, z)] + Ti[I(x, y - 1, z)]) + Biz * T_inf) /
                                     ((R)1 + (R)2 * az + (R)2 * bz + Biz);
                }
            }
}

// heatEquationTopOpen (src/headers/Tfield.h:113-221): one Jacobi sweep Tf <- J(Ti, Q).
// The y = 0 plane is never written (Dirichlet ovens).  NZ here is the heat grid depth.
// With the surrounded-oven boundary the sweep is jacobi_sweep_oven.
template <class R> void jacobi_sweep(const K<R> &k, int NZ, R dz, R *Tf, const R *Ti, R T_inf, const R *Q) {
    if (k.bc == SHG_BC_OVEN_SURROUNDED) {
        jacobi_sweep_oven<R>(k, NZ, dz, Tf, Ti, T_inf, Q);
        return;
    }
    const int NX = k.NX, NY = k.NY;
    const R dx = k.dx, dy = k.dy;
    auto I = [&](int x, int y, int z) { return (size_t)z * NX * NY + (size_t)y * NX + x; };
    const R thermal_cond = k.k_th, h = k.h_conv;
    const R Bix = h * dx / thermal_cond, Biy = h * dy / thermal_cond, Biz = h * dz / thermal_cond;
    const R ax = dx * dx / ((R)2 * dy * dy), ay = dy * dy / ((R)2 * dx * dx), az = dz * dz / ((R)2 * dx * dx);
    const R bx = dx * dx / ((R)2 * dz * dz), by = dy * dy / ((R)2 * dz * dz), bz = dz * dz / ((R)2 * dy * dy);
    const R av = dz / ((R)2 * dx), bv = dx / ((R)2 * dz), cv = dx * dz / ((R)4 * dy * dy);
    const R V = (R)((R)1 / (double)((double)(av + bv + (R)2 * cv) + 0.5 * (double)(Bix + Biz)));
    const R ah = dz / ((R)2 * dy), bh = dy / ((R)2 * dz), ch = dy * dz / ((R)4 * dx * dx);
    const R H = (R)((R)1 / (double)((double)(ah + bh + (R)2 * ch) + 0.5 * (double)(Biy + Biz)));
    const R ad = dx / ((R)2 * dy), bd = dy / ((R)2 * dx), cd = dx * dy / ((R)4 * dz * dz);
    const R D = (R)((R)1 / (double)((double)(ad + bd + (R)2 * cd) + 0.5 * (double)(Bix + Biy)));
    const R AV = dy * dz / dx, BV = dx * dz / dy, CV = dx * dy / dz, BiD = Bix * dy + Biy * dz + Biz * dx;
    const R GV = (R)1 / (AV + BV + CV + BiD);
    const R L2 = (R)1 / ((R)2 / (dx * dx) + (R)2 / (dy * dy) + (R)2 / (dz * dz));
    const R fdenx = (R)1 + (R)2 * ax + (R)2 * bx + Bix;
    const R fdeny = (R)1 + (R)2 * ay + (R)2 * by + Biy;
    const R fdenz = (R)1 + (R)2 * az + (R)2 * bz + Biz;
    const double hv = 0.5 * (double)T_inf, sv = hv * (double)(Bix + Biz), sh = hv * (double)(Biy + Biz),
                 sd = hv * (double)(Bix + Biy);
    for (int z = 0; z < NZ; z++)
        for (int y = 1; y < NY; y++)
            for (int x = 0; x < NX; x++) {
                const bool xi = x > 0 && x < NX - 1, zi = z > 0 && z < NZ - 1, yi = y < NY - 1;
                const int xn = x == 0 ? x + 1 : x - 1;          // the single x-neighbour on an x face
                const int zn = z == 0 ? z + 1 : z - 1;
                R out;
                if (yi) {
                    if (xi && zi) {
                        out = L2 * ((Ti[I(x + 1, y, z)] + Ti[I(x - 1, y, z)]) / (dx * dx) +
                                    (Ti[I(x, y + 1, z)] + Ti[I(x, y - 1, z)]) / (dy * dy) +
                                    (Ti[I(x, y, z + 1)] + Ti[I(x, y, z - 1)]) / (dz * dz) + Q[I(x, y, z)] / thermal_cond);
                    } else if (!xi && zi) {          // x faces (:137-143)
                        out = (Ti[I(xn, y, z)] + ax * (Ti[I(x, y + 1, z)] + Ti[I(x, y - 1, z)]) +
                               bx * (Ti[I(x, y, z + 1)] + Ti[I(x, y, z - 1)]) + Bix * T_inf) / fdenx;
                    } else if (xi && !zi) {          // z faces (:149-155)
                        out = (Ti[I(x, y, zn)] + az * (Ti[I(x + 1, y, z)] + Ti[I(x - 1, y, z)]) +
                               bz * (Ti[I(x, y + 1, z)] + Ti[I(x, y - 1, z)]) + Biz * T_inf) / fdenz;
                    } else {                          // vertical sides (:161-175)
                        out = (R)(((double)(av * Ti[I(xn, y, z)] + bv * Ti[I(x, y, zn)] +
                                            cv * (Ti[I(x, y + 1, z)] + Ti[I(x, y - 1, z)])) + sv) * (double)V);
                    }
                } else {                              // y = NY-1 (top, open to air)
                    if (xi && zi) {                   // top face (:145-147)
                        out = (Ti[I(x, y - 1, z)] + ay * (Ti[I(x + 1, y, z)] + Ti[I(x - 1, y, z)]) +
                               by * (Ti[I(x, y, z + 1)] + Ti[I(x, y, z - 1)]) + Biy * T_inf) / fdeny;
                    } else if (xi && !zi) {           // horizontal sides (:179-185)
                        out = (R)(((double)(ah * Ti[I(x, y - 1, z)] + bh * Ti[I(x, y, zn)] +
                                            ch * (Ti[I(x - 1, y, z)] + Ti[I(x + 1, y, z)])) + sh) * (double)H);
                    } else if (!xi && zi) {           // depth sides (:189-195)
                        out = (R)(((double)(ad * Ti[I(x, y - 1, z)] + bd * Ti[I(xn, y, z)] +
                                            cd * (Ti[I(x, y, z + 1)] + Ti[I(x, y, z - 1)])) + sd) * (double)D);
                    } else {                          // vertices (:199-214)
                        out = (AV * Ti[I(xn, y, z)] + BV * Ti[I(x, y - 1, z)] + CV * Ti[I(x, y, zn)] + T_inf * BiD) * GV;
                    }
                }
                Tf[I(x, y, z)] = out;
            }
}

// reducedSum (src/headers/Tfield.h:13-28): float sum of squared updates / TSIZE.
template <class R> R reduced_sum(const R *Tf, const R *Ti, size_t n) {
    R sum = (R)0;
    for (size_t i = 0; i < n; i++) sum += m_pow2(Tf[i] - Ti[i]);
    return sum / (R)(unsigned)n;
}

template <class R> void fields_in(std::vector<cx<R>> &v, const double *src) {
    for (size_t i = 0; i < v.size(); i++) v[i] = {(R)src[2 * i], (R)src[2 * i + 1]};
}
template <class R> void fields_out(const std::vector<cx<R>> &v, double *dst) {
    if (!dst) return;
    for (size_t i = 0; i < v.size(); i++) { dst[2 * i] = (double)v[i].re; dst[2 * i + 1] = (double)v[i].im; }
}

// Solver::CWES, plain build (src/headers/Solver.h:245-276, 343-348): `nsteps` steps of the
// z-loop with z accumulated in R.  The reference saves Pump_out/Signal_out before the
// last step, so callers wanting its output pass nsteps = nz - 1.
template <class R> void cwes_plain(State<R> &st, int scheme, int nsteps) {
    const R DK = dk_of(st.k, st.Temp);
    R z = (R)0;
    for (int s = 0; s < nsteps; s++) {
        ssfm(st, z, DK, scheme, false, s, 0);
        z += st.k.dz;
    }
}

// Maps field step s to heat slice: the reference couples them one to one (heat_nz = nz).
inline int heat_slice(int s, int nz, int hz) { return hz == nz ? s : (int)(((int64_t)s * hz) / nz); }

}  // namespace

// ===========================================================================
// C ABI for ctypes (tests and the CPU baseline).  Complex arrays are interleaved
// doubles (re, im); oracle32 rounds them to float on entry.
extern "C" {

void oracle_set_percell_index(int on) { g_percell_index = on; }

double oracle_n(int crystal, int prec, double L, double T) {
    if (prec == 32) return (double)n_index<float>(crystal, (float)L, (float)T);
    return n_index<double>(crystal, L, T);
}

double oracle_dk(const shg_params *p, int prec, double T) {
    if (prec == 32) { K<float> k(*p); return (double)dk_of(k, (float)T); }
    K<double> k(*p);
    return dk_of(k, T);
}

void oracle_pump(const shg_params *p, int prec, double t_centre, double *out) {
    const size_t n2 = (size_t)p->nx * p->ny;
    if (prec == 32) {
        K<float> k(*p);
        std::vector<cx<float>> P(n2);
        set_pump<float>(k, (float)p->power, (float)p->focus, (float)p->waist, (float)t_centre, P.data());
        fields_out(P, out);
    } else {
        K<double> k(*p);
        std::vector<cx<double>> P(n2);
        set_pump<double>(k, p->power, p->focus, p->waist, t_centre, P.data());
        fields_out(P, out);
    }
}

void oracle_propagators(const shg_params *p, int prec, double temp, double *hp, double *hs) {
    const size_t n2 = (size_t)p->nx * p->ny;
    if (prec == 32) {
        K<float> k(*p);
        std::vector<cx<float>> a(n2), b(n2);
        set_propagators<float>(k, (float)temp, a.data(), b.data());
        fields_out(a, hp); fields_out(b, hs);
    } else {
        K<double> k(*p);
        std::vector<cx<double>> a(n2), b(n2);
        set_propagators<double>(k, temp, a.data(), b.data());
        fields_out(a, hp); fields_out(b, hs);
    }
}

void oracle_fft2(int nx, int ny, int prec, int sign, double *data) {
    const size_t n2 = (size_t)nx * ny;
    if (prec == 32) {
        std::vector<cx<float>> a(n2);
        fields_in(a, data); fft2d(a.data(), nx, ny, sign); fields_out(a, data);
    } else {
        std::vector<cx<double>> a(n2);
        fields_in(a, data); fft2d(a.data(), nx, ny, sign); fields_out(a, data);
    }
}

// One solverRK4 application of the plain build at position z (src/headers/Solver.h:216-228).
void oracle_rk4(const shg_params *p, int prec, double z, double *fh, double *sh) {
    if (prec == 32) {
        State<float> st(*p);
        fields_in(st.P, fh); fields_in(st.S, sh);
        const float h = p->scheme == SHG_SCHEME_STRANG ? st.k.dz * 0.5f : st.k.dz;
        solver_rk4<float>(st, (float)z, dk_of(st.k, st.Temp), h, false, 0, 0);
        fields_out(st.P, fh); fields_out(st.S, sh);
    } else {
        State<double> st(*p);
        fields_in(st.P, fh); fields_in(st.S, sh);
        const double h = p->scheme == SHG_SCHEME_STRANG ? st.k.dz * 0.5 : st.k.dz;
        solver_rk4<double>(st, z, dk_of(st.k, st.Temp), h, false, 0, 0);
        fields_out(st.P, fh); fields_out(st.S, sh);
    }
}

// Plain z-march: `nsteps` < 0 means the reference's nz - 1 (the Pump_out/Signal_out state).
void oracle_cwes(const shg_params *p, int prec, int nsteps, double *fh, double *sh) {
    if (nsteps < 0) nsteps = p->nz - 1;
    if (prec == 32) {
        State<float> st(*p);
        set_propagators<float>(st.k, st.Temp, st.Hp.data(), st.Hs.data());
        fields_in(st.P, fh); fields_in(st.S, sh);
        cwes_plain(st, p->scheme, nsteps);
        fields_out(st.P, fh); fields_out(st.S, sh);
    } else {
        State<double> st(*p);
        set_propagators<double>(st.k, st.Temp, st.Hp.data(), st.Hs.data());
        fields_in(st.P, fh); fields_in(st.S, sh);
        cwes_plain(st, p->scheme, nsteps);
        fields_out(st.P, fh); fields_out(st.S, sh);
    }
}

// FP64 power of a plane: 0.5*eps0*c*n*sum|A|^2*dx*dy (new; no reference counterpart).
double oracle_power(const shg_params *p, int which /*0 FH, 1 SH*/, double temp, const double *a) {
    K<double> k(*p);
    const double lam = which == 0 ? k.lp : k.ls;
    const double n = k.n(lam, temp);
    double s = 0.0;
    const size_t n2 = (size_t)p->nx * p->ny;
    for (size_t i = 0; i < n2; i++) s += a[2 * i] * a[2 * i] + a[2 * i + 1] * a[2 * i + 1];
    return 0.5 * k.EPS0 * k.C * n * s * k.dx * k.dy;
}

// One Jacobi sweep on a heat grid of depth hz (float arithmetic of the reference).
void oracle_jacobi_sweep(const shg_params *p, int hz, const float *Ti, const float *Q, float *Tf) {
    K<float> k(*p);
    const float dz = k.LZ / (float)(unsigned)(hz - 1);
    jacobi_sweep<float>(k, hz, dz, Tf, Ti, (float)p->t_inf, Q);
}

double oracle_reduced_sum(const float *Tf, const float *Ti, int64_t n) {
    return (double)reduced_sum<float>(Tf, Ti, (size_t)n);
}

// setQ for one slice from given fields and temperature slice (float arithmetic).
void oracle_set_q(const shg_params *p, const float *Tslice, const double *fh, const double *sh, float *Qslice) {
    State<float> st(*p);
    const size_t n2 = (size_t)p->nx * p->ny;
    st.Tf.assign(Tslice, Tslice + n2);
    st.Q.assign(n2, 0.f);
    fields_in(st.P, fh); fields_in(st.S, sh);
    set_q<float>(st, 0);
    std::memcpy(Qslice, st.Q.data(), n2 * sizeof(float));
}

// One THERMAL CWES (src/headers/Solver.h:245-276 with setDKFromTemperature, PhaseMatching.h:
// 41-54) at a given temperature T (hz slices): the fields of the z = 0 plane in fh/sh are
// replaced by the Pump_out/Signal_out state (before the last step); Q (hz slices, in/out)
// receives the heat source of every step.  Propagators at T of the grid centre, slice 0.
extern "C++" {
template <class R> void thermal_cwes(const shg_params *p, const float *T, double *fh, double *sh, float *Q) {
    State<R> st(*p);
    const K<R> &k = st.k;
    const int NX = k.NX, NY = k.NY, NZ = k.NZ;
    const int HZ = p->heat_nz > 0 ? p->heat_nz : NZ;
    const size_t n2 = (size_t)NX * NY, n3 = n2 * HZ;
    st.Tf.assign(T, T + n3);
    st.Q.assign(Q, Q + n3);
    st.DK.assign(n2 * NZ, (R)0);
    st.DKint.assign(n2, (R)0);
    for (int s = 0; s < NZ; s++) {
        const int hs = heat_slice(s, NZ, HZ);
        for (size_t i = 0; i < n2; i++) {
            const R Tv = st.Tf[(size_t)hs * n2 + i];
            st.DK[(size_t)s * n2 + i] = 2 * k.PI * (2 * k.n(k.lp, Tv) / k.lp - k.n(k.ls, Tv) / k.ls + 1 / (k.Lambda));
        }
    }
    const R Tc = st.Tf[(size_t)(NY / 2) * NX + NX / 2];
    set_propagators<R>(k, Tc, st.Hp.data(), st.Hs.data());
    fields_in(st.P, fh); fields_in(st.S, sh);
    std::vector<cx<R>> outP, outS;
    R z = 0;
    for (int s = 0; s < NZ; s++) {
        if (s == NZ - 1) { outP = st.P; outS = st.S; }
        const int hs = heat_slice(s, NZ, HZ);
        ssfm<R>(st, z, (R)0, p->scheme, true, s, hs);
        set_q<R>(st, hs);
        z += k.dz;
    }
    fields_out(outP, fh); fields_out(outS, sh);
    for (size_t i = 0; i < n3; i++) Q[i] = (float)st.Q[i];
}
}  // extern "C++"

void oracle_thermal_cwes(const shg_params *p, const float *T, double *fh, double *sh, float *Q) {
    thermal_cwes<float>(p, T, fh, sh, Q);
}

// The same march in FP64 throughout (temperatures taken from the float field).
void oracle_thermal_cwes64(const shg_params *p, const float *T, double *fh, double *sh, float *Q) {
    thermal_cwes<double>(p, T, fh, sh, Q);
}

// THERMAL Solver::run (src/headers/Solver.h:289-340) in float.
//   fh0/sh0: z = 0 fields.  If pump_from_T != 0 the pump is rebuilt each outer iteration
//   by setPump at the grid-centre temperature of slice 0 (as the reference does) and fh0
//   is ignored; sh0 is reused each outer iteration (the reference redraws random noise).
//   On return fh/sh hold the last outer iteration's output (after nz-1 steps), T the final
//   temperature (nx*ny*hz floats), Q the final heat source; stats[0] = outer iterations,
//   stats[1] = Jacobi sweeps of the last outer iteration, stats[2] = total sweeps.
int oracle_thermal_run(const shg_params *p, int pump_from_T, const double *fh0, const double *sh0,
                       double *fh, double *sh, float *T, float *Qout, int64_t *stats) {
    State<float> st(*p);
    const K<float> &k = st.k;
    const int NX = k.NX, NY = k.NY, NZ = k.NZ;
    const int HZ = p->heat_nz > 0 ? p->heat_nz : NZ;
    const size_t n2 = (size_t)NX * NY, n3 = n2 * HZ;
    const float hdz = HZ == NZ ? k.dz : k.LZ / (float)(unsigned)(HZ - 1);
    const float Temp = (float)p->temp, T_inf = (float)p->t_inf;
    st.Ti.assign(n3, Temp);                                      // initTemperature (Tfield.h:31-45)
    const float nzf = (float)(unsigned)HZ;
    for (int z = 0; z < HZ; z++) {
        const float v = (z < nzf / 2.0) ? (float)p->t_peltier1 : (float)p->t_peltier2;
        for (int y = 0; y < NY; y++)
            for (int x = 0; x < NX; x++) {
                const bool oven = p->thermal_bc == SHG_BC_OVEN_SURROUNDED
                                      ? (y == 0 || y == NY - 1 || x == 0 || x == NX - 1)   // initOvenSurrounded (:71-110)
                                      : y == 0;                   // initBottomOvensTemperature (:48-68)
                if (oven) st.Ti[(size_t)z * n2 + (size_t)y * NX + x] = v;
            }
    }
    st.Tf = st.Ti;
    st.Q.assign(n3, 0.f);                                        // initQ0
    st.DK.assign(n2 * NZ, 0.f);
    st.DKint.assign(n2, 0.f);
    const int check = p->jacobi_check > 0 ? p->jacobi_check : 5000;
    const int num_of_iter = p->max_outer > 0 ? p->max_outer : 5;
    const float tol = p->jacobi_tol > 0 ? (float)p->jacobi_tol : (float)(5e-4 / (double)n3);
    int64_t total = 0, last = 0;
    int global_count = 0, outer = 0;
    while (global_count < num_of_iter) {
        float Reduced_sum = 1.0f, aux_sum = -1.0f;               // aux_sum: reference leaves it unset
        int64_t counter_T = 0;
        bool have_aux = false;
        while (Reduced_sum >= tol) {
            jacobi_sweep<float>(k, HZ, hdz, st.Tf.data(), st.Ti.data(), T_inf, st.Q.data());
            if (counter_T % check == 0) {
                if (counter_T > 0) { aux_sum = Reduced_sum; have_aux = true; }
                Reduced_sum = reduced_sum<float>(st.Tf.data(), st.Ti.data(), n3);
                if (have_aux && Reduced_sum == aux_sum) Reduced_sum = 0.5f * tol;
            }
            st.Ti = st.Tf;                                       // copyTensor
            counter_T++;
            if (p->jacobi_max > 0 && counter_T >= p->jacobi_max) break;
        }
        total += counter_T; last = counter_T;
        // setInicialDKThermal (PhaseMatching.h:41-54) on the field steps' heat slices
        for (int s = 0; s < NZ; s++) {
            const int hs = heat_slice(s, NZ, HZ);
            for (size_t i = 0; i < n2; i++) {
                const float Tv = st.Tf[(size_t)hs * n2 + i];
                st.DK[(size_t)s * n2 + i] =
                    2 * k.PI * (2 * k.n(k.lp, Tv) / k.lp - k.n(k.ls, Tv) / k.ls + 1 / (k.Lambda));
            }
        }
        const float Tc = st.Tf[(size_t)(NY / 2) * NX + NX / 2];
        if (pump_from_T) set_pump<float>(k, (float)p->power, (float)p->focus, (float)p->waist, Tc, st.P.data());
        else fields_in(st.P, fh0);
        fields_in(st.S, sh0);
        set_propagators<float>(k, Tc, st.Hp.data(), st.Hs.data());
        std::vector<cx<float>> outP, outS;
        float z = 0.f;
        for (int s = 0; s < NZ; s++) {                           // CWES, THERMAL (Solver.h:245-276)
            if (s == NZ - 1) { outP = st.P; outS = st.S; }
            const int hs = heat_slice(s, NZ, HZ);
            ssfm<float>(st, z, 0.f, p->scheme, true, s, hs);
            set_q<float>(st, hs);                                // upDateQ (Tfield.h:356-361)
            z += k.dz;
        }
        st.P = outP; st.S = outS;
        outer++;
        if (counter_T == 1 && global_count != 0) global_count = num_of_iter;
        global_count++;
    }
    fields_out(st.P, fh); fields_out(st.S, sh);
    if (T) std::memcpy(T, st.Tf.data(), n3 * sizeof(float));
    if (Qout) std::memcpy(Qout, st.Q.data(), n3 * sizeof(float));
    if (stats) { stats[0] = outer; stats[1] = last; stats[2] = total; }
    return 0;
}

}  // extern "C"
