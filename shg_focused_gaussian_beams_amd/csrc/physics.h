// physics.h — host-side constants of the SHG solve, evaluated once per context.
//
// The reference declares its material and grid constants as `const real_t` (float)
// initialised from double expressions (src/headers/Crystal.h:18-53) and evaluates the
// refractive index with float coefficients but double pow/sqrt (src/headers/Crystal.h:
// 55-71).  The phase mismatch is a float difference of three O(4) terms
// (src/headers/PhaseMatching.h:63), so reproducing its rounding matters at the 0.06 rad
// level over 30 mm: every value below is produced with the same float/double typing.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/shg.h"

namespace shg {
namespace phys {

// MgO:sPPLT extraordinary index (Bruner et al., Opt. Lett. 28, 194 (2003)), in the
// reference's typing: float coefficients, b(T)/c(T) rounded to float, double sum/sqrt.
inline float index_spplt(float lambda, float temp) {
    const float A = 4.502483f, B = 0.007294f, C = 0.185087f, D = -0.02357f;
    const float E = 0.073423f, F = 0.199595f, G = 0.001f, H = 7.99724f;
    const double tk = (double)temp + 273.15;
    const float b = (float)(3.483933e-8 * (tk * tk));
    const float c = (float)(1.607839e-8 * (tk * tk));
    const double l2 = (double)lambda * (double)lambda;
    const float bb = B + b, cc = C + c;
    const double s = (double)A + (double)bb / (l2 - (double)cc * (double)cc) + (double)E / (l2 - (double)F * (double)F) +
                     (double)G / (l2 - (double)H * (double)H) + (double)D * l2;
    return (float)std::sqrt(s);
}

// MgO:PPLN (5 mol% MgO:CLN) extraordinary index, Gayer et al., Appl. Phys. B 91, 343
// (2008): n^2 = a1 + b1 f + (a2 + b2 f)/(l^2 - (a3 + b3 f)^2) + (a4 + b4 f)/(l^2 - a5^2)
// - a6 l^2, f = (T - 24.5)(T + 570.82).  Not in the reference; double throughout.
__host__ __device__ inline double index_ppln(double lambda, double temp) {
    const double f = (temp - 24.5) * (temp + 570.82);
    const double l2 = lambda * lambda, p = 0.2020 + 6.113e-8 * f;
    return std::sqrt(5.756 + 2.860e-6 * f + (0.0983 + 4.700e-8 * f) / (l2 - p * p) +
                     (189.32 + 1.516e-4 * f) / (l2 - 12.52 * 12.52) - 1.32e-2 * l2);
}

// Double-precision index for the FP64 power reductions and the temperature derivatives
// of the adjoint (same formulas, no float rounding; also compiled for the device).
__host__ __device__ inline double index_spplt_d(double lambda, double temp) {
    const double tk = temp + 273.15, b = 3.483933e-8 * tk * tk, c = 1.607839e-8 * tk * tk;
    const double l2 = lambda * lambda, cc = 0.185087 + c;
    return std::sqrt(4.502483 + (0.007294 + b) / (l2 - cc * cc) + 0.073423 / (l2 - 0.199595 * 0.199595) +
                     0.001 / (l2 - 7.99724 * 7.99724) - 0.02357 * l2);
}
__host__ __device__ inline double index_d(int crystal, double lambda, double temp) {
    return crystal == SHG_CRYSTAL_PPLN ? index_ppln(lambda, temp) : index_spplt_d(lambda, temp);
}

inline float index(int crystal, float lambda, float temp) {
    return crystal == SHG_CRYSTAL_PPLN ? (float)index_ppln(lambda, temp) : index_spplt(lambda, temp);
}

// The reference's global constants at run-time sizes.
struct Consts {
    int nx, ny, nz, crystal;
    float pi, c, eps0, lp, ls, lx, ly, lz, dx, dy, dz, deff, dq, period;
    float alpha_p, alpha_s, beta_s, rho_p, rho_s;
    explicit Consts(const shg_params &p) {
        nx = p.nx; ny = p.ny; nz = p.nz; crystal = p.crystal;
        pi = (float)3.14159265358979323846;
        c = (float)(299792458 * 1E6 / 1E12);               // um/ps
        eps0 = (float)(8.8541878128E-12 * 1E12 / 1E6);     // W ps/(V^2 um)
        lp = (float)p.lambda_p;
        ls = (float)(0.5 * (double)lp);
        lx = (float)p.lx; ly = (float)p.ly; lz = (float)p.lz;
        dx = lx / (float)(unsigned)(nx - 1);
        dy = ly / (float)(unsigned)(ny - 1);
        dz = lz / (float)(unsigned)(nz - 1);
        deff = (float)p.deff;
        dq = (float)((2.0 / (double)pi) * (double)deff);
        period = (float)p.qpm_period;
        alpha_p = (float)p.alpha_p; alpha_s = (float)p.alpha_s; beta_s = (float)p.beta_s;
        rho_p = (float)p.rho_p; rho_s = (float)p.rho_s;
    }
    float n_p(float T) const { return index(crystal, lp, T); }
    float n_s(float T) const { return index(crystal, ls, T); }
    // dk = 2 pi (2 n_p/lp - n_s/ls + 1/Lambda), float (src/headers/PhaseMatching.h:63)
    float dk(float T) const { return 2.0f * pi * (2.0f * n_p(T) / lp - n_s(T) / ls + 1.0f / period); }
    // coupling magnitudes of dAdz: pi*dQ/(n_p*lp) and 0.5*pi*dQ/(n_s*ls) (Solver.h:75,77)
    float kappa_p(float T) const { return ((1.0f * pi) * dq) / (n_p(T) * lp); }
    float kappa_s(float T) const { return ((0.5f * pi) * dq) / (n_s(T) * ls); }
};

// setPump (src/headers/Efields.h:74-104), evaluated in the reference's float arithmetic:
// A(x,y) = Ap0/M * exp(-(x'^2 dx^2 + y'^2 dy^2)/(w0^2 M)), M = 1 - i f/zR.
inline void make_pump(const Consts &k, float power, float focus, float waist, float temp, float *out) {
    const float ux = (float)(0.5 * k.nx), uy = (float)(0.5 * k.ny);
    const float w2 = waist * waist;
    const float np = k.n_p(temp);
    const float zr = k.pi * np * w2 / k.lp;
    const float ap0 = sqrtf(4.0f * power / (k.eps0 * k.c * k.pi * np * w2));
    const float eta = focus / zr;
    const float mr = 1.0f - 0.0f * eta, mi = -(1.0f * eta);       // M = 1 - i*eta
    // w2*M and its reciprocal as the reference's real/complex operator computes it
    const float qr = w2 * mr, qi = w2 * mi;
    const float qden = qr * qr + qi * qi;
    const float mden = mr * mr + mi * mi;
    const float ar = (ap0 * mr) / mden, ai = (-ap0 * mi) / mden;  // Ap0 / M
    for (int iy = 0; iy < k.ny; iy++)
        for (int ix = 0; ix < k.nx; ix++) {
            const float xx = ((float)(unsigned)ix - ux) * k.dx, yy = ((float)(unsigned)iy - uy) * k.dy;
            const float e = -powf(xx, 2.0f) - powf(yy, 2.0f);
            const float er = (e * qr) / qden, ei = (-e * qi) / qden;   // e / (w2*M)
            const float mag = expf(er);
            const float gr = mag * cosf(ei), gi = mag * sinf(ei);
            const size_t o = 2 * ((size_t)iy * k.nx + ix);
            out[o] = ar * gr - ai * gi;
            out[o + 1] = ar * gi + ai * gr;
        }
}

// Separable diffraction operator for one dz on the fftshifted grid (DC at index 0),
// src/headers/Efields.h:107-126: H = exp(-i dz (2 pi^2/k (fx^2 + fy^2) + 2 pi fx tan rho))
// = Gx(fx) Gy(fy).  Evaluated in double; Gy also carries the 1/(nx*ny) of FFTscale.
// Layout: Gx_p[nx], Gy_p[ny], Gx_s[nx], Gy_s[ny].
inline std::vector<float> propagator_factors(const Consts &k, float temp, double dz_prop) {
    const double PI = 3.14159265358979323846;
    const int n = k.nx;
    std::vector<float> g(8 * (size_t)n);
    const double kk[2] = {2.0 * PI * (double)k.n_p(temp) / (double)k.lp, 2.0 * PI * (double)k.n_s(temp) / (double)k.ls};
    const double rho[2] = {(double)k.rho_p, (double)k.rho_s};
    const double dfx = 1.0 / (double)k.dx / n, dfy = 1.0 / (double)k.dy / n;
    const double scale = 1.0 / ((double)n * (double)n);
    for (int f = 0; f < 2; f++)
        for (int i = 0; i < n; i++) {
            const int orig = (i + n / 2) % n;                      // index before fftshift
            const double fx = (orig - 0.5 * n) * dfx, fy = (orig - 0.5 * n) * dfy;
            const double phx = -dz_prop * (2.0 * PI * PI / kk[f] * fx * fx + 2.0 * PI * fx * std::tan(rho[f]));
            const double phy = -dz_prop * (2.0 * PI * PI / kk[f] * fy * fy);
            float *gx = g.data() + (size_t)(2 * f) * 2 * n, *gy = gx + 2 * n;
            gx[2 * i] = (float)std::cos(phx);
            gx[2 * i + 1] = (float)std::sin(phx);
            gy[2 * i] = (float)(std::cos(phy) * scale);
            gy[2 * i + 1] = (float)(std::sin(phy) * scale);
        }
    return g;
}

// Twiddles W_n^j = exp(-2 pi i j/n), j < n, from double.
inline std::vector<float> twiddle_table(int n) {
    std::vector<float> w(2 * (size_t)n);
    for (int j = 0; j < n; j++) {
        const double th = -2.0 * 3.14159265358979323846 * j / n;
        w[2 * j] = (float)std::cos(th);
        w[2 * j + 1] = (float)std::sin(th);
    }
    return w;
}

// Per-step coefficients (c1, c2) of the plain build: z accumulated in float (z += dz,
// src/headers/Solver.h:271), theta = fl(z*dk), e = (cosf, sinf) as CpxExp does;
// c1 = i kappa_p e^{-i theta}, c2 = i kappa_s e^{+i theta}.
inline std::vector<float> step_coefficients(const Consts &k, float temp, int nsteps) {
    std::vector<float> c(4 * (size_t)(nsteps > 0 ? nsteps : 1));
    const float dk = k.dk(temp), kp = k.kappa_p(temp), ks = k.kappa_s(temp);
    float z = 0.0f;
    for (int s = 0; s < nsteps; s++) {
        const float th = z * dk;
        const float co = cosf(th), si = sinf(th);
        c[4 * s + 0] = kp * si;
        c[4 * s + 1] = kp * co;
        c[4 * s + 2] = -ks * si;
        c[4 * s + 3] = ks * co;
        z += k.dz;
    }
    return c;
}

// Heat-grid z-spacing: the reference couples the heat grid to the field grid (dz); a
// heat grid of hz slices spans the same crystal length.
inline float heat_dz(const Consts &k, int hz) { return hz == k.nz ? k.dz : k.lz / (float)(unsigned)(hz - 1); }

// Field step s -> heat slice (one to one when hz == nz).
inline int heat_slice(int s, int nz, int hz) { return hz == nz ? s : (int)(((int64_t)s * hz) / nz); }

// Jacobi weights of heatEquationTopOpen (src/headers/Tfield.h:118-125) with the
// reference's typing: float products/quotients, the edge denominators and 0.5*T_inf*Bi
// terms in double, V/H/D rounded to float.
template <class JC> inline JC jacobi_const(const Consts &k, float hdz, float k_th, float h_conv, float t_inf) {
    JC c{};
    const float dx = k.dx, dy = k.dy, dz = hdz;
    const float Bix = h_conv * dx / k_th, Biy = h_conv * dy / k_th, Biz = h_conv * dz / k_th;
    c.ax = dx * dx / (2.0f * dy * dy); c.ay = dy * dy / (2.0f * dx * dx); c.az = dz * dz / (2.0f * dx * dx);
    c.bx = dx * dx / (2.0f * dz * dz); c.by = dy * dy / (2.0f * dz * dz); c.bz = dz * dz / (2.0f * dy * dy);
    c.av = dz / (2.0f * dx); c.bv = dx / (2.0f * dz); c.cv = dx * dz / (4.0f * dy * dy);
    c.V = (double)(float)(1.0 / ((double)(c.av + c.bv + 2.0f * c.cv) + 0.5 * (double)(Bix + Biz)));
    c.ah = dz / (2.0f * dy); c.bh = dy / (2.0f * dz); c.ch = dy * dz / (4.0f * dx * dx);
    c.H = (double)(float)(1.0 / ((double)(c.ah + c.bh + 2.0f * c.ch) + 0.5 * (double)(Biy + Biz)));
    c.ad = dx / (2.0f * dy); c.bd = dy / (2.0f * dx); c.cd = dx * dy / (4.0f * dz * dz);
    c.D = (double)(float)(1.0 / ((double)(c.ad + c.bd + 2.0f * c.cd) + 0.5 * (double)(Bix + Biy)));
    c.AV = dy * dz / dx; c.BV = dx * dz / dy; c.CV = dx * dy / dz;
    const float BiD = Bix * dy + Biy * dz + Biz * dx;
    c.GV = 1.0f / (c.AV + c.BV + c.CV + BiD);
    c.L2 = 1.0f / (2.0f / (dx * dx) + 2.0f / (dy * dy) + 2.0f / (dz * dz));
    c.dxx = dx * dx; c.dyy = dy * dy; c.dzz = dz * dz; c.kth = k_th;
    c.fdenx = 1.0f + 2.0f * c.ax + 2.0f * c.bx + Bix;
    c.fdeny = 1.0f + 2.0f * c.ay + 2.0f * c.by + Biy;
    c.fdenz = 1.0f + 2.0f * c.az + 2.0f * c.bz + Biz;
    c.bixT = Bix * t_inf; c.biyT = Biy * t_inf; c.bizT = Biz * t_inf;
    c.tinfBiD = t_inf * BiD;
    c.rdxx = 1.0f / c.dxx; c.rdyy = 1.0f / c.dyy; c.rdzz = 1.0f / c.dzz; c.rkth = 1.0f / c.kth;
    c.rfdenx = 1.0f / c.fdenx; c.rfdeny = 1.0f / c.fdeny; c.rfdenz = 1.0f / c.fdenz;
    const double hv = 0.5 * (double)t_inf;
    c.sv = hv * (double)(Bix + Biz);
    c.sh = hv * (double)(Biy + Biz);
    c.sd = hv * (double)(Bix + Biy);
    return c;
}

}  // namespace phys
}  // namespace shg
