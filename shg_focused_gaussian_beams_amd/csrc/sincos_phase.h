// sincos_phase.h — sin and cos of a float argument for the THERMAL stage-group phases:
// the argument is reduced by pi/2 in double (x - q pi/2 with q = rint(2x/pi), a fused
// multiply-add whose error stays below 1e-10 for |x| < 1e6), and the reduced argument
// r in [-pi/4, pi/4] goes through the single-precision minimax polynomials of the Cephes
// library (S. L. Moshier, Cephes Mathematical Library, sinf.c / cosf.c).  The result is
// within 2 ulp of the correctly rounded sin/cos of the float argument (checked on the
// host against double precision), with no branches: every lane of a wave runs the same
// ~25 instructions whatever its argument.
#pragma once
#include <cmath>
#include <cstdint>

#ifndef SHG_HOST_EMULATION
#include <hip/hip_runtime.h>
#define SHG_HD __host__ __device__
#else
#define SHG_HD
#endif

namespace shg {

SHG_HD inline void sincos_phase(float x, float &s, float &c) {
    const double xd = (double)x;
    const double q = rint(xd * 0.63661977236758134308);          // 2/pi
    const float r = (float)fma(-q, 1.5707963267948966192, xd);    // x - q pi/2
    const int quad = (int)q & 3;                                  // q exact, |q| < 2^31 for |x| < 3e9
    const float z = r * r;
    const float ps = fmaf(fmaf(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f);
    const float pc = fmaf(fmaf(2.443315711809948e-5f, z, -1.388731625493765e-3f), z, 4.166664568298827e-2f);
    const float sr = fmaf(ps * z, r, r);                          // sin r
    const float cr = 1.0f + fmaf(pc, z * z, -0.5f * z);           // cos r
    // sin(r + quad pi/2), cos(r + quad pi/2)
    float sv = (quad & 1) ? cr : sr;
    float cv = (quad & 1) ? sr : cr;
    s = (quad & 2) ? -sv : sv;
    c = ((quad + 1) & 2) ? -cv : cv;
}

}  // namespace shg
