"""The Jacobi kernels divide by the sweep's constant weights in three operations:
q = RN(a r) with r = RN(1/d), e = fma(-q, d, a) (exact), q' = fma(e, r, q)
(div_const, shg_focused_gaussian_beams_amd/csrc/thermal_kernels.hip), which must equal the
IEEE float quotient RN(a/d) that the reference's sweep computes (Markstein's theorem).
CPU check: the sequence with its fused multiply-adds emulated in extended precision
against numpy's float32 division, for the weights of the heat grids in use (dx^2, dy^2,
dz^2, k_th and the face denominators, from the same float formulas) and random divisors,
each over numerators spread across the magnitudes a sweep meets (temperatures ~1e1,
weighted neighbour sums ~1e5, heat sources ~1e-8 and below) and random bit patterns.
"""
import numpy as np
import pytest

from shg_focused_gaussian_beams_amd.params import thermal256_params, inverse512_params, default_params

f32, ld = np.float32, np.longdouble


def _div_const(a, d):
    r = f32(1) / d
    q = (a * r).astype(f32)
    e = (ld(a) - ld(q) * ld(d)).astype(f32)            # fma(-q, d, a): exact in extended precision
    return (ld(e) * ld(r) + ld(q)).astype(f32)         # fma(e, r, q): one rounding to float


def _weights(p, hz):
    dx = f32(p.lx) / f32(p.nx - 1)
    dy = f32(p.ly) / f32(p.ny - 1)
    dz = f32(p.lz) / f32(hz - 1)
    kth, h = f32(p.k_th), f32(p.h_conv)
    Bix, Biy, Biz = h * dx / kth, h * dy / kth, h * dz / kth
    two = f32(2)
    ax, bx = dx * dx / (two * dy * dy), dx * dx / (two * dz * dz)
    ay, by = dy * dy / (two * dx * dx), dy * dy / (two * dz * dz)
    az, bz = dz * dz / (two * dx * dx), dz * dz / (two * dy * dy)
    one = f32(1)
    return [dx * dx, dy * dy, dz * dz, kth, one + two * ax + two * bx + Bix, one + two * ay + two * by + Biy,
            one + two * az + two * bz + Biz]


@pytest.mark.parametrize("which", ["thermal256_heat100", "inverse512_heat200", "reference_default"])
def test_division_by_known_divisor_is_correctly_rounded(which):
    if which == "thermal256_heat100":
        p, hz = thermal256_params(heat_nz=100), 100
    elif which == "inverse512_heat200":
        p, hz = inverse512_params(), 200
    else:
        p, hz = default_params(thermal=1), 200
    rng = np.random.default_rng(hz)
    divisors = _weights(p, hz) + list(f32(10.0) ** rng.uniform(-8, 6, 40).astype(f32))
    mags = f32(10.0) ** rng.uniform(-30, 8, 400_000).astype(f32)
    a = (mags * rng.uniform(0.5, 2.0, mags.size).astype(f32)).astype(f32)
    bits = rng.integers(0x0D000000, 0x4F000000, 200_000, dtype=np.uint32).view(f32)   # normal, |a| < 2^31
    a = np.concatenate([a, bits, -a[:1000], np.array([0.0, 25.0, 27.0, 31.0], f32)])
    for d in divisors:
        d = f32(d)
        got = _div_const(a, d)
        ref = (a / d).astype(f32)
        bad = np.nonzero(got != ref)[0]
        assert bad.size == 0, (which, float(d), a[bad[:5]], got[bad[:5]], ref[bad[:5]])
