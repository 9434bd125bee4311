"""Seeded random draws over the parameter space, each solved on the GPU and by the oracle
(FP32 restatement): grid size, steps, crystal, grating period, scheme, walk-off, losses,
two-photon absorption, power, waist, focus, window, temperature; THERMAL draws also pick
the heat depth, the heat boundary and a random smooth temperature field.  Every draw must
meet the parity bar (fields within relative L2 1e-4; THERMAL also the heat source),
whatever layout the tuner picks for it.  The draws are fixed by their seeds, so a failure
names a reproducible case.  Inputs are the workloads' (pump with 1e-6 relative noise,
SH = 0).

White-noise SH input (test_white_noise_signal_draws): an SH of white noise at 1e-3 of the
pump with weak conversion leaves an output SH dominated by the noise's highest spatial
frequencies, whose diffraction phase turns by 400-1500 rad per step in these draws.  The
reference evaluates that phase in float per cell (src/headers/Efields.h:121-122); this
build evaluates the separable factors in FP64 and rounds them once.  The draws run at a
temperature where the float and double phase mismatch agree, and print the three
distances GPU-oracle32, GPU-oracle64 and oracle32-oracle64: all three sit at float
rounding (measured <= 1e-5), so the draws are held to the same 1e-4 bar as the others.
"""
import numpy as np
import pytest

import oracle_lib as orc
from shg_focused_gaussian_beams_amd.params import (plain128_params, thermal256_params, SHG_CRYSTAL_SPPLT,
                                                   SHG_CRYSTAL_PPLN, SHG_SCHEME_REFERENCE, SHG_SCHEME_STRANG,
                                                   SHG_BC_TOP_OPEN, SHG_BC_OVEN_SURROUNDED)

TOL = 1e-4


def _draw(seed, thermal):
    rng = np.random.default_rng(seed)
    n = int(rng.choice([64, 128]))
    crystal = int(rng.choice([SHG_CRYSTAL_SPPLT, SHG_CRYSTAL_PPLN]))
    waist = float(rng.uniform(15.0, 40.0))
    kw = dict(nx=n, ny=n, nz=int(rng.integers(6, 24)), crystal=crystal,
              qpm_period=float(rng.uniform(7.95, 8.05) if crystal == SHG_CRYSTAL_SPPLT else rng.uniform(6.9, 6.96)),
              scheme=int(rng.choice([SHG_SCHEME_REFERENCE, SHG_SCHEME_STRANG])),
              power=float(rng.uniform(1.0, 60.0)), waist=waist, lx=waist * float(rng.uniform(5.0, 9.0)),
              temp=float(rng.uniform(25.0, 60.0)), focus=float(rng.uniform(0.25, 0.75)) * 30000.0,
              alpha_p=float(rng.uniform(0.0, 2e-7)), alpha_s=float(rng.uniform(0.0, 5e-6)),
              beta_s=float(rng.choice([0.0, 1e-7, 1e-4])))
    kw["ly"] = kw["lx"]
    if rng.random() < 0.5:
        kw["rho_p"], kw["rho_s"] = float(rng.uniform(0, 3e-3)), float(rng.uniform(0, 3e-3))
    if not thermal:
        return plain128_params(**kw), rng
    kw["heat_nz"] = int(rng.integers(2, kw["nz"] + 1))
    kw["thermal_bc"] = int(rng.choice([SHG_BC_TOP_OPEN, SHG_BC_OVEN_SURROUNDED]))
    kw["t_peltier1"] = kw["t_peltier2"] = kw["temp"]
    return thermal256_params(**kw), rng


def _fields(gpu, p, rng):
    pump = gpu.make_pump(p)
    amp = 1e-6 * np.abs(pump).max()
    fh = (pump + amp * (rng.standard_normal(pump.shape) + 1j * rng.standard_normal(pump.shape))).astype(np.complex64)
    return fh, np.zeros_like(fh)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(100, 112))
def test_random_plain_draw_matches_oracle(gpu, seed):
    p, rng = _draw(seed, thermal=False)
    print(f"\nplain draw {seed}: n={p.nx} nz={p.nz} crystal={p.crystal} scheme={p.scheme} P={p.power:.1f} "
          f"w0={p.waist:.1f} rho=({p.rho_p:.1e},{p.rho_s:.1e})", flush=True)
    fh0, sh0 = _fields(gpu, p, rng)
    with gpu.Context(p) as ctx:
        ctx.set_fields(fh0, sh0)
        ctx.forward()
        fh, sh = ctx.get_fields()
    rfh, rsh = orc.cwes(p, fh0, sh0, prec=32)
    e = (orc.rel_l2(fh, rfh), orc.rel_l2(sh, rsh))
    print(f"  relL2 FH {e[0]:.2e} SH {e[1]:.2e}")
    assert max(e) <= TOL, (seed, e)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(200, 208))
def test_random_thermal_draw_matches_oracle(gpu, seed):
    p, rng = _draw(seed, thermal=True)
    print(f"\nTHERMAL draw {seed}: n={p.nx} nz={p.nz} heat_nz={p.heat_nz} bc={p.thermal_bc} crystal={p.crystal} "
          f"scheme={p.scheme} P={p.power:.1f} w0={p.waist:.1f}", flush=True)
    hz = p.heat_nz
    z = np.linspace(0, 1, hz)[:, None, None]
    y = np.linspace(0, 1, p.ny)[None, :, None]
    x = np.linspace(0, 1, p.nx)[None, None, :]
    a, b, c = rng.uniform(0.5, 3.0, 3)
    T = (p.temp + rng.uniform(0.2, 2.0) * np.sin(a * z + b * x) * np.cos(c * y)).astype(np.float32)
    fh0, sh0 = _fields(gpu, p, rng)
    with gpu.Context(p) as ctx:
        ctx.set_temperature(T)
        ctx.set_fields(fh0, sh0)
        ctx.forward()
        fh, sh = ctx.get_fields()
        Q = ctx.get_heat_source()
    rfh, rsh, rQ = orc.thermal_cwes(p, T, fh0, sh0)
    e = (orc.rel_l2(fh, rfh), orc.rel_l2(sh, rsh), float(np.linalg.norm(Q - rQ) / np.linalg.norm(rQ)))
    print(f"  relL2 FH {e[0]:.2e} SH {e[1]:.2e} Q {e[2]:.2e}")
    assert max(e) <= TOL, (seed, e)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(300, 304))
def test_white_noise_signal_draws(gpu, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.choice([64, 128]))
    waist = float(rng.uniform(15.0, 30.0))
    # Lambda = 8.0265 um at 26.913 C: the float and double phase mismatch agree there
    # a fine window and long steps: the highest SH frequencies turn by 400-1500 rad per step
    p = plain128_params(nx=n, ny=n, nz=int(rng.integers(8, 16)), power=float(rng.uniform(0.5, 3.0)), waist=waist,
                        lx=waist * float(rng.uniform(3.0, 4.5)), temp=26.913, alpha_p=1e-7, alpha_s=4e-6, beta_s=1e-7)
    p.ly = p.lx
    assert abs(orc.dk(p, 32, p.temp) - orc.dk(p, 64, p.temp)) * p.lz < 1e-5
    pump = gpu.make_pump(p)
    amp = 1e-3 * np.abs(pump).max()
    sh0 = (amp * (rng.standard_normal(pump.shape) + 1j * rng.standard_normal(pump.shape))).astype(np.complex64)
    print(f"\nwhite-noise SH draw {seed}: n={n} nz={p.nz} P={p.power:.2f} w0={waist:.1f}", flush=True)
    with gpu.Context(p) as ctx:
        ctx.set_fields(pump, sh0)
        ctx.forward()
        fh, sh = ctx.get_fields()
    rfh, rsh = orc.cwes(p, pump, sh0, prec=32)
    dfh, dsh = orc.cwes(p, pump.astype(np.complex128), sh0.astype(np.complex128), prec=64)
    e32 = max(orc.rel_l2(fh, rfh), orc.rel_l2(sh, rsh))
    e64 = max(orc.rel_l2(fh, dfh), orc.rel_l2(sh, dsh))
    d32 = max(orc.rel_l2(rfh, dfh), orc.rel_l2(rsh, dsh))
    print(f"  relL2 GPU-oracle32 {e32:.2e}, GPU-oracle64 {e64:.2e}, oracle32-oracle64 {d32:.2e}")
    assert e32 <= TOL, (seed, e32)
