"""The two-sweep Jacobi pass (z-marching LDS tiles, k_jacobi2) against the oracle's float
sweep (heatEquationTopOpen, src/headers/Tfield.h:113-221; the surrounded ovens,
:224-258): every layout bit-exact over even and odd sweep counts (odd counts end with a
single sweep), on grids that tiles and z-chunks do not divide evenly, at the THERMAL
workload's heat grid and the inverse-design one; the residual of the last sweep equals
the FP64 mean squared update; the tuner's pick is one of the layouts and leaves the
temperature untouched.
"""
import numpy as np
import pytest

import oracle_lib as orc
from shg_focused_gaussian_beams_amd.params import (thermal256_params, inverse512_params, SHG_BC_TOP_OPEN,
                                                   SHG_BC_OVEN_SURROUNDED)


def _field(p, hz, seed):
    rng = np.random.default_rng(seed)
    z = np.linspace(0, 1, hz)[:, None, None]
    y = np.linspace(0, 1, p.ny)[None, :, None]
    x = np.linspace(0, 1, p.nx)[None, None, :]
    T = (27.0 + 2.0 * np.exp(-((x - .5) ** 2 + (y - .5) ** 2) / 0.05) * (1 - 0.5 * z)
         + 0.05 * rng.standard_normal((hz, p.ny, p.nx))).astype(np.float32)
    Q = (1e-8 * rng.random((hz, p.ny, p.nx))).astype(np.float32)
    return T, Q


def _oracle(p, hz, T, Q, n):
    o = T
    prev = T
    for _ in range(n):
        prev = o
        o = orc.jacobi_sweep(p, hz, o, Q)
    r = float(np.mean((o.astype(np.float64) - prev.astype(np.float64)) ** 2))
    return o, r


@pytest.mark.gpu
@pytest.mark.parametrize("n,hz,bc", [(64, 12, SHG_BC_TOP_OPEN), (64, 37, SHG_BC_OVEN_SURROUNDED),
                                     (128, 23, SHG_BC_TOP_OPEN), (256, 9, SHG_BC_TOP_OPEN),
                                     (512, 5, SHG_BC_OVEN_SURROUNDED),
                                     # interior tiles (unrolled march) with chunks that start
                                     # off a multiple of four planes, both boundary kinds
                                     (256, 37, SHG_BC_TOP_OPEN), (256, 30, SHG_BC_OVEN_SURROUNDED)])
def test_every_jacobi_layout_bit_exact(gpu, n, hz, bc):
    p = thermal256_params(nx=n, ny=n, nz=max(hz, 2), heat_nz=hz, lx=2.0 * n, ly=1.5 * n, thermal_bc=bc)
    print(f"\njacobi layouts n={n} hz={hz} bc={bc}", flush=True)
    T, Q = _field(p, hz, n + hz)
    o4, r4 = _oracle(p, hz, T, Q, 4)
    o5, r5 = _oracle(p, hz, T, Q, 5)
    with gpu.Context(p) as ctx:
        ctx.set_heat_source(Q)
        for layout in range(gpu.SHG_JACOBI_LAYOUTS):
            ctx.set_jacobi_layout(layout)
            for ns, o, r_ref in ((4, o4, r4), (5, o5, r5)):
                ctx.set_temperature(T)
                r = ctx.jacobi(ns)
                g = ctx.get_temperature()
                print(f"  layout {layout} sweeps {ns}: max diff {np.abs(g - o).max():.3e} resid {r:.6e} ref {r_ref:.6e}")
                assert np.array_equal(g, o), (layout, ns)
                assert abs(r - r_ref) <= 1e-9 * r_ref, (layout, ns, r, r_ref)


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["256x256x100", "512x512x200"])
def test_tuned_jacobi_full_size_bit_exact(gpu, which):
    p = thermal256_params(heat_nz=100) if which == "256x256x100" else inverse512_params()
    hz = p.heat_nz
    T, Q = _field(p, hz, 5)
    with gpu.Context(p) as ctx:
        ctx.set_temperature(T)
        ctx.set_heat_source(Q)
        r = ctx.jacobi(4)
        layout = ctx.jacobi_layout()
        g = ctx.get_temperature()
    o, r_ref = _oracle(p, hz, T, Q, 4)
    print(f"\n{which}: tuned layout {layout}, 4 sweeps bit-exact {np.array_equal(g, o)}", flush=True)
    assert 0 <= layout < gpu.SHG_JACOBI_LAYOUTS
    assert np.array_equal(g, o)
    assert abs(r - r_ref) <= 1e-9 * r_ref
