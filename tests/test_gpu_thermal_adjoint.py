"""Gradients of the SH efficiency of a THERMAL march with the temperature field held
fixed (shg_adjoint on a THERMAL context) against central finite differences.

- field: against FD of the FP32 THERMAL restatement (oracle32, the GPU forward's
  arithmetic), step 1e-2 relative, per direction.
- temperature: the float phase mismatch is a staircase in T, so FD needs the FP64
  restatement (oracle64); at 26.913 C the float and double dk agree to 7e-7 rad over
  the crystal, so both solve the same problem there.  A uniform shift of the field is
  compared with d_eta_d_temp_crystal, a shift of one heat slice with the sum of that
  slice of the per-voxel map.
- the checkpoint interval changes nothing; the temperature and heat source are untouched.
- full size (256x256, nz 400, heat 256x256x100): the field gradient against FD of the
  GPU THERMAL forward along smooth modulations of the pump.
"""
import numpy as np
import pytest

import oracle_lib as orc
from shg_focused_gaussian_beams_amd.params import thermal256_params


def _params(**kw):
    base = dict(nx=64, ny=64, nz=16, heat_nz=4, power=40.0, waist=30.0, lx=240.0, ly=240.0)
    base.update(kw)
    return thermal256_params(**base)


def _smooth_T(p, lo=26.5, hi=27.5):
    hz = p.heat_nz
    z = np.linspace(0, 1, hz)[:, None, None]
    y = np.linspace(0, 1, p.ny)[None, :, None]
    x = np.linspace(0, 1, p.nx)[None, None, :]
    return (lo + (hi - lo) * (0.5 + 0.4 * np.sin(2 * z + 3 * x) * np.cos(2 * y))).astype(np.float32)


def _inputs(gpu, p, T, seed=1):
    rng = np.random.default_rng(seed)
    pump = gpu.make_pump(p, float(T[0, p.ny // 2, p.nx // 2]))
    pump = (pump + 1e-3 * np.abs(pump).max() * (rng.standard_normal(pump.shape) + 1j * rng.standard_normal(pump.shape))
            ).astype(np.complex64)
    sh = (1e-4 * np.abs(pump).max() * (rng.standard_normal(pump.shape) + 1j * rng.standard_normal(pump.shape))
          ).astype(np.complex64)
    return pump, sh


def _eta_orc(p, T, fh, sh, prec):
    tc = float(T[0, p.ny // 2, p.nx // 2])
    _, s, _ = orc.thermal_cwes(p, T, fh, sh, prec=prec)
    return orc.power(p, 1, s, tc) / orc.power(p, 0, fh, tc)


def _adjoint(gpu, p, T, fh, sh, ckpt=0):
    with gpu.Context(p) as ctx:
        ctx.set_temperature(T)
        ctx.set_fields(fh, sh)
        Q0 = ctx.get_heat_source()
        res, g = ctx.adjoint(checkpoint_interval=ckpt)
        m = ctx.adjoint_temperature_map()
        assert np.array_equal(ctx.get_temperature(), T)
        assert np.array_equal(ctx.get_heat_source(), Q0)
        ctx.forward()
        eta_fwd = ctx.powers()[2]
    return res, g, m, eta_fwd


@pytest.mark.gpu
def test_thermal_field_gradient_vs_oracle32_fd(gpu):
    p = _params()
    T = _smooth_T(p)
    fh, sh = _inputs(gpu, p, T)
    res, g, m, eta_fwd = _adjoint(gpu, p, T, fh, sh)
    print(f"\neta adjoint {res['eta']:.8e} forward {eta_fwd:.8e}", flush=True)
    assert abs(res["eta"] - eta_fwd) <= 1e-6 * eta_fwd
    g = g.astype(np.complex128)
    rng = np.random.default_rng(7)
    e = 1e-2
    for _ in range(3):
        v = rng.standard_normal(fh.shape) + 1j * rng.standard_normal(fh.shape)
        v *= np.linalg.norm(fh) / np.linalg.norm(v)
        fd = (_eta_orc(p, T, fh + e * v, sh, 32) - _eta_orc(p, T, fh - e * v, sh, 32)) / (2 * e)
        ad = float(np.real(np.sum(np.conj(g) * v)))
        print(f"  field: fd32 {fd:+.6e} adjoint {ad:+.6e}")
        assert abs(fd - ad) / abs(fd) < 3e-3


@pytest.mark.gpu
def test_thermal_temperature_gradient_vs_oracle64_fd(gpu):
    p = _params(temp=26.913)
    assert abs(orc.dk(p, 32, p.temp) - orc.dk(p, 64, p.temp)) * p.lz < 1e-5
    T = np.full((p.heat_nz, p.ny, p.nx), 26.913, np.float32)
    fh, sh = _inputs(gpu, p, T)
    res, g, m, _ = _adjoint(gpu, p, T, fh, sh)
    assert abs(m.sum() - res["d_eta_d_temp_crystal"]) <= 1e-9 * abs(res["d_eta_d_temp_crystal"]) + 1e-15
    h = 1e-2
    fh64, sh64 = fh.astype(np.complex128), sh.astype(np.complex128)

    def eta_shift(mask, d):
        return _eta_orc(p, (T + d * mask).astype(np.float32), fh64, sh64, 64)
    allv = np.ones_like(T)
    fdT = (eta_shift(allv, h) - eta_shift(allv, -h)) / (2 * h)
    print(f"\n  uniform T: fd64 {fdT:+.6e} adjoint {res['d_eta_d_temp_crystal']:+.6e}", flush=True)
    assert abs(fdT - res["d_eta_d_temp_crystal"]) / abs(fdT) < 2e-2
    for j in (0, 2):
        mask = np.zeros_like(T)
        mask[j] = 1.0
        fd = (eta_shift(mask, h) - eta_shift(mask, -h)) / (2 * h)
        print(f"  slice {j}: fd64 {fd:+.6e} map sum {m[j].sum():+.6e}")
        assert abs(fd - m[j].sum()) / abs(fdT) < 2e-2


@pytest.mark.gpu
def test_thermal_checkpoint_interval_does_not_change_gradients(gpu):
    p = _params(nz=13)
    T = _smooth_T(p)
    fh, sh = _inputs(gpu, p, T)
    out = [_adjoint(gpu, p, T, fh, sh, ckpt=k) for k in (1, 4, 0)]
    assert out[2][0]["checkpoint_interval"] == 12 and out[2][0]["segments"] == 1  # default: whole-march tape
    r0, g0, m0, _ = out[0]
    for r, g, m, _ in out[1:]:
        assert abs(r["eta"] - r0["eta"]) <= 1e-6 * r0["eta"]
        assert np.abs(g - g0).max() <= 1e-5 * np.abs(g0).max()
        assert np.abs(m - m0).max() <= 1e-5 * np.abs(m0).max()
        assert abs(r["d_eta_d_temp"] - r0["d_eta_d_temp"]) <= 1e-5 * abs(r0["d_eta_d_temp"])


@pytest.mark.gpu
def test_phase_mask_ascent_increases_efficiency(gpu):
    from shg_focused_gaussian_beams_amd.inverse import ascend_phase_mask, mask_gradient
    p = _params(nz=20)
    T = _smooth_T(p)
    pump = gpu.make_pump(p, float(T[0, p.ny // 2, p.nx // 2]))
    mask0 = np.random.default_rng(5).uniform(-0.1, 0.1, pump.shape)
    with gpu.Context(p) as ctx:
        ctx.set_temperature(T)
        # the mask gradient is the chain rule of the field gradient: check it by FD once
        fld = (pump * np.exp(1j * mask0)).astype(np.complex64)
        ctx.set_fields(fld, np.zeros_like(fld))
        res, g = ctx.adjoint()
        d = mask_gradient(g, fld)
        v = np.random.default_rng(6).standard_normal(pump.shape)
        e = 1e-2
        etas = []
        for sgn in (1, -1):
            f = (pump * np.exp(1j * (mask0 + sgn * e * v))).astype(np.complex64)
            ctx.set_fields(f, np.zeros_like(f))
            ctx.forward()
            etas.append(ctx.powers()[2])
        fd = (etas[0] - etas[1]) / (2 * e)
        print(f"\n  mask: fd {fd:+.6e} adjoint {np.sum(d * v):+.6e}", flush=True)
        assert abs(fd - np.sum(d * v)) <= 5e-3 * abs(fd)
        mask, hist = ascend_phase_mask(ctx, pump, mask0, 4, step=0.05)
    print("  eta history", ["%.6e" % x for x in hist])
    assert hist[-1] > hist[0]


@pytest.mark.gpu
def test_concurrent_phase_mask_ascents_match_one_at_a_time(gpu):
    # the batched ascent (one context and host thread per instance, all at once) follows
    # the same iterates as each instance ascending alone
    from shg_focused_gaussian_beams_amd.inverse import ascend_phase_mask, ascend_phase_masks
    p = _params(nz=20)
    T = _smooth_T(p)
    pump = gpu.make_pump(p, float(T[0, p.ny // 2, p.nx // 2]))
    masks = [np.random.default_rng(10 + k).uniform(-0.1, 0.1, pump.shape) for k in range(3)]
    ctxs = [gpu.Context(p) for _ in range(3)]
    try:
        for c in ctxs:
            c.set_temperature(T)
        together = ascend_phase_masks(ctxs, [pump] * 3, masks, 3, step=0.05)
        for k, c in enumerate(ctxs):
            mask, etas = ascend_phase_mask(c, pump, masks[k], 3, step=0.05)
            print(f"\n  instance {k}: eta {etas[0]:.6e} -> {etas[-1]:.6e}", flush=True)
            assert np.allclose(together[k][0], mask, rtol=0, atol=1e-5), k
            assert np.allclose(together[k][1], etas, rtol=1e-5, atol=0), k
    finally:
        for c in ctxs:
            c.close()


@pytest.mark.gpu
def test_thermal_field_gradient_full_size_vs_gpu_fd(gpu):
    # THERMAL at its full size (256x256, nz 400, heat grid 256x256x100) on a temperature
    # field with an on-axis hot spot: the frozen-T adjoint against central FD of the GPU
    # THERMAL forward, step 1e-3 relative, along 32 smooth modulations of the pump (the
    # pump times random 2D Fourier modes |k| <= 2), along which eta moves at O(eta)
    p = thermal256_params(heat_nz=100)
    z = np.linspace(0, 1, p.heat_nz)[:, None, None]
    y = np.linspace(0, 1, p.ny)[None, :, None]
    x = np.linspace(0, 1, p.nx)[None, None, :]
    T = (27.0 + 0.5 * np.exp(-((x - .5) ** 2 + (y - .5) ** 2) / 0.02) * (1 - 0.5 * z)).astype(np.float32)
    fh, sh = _inputs(gpu, p, T)
    rng = np.random.default_rng(8)
    yy, xx = np.meshgrid(np.arange(p.ny) / p.ny, np.arange(p.nx) / p.nx, indexing="ij")
    with gpu.Context(p) as ctx:
        ctx.set_temperature(T)
        ctx.set_fields(fh, sh)
        res, g = ctx.adjoint()
        g = g.astype(np.complex128)
        fd, ad = [], []
        for _ in range(32):
            m = sum((rng.standard_normal() + 1j * rng.standard_normal()) * np.exp(2j * np.pi * (kx * xx + ky * yy))
                    for ky in range(-2, 3) for kx in range(-2, 3))
            v = fh * m
            v *= np.linalg.norm(fh) / np.linalg.norm(v)
            ad.append(float(np.real(np.sum(np.conj(g) * v))))
            e = 1e-3
            etas = []
            for sgn in (1, -1):
                ctx.set_fields((fh + sgn * e * v).astype(np.complex64), sh)
                ctx.forward()
                etas.append(ctx.powers()[2])
            fd.append((etas[0] - etas[1]) / (2 * e))
    fd, ad = np.array(fd), np.array(ad)
    rel = np.linalg.norm(fd - ad) / np.linalg.norm(ad)
    print(f"\neta {res['eta']:.6e}: 32 smooth directions, FD step 1e-3 vs adjoint relL2 {rel:.2e}", flush=True)
    assert rel < 5e-3


@pytest.mark.gpu
def test_thermal_field_only_adjoint_equals_full_adjoint(gpu):
    p = _params(nz=20)
    T = _smooth_T(p)
    fh, sh = _inputs(gpu, p, T)
    with gpu.Context(p) as ctx:
        ctx.set_temperature(T)
        ctx.set_fields(fh, sh)
        res, g_full = ctx.adjoint()
        m = ctx.adjoint_temperature_map()
        eta, g_field = ctx.adjoint_field()
        with pytest.raises(gpu.ShgError):                              # no map from a field-only run
            ctx.adjoint_temperature_map()
        res2, _ = ctx.adjoint()
        m2 = ctx.adjoint_temperature_map()
    # the same lambda to rounding (the field-only kernels are compiled differently)
    assert orc.rel_l2(g_field, g_full) <= 1e-5 and eta == res["eta"]
    assert np.array_equal(m, m2) and res2["d_eta_d_temp"] == res["d_eta_d_temp"]
