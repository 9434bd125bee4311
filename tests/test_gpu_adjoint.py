"""Gradients of the SH efficiency from the hand-written adjoint (shg_adjoint) against
central finite differences.

- field direction derivatives vs central FD of the GPU forward itself (step 1e-3 relative
  to |A0|) on seeded random directions — same operating point, FP32 FD noise ~1e-4;
- field, focus and temperature derivatives vs central FD of the FP64 restatement
  (oracle64).  The FP32 reference arithmetic carries a float phase-mismatch offset
  (0.008 rad over 30 mm at 27 C) that the FP64 model does not have, so these agree to the
  percent level, not to FP32 rounding;
- the checkpoint interval changes nothing;
- the torch.autograd wrapper returns the adjoint field.
"""
import numpy as np
import pytest

import oracle_lib as orc
from shg_focused_gaussian_beams_amd.params import config1_params


def _params(**kw):
    base = dict(nx=64, ny=64, nz=24, power=50.0, waist=20.0, lx=160.0, ly=160.0,
                alpha_p=1e-7, alpha_s=4e-6, beta_s=1e-4)
    base.update(kw)
    return config1_params(**base)


def _inputs(gpu, p, seed=1):
    rng = np.random.default_rng(seed)
    pump = gpu.make_pump(p)
    noise = (1e-3 * np.abs(pump).max() * (rng.standard_normal(pump.shape) + 1j * rng.standard_normal(pump.shape))
             ).astype(np.complex64)
    sh = (1e-4 * np.abs(pump).max() * (rng.standard_normal(pump.shape) + 1j * rng.standard_normal(pump.shape))
          ).astype(np.complex64)
    return pump, noise, sh


def _eta_gpu(gpu, p, fh, sh):
    with gpu.Context(p) as ctx:
        ctx.set_fields(fh, sh)
        ctx.forward()
        return ctx.powers()[2]


def _eta64(p, fh, sh, temp=None):
    t = p.temp if temp is None else temp
    _, s = orc.cwes(p, fh, sh, prec=64)
    return orc.power(p, 1, s, t) / orc.power(p, 0, fh, t)


@pytest.mark.gpu
@pytest.mark.parametrize("n,nz", [(64, 24), (128, 40)])
def test_field_gradient_vs_gpu_fd(gpu, n, nz):
    p = _params(nx=n, ny=n, nz=nz)
    pump, noise, sh = _inputs(gpu, p)
    a0 = pump + noise
    with gpu.Context(p) as ctx:
        ctx.set_fields(a0, sh)
        res, g = ctx.adjoint()
    print(f"\nn={n} nz={nz} eta={res['eta']:.6e} K={res['checkpoint_interval']}", flush=True)
    rng = np.random.default_rng(4)
    eps = 1e-3
    errs = []
    for k in range(8):
        v = rng.standard_normal(a0.shape) + 1j * rng.standard_normal(a0.shape)
        v *= np.linalg.norm(a0) / np.linalg.norm(v)
        fd = (_eta_gpu(gpu, p, (a0 + eps * v).astype(np.complex64), sh) -
              _eta_gpu(gpu, p, (a0 - eps * v).astype(np.complex64), sh)) / (2 * eps)
        ad = float(np.real(np.sum(np.conj(g.astype(np.complex128)) * v)))
        errs.append(abs(fd - ad) / max(abs(fd), 1e-3 * res["eta"]))
        print(f"  dir {k}: fd {fd:+.6e} adjoint {ad:+.6e} rel {errs[-1]:.1e}")
    assert max(errs) < 5e-3, errs


@pytest.mark.gpu
def test_gradients_vs_oracle64_fd(gpu):
    p = _params(nz=30)
    pump, noise, sh = _inputs(gpu, p)
    a0 = (pump + noise).astype(np.complex64)
    with gpu.Context(p) as ctx:
        ctx.set_fields(a0, sh)
        res, g = ctx.adjoint()
    rng = np.random.default_rng(5)
    v = rng.standard_normal(a0.shape) + 1j * rng.standard_normal(a0.shape)
    v *= np.linalg.norm(a0) / np.linalg.norm(v)
    e = 1e-4
    a064 = a0.astype(np.complex128)
    fd = (_eta64(p, a064 + e * v, sh) - _eta64(p, a064 - e * v, sh)) / (2 * e)
    ad = float(np.real(np.sum(np.conj(g.astype(np.complex128)) * v)))
    print(f"\n  field: fd64 {fd:+.6e} adjoint {ad:+.6e}")
    assert abs(fd - ad) / abs(fd) < 2e-2

    # focus: the pump is rebuilt by setPump (FP64 oracle) at f +- h, the noise kept
    h = 1e-3 * p.focus
    def eta_focus(f):
        q = p.copy(focus=f)
        return _eta64(q, orc.pump(q, 64) + noise, sh)
    fdf = (eta_focus(p.focus + h) - eta_focus(p.focus - h)) / (2 * h)
    print(f"  focus: fd64 {fdf:+.6e} adjoint {res['d_eta_d_focus']:+.6e}")
    assert abs(fdf - res["d_eta_d_focus"]) / abs(fdf) < 2e-2

    # temperature: crystal and pump at T +- h
    hT = 1e-3 * p.temp
    def eta_T(t):
        q = p.copy(temp=t)
        return _eta64(q, orc.pump(q, 64) + noise, sh, t)
    fdT = (eta_T(p.temp + hT) - eta_T(p.temp - hT)) / (2 * hT)
    print(f"  temp: fd64 {fdT:+.6e} adjoint {res['d_eta_d_temp']:+.6e} (crystal part {res['d_eta_d_temp_crystal']:+.6e})")
    assert abs(fdT - res["d_eta_d_temp"]) / abs(fdT) < 3e-2


@pytest.mark.gpu
def test_focus_gradient_vs_gpu_fd(gpu):
    p = _params(nz=30)
    pump, noise, sh = _inputs(gpu, p)
    with gpu.Context(p) as ctx:
        ctx.set_fields(pump + noise, sh)
        res, _ = ctx.adjoint(want_field=False)
    h = 1e-3 * p.focus
    def eta_focus(f):
        q = p.copy(focus=f)
        return _eta_gpu(gpu, q, (gpu.make_pump(q) + noise).astype(np.complex64), sh)
    fd = (eta_focus(p.focus + h) - eta_focus(p.focus - h)) / (2 * h)
    print(f"\n  focus: fd(gpu) {fd:+.6e} adjoint {res['d_eta_d_focus']:+.6e}")
    assert abs(fd - res["d_eta_d_focus"]) / abs(fd) < 5e-3


@pytest.mark.gpu
def test_checkpoint_interval_does_not_change_gradients(gpu):
    p = _params(nz=25)
    pump, noise, sh = _inputs(gpu, p)
    out = []
    with gpu.Context(p) as ctx:
        ctx.set_fields(pump + noise, sh)
        for k in (1, 5, 7, 24):
            res, g = ctx.adjoint(checkpoint_interval=k)
            out.append((res, g))
    r0, g0 = out[0]
    for r, g in out[1:]:
        assert abs(r["eta"] - r0["eta"]) <= 1e-6 * r0["eta"]
        assert np.abs(g - g0).max() <= 1e-5 * np.abs(g0).max()
        for key in ("d_eta_d_focus", "d_eta_d_temp"):
            assert abs(r[key] - r0[key]) <= 1e-5 * abs(r0[key]), key


@pytest.mark.gpu
def test_torch_autograd_wrapper(gpu):
    import torch
    from shg_focused_gaussian_beams_amd.autograd import efficiency
    p = _params(nz=12)
    pump, noise, sh = _inputs(gpu, p)
    a0 = torch.from_numpy((pump + noise).astype(np.complex64)).requires_grad_(True)
    with gpu.Context(p) as ctx:
        eta = efficiency(ctx, a0, sh)
        eta.backward()
        res, g = ctx.adjoint()
    assert abs(eta.item() - res["eta"]) <= 1e-12
    assert np.array_equal(a0.grad.numpy(), g)
