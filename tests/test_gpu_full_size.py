"""GPU parity at the workloads' full sizes.

- Jacobi sweeps bit-exact with the oracle on the THERMAL workload's heat grid
  (256x256x100) and on the inverse-design heat grid (512x512x200).
- The THERMAL workload's whole outer loop (256x256, nz 400, heat 256x256x100, mean
  squared update < 1e-6; src/headers/Solver.h:289-340) against the oracle's outer loop:
  sweep counts equal, temperature, fields and efficiency within 1e-4; the distance of the
  final fields to the FP64 restatement (its march at the GPU's final temperature) is
  printed.
- shg_thermal_run stages its own pump (setPump at the grid-centre temperature) in every
  outer iteration, also when the context is tuned inside the call: a caller pump of a
  different power must not leak into the first march.
- The inverse-design grid: the 512x512 THERMAL march with its 512x512x200 heat grid
  (nz 200, one field step per heat slice) against oracle32 (and the oracle64 distance);
  at nz 1000 the field-only adjoint against central FD of the GPU forward along smooth
  pump modulations, and one ascent step of the phase mask raises eta.
- The gradient workload (plain 256x256, nz 400, P 10 W, w0 35 um): the focus derivative
  against central FD of the GPU forward, the temperature derivative against central FD
  of oracle64 at 26.913 C (where the float and double phase mismatch agree: the float
  model's eta(T) is a staircase on the scale of an FD step, the FP64 model is smooth),
  and the field gradient against oracle64 central FD along smooth directions.

The oracle runs on the host; independent oracle calls run in threads (ctypes releases
the GIL), so these cases take tens of seconds, not minutes.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_lib as orc
from shg_focused_gaussian_beams_amd.params import thermal256_params, inverse512_params, plain128_params

TOL = 1e-4


def _hot_spot_T(p, amp=0.5, base=27.0):
    # an on-axis hot spot that cools along z: dk(T) varies over every heat slice
    hz = p.heat_nz if p.heat_nz > 0 else p.nz
    z = np.linspace(0, 1, hz)[:, None, None]
    y = np.linspace(0, 1, p.ny)[None, :, None]
    x = np.linspace(0, 1, p.nx)[None, None, :]
    return (base + amp * np.exp(-((x - .5) ** 2 + (y - .5) ** 2) / 0.02) * (1 - 0.5 * z)).astype(np.float32)


def _smooth_modulations(a0, count, seed, kmax=2):
    # the field times a random low-order modulation (2D Fourier modes |k| <= kmax with
    # complex normal weights), scaled to |a0|: eta moves at O(eta) along these
    rng = np.random.default_rng(seed)
    ny, nx = a0.shape
    y, x = np.meshgrid(np.arange(ny) / ny, np.arange(nx) / nx, indexing="ij")
    out = []
    for _ in range(count):
        m = np.zeros(a0.shape, np.complex128)
        for ky in range(-kmax, kmax + 1):
            for kx in range(-kmax, kmax + 1):
                m += (rng.standard_normal() + 1j * rng.standard_normal()) * np.exp(2j * np.pi * (kx * x + ky * y))
        v = a0 * m
        out.append(v * (np.linalg.norm(a0) / np.linalg.norm(v)))
    return out


def _parallel(calls, workers=8):
    with ThreadPoolExecutor(max_workers=workers) as pool:
        futs = [pool.submit(f, *a) for f, *a in calls]
        return [f.result() for f in futs]


# ---- Jacobi at the heat grids of the workloads -----------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("which", ["256x256x100", "512x512x200"])
def test_jacobi_full_size_bit_exact(gpu, which):
    p = thermal256_params(heat_nz=100) if which == "256x256x100" else inverse512_params()
    hz = p.heat_nz
    print(f"\njacobi {which}: sweeps vs oracle", flush=True)
    T = _hot_spot_T(p, amp=2.0)
    T[:, 0, :] = 27.0                                            # the ovens under y = 0
    rng = np.random.default_rng(12)
    Q = (1e-8 * rng.random((hz, p.ny, p.nx))).astype(np.float32)
    with gpu.Context(p) as ctx:
        ctx.set_temperature(T)
        ctx.set_heat_source(Q)
        r1 = ctx.jacobi(1)
        g1 = ctx.get_temperature()
        ctx.jacobi(2, residual=False)
        g3 = ctx.get_temperature()
    o1 = orc.jacobi_sweep(p, hz, T, Q)
    o3 = orc.jacobi_sweep(p, hz, orc.jacobi_sweep(p, hz, o1, Q), Q)
    assert np.array_equal(g1, o1), np.abs(g1 - o1).max()
    assert np.array_equal(g3, o3), np.abs(g3 - o3).max()
    r_ref = float(np.mean((o1.astype(np.float64) - T.astype(np.float64)) ** 2))
    print(f"  bit-exact after 1 and 3 sweeps; mean squared update {r1:.6e} (FP64 host {r_ref:.6e})")
    assert abs(r1 - r_ref) <= 1e-9 * r_ref


# ---- the THERMAL workload's outer loop ---------------------------------------------------

@pytest.mark.gpu
def test_thermal_outer_loop_headline_matches_oracle(gpu):
    p = thermal256_params(heat_nz=100, jacobi_tol=1e-6)
    fh0 = gpu.make_pump(p)
    rng = np.random.default_rng(1)
    amp = 1e-6 * np.abs(fh0).max()
    fh0 = (fh0 + amp * (rng.standard_normal(fh0.shape) + 1j * rng.standard_normal(fh0.shape))).astype(np.complex64)
    sh0 = np.zeros_like(fh0)
    print("\nTHERMAL outer loop 256x256, nz 400, heat 256x256x100, tol 1e-6", flush=True)
    with gpu.Context(p) as ctx:
        ctx.set_fields(fh0, sh0)
        st = ctx.thermal_run(pump_from_temperature=True)
        fh, sh = ctx.get_fields()
        T = ctx.get_temperature()
        Q = ctx.get_heat_source()
        pin, _, eta = ctx.powers()
    print(f"  gpu stats {st}; oracle outer loop and an oracle64 march at the GPU's temperature ...", flush=True)
    tcg = float(T[0, p.ny // 2, p.nx // 2])
    (rfh, rsh, rT, rQ, rst), (f64, s64, _) = _parallel(
        [(orc.thermal_run, p, sh0), (orc.thermal_cwes, p, T, orc.pump(p, 32, tcg), sh0, None, 64)])
    print(f"  oracle stats {list(rst)}", flush=True)
    assert st["outer_iterations"] == rst[0] and st["sweeps_total"] == rst[2] and st["sweeps_last"] == rst[1]
    eT = float(np.abs(T - rT).max() / max(np.abs(rT - p.t_inf).max(), 1e-30))
    eQ = float(np.linalg.norm(Q - rQ) / np.linalg.norm(rQ))
    tc = float(rT[0, p.ny // 2, p.nx // 2])
    rpump = orc.pump(p, 32, tc)
    eta_ref = orc.power(p, 1, rsh, tc) / orc.power(p, 0, rpump, tc)
    e = (orc.rel_l2(fh, rfh), orc.rel_l2(sh, rsh), eT, eQ)
    print(f"  relL2 FH {e[0]:.2e} SH {e[1]:.2e} T {e[2]:.2e} Q {e[3]:.2e}; eta {eta:.8f} oracle32 {eta_ref:.8f}; "
          f"vs oracle64: GPU FH {orc.rel_l2(fh, f64):.2e} SH {orc.rel_l2(sh, s64):.2e}, oracle32 SH "
          f"{orc.rel_l2(rsh, s64):.2e}")
    assert max(e) <= TOL, e
    assert abs(eta - eta_ref) <= TOL * eta_ref


@pytest.mark.gpu
def test_thermal_run_stages_its_own_pump_on_an_untuned_context(gpu):
    # pump_from_temperature: every outer iteration marches setPump at the grid-centre
    # temperature, whatever pump the caller set; the first march of an untuned context
    # tunes first and must not fall back to the caller's fields afterwards
    p = thermal256_params(nx=64, ny=64, nz=16, heat_nz=8, lx=240.0, ly=240.0, power=40.0, waist=30.0,
                          max_outer=1, jacobi_check=20, jacobi_max=40, jacobi_tol=1e-30, temp=31.0,
                          t_peltier1=27.0, t_peltier2=27.5)
    pump = gpu.make_pump(p)
    rng = np.random.default_rng(3)
    sh0 = (1e-9 * (rng.standard_normal(pump.shape) + 1j * rng.standard_normal(pump.shape))).astype(np.complex64)
    with gpu.Context(p) as ctx:
        ctx.set_fields((0.5 * pump).astype(np.complex64), sh0)
        st = ctx.thermal_run(pump_from_temperature=True)
        fh, sh = ctx.get_fields()
        pin, _, eta = ctx.powers()
        variants = ctx.variants()
    with gpu.Context(p) as ctx:                                  # the same layouts, pinned up front
        ctx.set_variants(*variants)
        ctx.set_fields(pump, sh0)
        ctx.thermal_run(pump_from_temperature=True)
        fh2, sh2 = ctx.get_fields()
    rfh, rsh, rT, rQ, rst = orc.thermal_run(p, sh0)
    tc = float(rT[0, p.ny // 2, p.nx // 2])
    pin_ref = orc.power(p, 0, orc.pump(p, 32, tc), tc)
    e = (orc.rel_l2(fh, rfh), orc.rel_l2(sh, rsh))
    print(f"\n  stats {st} oracle {list(rst)}; relL2 FH {e[0]:.2e} SH {e[1]:.2e}; P_in {pin:.6f} oracle {pin_ref:.6f}")
    assert st["outer_iterations"] == rst[0] == 1
    assert max(e) <= TOL, e
    assert abs(pin - pin_ref) <= TOL * pin_ref
    assert np.array_equal(fh, fh2) and np.array_equal(sh, sh2)


@pytest.mark.gpu
def test_thermal_run_reports_the_reference_progress_events(gpu):
    # the events behind the reference's stdout lines (src/headers/Solver.h:299-333)
    p = thermal256_params(nx=64, ny=64, nz=10, lx=240.0, ly=240.0, jacobi_check=50)
    ev = []
    with gpu.Context(p) as ctx:
        ctx.set_fields(gpu.make_pump(p), np.zeros((64, 64), np.complex64))
        st = ctx.thermal_run(log=lambda e, outer, sweeps, value: ev.append((e, outer, sweeps, value)))
    outers = [x for x in ev if x[0] == gpu.SHG_THERMAL_EVENT_OUTER]
    checks = [x for x in ev if x[0] == gpu.SHG_THERMAL_EVENT_CHECK]
    steady = [x for x in ev if x[0] == gpu.SHG_THERMAL_EVENT_STEADY]
    marches = [x for x in ev if x[0] == gpu.SHG_THERMAL_EVENT_MARCH]
    print(f"\n  stats {st}; {len(outers)} outer, {len(checks)} checks, steady {[x[2] for x in steady]}")
    assert len(outers) == len(steady) == len(marches) == st["outer_iterations"]
    assert [x[1] for x in outers] == list(range(st["outer_iterations"]))
    assert all(x[2] % 50 == 0 for x in checks)                  # checks at sweeps 0, 50, 100, ...
    assert steady[-1][2] == st["sweeps_last"] and sum(x[2] for x in steady) == st["sweeps_total"]
    assert abs(outers[0][3] - np.float32(5e-4 / (64 * 64 * 10))) <= 1e-12


# ---- the inverse-design grid -------------------------------------------------------------

@pytest.mark.gpu
def test_inverse_design_grid_thermal_march_matches_oracle(gpu):
    p = inverse512_params(nz=200)
    T = _hot_spot_T(p, amp=1.0)
    fh0 = gpu.make_pump(p, float(T[0, p.ny // 2, p.nx // 2]))
    sh0 = np.zeros_like(fh0)
    print("\nTHERMAL march 512x512, nz 200, heat 512x512x200", flush=True)
    with gpu.Context(p) as ctx:
        ctx.set_temperature(T)
        ctx.set_fields(fh0, sh0)
        ctx.forward()
        fh, sh = ctx.get_fields()
        Q = ctx.get_heat_source()
        _, _, eta = ctx.powers()
    (rfh, rsh, rQ), (f64, s64, _) = _parallel([(orc.thermal_cwes, p, T, fh0, sh0, None, 32),
                                               (orc.thermal_cwes, p, T, fh0, sh0, None, 64)])
    t_c = float(T[0, p.ny // 2, p.nx // 2])
    r_eta = orc.power(p, 1, rsh, t_c) / orc.power(p, 0, fh0, t_c)
    e = (orc.rel_l2(fh, rfh), orc.rel_l2(sh, rsh), float(np.linalg.norm(Q - rQ) / np.linalg.norm(rQ)))
    print(f"  relL2 FH {e[0]:.2e} SH {e[1]:.2e} Q {e[2]:.2e}; eta {eta:.8f} oracle {r_eta:.8f}; "
          f"vs oracle64: GPU SH {orc.rel_l2(sh, s64):.2e}, oracle32 SH {orc.rel_l2(rsh, s64):.2e}")
    assert max(e) <= TOL, e
    assert abs(eta - r_eta) <= TOL * r_eta


@pytest.mark.gpu
def test_inverse_design_full_size_field_gradient_and_ascent(gpu):
    # 512x512, nz 1000, heat 512x512x200, P = 50 W: the field-only adjoint against central
    # FD (step 1e-3) of the GPU THERMAL forward along 8 smooth pump modulations, and of
    # the phase-mask gradient along one smooth mask direction; one ascent step raises eta
    from shg_focused_gaussian_beams_amd.inverse import ascend_phase_mask, mask_gradient
    p = inverse512_params()
    T = _hot_spot_T(p, amp=1.0)
    pump = gpu.make_pump(p, float(T[0, p.ny // 2, p.nx // 2]))
    mask0 = np.random.default_rng(5).uniform(-0.1, 0.1, pump.shape)
    fld = (pump * np.exp(1j * mask0)).astype(np.complex64)
    sh0 = np.zeros_like(fld)
    print("\ninverse design 512x512, nz 1000, heat 512x512x200: adjoint vs FD", flush=True)
    with gpu.Context(p) as ctx:
        ctx.set_temperature(T)
        ctx.set_fields(fld, sh0)
        eta0, g = ctx.adjoint_field(checkpoint_interval=p.nz - 1)
        g = g.astype(np.complex128)

        def eta_of(f):
            ctx.set_fields(f.astype(np.complex64), sh0)
            ctx.forward()
            return ctx.powers()[2]
        fd, ad = [], []
        for v in _smooth_modulations(fld.astype(np.complex128), 8, 21):
            ad.append(float(np.real(np.sum(np.conj(g) * v))))
            fd.append((eta_of(fld + 1e-3 * v) - eta_of(fld - 1e-3 * v)) / 2e-3)
        fd, ad = np.array(fd), np.array(ad)
        rel = np.linalg.norm(fd - ad) / np.linalg.norm(ad)
        # the mask chain rule along a smooth real mask direction
        yy, xx = np.meshgrid(np.arange(p.ny) / p.ny, np.arange(p.nx) / p.nx, indexing="ij")
        w = np.cos(2 * np.pi * xx) + 0.5 * np.sin(2 * np.pi * (xx + 2 * yy))
        dm = float(np.sum(mask_gradient(g, fld) * w))
        e = 1e-2
        fdm = (eta_of(pump * np.exp(1j * (mask0 + e * w))) - eta_of(pump * np.exp(1j * (mask0 - e * w)))) / (2 * e)
        mask, hist = ascend_phase_mask(ctx, pump, mask0, 1, checkpoint_interval=p.nz - 1)
    print(f"  eta {eta0:.6e}; 8 smooth directions relL2 {rel:.2e}; mask direction fd {fdm:+.6e} adjoint {dm:+.6e}; "
          f"ascent {hist[0]:.8f} -> {hist[-1]:.8f}")
    assert rel < 5e-3
    assert abs(fdm - dm) <= 5e-3 * abs(fdm)
    assert hist[-1] > hist[0]


# ---- the gradient workload at full size --------------------------------------------------

def _grad_params(**kw):
    return plain128_params(nx=256, ny=256, nz=400, lx=8 * 35.0, ly=8 * 35.0, power=10.0, waist=35.0, **kw)


def _grad_inputs(gpu, p, seed=4):
    pump = gpu.make_pump(p)
    rng = np.random.default_rng(seed)
    noise = (1e-6 * np.abs(pump).max() * (rng.standard_normal(pump.shape) + 1j * rng.standard_normal(pump.shape)))
    return pump.astype(np.complex128), noise


def _eta_orc(p, fh, prec, temp=None):
    t = p.temp if temp is None else temp
    _, s = orc.cwes(p, fh, np.zeros_like(fh), prec=prec)
    return orc.power(p, 1, s, t) / orc.power(p, 0, fh, t)


@pytest.mark.gpu
def test_gradient_workload_focus_derivative_vs_gpu_fd(gpu):
    # the focus enters through the z = 0 pump only (setPump, src/headers/Efields.h:74-104)
    p = _grad_params()
    pump, noise = _grad_inputs(gpu, p)
    sh0 = np.zeros(pump.shape, np.complex64)
    with gpu.Context(p) as ctx:
        ctx.set_fields((pump + noise).astype(np.complex64), sh0)
        res, _ = ctx.adjoint()
        out = []
        for rel_step in (1e-3, 5e-3, 1e-2):
            h = rel_step * p.focus
            etas = []
            for f in (p.focus + h, p.focus - h):
                ctx.set_fields((gpu.make_pump(p.copy(focus=f)) + noise).astype(np.complex64), sh0)
                ctx.forward()
                etas.append(ctx.powers()[2])
            out.append((etas[0] - etas[1]) / (2 * h))
    # the beam is focused near its optimum (the derivative is ~3% of eta over the crystal
    # length), so the central FD carries an O(h^2) curvature term: Richardson-extrapolate
    rich = (4 * out[1] - out[2]) / 3
    adj = res["d_eta_d_focus"]
    print(f"\n  d eta / d focus: adjoint {adj:+.6e}, GPU FD step 1e-3 {out[0]:+.6e}, 5e-3 {out[1]:+.6e}, "
          f"1e-2 {out[2]:+.6e}, Richardson(5e-3, 1e-2) {rich:+.6e}")
    assert abs(rich - adj) <= 1e-2 * abs(adj)
    assert abs(out[0] - adj) <= 5e-2 * abs(adj)


@pytest.mark.gpu
def test_gradient_workload_temperature_and_field_vs_oracle64_fd(gpu):
    # at 26.913 C the float and double phase mismatch agree (7e-7 rad over the crystal):
    # the FP64 model is the smooth function the FP32 reference rounds
    p = _grad_params(temp=26.913)
    assert abs(orc.dk(p, 32, p.temp) - orc.dk(p, 64, p.temp)) * p.lz < 1e-5
    pump, noise = _grad_inputs(gpu, p)
    a0 = (pump + noise).astype(np.complex64)
    sh0 = np.zeros_like(a0)
    with gpu.Context(p) as ctx:
        ctx.set_fields(a0, sh0)
        res, g = ctx.adjoint()
    g = g.astype(np.complex128)
    a064 = a0.astype(np.complex128)
    noise64 = a064 - orc.pump(p, 64)
    dirs = _smooth_modulations(a064, 4, 31)
    hT = 1e-3 * p.temp

    def eta_T(t):
        q = p.copy(temp=t)
        return _eta_orc(q, orc.pump(q, 64) + noise64, 64, t)
    hf = 1e-2 * p.focus

    def eta_f(f):
        q = p.copy(focus=f)
        return _eta_orc(q, orc.pump(q, 64) + noise64, 64)
    e = 1e-3
    calls = [(eta_T, p.temp + hT), (eta_T, p.temp - hT), (eta_f, p.focus + hf), (eta_f, p.focus - hf)]
    for v in dirs:
        calls += [(_eta_orc, p, a064 + e * v, 64), (_eta_orc, p, a064 - e * v, 64)]
    print("\n  oracle64 FD: 12 forwards at 256x256, nz 400 ...", flush=True)
    r = _parallel(calls, workers=12)
    fdT = (r[0] - r[1]) / (2 * hT)
    fdf = (r[2] - r[3]) / (2 * hf)
    fd = np.array([(r[4 + 2 * k] - r[5 + 2 * k]) / (2 * e) for k in range(len(dirs))])
    ad = np.array([float(np.real(np.sum(np.conj(g) * v))) for v in dirs])
    rel = np.abs(fd - ad) / np.abs(fd)
    print(f"  d eta / d T: adjoint {res['d_eta_d_temp']:+.6e} oracle64 FD {fdT:+.6e}; d eta / d focus: adjoint "
          f"{res['d_eta_d_focus']:+.6e} oracle64 FD {fdf:+.6e}; field directions oracle64 FD {fd} adjoint {ad} "
          f"rel {rel}")
    assert abs(fdT - res["d_eta_d_temp"]) <= 2e-2 * abs(fdT)
    assert abs(fdf - res["d_eta_d_focus"]) <= 2e-2 * abs(fdf)
    assert rel.max() <= 2e-2


@pytest.mark.gpu
def test_gradient_workload_64_joint_directions(gpu):
    # 64 seeded directions (seed 4) in the joint parameters (A0, focus, T) of the gradient
    # workload, each block's FD step 1e-3 relative to its norm (|A0|, the focus, T).  The
    # adjoint's directional derivative Re<g, dA> + d_focus df + d_T dT is compared with
    # central FD: the (A0, focus) part on the GPU forward (float; pump rebuilt at each
    # focus), the T part from central FD of the FP64 restatement at 26.913 C, where the
    # float and double phase mismatch agree (the float model's eta(T) is a staircase on
    # the scale of a 1e-3 relative T step, so T differences run on the smooth FP64 model).
    # The focus FD carries an O(h^2) curvature term (the beam sits near its optimal focus),
    # so every direction is also Richardson-extrapolated from steps 1e-3 and 2e-3.
    p = _grad_params(temp=26.913)
    pump, noise = _grad_inputs(gpu, p)
    a0 = (pump + noise).astype(np.complex64)
    sh0 = np.zeros_like(a0)
    rng = np.random.default_rng(4)
    ny, nx = a0.shape
    y, x = np.meshgrid(np.arange(ny) / ny, np.arange(nx) / nx, indexing="ij")
    dirs = []
    for _ in range(64):
        m = np.zeros(a0.shape, np.complex128)
        for ky in range(-2, 3):
            for kx in range(-2, 3):
                m += (rng.standard_normal() + 1j * rng.standard_normal()) * np.exp(2j * np.pi * (kx * x + ky * y))
        dA = a0 * m
        dA *= np.linalg.norm(a0) / np.linalg.norm(dA)
        dirs.append((dA, p.focus * rng.standard_normal(), p.temp * rng.standard_normal()))
    e = 1e-3
    with gpu.Context(p) as ctx:
        ctx.set_fields(a0, sh0)
        res, g = ctx.adjoint()
        g = g.astype(np.complex128)

        def eta_gpu(s, k):
            dA, df, _ = dirs[k]
            f = (gpu.make_pump(p.copy(focus=p.focus + s * df)) + noise + s * dA).astype(np.complex64)
            ctx.set_fields(f, sh0)
            ctx.forward()
            return ctx.powers()[2]
        fd1 = np.array([(eta_gpu(e, k) - eta_gpu(-e, k)) / (2 * e) for k in range(64)])
        fd2 = np.array([(eta_gpu(2 * e, k) - eta_gpu(-2 * e, k)) / (4 * e) for k in range(64)])
    rich = (4 * fd1 - fd2) / 3
    noise64 = a0.astype(np.complex128) - orc.pump(p, 64)
    hT = e * p.temp

    def eta_T(t):
        q = p.copy(temp=t)
        return _eta_orc(q, orc.pump(q, 64) + noise64, 64, t)
    print("\n  oracle64 FD in T: 2 forwards at 256x256, nz 400 ...", flush=True)
    rT = _parallel([(eta_T, p.temp + hT), (eta_T, p.temp - hT)], workers=2)
    fdT = (rT[0] - rT[1]) / (2 * hT)
    ad_af = np.array([float(np.real(np.sum(np.conj(g) * dA))) + res["d_eta_d_focus"] * df for dA, df, _ in dirs])
    dT = np.array([d[2] for d in dirs])
    ad = ad_af + res["d_eta_d_temp"] * dT
    ref1, refr = fd1 + fdT * dT, rich + fdT * dT
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    print(f"  64 joint directions (A0, focus, T): relL2 adjoint vs FD step 1e-3 {rel(ad, ref1):.2e}, "
          f"vs Richardson {rel(ad, refr):.2e}; (A0, focus) part: step 1e-3 {rel(ad_af, fd1):.2e}, "
          f"Richardson {rel(ad_af, rich):.2e}; T part: adjoint {res['d_eta_d_temp']:+.6e} oracle64 FD {fdT:+.6e}")
    assert rel(ad, refr) <= 5e-3
    assert rel(ad_af, rich) <= 5e-3
    assert rel(ad, ref1) <= 1e-2
