"""ctypes mirror of ``shg_params`` (include/shg.h) and the reference's default values.

The defaults are the constants the reference hard-codes: grid and box
(src/SHG.cpp:11-13, src/headers/Crystal.h:31-38), material constants
(src/headers/Crystal.h:41-55), the fixed focal point and ambient temperature
(src/SHG.cpp:39,43) and the SHG.sh default run (src/SHG.sh:31-37).
"""
import ctypes

SHG_CRYSTAL_SPPLT = 0
SHG_CRYSTAL_PPLN = 1
SHG_SCHEME_REFERENCE = 0
SHG_SCHEME_STRANG = 1
SHG_BC_TOP_OPEN = 0
SHG_BC_OVEN_SURROUNDED = 1


class ShgParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("crystal", ctypes.c_int32), ("thermal", ctypes.c_int32), ("scheme", ctypes.c_int32),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("lambda_p", ctypes.c_double), ("deff", ctypes.c_double), ("qpm_period", ctypes.c_double),
        ("alpha_p", ctypes.c_double), ("alpha_s", ctypes.c_double), ("beta_s", ctypes.c_double),
        ("rho_p", ctypes.c_double), ("rho_s", ctypes.c_double),
        ("k_th", ctypes.c_double), ("h_conv", ctypes.c_double),
        ("power", ctypes.c_double), ("waist", ctypes.c_double), ("focus", ctypes.c_double),
        ("temp", ctypes.c_double), ("t_inf", ctypes.c_double),
        ("t_peltier1", ctypes.c_double), ("t_peltier2", ctypes.c_double),
        ("heat_nz", ctypes.c_int32), ("jacobi_check", ctypes.c_int32),
        ("max_outer", ctypes.c_int32), ("jacobi_max", ctypes.c_int32),
        ("jacobi_tol", ctypes.c_double),
        ("thermal_bc", ctypes.c_int32),
    ]

    def copy(self, **kw):
        q = ShgParams()
        ctypes.pointer(q)[0] = self
        for k, v in kw.items():
            setattr(q, k, v)
        return q

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def default_params(**overrides):
    """The reference run: NX=NY=64, NZ=200, 2000x1000x30000 um MgO:sPPLT box,
    P=30 W, w0=29 um, T=55.66 C (src/SHG.sh:31-37), focus at L/2."""
    p = ShgParams()
    p.nx, p.ny, p.nz = 64, 64, 200
    p.crystal, p.thermal, p.scheme = SHG_CRYSTAL_SPPLT, 0, SHG_SCHEME_REFERENCE
    p.lx, p.ly, p.lz = 2000.0, 1000.0, 30000.0
    p.lambda_p = 1.064
    p.deff = 11.00e-6
    p.qpm_period = 7.97
    p.alpha_p, p.alpha_s, p.beta_s = 0.17e-6, 1.57e-6, 5e-5
    p.rho_p, p.rho_s = 0.0, 0.0
    p.k_th, p.h_conv = 8.4e-6, 10e-12
    p.power, p.waist = 30.0, 29.0
    p.focus = 0.5 * 30000.0
    p.temp, p.t_inf = 55.66, 25.0
    p.t_peltier1, p.t_peltier2 = 55.66, 55.66
    p.heat_nz, p.jacobi_check, p.max_outer, p.jacobi_max = 0, 5000, 5, 0
    p.jacobi_tol = 0.0
    p.thermal_bc = SHG_BC_TOP_OPEN
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def plain128_params(**overrides):
    """Plain SSFM workload, 1064->532 nm CW SHG in MgO:sPPLT: L = 30 mm, 128x128 grid over
    an 8*w0 window, nz = 200, P = 5 W, w0 = 31 um focused at the crystal centre,
    deff = 10 pm/V, T = 27 C with the grating period that phase-matches at 27 C."""
    w0 = 31.0
    p = default_params(nx=128, ny=128, nz=200, lx=8 * w0, ly=8 * w0, deff=10e-6,
                       qpm_period=8.0265, power=5.0, waist=w0, temp=27.0,
                       t_peltier1=27.0, t_peltier2=27.0)
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def uneven_sweep_params(count=256, seed=3, sizes=(64, 128, 256, 512), nz_range=(100, 400), **overrides):
    """Batched sweep: per instance, grid size uniform over `sizes`, w0 ~ U(20, 80) um,
    P ~ U(1, 50) W, T ~ 27 C (QPM) + U(-2, 2), focus ~ L/2 + U(-L/4, L/4),
    nz ~ U{nz_range}, window 8*w0; otherwise as plain128_params (seeded)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    L = 30000.0
    for _ in range(count):
        n = int(rng.choice(sizes))
        w0 = float(rng.uniform(20.0, 80.0))
        p = plain128_params(nx=n, ny=n, nz=int(rng.integers(nz_range[0], nz_range[1] + 1)),
                           lx=8 * w0, ly=8 * w0, waist=w0, power=float(rng.uniform(1.0, 50.0)),
                           temp=27.0 + float(rng.uniform(-2.0, 2.0)),
                           focus=L / 2 + float(rng.uniform(-L / 4, L / 4)))
        p.t_peltier1 = p.t_peltier2 = p.temp
        for k, v in overrides.items():
            setattr(p, k, v)
        out.append(p)
    return out


def thermal256_params(**overrides):
    """THERMAL workload: as plain128_params but P = 30 W, w0 = 40 um, 256x256 grid, nz = 400,
    alpha_p = 0.1/m, alpha_s = 4/m, beta_s = 1e-13 m/W, 25 C ambient."""
    w0 = 40.0
    p = default_params(nx=256, ny=256, nz=400, lx=8 * w0, ly=8 * w0, deff=10e-6,
                       qpm_period=8.0265, power=30.0, waist=w0, temp=27.0,
                       t_peltier1=27.0, t_peltier2=27.0, alpha_p=1e-7, alpha_s=4e-6,
                       beta_s=1e-7, thermal=1)
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def inverse512_params(**overrides):
    """Large inverse-design instance: as thermal256_params (THERMAL, w0 = 40 um, 8*w0 window) but
    512x512 grid, nz = 1000, heat grid 512x512x200, P = 50 W."""
    p = thermal256_params(nx=512, ny=512, nz=1000, heat_nz=200, power=50.0)
    for k, v in overrides.items():
        setattr(p, k, v)
    return p
