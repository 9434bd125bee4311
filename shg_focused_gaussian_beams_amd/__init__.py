"""MI355X-native focused-Gaussian-beam SHG split-step solver.

The compute path is libshg.so (hand-written HIP for gfx950 behind the C ABI of
include/shg.h).  This module is a thin ctypes binding that mirrors the reference's
entry point, ``Solver::run(Power, waist, focalpoint, Temp, T_inf, Tpeltier1,
Tpeltier2, save_only_last, save_temperature)`` (src/headers/Solver.h:152,279), and
exposes the context calls for tests and benchmarks.  There is no CPU fallback: if the
library is missing, importing the solver raises.
"""
import ctypes
import os

import numpy as np

from .params import (ShgParams, default_params, plain128_params, thermal256_params, uneven_sweep_params, inverse512_params,  # noqa: F401
                     SHG_CRYSTAL_SPPLT, SHG_CRYSTAL_PPLN, SHG_SCHEME_REFERENCE, SHG_SCHEME_STRANG)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libshg.so")
_lib = None

SHG_OK = 0
ERRORS = {-1: "SHG_E_INVALID", -2: "SHG_E_DEVICE", -3: "SHG_E_NOMEM", -4: "SHG_E_STATE",
          -5: "SHG_E_NONFINITE", -6: "SHG_E_NOCONVERGE", -7: "SHG_E_INTERNAL"}


class ThermalStats(ctypes.Structure):
    _fields_ = [("outer_iterations", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("sweeps_last", ctypes.c_int64), ("sweeps_total", ctypes.c_int64),
                ("residual_last", ctypes.c_double), ("jacobi_ms", ctypes.c_double),
                ("march_ms", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class AdjointResult(ctypes.Structure):
    _fields_ = [("eta", ctypes.c_double), ("d_eta_d_focus", ctypes.c_double),
                ("d_eta_d_temp", ctypes.c_double), ("d_eta_d_temp_crystal", ctypes.c_double),
                ("forward_ms", ctypes.c_double), ("adjoint_ms", ctypes.c_double),
                ("checkpoint_interval", ctypes.c_int32), ("segments", ctypes.c_int32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


# shg_thermal_log_fn and its events (include/shg.h)
THERMAL_LOG_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_int32, ctypes.c_int64, ctypes.c_double)
SHG_THERMAL_EVENT_OUTER, SHG_THERMAL_EVENT_CHECK, SHG_THERMAL_EVENT_STEADY, SHG_THERMAL_EVENT_MARCH = 0, 1, 2, 3
SHG_JACOBI_LAYOUTS = 10                 # shg_set_jacobi_layout range (include/shg.h)


class ShgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def load_library(path=LIB_PATH):
    """Loads libshg.so (built in-tree by build.py).  Raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: run `python -m shg_focused_gaussian_beams_amd.build` "
                          "(the HIP library is the only compute path)")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER(ShgParams)
    vp = ctypes.c_void_p
    fp = ctypes.POINTER(ctypes.c_float)
    dp = ctypes.POINTER(ctypes.c_double)
    sigs = {
        "shg_default_params": (None, [P]),
        "shg_last_error": (ctypes.c_char_p, []),
        "shg_create": (ctypes.c_int, [P, ctypes.POINTER(vp)]),
        "shg_destroy": (None, [vp]),
        "shg_make_pump": (ctypes.c_int, [P, ctypes.c_double, fp]),
        "shg_set_fields": (ctypes.c_int, [vp, fp, fp]),
        "shg_forward": (ctypes.c_int, [vp]),
        "shg_reset_fields": (ctypes.c_int, [vp]),
        "shg_march": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
        "shg_get_fields": (ctypes.c_int, [vp, fp, fp]),
        "shg_forward_snapshots": (ctypes.c_int, [vp, fp, fp]),
        "shg_powers": (ctypes.c_int, [vp, dp, dp, dp]),
        "shg_last_forward_ms": (ctypes.c_int, [vp, dp]),
        "shg_cell_steps": (ctypes.c_int64, [vp]),
        "shg_profile_kernels": (ctypes.c_int, [vp, ctypes.c_int, dp, dp]),
        "shg_thermal_run": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ThermalStats)]),
        "shg_jacobi": (ctypes.c_int, [vp, ctypes.c_int64, dp]),
        "shg_set_thermal_log": (ctypes.c_int, [vp, THERMAL_LOG_FN, vp]),
        "shg_get_temperature": (ctypes.c_int, [vp, fp]),
        "shg_set_temperature": (ctypes.c_int, [vp, fp]),
        "shg_get_heat_source": (ctypes.c_int, [vp, fp]),
        "shg_set_heat_source": (ctypes.c_int, [vp, fp]),
        "shg_adjoint": (ctypes.c_int, [vp, ctypes.c_int, fp, ctypes.POINTER(AdjointResult)]),
        "shg_adjoint_field": (ctypes.c_int, [vp, ctypes.c_int, fp, dp]),
        "shg_adjoint_temperature_map": (ctypes.c_int, [vp, dp]),
        "shg_tune": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
        "shg_set_variants": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
        "shg_get_variants": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
        "shg_set_jacobi_layout": (ctypes.c_int, [vp, ctypes.c_int]),
        "shg_jacobi_noalloc": (ctypes.c_int, [P, vp, vp, vp, ctypes.c_int64, ctypes.c_int, vp, vp]),
        "shg_get_jacobi_layout": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int)]),
        "shg_batch_create": (ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(vp)]),
        "shg_batch_destroy": (None, [vp]),
        "shg_batch_init_fields": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_uint64]),
        "shg_batch_set_fields": (ctypes.c_int, [vp, ctypes.c_int, fp, fp]),
        "shg_batch_run": (ctypes.c_int, [vp]),
        "shg_batch_set_schedule": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
        "shg_batch_get_fields": (ctypes.c_int, [vp, ctypes.c_int, fp, fp]),
        "shg_batch_get_input": (ctypes.c_int, [vp, ctypes.c_int, fp, fp]),
        "shg_batch_powers": (ctypes.c_int, [vp, dp, dp, dp]),
        "shg_batch_cell_steps": (ctypes.c_int64, [vp]),
        "shg_batch_last_ms": (ctypes.c_int, [vp, dp]),
        "shg_thermal_batch_create": (ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(vp)]),
        "shg_thermal_batch_destroy": (None, [vp]),
        "shg_thermal_batch_set_signal": (ctypes.c_int, [vp, ctypes.c_int, fp]),
        "shg_thermal_batch_set_variants": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
        "shg_thermal_batch_get_variants": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int),
                                                          ctypes.POINTER(ctypes.c_int)]),
        "shg_thermal_batch_set_jacobi_layout": (ctypes.c_int, [vp, ctypes.c_int]),
        "shg_thermal_batch_get_jacobi_layout": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int)]),
        "shg_thermal_batch_run": (ctypes.c_int, [vp, ctypes.POINTER(ThermalStats)]),
        "shg_thermal_batch_get_fields": (ctypes.c_int, [vp, ctypes.c_int, fp, fp]),
        "shg_thermal_batch_get_temperature": (ctypes.c_int, [vp, ctypes.c_int, fp]),
        "shg_thermal_batch_powers": (ctypes.c_int, [vp, dp, dp, dp]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != SHG_OK:
        raise ShgError(rc, load_library().shg_last_error().decode())


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def jacobi_noalloc(params, T_ptr, T_work_ptr, Q_ptr, nsweeps, layout=-1, mean_sq_update_ptr=None, stream=None):
    """shg_jacobi_noalloc: Jacobi sweeps on caller-owned device buffers (integer device
    addresses, e.g. torch tensors' data_ptr()) with no device allocation; enqueued on
    `stream` (a raw hipStream_t handle) and not waited for."""
    _check(load_library().shg_jacobi_noalloc(ctypes.byref(params), T_ptr, T_work_ptr, Q_ptr, nsweeps, layout,
                                             mean_sq_update_ptr, stream))


def make_pump(params, t_centre=None):
    """The reference's initial Gaussian pump (setPump, src/headers/Efields.h:74-104),
    evaluated in its float arithmetic by the library's host code."""
    lib = load_library()
    out = np.zeros((params.ny, params.nx), np.complex64)
    _check(lib.shg_make_pump(ctypes.byref(params), params.temp if t_centre is None else t_centre, _fptr(out)))
    return out


class Context:
    """One solver context: device state, stream and precomputed tables."""

    def __init__(self, params):
        self.lib = load_library()
        self.params = params.copy()
        h = ctypes.c_void_p()
        _check(self.lib.shg_create(ctypes.byref(self.params), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.shg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_fields(self, fh, sh):
        fh = np.ascontiguousarray(fh, np.complex64)
        sh = np.ascontiguousarray(sh, np.complex64)
        n = self.params.nx * self.params.ny
        if fh.size != n or sh.size != n:
            raise ValueError(f"fields must hold {n} cells")
        _check(self.lib.shg_set_fields(self.h, _fptr(fh), _fptr(sh)))

    def forward(self):
        _check(self.lib.shg_forward(self.h))

    def reset_fields(self):
        _check(self.lib.shg_reset_fields(self.h))

    def march(self, first_step, nsteps):
        _check(self.lib.shg_march(self.h, first_step, nsteps))

    def get_fields(self):
        shape = (self.params.ny, self.params.nx)
        fh = np.empty(shape, np.complex64)
        sh = np.empty(shape, np.complex64)
        _check(self.lib.shg_get_fields(self.h, _fptr(fh), _fptr(sh)))
        return fh, sh

    def forward_snapshots(self):
        """The march of forward(), keeping the fields before every step: (fh, sh) of
        shape (nz, ny, nx) complex64 (the reference's Pump_<s>, Signal_<s>)."""
        shape = (self.params.nz, self.params.ny, self.params.nx)
        fh = np.empty(shape, np.complex64)
        sh = np.empty(shape, np.complex64)
        _check(self.lib.shg_forward_snapshots(self.h, _fptr(fh), _fptr(sh)))
        return fh, sh

    def powers(self):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(self.lib.shg_powers(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def last_forward_ms(self):
        v = ctypes.c_double()
        _check(self.lib.shg_last_forward_ms(self.h, ctypes.byref(v)))
        return v.value

    def cell_steps(self):
        return int(self.lib.shg_cell_steps(self.h))

    def tune(self, reps=10):
        """Times every row/column-pass layout variant and keeps the fastest."""
        r, c = ctypes.c_int(), ctypes.c_int()
        _check(self.lib.shg_tune(self.h, reps, ctypes.byref(r), ctypes.byref(c)))
        return r.value, c.value

    def set_variants(self, row_variant, col_variant):
        _check(self.lib.shg_set_variants(self.h, row_variant, col_variant))

    def variants(self):
        """(row, column) layouts the next march runs; (-1, -1) before tuning."""
        r, c = ctypes.c_int(), ctypes.c_int()
        _check(self.lib.shg_get_variants(self.h, ctypes.byref(r), ctypes.byref(c)))
        return r.value, c.value

    def adjoint(self, checkpoint_interval=0, want_field=True):
        """Efficiency and its gradients (discrete adjoint).  Returns (result dict,
        d eta / d FH(z=0) as complex64 (ny, nx) or None)."""
        r = AdjointResult()
        g = np.empty((self.params.ny, self.params.nx), np.complex64) if want_field else None
        _check(self.lib.shg_adjoint(self.h, checkpoint_interval, _fptr(g) if want_field else None, ctypes.byref(r)))
        return r.as_dict(), g

    def adjoint_field(self, checkpoint_interval=0):
        """eta and d eta / d FH(z=0) only (shg_adjoint_field): the field gradient of
        adjoint() without its focus and temperature sensitivities."""
        g = np.empty((self.params.ny, self.params.nx), np.complex64)
        eta = ctypes.c_double()
        _check(self.lib.shg_adjoint_field(self.h, checkpoint_interval, _fptr(g), ctypes.byref(eta)))
        return eta.value, g

    def adjoint_temperature_map(self):
        """After adjoint() on a THERMAL context: d eta / d T per heat voxel (heat_shape())."""
        m = np.empty(self.heat_shape(), np.float64)
        _check(self.lib.shg_adjoint_temperature_map(self.h, _dptr(m)))
        return m

    # THERMAL
    def heat_shape(self):
        hz = self.params.heat_nz if self.params.heat_nz > 0 else self.params.nz
        return (hz, self.params.ny, self.params.nx)

    def thermal_run(self, pump_from_temperature=True, log=None):
        """THERMAL Solver::run (src/headers/Solver.h:289-340).  `log`, if given, receives
        the progress events (event, outer, sweeps, value) of shg_set_thermal_log."""
        st = ThermalStats()
        cb = None
        if log is not None:
            cb = THERMAL_LOG_FN(lambda user, ev, outer, sweeps, value: log(ev, outer, sweeps, value))
            _check(self.lib.shg_set_thermal_log(self.h, cb, None))
        try:
            _check(self.lib.shg_thermal_run(self.h, 1 if pump_from_temperature else 0, ctypes.byref(st)))
        finally:
            if cb is not None:
                _check(self.lib.shg_set_thermal_log(self.h, THERMAL_LOG_FN(), None))
        return st.as_dict()

    def set_jacobi_layout(self, layout):
        _check(self.lib.shg_set_jacobi_layout(self.h, layout))

    def jacobi_layout(self):
        v = ctypes.c_int()
        _check(self.lib.shg_get_jacobi_layout(self.h, ctypes.byref(v)))
        return v.value

    def jacobi(self, nsweeps, residual=True):
        r = ctypes.c_double()
        _check(self.lib.shg_jacobi(self.h, nsweeps, ctypes.byref(r) if residual else None))
        return r.value

    def get_temperature(self):
        t = np.empty(self.heat_shape(), np.float32)
        _check(self.lib.shg_get_temperature(self.h, _fptr(t)))
        return t

    def set_temperature(self, T):
        T = np.ascontiguousarray(T, np.float32)
        if T.size != int(np.prod(self.heat_shape())):
            raise ValueError("temperature must hold nx*ny*heat_nz values")
        _check(self.lib.shg_set_temperature(self.h, _fptr(T)))

    def get_heat_source(self):
        q = np.empty(self.heat_shape(), np.float32)
        _check(self.lib.shg_get_heat_source(self.h, _fptr(q)))
        return q

    def set_heat_source(self, Q):
        Q = np.ascontiguousarray(Q, np.float32)
        if Q.size != int(np.prod(self.heat_shape())):
            raise ValueError("heat source must hold nx*ny*heat_nz values")
        _check(self.lib.shg_set_heat_source(self.h, _fptr(Q)))

    def profile_kernels(self, reps=50):
        r, c = ctypes.c_double(), ctypes.c_double()
        _check(self.lib.shg_profile_kernels(self.h, reps, ctypes.byref(r), ctypes.byref(c)))
        return r.value, c.value


class Batch:
    """A batched sweep: many independent plain instances marched in one device schedule
    (replaces the one-process-per-point loop of src/SHG.sh:39-81)."""

    def __init__(self, params_list):
        self.lib = load_library()
        self.params = [p.copy() for p in params_list]
        arr = (ShgParams * len(self.params))(*self.params)
        h = ctypes.c_void_p()
        _check(self.lib.shg_batch_create(arr, len(self.params), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.shg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def init_fields(self, noise_rel=1e-6, seed=1):
        _check(self.lib.shg_batch_init_fields(self.h, noise_rel, seed))

    def set_fields(self, i, fh, sh):
        fh = np.ascontiguousarray(fh, np.complex64)
        sh = np.ascontiguousarray(sh, np.complex64)
        n = self.params[i].nx * self.params[i].ny
        if fh.size != n or sh.size != n:
            raise ValueError(f"instance {i}: fields must hold {n} cells")
        _check(self.lib.shg_batch_set_fields(self.h, i, _fptr(fh), _fptr(sh)))

    def run(self):
        _check(self.lib.shg_batch_run(self.h))

    def set_schedule(self, group_mib=128, streams=3):
        """Group size (MiB of fields marched together) and number of concurrent streams."""
        _check(self.lib.shg_batch_set_schedule(self.h, group_mib, streams))

    def _get(self, fn, i):
        shape = (self.params[i].ny, self.params[i].nx)
        fh = np.empty(shape, np.complex64)
        sh = np.empty(shape, np.complex64)
        _check(fn(self.h, i, _fptr(fh), _fptr(sh)))
        return fh, sh

    def get_fields(self, i):
        return self._get(self.lib.shg_batch_get_fields, i)

    def get_input(self, i):
        return self._get(self.lib.shg_batch_get_input, i)

    def powers(self):
        n = len(self.params)
        a, b, c = (np.zeros(n) for _ in range(3))
        dp = ctypes.POINTER(ctypes.c_double)
        _check(self.lib.shg_batch_powers(self.h, a.ctypes.data_as(dp), b.ctypes.data_as(dp), c.ctypes.data_as(dp)))
        return a, b, c

    def cell_steps(self):
        return int(self.lib.shg_batch_cell_steps(self.h))

    def last_ms(self):
        v = ctypes.c_double()
        _check(self.lib.shg_batch_last_ms(self.h, ctypes.byref(v)))
        return v.value


class ThermalBatch:
    """Batched THERMAL sweep points: the THERMAL Solver::run (src/headers/Solver.h:289-340)
    of every instance, run in lockstep on one device state (replaces the per-point
    processes of src/SHG.sh:39-81).  Instances share everything but power, waist, focus,
    temp, t_peltier1 and t_peltier2."""

    def __init__(self, params_list):
        self.lib = load_library()
        self.params = [p.copy() for p in params_list]
        arr = (ShgParams * len(self.params))(*self.params)
        h = ctypes.c_void_p()
        _check(self.lib.shg_thermal_batch_create(arr, len(self.params), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.shg_thermal_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def heat_shape(self):
        p = self.params[0]
        return (p.heat_nz if p.heat_nz > 0 else p.nz, p.ny, p.nx)

    def set_signal(self, i, sh):
        sh = np.ascontiguousarray(sh, np.complex64)
        if sh.size != self.params[0].nx * self.params[0].ny:
            raise ValueError("the signal must hold nx*ny cells")
        _check(self.lib.shg_thermal_batch_set_signal(self.h, i, _fptr(sh)))

    def set_variants(self, row_variant, col_variant):
        _check(self.lib.shg_thermal_batch_set_variants(self.h, row_variant, col_variant))

    def variants(self):
        r, c = ctypes.c_int(), ctypes.c_int()
        _check(self.lib.shg_thermal_batch_get_variants(self.h, ctypes.byref(r), ctypes.byref(c)))
        return r.value, c.value

    def set_jacobi_layout(self, layout):
        _check(self.lib.shg_thermal_batch_set_jacobi_layout(self.h, layout))

    def jacobi_layout(self):
        v = ctypes.c_int()
        _check(self.lib.shg_thermal_batch_get_jacobi_layout(self.h, ctypes.byref(v)))
        return v.value

    def run(self):
        """Runs every instance; returns one stats dict per instance."""
        st = (ThermalStats * len(self.params))()
        _check(self.lib.shg_thermal_batch_run(self.h, st))
        return [s.as_dict() for s in st]

    def get_fields(self, i):
        shape = (self.params[0].ny, self.params[0].nx)
        fh = np.empty(shape, np.complex64)
        sh = np.empty(shape, np.complex64)
        _check(self.lib.shg_thermal_batch_get_fields(self.h, i, _fptr(fh), _fptr(sh)))
        return fh, sh

    def get_temperature(self, i):
        t = np.empty(self.heat_shape(), np.float32)
        _check(self.lib.shg_thermal_batch_get_temperature(self.h, i, _fptr(t)))
        return t

    def powers(self):
        n = len(self.params)
        a, b, c = (np.zeros(n) for _ in range(3))
        _check(self.lib.shg_thermal_batch_powers(self.h, _dptr(a), _dptr(b), _dptr(c)))
        return a, b, c


class Solver:
    """Mirror of the reference's ``Solver`` (src/headers/Solver.h:120-153).

    ``run`` has the reference's argument list and meaning.  Instead of writing .dat
    files it returns the output fields (the state the reference saves as Pump_out /
    Signal_out) and the FP64 powers; the SHG command-line program writes the files.
    ``signal`` replaces the reference's clock-seeded N(0, 1e-20) SH noise
    (src/headers/Efields.h:129-144); by default it is that noise with a fixed seed.
    """

    def __init__(self, params=None):
        self.params = (params or default_params()).copy()

    def run(self, Power, waist, focalpoint, Temp, T_inf, Tpeltier1, Tpeltier2,
            save_only_last=True, save_temperature=False, signal=None, seed=0):
        p = self.params.copy(power=Power, waist=waist, focus=focalpoint, temp=Temp, t_inf=T_inf,
                             t_peltier1=Tpeltier1, t_peltier2=Tpeltier2)
        pump = make_pump(p, Temp)
        if signal is None:
            rng = np.random.default_rng(seed)
            signal = (rng.normal(0, 1e-20, pump.shape) + 1j * rng.normal(0, 1e-20, pump.shape)).astype(np.complex64)
        with Context(p) as ctx:
            ctx.set_fields(pump, signal)
            if p.thermal:
                self.thermal_stats = ctx.thermal_run(pump_from_temperature=True)
                self.temperature = ctx.get_temperature() if save_temperature else None
            else:
                ctx.forward()
            fh, sh = ctx.get_fields()
            powers = ctx.powers()
        return fh, sh, powers
