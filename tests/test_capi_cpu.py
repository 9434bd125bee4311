"""CPU-side checks of the product: the C ABI library loads and exports every function
include/shg.h declares, argument validation fails before any device call, the host-side
reference arithmetic (initial pump) matches the oracle bit for bit, the device FFT's
index logic is correct in a host build, and the SHG program rejects bad command lines."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as orc
import shg_focused_gaussian_beams_amd as shg
from shg_focused_gaussian_beams_amd.params import default_params, plain128_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(shg_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(shg.LIB_PATH)
    names = _declared_functions()
    assert len(names) >= 20, names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_default_params_match_reference_constants():
    lib = shg.load_library()
    p = shg.ShgParams()
    lib.shg_default_params(ctypes.byref(p))
    q = default_params()
    for name, _ in shg.ShgParams._fields_:
        assert getattr(p, name) == getattr(q, name), name


@pytest.mark.parametrize("bad", [dict(nx=64, ny=128), dict(nx=100, ny=100), dict(nz=1), dict(crystal=7),
                                 dict(scheme=3), dict(lz=0.0), dict(thermal=1, heat_nz=500),
                                 dict(thermal=1, thermal_bc=7)])
def test_create_rejects_bad_parameters_without_device(bad):
    lib = shg.load_library()
    p = default_params(**bad)
    h = ctypes.c_void_p()
    rc = lib.shg_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == -1 and not h.value
    assert lib.shg_last_error().decode().startswith("shg_create")


@pytest.mark.parametrize("case", ["plain", "nz", "lx", "bc", "heat_nz", "count"])
def test_thermal_batch_rejects_mixed_or_bad_builds_without_device(case):
    # instances may differ only in the beam and the ovens; every check runs before any
    # device call
    lib = shg.load_library()
    a = default_params(thermal=1, nx=64, ny=64, nz=12)
    b = a.copy(power=3.0, waist=40.0, focus=100.0, temp=30.0, t_peltier1=30.0, t_peltier2=31.0)
    count = 2
    if case == "plain":
        a, b = a.copy(thermal=0), b.copy(thermal=0)
    elif case == "nz":
        b = b.copy(nz=14)
    elif case == "lx":
        b = b.copy(lx=a.lx * 2)
    elif case == "bc":
        b = b.copy(thermal_bc=1)
    elif case == "heat_nz":
        a, b = a.copy(heat_nz=500), b.copy(heat_nz=500)
    else:
        count = 0
    arr = (shg.ShgParams * 2)(a, b)
    h = ctypes.c_void_p()
    rc = lib.shg_thermal_batch_create(arr, count, ctypes.byref(h))
    assert rc == -1 and not h.value
    msg = lib.shg_last_error().decode()
    assert msg.startswith("shg_thermal_batch_create") or (case == "heat_nz" and msg.startswith("shg_create")), msg


@pytest.mark.parametrize("mk", [default_params, plain128_params])
def test_host_pump_is_bit_exact_with_oracle32(mk):
    p = mk()
    a = shg.make_pump(p)
    ref = orc.pump(p, 32)
    assert np.array_equal(a.astype(np.complex128), ref)


def test_device_fft_index_logic_in_host_build(tmp_path):
    exe = tmp_path / "fft_emul"
    subprocess.run(["g++", "-std=c++20", "-O1", "-pthread", "-o", str(exe),
                    os.path.join(ROOT, "tests", "cpu_emul", "fft_emul.cpp")], check=True)
    for n, e in [(n, e) for n in (64, 128, 256, 512) for e in (0, 4, 8, 16) if n // max(e, 1) <= 128]:
        for sign in (1, -1):
            for mode in ("row", "col"):
                if mode == "col" and e == 4 and n == 512:
                    continue
                out = subprocess.run([str(exe), str(n), str(sign), mode, "11", str(e)], check=True,
                                     capture_output=True, text=True).stdout
                d = np.array([list(map(float, l.split())) for l in out.strip().split("\n")])
                cw = 1 if mode == "row" else (4 if n == 512 else 8)
                for k in (0, 2):
                    a = (d[:, k] + 1j * d[:, k + 1]).reshape(n, cw)
                    o = (d[:, k + 4] + 1j * d[:, k + 5]).reshape(n, cw)
                    ref = np.fft.ifft(a, axis=0) * n if sign > 0 else np.fft.fft(a, axis=0)
                    assert np.abs(o - ref).max() / np.abs(ref).max() < 1e-6, (n, sign, mode)
            # the single-field transform (fft1fe): first field transformed, second untouched
            if e in (4, 8) and n // e <= 64:
                out = subprocess.run([str(exe), str(n), str(sign), "row", "12", str(e), "one"], check=True,
                                     capture_output=True, text=True).stdout
                d = np.array([list(map(float, l.split())) for l in out.strip().split("\n")])
                a, o = d[:, 0] + 1j * d[:, 1], d[:, 4] + 1j * d[:, 5]
                ref = np.fft.ifft(a) * n if sign > 0 else np.fft.fft(a)
                assert np.abs(o - ref).max() / np.abs(ref).max() < 1e-6, (n, sign, "one")
                assert np.array_equal(d[:, 2:4], d[:, 6:8])


def test_sweep_rejects_bad_command_lines(tmp_path):
    exe = os.path.join(ROOT, "shg_focused_gaussian_beams_amd", "SHG_sweep")
    if not os.path.exists(exe):
        pytest.skip("SHG_sweep program not built")
    for bad in (["--power=a:b"], ["--power=5:1:1"], ["--nope=1"], ["stray"],
                ["--power=30", "--waist=20", "--t1=27", "--workers=0"],
                ["--power=30", "--waist=20", "--t1=27", "--workers=65"]):
        r = subprocess.run([exe] + bad, capture_output=True, text=True, cwd=tmp_path)
        assert r.returncode == 1 and ("usage" in r.stderr or "option" in r.stderr), (bad, r.stderr)


def test_cli_rejects_bad_command_lines(tmp_path):
    exe = os.path.join(ROOT, "shg_focused_gaussian_beams_amd", "SHG")
    if not os.path.exists(exe):
        pytest.skip("SHG program not built")
    r = subprocess.run([exe, "30", "29"], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 1 and "usage" in r.stderr
    r = subprocess.run([exe, "30", "29", "55", "55", "1", "0", "--nope=1"], capture_output=True, text=True,
                       cwd=tmp_path)
    assert r.returncode == 1 and "unknown option" in r.stderr
    # parameter checks run before any device call
    r = subprocess.run([exe, "30", "29", "55", "55", "1", "0", "--thermal=1", "--thermal_bc=sideways"],
                       capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 1 and "thermal_bc" in r.stderr


def test_phase_sincos_within_two_ulp(tmp_path):
    # sincos_phase (the THERMAL stage-group phases): double-precision reduction by pi/2 and
    # Cephes polynomials; against long-double sin/cos of the same float argument
    exe = tmp_path / "sincos_emul"
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", str(exe), os.path.join(ROOT, "tests", "cpu_emul", "sincos_emul.cpp")],
                   check=True)
    for xmax in (1.0, 30.0, 1e3, 1e5):
        out = subprocess.run([str(exe), str(xmax), "300000", "7"], check=True, capture_output=True, text=True).stdout
        max_abs, max_ulp = map(float, out.split())
        assert max_ulp <= 2.0 and max_abs <= 1.5e-7, (xmax, max_abs, max_ulp)


def test_entry_point_guard_turns_exceptions_into_codes(tmp_path):
    # every int entry point runs its body through shg_internal_guard (csrc/guard.h)
    exe = tmp_path / "guard_emul"
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", str(exe), os.path.join(ROOT, "tests", "cpu_emul", "guard_emul.cpp")],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines()
    assert out[0] == "-3 entry_a: host allocation failed"
    assert out[1] == "-7 entry_b: vector"
    assert out[2] == "-7 entry_c: unexpected error"
    assert out[3] == "-4"
    # and every exported int entry point of the library is wrapped
    import re
    wrapped = 0
    for name in ("shg_capi.cpp", "shg_batch.cpp", "shg_adjoint.cpp"):
        src = open(os.path.join(ROOT, "shg_focused_gaussian_beams_amd", "csrc", name)).read()
        for m in re.finditer(r"^int (shg_[a-z_0-9]+)\(.*\) \{\n(.*)\n", src, re.M):
            if not m.group(1).startswith("shg_internal"):
                assert "shg_internal_guard(__func__" in m.group(2), (name, m.group(1))
                wrapped += 1
    assert wrapped >= 31, wrapped
