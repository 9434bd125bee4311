"""Throughput of the split-step (SSFM) z-march on one MI355X.

A "step" of this benchmark is one full solve of the workload: the fields enter the
crystal and are marched through all z-steps to the output plane (the state the
reference writes as Pump_out/Signal_out).  Inputs are resident in HBM before timing.

metric: SSFM cell-steps/s = nx*ny*(z-steps executed) per second of wall time.
roofline: the fused z-step (row pass + column pass) at 32 algorithmic bytes per
cell-step (FH + SH complex64, each read once and written once per step), timed with
HIP events on the library's stream; peak 8 TB/s HBM.
cpu_baseline: the FP32 CPU restatement of the reference (oracle/, -O3, per-cell n()
as the reference evaluates it) on one host core, over a bounded number of z-steps of
the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SSFM cell-steps/s (plain, batched, fwd+adjoint) on 1 GPU; % HBM roofline"
HBM_PEAK_GBPS = 8000.0
BYTES_PER_CELL_STEP = 32


def workload_params(name):
    import shg_focused_gaussian_beams_amd as shg
    if name == "config2-plain":
        return shg.config2_params(thermal=0)
    if name == "config1":
        return shg.config1_params()
    raise SystemExit(f"unknown workload {name}")


def make_inputs(p, seed=1):
    import numpy as np
    import shg_focused_gaussian_beams_amd as shg
    pump = shg.make_pump(p)
    rng = np.random.default_rng(seed)
    amp = 1e-6 * np.abs(pump).max()
    pump = (pump + amp * (rng.standard_normal(pump.shape) + 1j * rng.standard_normal(pump.shape))).astype(np.complex64)
    return pump, np.zeros_like(pump)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p, fh0, sh0, budget_s=12.0):
    """oracle32 (-O3, per-cell n()) on one core over a bounded number of z-steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    lib = orc.load()
    lib.oracle_set_percell_index(1)
    try:
        t0 = time.perf_counter()
        orc.cwes(p, fh0, sh0, prec=32, nsteps=1)
        one = time.perf_counter() - t0
        k = max(1, min(p.nz - 1, int(budget_s / max(one, 1e-6))))
        t0 = time.perf_counter()
        orc.cwes(p, fh0, sh0, prec=32, nsteps=k)
        dt = time.perf_counter() - t0
    finally:
        lib.oracle_set_percell_index(0)
    cells = p.nx * p.ny * k
    return {"value": cells / dt, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle32 -O3 (per-cell n(), own radix-2 FFT) on {p.nx}x{p.ny}, {k} of {p.nz - 1} "
                      f"z-steps, {dt:.1f} s; host CPU: {cpu_model()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="config2-plain")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-reps", type=int, default=200)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the solver has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import shg_focused_gaussian_beams_amd as shg
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)

    p = workload_params(args.workload)
    fh0, sh0 = make_inputs(p)
    ctx = shg.Context(p)
    ctx.set_fields(fh0, sh0)
    for _ in range(args.warmup):
        ctx.forward()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev_ms = 0.0
    for _ in range(args.steps):
        ctx.forward()
        dev_ms += ctx.last_forward_ms()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    cell_steps = ctx.cell_steps()
    value = cell_steps / (wall / args.steps)
    _, _, eta = ctx.powers()

    row_ms, col_ms = ctx.profile_kernels(args.profile_reps)
    step_s = (row_ms + col_ms) * 1e-3
    n2 = p.nx * p.ny
    achieved = BYTES_PER_CELL_STEP * n2 / step_s / 1e9
    ctx.close()

    out = {
        "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Gaussian pump + 1e-6 relative complex noise, SH = 0)",
        "config": {"workload": args.workload, "nx": p.nx, "ny": p.ny, "nz": p.nz,
                   "z_steps": p.nz - 1, "cell_steps_per_solve": cell_steps, "thermal": bool(p.thermal),
                   "eta": eta, "device_ms_per_solve": dev_ms / args.steps,
                   "algorithmic_GBps": value * BYTES_PER_CELL_STEP / 1e9},
        "roofline": {"bound": "hbm", "kernel": "fused z-step = k_row_pass + k_col_pass",
                     "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
                     "row_pass_us": row_ms * 1e3, "col_pass_us": col_ms * 1e3},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, fh0, sh0)
        out["cpu_baseline"]["gpu_speedup"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
