"""Throughput of the split-step (SSFM) z-march on one MI355X.

Headline workload: THERMAL SSFM, 1064 -> 532 nm in MgO:sPPLT,
P = 30 W, w0 = 40 um, 256x256 grid, nz = 400, heat grid 256x256x100, Jacobi to a mean
squared update < 1e-6, dk(T) fed back.  A "step" of this benchmark is one THERMAL
z-march of the field pair through the crystal at the converged temperature (nz steps:
the last one feeds the heat source), inputs resident in HBM.
  value  = SSFM cell-steps/s = nx*ny*nz / time per march
  roofline = the fused THERMAL z-step (row pass + column pass): 56 algorithmic bytes
           per cell-step as the metric defines them (FH+SH read+write 32, T slice 4, Q
           slice 4, phase accumulator read+write 16), over the device time per z-step of
           the timed marches (HIP events on the library stream); 8 TB/s peak, and the
           copy bandwidth this GPU reaches (a device copy timed in the same run).
Also reported (same run): the THERMAL outer loop, Jacobi voxel-sweeps/s, the same outer
loop under the reference's own stopping rule on a 256x256x400 heat grid, the reference's
default THERMAL run (64x64x200), the plain 128x128, nz 200 march, the batched
sweep of 256 uneven instances (64^2..512^2, nz 100..400, in one hipGraph) against
one-at-a-time solves, forward+adjoint at 256x256, nz 400 against one forward, the plain
256x256, nz 400 march, and a timed sample of the inverse design (THERMAL 512x512,
nz 1000, adjoint ascent
of a pump phase mask), projected to its 32 instances x 100 iterations.
cpu_baseline: the FP32 CPU restatement of the reference's THERMAL march (oracle/, -O3,
per-cell n() as the reference evaluates it) on one host core over a bounded number of
z-steps of the same grid; beside it the same restatement on the plain 128x128 and
256x256 marches, the Jacobi sweep and the reference's default run.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Kernel of each tuned layout at 256 x 256 (the launch tables of ssfm_kernels.hip and
# thermal_kernels.hip), for the roofline line's label.
THERM_ROW_256 = {0: "k_therm_row_pass<256,16,1>", 1: "k_therm_row_pass<256,8,1>", 2: "k_therm_row_pass<256,4,1>",
                 3: "k_therm_row_pass<256,4,4>", 4: "k_therm_row_wide<256,4,4>", 5: "k_therm_row_wide<256,4,2>",
                 6: "k_therm_row_split<256,4,4>"}
COL_256 = {0: "k_col_pass<256,16,8>", 1: "k_col_pass<256,8,8>", 2: "k_col_pass<256,4,8>", 3: "k_col_pass<256,4,1>",
           4: "k_col_pass<256,4,2>", 5: "k_col_split<256,4,1>", 6: "k_col_split<256,4,2>",
           7: "k_col_split<256,4,4>", 8: "k_col_fsplit<256,4,1>", 9: "k_col_fsplit<256,4,2>"}

METRIC = "SSFM cell-steps/s (plain, batched, fwd+adjoint) on 1 GPU; % HBM roofline"
HBM_PEAK_GBPS = 8000.0
# HBM bytes per THERMAL z-step (row + column pass, 256x256, the tuned layouts) from
# rocprofv3 --pmc FETCH_SIZE (x 2.0) and WRITE_SIZE (x 1.0), the factors read for a copy
# of known size.
TRAFFIC_THERMAL_STEP = 5369 * 1024
B_PLAIN, B_THERMAL, B_FWDADJ, B_JACOBI = 32, 56, 80, 12
# VALU lane-operations per plain cell-step of the batch's passes (rocprofv3 SQ_INSTS_VALU
# x 64 per cell: 393 row pass + 141 column pass at 512^2) and the chip's FP32 issue rate
# without packed FP32 (256 CU x 4 SIMD x 16 lanes x 2.4 GHz).
VALU_OPS_PER_CELL_STEP, VALU_PEAK_LANE_OPS = 534, 39.3e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(p, seed=1, noise=1e-6):
    import numpy as np
    import shg_focused_gaussian_beams_amd as shg
    pump = shg.make_pump(p)
    rng = np.random.default_rng(seed)
    amp = noise * np.abs(pump).max()
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


def _thermal_sample(orc, p, T, fh0, sh0, budget_s, lib=None):
    """Times oracle32's THERMAL march over the first k z-steps of the workload grid, k
    sized so that the run takes about budget_s; returns (k, seconds)."""
    hz = p.heat_nz
    t0 = time.perf_counter()
    q = p.copy(nz=2, heat_nz=2)
    orc.thermal_cwes(q, T[[0, -1]], fh0, sh0, lib=lib)
    one = (time.perf_counter() - t0) / 2
    k = max(2, min(p.nz, int(budget_s / max(one, 1e-6))))
    q = p.copy(nz=k, heat_nz=max(2, min(k, hz)), lz=p.lz * (k - 1) / (p.nz - 1))
    idx = [min(hz - 1, int(i * hz / k)) for i in range(q.heat_nz)]
    t0 = time.perf_counter()
    orc.thermal_cwes(q, T[idx], fh0, sh0, lib=lib)
    return k, time.perf_counter() - t0


def cpu_baseline_thermal(p, T, fh0, sh0, budget_s=12.0, budget_o0_s=4.0):
    """oracle32 THERMAL march (per-cell n(), the reference's O(s) dk re-integration) on
    one core (this process pinned to one CPU for the measurement), over the first k
    z-steps of the workload grid: the -O3 build (the baseline) and the -O0 build (the
    optimisation level of the reference's own build line)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    affinity = os.sched_getaffinity(0)
    os.sched_setaffinity(0, {min(affinity)})
    try:
        k, dt = _thermal_sample(orc, p, T, fh0, sh0, budget_s)
        k0, dt0 = _thermal_sample(orc, p, T, fh0, sh0, budget_o0_s, lib=orc.load(orc.ORACLE_O0_SO))
    finally:
        os.sched_setaffinity(0, affinity)
    cells = p.nx * p.ny
    return {"value": cells * k / dt, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle32 -O3 THERMAL march (per-cell n(), own radix-2 FFT, no FFTW on the host) on "
                      f"{p.nx}x{p.ny}, first {k} of {p.nz} z-steps, {dt:.1f} s, pinned to one core; host CPU: "
                      f"{cpu_model()}, {os.cpu_count()} CPUs visible",
            "value_O0": cells * k0 / dt0,
            "sample_O0": f"the same at -O0: first {k0} z-steps, {dt0:.1f} s"}


def cpu_batch_projection(ps, budget_s=1.5, per_size=2):
    """The sweep as the reference's SHG.sh runs it: one serial solve per point.  oracle32's
    plain march is timed on one core (pinned) for the first `per_size` instances of every
    grid size in the sweep, over as many z-steps as fit in about budget_s each, and the
    per-size rate (cells x steps over the summed time) is extrapolated to every instance's
    cells x steps."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    affinity = os.sched_getaffinity(0)
    os.sched_setaffinity(0, {min(affinity)})
    rates = {}
    try:
        for n in sorted({p.nx for p in ps}):
            work, secs = 0.0, 0.0
            for q in [p for p in ps if p.nx == n][:per_size]:
                fh = orc.pump(q, 32)
                sh = np.zeros_like(fh)
                t0 = time.perf_counter()
                orc.cwes(q, fh, sh, prec=32, nsteps=1)
                one = time.perf_counter() - t0
                k = max(1, min(q.nz - 1, int(budget_s / max(one, 1e-6))))
                t0 = time.perf_counter()
                orc.cwes(q, fh, sh, prec=32, nsteps=k)
                secs += time.perf_counter() - t0
                work += n * n * k
            rates[n] = work / secs
    finally:
        os.sched_setaffinity(0, affinity)
    serial_s = sum(p.nx * p.ny * (p.nz - 1) / rates[p.nx] for p in ps)
    return {"serial_s_projected": serial_s, "cell_steps_per_s_by_size": {str(n): r for n, r in rates.items()},
            "cores": 1, "kind": "port",
            "sample": f"oracle32 -O3 plain march, {per_size} instances per grid size for about "
                      f"{budget_s:g} s each, pinned to one core, extrapolated by cells x steps"}


def copy_peak_gbps(mib=1024, reps=5):
    """Read+write bandwidth of a device-to-device copy (the practical HBM ceiling)."""
    import torch
    x = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        y.copy_(x)
    e1.record()
    e1.synchronize()
    gbps = 2 * x.numel() * 4 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del x, y
    torch.cuda.empty_cache()
    return gbps


def bench_thermal(shg, args):
    import numpy as np
    p = shg.thermal256_params(heat_nz=100, jacobi_tol=1e-6)
    fh0, sh0 = make_inputs(p)
    ctx = shg.Context(p)
    ctx.set_fields(fh0, sh0)
    if args.variants:
        ctx.set_variants(*(int(v) for v in args.variants.split(",")))
    log("thermal outer loop ...")
    t0 = time.perf_counter()
    st = ctx.thermal_run(pump_from_temperature=True)
    loop_s = time.perf_counter() - t0
    T = ctx.get_temperature()
    for _ in range(args.warmup):
        ctx.forward()
    import torch
    torch.cuda.synchronize()
    log("timed THERMAL marches ...")
    t0 = time.perf_counter()
    dev = 0.0
    for _ in range(args.steps):
        ctx.forward()
        dev += ctx.last_forward_ms()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cells = ctx.cell_steps()
    _, _, eta = ctx.powers()
    variants = ctx.variants()
    row_ms, col_ms = ctx.profile_kernels(args.profile_reps)
    # Jacobi throughput on the heat grid (two sweeps per pass; the first call tunes the layout)
    nj = 200
    ctx.jacobi(2, residual=False)
    torch.cuda.synchronize()
    jac_layout = ctx.jacobi_layout()
    t1 = time.perf_counter()
    ctx.jacobi(nj, residual=False)
    torch.cuda.synchronize()
    jac_s = (time.perf_counter() - t1) / nj
    vox = p.nx * p.ny * p.heat_nz
    ctx.close()
    return dict(p=p, T=T, fh0=fh0, sh0=sh0, wall=wall, steps=args.steps, dev_ms=dev / args.steps, cells=cells,
                eta=eta, row_ms=row_ms, col_ms=col_ms, stats=st, loop_s=loop_s, jac_s=jac_s, vox=vox,
                variants=variants, jac_layout=jac_layout)


def bench_plain(shg, args, p):
    import torch
    fh0, sh0 = make_inputs(p)
    ctx = shg.Context(p)
    ctx.set_fields(fh0, sh0)
    if args.plain_variants:
        ctx.set_variants(*(int(v) for v in args.plain_variants.split(",")))
    for _ in range(args.warmup):
        ctx.forward()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.forward()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    variants = ctx.variants()
    row_ms, col_ms = ctx.profile_kernels(args.profile_reps)
    cells = ctx.cell_steps()
    ctx.close()
    return dict(s_per_solve=wall, cells=cells, row_ms=row_ms, col_ms=col_ms, variants=variants)


def bench_batch(shg, args):
    import numpy as np
    import torch
    ps = shg.uneven_sweep_params(count=args.batch, seed=3)
    b = shg.Batch(ps)
    b.init_fields(noise_rel=1e-6, seed=3)
    log("batch: first run (graph capture) ...")
    b.run()
    torch.cuda.synchronize()
    reps = 2
    t0 = time.perf_counter()
    for _ in range(reps):
        b.run()
    torch.cuda.synchronize()
    batch_s = (time.perf_counter() - t0) / reps
    cells = b.cell_steps()
    _, _, eta = b.powers()
    # one-at-a-time: every instance solved alone by a single context (graph replayed)
    log("batch: one-at-a-time reference solves ...")
    single_s = 0.0
    sample = list(range(0, len(ps), max(1, len(ps) // args.single_sample)))
    per = []
    for i in sample:
        ctx = shg.Context(ps[i])
        fh, sh = b.get_input(i)
        ctx.set_fields(fh, sh)
        ctx.forward()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.forward()
        torch.cuda.synchronize()
        per.append((ps[i].nx * ps[i].ny * (ps[i].nz - 1), time.perf_counter() - t0))
        ctx.close()
    # extrapolate by cell-steps within each grid size
    by_n = {}
    for i, (c, t) in zip(sample, per):
        by_n.setdefault(ps[i].nx, []).append(t / c)
    for p in ps:
        rate = np.mean(by_n.get(p.nx, [np.mean([t / c for c, t in per])]))
        single_s += rate * p.nx * p.ny * (p.nz - 1)
    b.close()
    return dict(count=len(ps), s=batch_s, cells=cells, single_s=single_s, eta_mean=float(np.mean(eta)),
                sampled=len(sample))


def bench_adjoint(shg, args):
    import torch
    p = shg.plain128_params(nx=256, ny=256, nz=400, lx=8 * 35.0, ly=8 * 35.0, power=10.0, waist=35.0)
    fh0, sh0 = make_inputs(p)
    ctx = shg.Context(p)
    ctx.set_fields(fh0, sh0)
    ctx.forward()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        ctx.forward()
    torch.cuda.synchronize()
    fwd_s = (time.perf_counter() - t0) / 3
    res, _ = ctx.adjoint()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(2):
        res, _ = ctx.adjoint()
    torch.cuda.synchronize()
    adj_s = (time.perf_counter() - t0) / 2
    k_sqrt = math.ceil(math.sqrt(p.nz - 1))
    ctx.adjoint(checkpoint_interval=k_sqrt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(2):
        ctx.adjoint(checkpoint_interval=k_sqrt)
    torch.cuda.synchronize()
    adj_sqrt_s = (time.perf_counter() - t0) / 2
    cells = ctx.cell_steps()
    ctx.close()
    return dict(fwd_s=fwd_s, adj_s=adj_s, adj_sqrt_s=adj_sqrt_s, k_sqrt=k_sqrt, cells=cells, res=res)


def bench_thermal_reference_rule(shg, args):
    """The THERMAL outer loop as the reference runs it (src/headers/Solver.h:289-340): heat
    grid = field grid (256x256x400), Jacobi to a mean squared update below 5e-4/TSIZE
    checked every 5000 sweeps, at most 5 outer iterations, on the THERMAL workload's
    crystal and beam.  Wall time of the whole loop and its Jacobi / march split."""
    import torch
    p = shg.thermal256_params(heat_nz=0, jacobi_tol=0.0)
    fh0, sh0 = make_inputs(p)
    ctx = shg.Context(p)
    ctx.set_fields(fh0, sh0)
    ctx.tune()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = ctx.thermal_run(pump_from_temperature=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _, _, eta = ctx.powers()
    ctx.close()
    return dict(p=p, wall=wall, stats=st, eta=eta)


def bench_reference_default_run(shg, args):
    """The reference's own default run (src/SHG.sh:23,31-37; src/SHG.cpp:11-13): THERMAL,
    64x64x200 over its 2000x1000x30000 um box, P = 30 W, w0 = 29 um, T = 55.66 C, the
    reference's Jacobi stopping rule; wall time of the whole run on the GPU."""
    import numpy as np
    import torch
    p = shg.default_params(thermal=1)
    pump = shg.make_pump(p)
    ctx = shg.Context(p)
    ctx.set_fields(pump, np.zeros_like(pump))
    ctx.tune()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = ctx.thermal_run(pump_from_temperature=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _, _, eta = ctx.powers()
    ctx.close()
    return dict(p=p, wall=wall, stats=st, eta=eta)


def bench_thermal_sweep(shg, args):
    """The temperature sweep of the reference's sweep script (src/SHG.sh:40, T1 = 54..59 C
    in steps of 0.2, T2 = T1) at its default THERMAL run (64x64x200, P = 30 W, w0 = 29 um,
    the reference's Jacobi stopping rule): 26 points as one lockstep batch, as 8
    concurrent contexts (one host thread and stream each), and one at a time.  The
    batch's and the contexts' efficiencies must agree bit for bit (same layouts)."""
    import threading
    import numpy as np
    import torch
    base = shg.default_params(thermal=1)
    t1s = [round(54.0 + 0.2 * k, 6) for k in range(26)]
    ps = [base.copy(temp=t, t_peltier1=t, t_peltier2=t) for t in t1s]
    n = base.nx
    zero = np.zeros((n, n), np.complex64)
    ctx = shg.Context(ps[0])
    ctx.set_fields(shg.make_pump(ps[0]), zero)
    rv, cv = ctx.tune()
    ctx.close()

    def one(k, out):
        c = shg.Context(ps[k])
        c.set_fields(shg.make_pump(ps[k]), zero)
        c.set_variants(rv, cv)
        st = c.thermal_run()
        out[k] = (c.powers()[2], st)
        c.close()

    b = shg.ThermalBatch(ps)
    b.set_variants(rv, cv)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bst = b.run()
    torch.cuda.synchronize()
    batch_s = time.perf_counter() - t0
    eta_b = b.powers()[2]
    jac_layout = b.jacobi_layout()
    b.close()
    res = [None] * len(ps)
    nxt = iter(range(len(ps)))
    lock = threading.Lock()

    def work():
        while True:
            with lock:
                k = next(nxt, None)
            if k is None:
                return
            one(k, res)
    t0 = time.perf_counter()
    th = [threading.Thread(target=work) for _ in range(8)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    torch.cuda.synchronize()
    ctx_s = time.perf_counter() - t0
    eta_c = np.array([r[0] for r in res])
    seq = [None] * len(ps)
    t0 = time.perf_counter()
    for k in range(len(ps)):
        one(k, seq)
    torch.cuda.synchronize()
    seq_s = time.perf_counter() - t0
    return dict(points=len(ps), batch_s=batch_s, contexts8_s=ctx_s, one_at_a_time_s=seq_s,
                batch_jacobi_ms=bst[0]["jacobi_ms"], batch_march_ms=bst[0]["march_ms"], jacobi_layout=jac_layout,
                variants=[rv, cv], sweeps_total=[s["sweeps_total"] for s in bst],
                outer_iterations=[s["outer_iterations"] for s in bst],
                eta_max_abs_diff_batch_vs_contexts=float(np.abs(eta_b - eta_c).max()),
                eta_min=float(eta_b.min()), eta_max=float(eta_b.max()), t1_of_eta_max=t1s[int(np.argmax(eta_b))])


def bench_jacobi_layouts(shg, args):
    """Every Jacobi layout (the nine two-sweep passes and single sweeps) on four heat
    grids: the reference's default run (64x64x200), the THERMAL workload (256x256x100),
    heat grid = field grid (256x256x400) and the inverse design's (512x512x200); us per
    sweep over 200 sweeps each, and the tuner's pick."""
    import torch
    grids = [("64x64x200", shg.default_params(thermal=1)), ("256x256x100", shg.thermal256_params(heat_nz=100)),
             ("256x256x400", shg.thermal256_params(heat_nz=0)), ("512x512x200", shg.inverse512_params())]
    out = {}
    for name, p in grids:
        if args.jacobi_grid and name != args.jacobi_grid:
            continue
        n = 200
        with shg.Context(p) as ctx:
            per = []
            for layout in range(shg.SHG_JACOBI_LAYOUTS):
                ctx.set_jacobi_layout(layout)
                ctx.jacobi(4, residual=False)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                ctx.jacobi(n, residual=False)
                torch.cuda.synchronize()
                per.append((time.perf_counter() - t0) / n * 1e6)
        with shg.Context(p) as ctx:
            ctx.jacobi(2, residual=False)
            tuned = ctx.jacobi_layout()
        out[name] = {"us_per_sweep": per, "single_sweep_us": per[-1], "best_two_sweep_us": min(per[:-1]),
                     "best_two_sweep_layout": per.index(min(per[:-1])), "tuned_layout": tuned,
                     "two_sweep_speedup": per[-1] / min(per[:-1])}
        log(f"jacobi {name}: single {per[-1]:.2f} us, best two-sweep {min(per[:-1]):.2f} us, tuned {tuned}")
    return out


def cpu_small_workloads(shg, default_run_stats):
    """oracle32 -O3 on one pinned core: the plain 128x128, nz 200 workload in full, the
    plain forward of the gradient workload (256x256, nz 400) in full, the Jacobi sweep on
    the THERMAL workload's heat grid (256x256x100), and the reference's default THERMAL
    run (64x64x200) projected from one of its marches and 200 of its sweeps to the sweep
    and outer-iteration counts of the run (its Jacobi is bit-exact with the GPU's, and a
    full oracle run takes ~80 s: 20 005 sweeps and 5 marches)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    affinity = os.sched_getaffinity(0)
    os.sched_setaffinity(0, {min(affinity)})
    out = {"cores": 1, "kind": "port", "host_cpu": cpu_model()}
    try:
        p1 = shg.plain128_params()
        fh = orc.pump(p1, 32)
        orc.cwes(p1, fh, np.zeros_like(fh), prec=32, nsteps=4)          # warm-up (first-call costs)
        t0 = time.perf_counter()
        orc.cwes(p1, fh, np.zeros_like(fh), prec=32)
        dt = time.perf_counter() - t0
        out["plain_128x128_nz200"] = {"s_per_solve": dt, "cell_steps_per_s": p1.nx * p1.ny * (p1.nz - 1) / dt}
        p4 = shg.plain128_params(nx=256, ny=256, nz=400, lx=8 * 35.0, ly=8 * 35.0, power=10.0, waist=35.0)
        fh = orc.pump(p4, 32)
        t0 = time.perf_counter()
        orc.cwes(p4, fh, np.zeros_like(fh), prec=32)
        dt = time.perf_counter() - t0
        out["plain_256x256_nz400_forward"] = {
            "s_per_solve": dt, "cell_steps_per_s": p4.nx * p4.ny * (p4.nz - 1) / dt,
            "note": "the CPU path has no adjoint: a central-FD directional derivative costs 2 such solves"}
        ph = shg.thermal256_params(heat_nz=100)
        T = np.full((ph.heat_nz, ph.ny, ph.nx), 27.0, np.float32)
        Q = np.zeros_like(T)
        orc.jacobi_sweep(ph, ph.heat_nz, T, Q)
        nj = 5
        t0 = time.perf_counter()
        for _ in range(nj):
            T = orc.jacobi_sweep(ph, ph.heat_nz, T, Q)
        dt = (time.perf_counter() - t0) / nj
        vox = ph.nx * ph.ny * ph.heat_nz
        out["jacobi_256x256x100"] = {"s_per_sweep": dt, "voxel_sweeps_per_s": vox / dt}
        pd = shg.default_params(thermal=1)
        Td = np.full((pd.nz, pd.ny, pd.nx), np.float32(pd.temp), np.float32)
        Qd = np.zeros_like(Td)
        nj = 200
        t0 = time.perf_counter()
        for _ in range(nj):
            Td = orc.jacobi_sweep(pd, pd.nz, Td, Qd)
        sweep_s = (time.perf_counter() - t0) / nj
        fh = orc.pump(pd, 32)
        t0 = time.perf_counter()
        orc.thermal_cwes(pd, Td, fh, np.zeros_like(fh))
        march_s = time.perf_counter() - t0
        sw, outer = default_run_stats["sweeps_total"], default_run_stats["outer_iterations"]
        out["reference_default_run"] = {
            "s_projected": sw * sweep_s + outer * march_s, "s_per_sweep": sweep_s, "s_per_march": march_s,
            "sweeps_total": sw, "outer_iterations": outer,
            "method": "one march and 200 sweeps timed, scaled to the run's sweep and outer-iteration counts"}
    finally:
        os.sched_setaffinity(0, affinity)
    return out


def bench_inverse(shg, args):
    """Inverse-design sample: THERMAL 512x512, nz = 1000, heat 512x512x200, P = 50 W; the
    temperature field from a bounded THERMAL outer loop, then adjoint ascent of a seeded
    +-0.1 rad pump phase mask with the field held fixed, checked by a plain forward.
    `inverse_instances` instances (16 by default) ascend at once (one context and host thread each) for
    `inverse_iterations` iterations; instance 0 is also timed alone.  Both are projected
    to 32 instances x 100 iterations.  The adjoint keeps a full tape (checkpoint interval
    nz - 1: ~9.4 GB per 512x512, nz 1000 instance) instead of recomputing sqrt(nz)
    segments: HBM holds it, and the recomputation competes with the adjoint march."""
    import numpy as np
    import torch
    from shg_focused_gaussian_beams_amd.inverse import ascend_phase_mask, ascend_phase_masks
    n = max(1, args.inverse_instances)
    ctxs, pumps, masks, th = [], [], [], []
    for i in range(n):
        p = shg.inverse512_params(max_outer=1, jacobi_check=500, jacobi_max=1000, jacobi_tol=1e-6)
        rng = np.random.default_rng(5 + i)
        masks.append(rng.uniform(-0.1, 0.1, (p.ny, p.nx)))
        ctx = shg.Context(p)
        pump = shg.make_pump(p)
        ctx.set_fields(pump, np.zeros_like(pump))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = ctx.thermal_run(pump_from_temperature=True)
        th.append(dict(thermal_s=time.perf_counter() - t0, jacobi_sweeps=st["sweeps_total"]))
        T = ctx.get_temperature()
        pumps.append(shg.make_pump(p, float(T[0, p.ny // 2, p.nx // 2])))
        ctxs.append(ctx)
    it = args.inverse_iterations
    ck = ctxs[0].params.nz - 1                                         # full tape
    ascend_phase_mask(ctxs[0], pumps[0], masks[0], 1, checkpoint_interval=ck)   # warm-up: graphs, tuning, buffers
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ascend_phase_mask(ctxs[0], pumps[0], masks[0], it, checkpoint_interval=ck)
    torch.cuda.synchronize()
    alone_s = (time.perf_counter() - t0) / it
    ascend_phase_masks(ctxs, pumps, masks, 1, checkpoint_interval=ck)             # warm-up of the others
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = ascend_phase_masks(ctxs, pumps, masks, it, checkpoint_interval=ck)
    torch.cuda.synchronize()
    conc_s = (time.perf_counter() - t0) / it / n
    out = []
    for i, (mask, etas) in enumerate(res):
        ctxs[i].close()
        p = ctxs[i].params
        check = shg.Context(p.copy(thermal=0))
        fld = (pumps[i] * np.exp(1j * mask)).astype(np.complex64)
        check.set_fields(fld, np.zeros_like(fld))
        check.forward()
        out.append(dict(**th[i], eta_first=etas[0], eta_last=etas[-1], eta_plain_check=check.powers()[2]))
        check.close()
    th_s = float(np.mean([o["thermal_s"] for o in out]))
    cells = 512 * 512 * 999
    return dict(instances_concurrent=n, iterations_timed=it, checkpoint_interval=ck, per_instance=out,
                s_per_iteration_alone=alone_s, s_per_instance_iteration_concurrent=conc_s,
                fwd_adj_cell_steps_per_s_alone=cells / alone_s,
                fwd_adj_cell_steps_per_s_concurrent=cells / conc_s,
                projected_s_32x100_alone=32 * (th_s + 100 * alone_s),
                projected_s_32x100_concurrent=32 * (th_s + 100 * conc_s))


def bench_inverse_full(shg, args):
    """The large batched inverse design end to end: 32 instances of THERMAL 512x512,
    nz = 1000, heat grid 512x512x200, P = 50 W, each with its temperature field from a
    bounded THERMAL outer loop and its seeded +-0.1 rad pump phase mask, 100 ascent
    iterations each, checked by a plain forward.  The instances run `inverse_instances` at
    a time (one context and host thread each, full tapes), wave after wave."""
    import numpy as np
    import torch
    from shg_focused_gaussian_beams_amd.inverse import ascend_phase_masks
    total, per_wave, it = 32, max(1, args.inverse_instances), 100
    out = []
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    for w0 in range(0, total, per_wave):
        ids = list(range(w0, min(total, w0 + per_wave)))
        ctxs, pumps, masks = [], [], []
        for i in ids:
            p = shg.inverse512_params(max_outer=1, jacobi_check=500, jacobi_max=1000, jacobi_tol=1e-6)
            masks.append(np.random.default_rng(5 + i).uniform(-0.1, 0.1, (p.ny, p.nx)))
            ctx = shg.Context(p)
            pump = shg.make_pump(p)
            ctx.set_fields(pump, np.zeros_like(pump))
            ctx.thermal_run(pump_from_temperature=True)
            T = ctx.get_temperature()
            pumps.append(shg.make_pump(p, float(T[0, p.ny // 2, p.nx // 2])))
            ctxs.append(ctx)
        res = ascend_phase_masks(ctxs, pumps, masks, it, checkpoint_interval=999)
        for i, c, pump, (mask, etas) in zip(ids, ctxs, pumps, res):
            c.close()
            check = shg.Context(shg.inverse512_params(thermal=0))
            fld = (pump * np.exp(1j * mask)).astype(np.complex64)
            check.set_fields(fld, np.zeros_like(fld))
            check.forward()
            out.append(dict(instance=i, eta_first=etas[0], eta_last=etas[-1], eta_plain_check=check.powers()[2]))
            check.close()
        torch.cuda.synchronize()
        log(f"inverse design: instances {ids[0]}..{ids[-1]} done at {time.perf_counter() - t_all:.1f} s")
    wall = time.perf_counter() - t_all
    gains = [o["eta_last"] / o["eta_first"] for o in out]
    return dict(instances=total, iterations=it, instances_concurrent=per_wave, wall_s=wall,
                s_per_instance_iteration=wall / (total * it),
                fwd_adj_cell_steps_per_s=total * it * 512 * 512 * 999 / wall,
                eta_first_mean=float(np.mean([o["eta_first"] for o in out])),
                eta_last_mean=float(np.mean([o["eta_last"] for o in out])),
                eta_gain_min=float(min(gains)), eta_gain_max=float(max(gains)),
                check_vs_last_max_rel=float(max(abs(o["eta_plain_check"] - o["eta_last"]) / o["eta_last"] for o in out)),
                per_instance=out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="headline workload only")
    ap.add_argument("--profile-reps", type=int, default=200)
    ap.add_argument("--variants", default="",
                    help="ROW,COL: pin the THERMAL march's layouts instead of tuning (profiler runs, "
                         "whose serialised launches mislead the tuner)")
    ap.add_argument("--plain-variants", default="", help="ROW,COL: the same for the plain 256x256 march")
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--single-sample", type=int, default=32)
    ap.add_argument("--inverse-instances", type=int, default=16)
    ap.add_argument("--inverse-iterations", type=int, default=2)
    ap.add_argument("--section", choices=["all", "plain", "batch", "adjoint", "inverse", "inverse-full",
                                          "thermal-sweep", "jacobi"], default="all",
                    help="run one extra workload alone and print its figures (for profiler runs)")
    ap.add_argument("--jacobi-grid", default="",
                    help="with --section jacobi: one heat grid only (64x64x200, 256x256x100, 256x256x400, 512x512x200)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device copy of this many MiB 5 times (calibrates rocprof byte counters)")
    args = ap.parse_args()
    if args.copy_probe_mib > 0:
        import torch
        x = torch.ones(args.copy_probe_mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read_per_copy": x.numel() * 4, "copies": 5}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the solver has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import shg_focused_gaussian_beams_amd as shg
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)

    if args.section != "all":
        if args.section == "plain":
            r = bench_plain(shg, args, shg.thermal256_params(thermal=0))
        elif args.section == "batch":
            r = bench_batch(shg, args)
        elif args.section == "inverse":
            r = bench_inverse(shg, args)
        elif args.section == "inverse-full":
            r = bench_inverse_full(shg, args)
        elif args.section == "thermal-sweep":
            r = bench_thermal_sweep(shg, args)
        elif args.section == "jacobi":
            r = bench_jacobi_layouts(shg, args)
        else:
            r = bench_adjoint(shg, args)
            r["res"] = {k: v for k, v in r["res"].items()}
        print(json.dumps({"section": args.section,
                          **{k: v for k, v in r.items() if isinstance(v, (int, float, dict, list))}}), flush=True)
        return

    th = bench_thermal(shg, args)
    p = th["p"]
    ms_per_step = th["wall"] * 1e3 / th["steps"]
    value = th["cells"] / (th["wall"] / th["steps"])
    step_s = th["dev_ms"] * 1e-3 / p.nz                 # device time per z-step of the timed marches
    achieved = B_THERMAL * p.nx * p.ny / step_s / 1e9
    copy_gbps = copy_peak_gbps()
    out = {
        "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Gaussian pump + 1e-6 relative complex noise, SH = 0)",
        "config": {"workload": "THERMAL SSFM march: MgO:sPPLT 30 mm, P=30 W, w0=40 um, 256x256, nz=400, "
                               "heat grid 256x256x100, Jacobi to mean squared update < 1e-6 (the reference's "
                               "residual metric; met after one sweep per outer iteration here, so the outer loop "
                               "under the reference's own stopping rule is timed beside it)",
                   "nx": p.nx, "ny": p.ny, "nz": p.nz, "heat_nz": p.heat_nz,
                   "cell_steps_per_march": th["cells"], "eta": th["eta"],
                   "device_ms_per_march": th["dev_ms"],
                   "algorithmic_GBps": value * B_THERMAL / 1e9,
                   "thermal_loop": {**th["stats"], "wall_s": th["loop_s"]},
                   "jacobi_voxel_sweeps_per_s": th["vox"] / th["jac_s"], "jacobi_layout": th["jac_layout"],
                   "jacobi_GBps": th["vox"] * B_JACOBI / th["jac_s"] / 1e9},
        "roofline": {"bound": "hbm", "kernel": "fused THERMAL z-step = %s + %s (256x256)" % (
                         THERM_ROW_256.get(th["variants"][0], "?"), COL_256.get(th["variants"][1], "?")),
                     "variants": th["variants"],
                     "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": TRAFFIC_THERMAL_STEP,
                     "bytes_per_cell_step": B_THERMAL, "step_us": step_s * 1e6,
                     "row_pass_us": th["row_ms"] * 1e3, "col_pass_us": th["col_ms"] * 1e3,
                     "measured_copy_GBps": copy_gbps, "frac_of_copy": achieved / copy_gbps},
    }
    if not args.no_extra:
        log("THERMAL outer loop under the reference's stopping rule ...")
        rr = bench_thermal_reference_rule(shg, args)
        st = rr["stats"]
        out["config"]["thermal_loop_reference_rule_256x256x400"] = {
            **st, "wall_s": rr["wall"], "eta": rr["eta"], "heat_nz": rr["p"].nz,
            "jacobi_voxel_sweeps_per_s": rr["p"].nx * rr["p"].ny * rr["p"].nz * st["sweeps_total"] /
                                         (st["jacobi_ms"] * 1e-3),
            "march_cell_steps_per_s": rr["p"].nx * rr["p"].ny * rr["p"].nz * st["outer_iterations"] /
                                      (st["march_ms"] * 1e-3)}
        log("the reference's default run (THERMAL 64x64x200) ...")
        dr = bench_reference_default_run(shg, args)
        out["config"]["reference_default_run_64x64x200"] = {**dr["stats"], "wall_s": dr["wall"], "eta": dr["eta"]}
        log("temperature sweep of the default run (26 points, batch / contexts / one at a time) ...")
        out["config"]["thermal_sweep_26_points_64x64x200"] = bench_thermal_sweep(shg, args)
        log("plain 128x128, nz 200 ...")
        p1 = bench_plain(shg, args, shg.plain128_params())
        out["config"]["plain_128x128_nz200"] = {
            "cell_steps_per_s": p1["cells"] / p1["s_per_solve"], "ms_per_solve": p1["s_per_solve"] * 1e3,
            "row_pass_us": p1["row_ms"] * 1e3, "col_pass_us": p1["col_ms"] * 1e3, "variants": p1["variants"]}
        log("plain 256x256, nz 400 ...")
        pl = bench_plain(shg, args, shg.thermal256_params(thermal=0))
        out["config"]["plain_256x256_nz400"] = {
            "cell_steps_per_s": pl["cells"] / pl["s_per_solve"], "ms_per_solve": pl["s_per_solve"] * 1e3,
            "row_pass_us": pl["row_ms"] * 1e3, "col_pass_us": pl["col_ms"] * 1e3, "variants": pl["variants"],
            "step_GBps_algorithmic": B_PLAIN * 65536 / ((pl["row_ms"] + pl["col_ms"]) * 1e-3) / 1e9}
        log("batched sweep of 256 uneven instances ...")
        bt = bench_batch(shg, args)
        out["config"]["batched_sweep_256_uneven"] = {
            "instances": bt["count"], "s_per_batch": bt["s"], "cell_steps_per_s": bt["cells"] / bt["s"],
            "instances_per_s": bt["count"] / bt["s"], "algorithmic_GBps": bt["cells"] * B_PLAIN / bt["s"] / 1e9,
            "hbm_frac": bt["cells"] * B_PLAIN / bt["s"] / 1e9 / HBM_PEAK_GBPS,
            "valu_frac_est": bt["cells"] * VALU_OPS_PER_CELL_STEP / bt["s"] / VALU_PEAK_LANE_OPS,
            "one_at_a_time_s": bt["single_s"], "speedup_vs_one_at_a_time": bt["single_s"] / bt["s"],
            "one_at_a_time_sampled_instances": bt["sampled"], "mean_eta": bt["eta_mean"]}
        log("forward+adjoint 256x256, nz 400 ...")
        ad = bench_adjoint(shg, args)
        out["config"]["fwd_adjoint_256x256_nz400"] = {
            "forward_ms": ad["fwd_s"] * 1e3, "fwd_adj_ms": ad["adj_s"] * 1e3,
            "fwd_adj_over_fwd": ad["adj_s"] / ad["fwd_s"],
            "fwd_adj_sqrt_checkpoints_ms": ad["adj_sqrt_s"] * 1e3,
            "fwd_adj_sqrt_checkpoints_over_fwd": ad["adj_sqrt_s"] / ad["fwd_s"], "sqrt_interval": ad["k_sqrt"],
            "cell_steps_per_s": ad["cells"] / ad["adj_s"],
            "checkpoint_interval": ad["res"]["checkpoint_interval"], "eta": ad["res"]["eta"],
            "d_eta_d_temp": ad["res"]["d_eta_d_temp"], "d_eta_d_focus": ad["res"]["d_eta_d_focus"]}
        log("inverse design 512x512, nz 1000 (sample) ...")
        out["config"]["inverse_design_512x512_nz1000_sample"] = bench_inverse(shg, args)
    if not args.no_cpu:
        log("cpu baseline ...")
        cb = cpu_baseline_thermal(p, th["T"], th["fh0"], th["sh0"])
        cb["gpu_speedup"] = value / cb["value"]
        out["cpu_baseline"] = cb
        bsw = out["config"].get("batched_sweep_256_uneven")
        if bsw is not None:
            log("cpu baseline of the sweep (sampled) ...")
            cs = cpu_batch_projection(shg.uneven_sweep_params(count=args.batch, seed=3))
            cs["gpu_speedup"] = cs["serial_s_projected"] / bsw["s_per_batch"]
            bsw["cpu_serial"] = cs
        if not args.no_extra:
            log("cpu baselines of the small workloads ...")
            c = out["config"]
            cw = cpu_small_workloads(shg, c["reference_default_run_64x64x200"])
            cw["plain_128x128_nz200"]["gpu_speedup"] = (c["plain_128x128_nz200"]["cell_steps_per_s"] /
                                                        cw["plain_128x128_nz200"]["cell_steps_per_s"])
            cw["jacobi_256x256x100"]["gpu_speedup"] = (c["jacobi_voxel_sweeps_per_s"] /
                                                       cw["jacobi_256x256x100"]["voxel_sweeps_per_s"])
            cw["reference_default_run"]["gpu_speedup"] = (cw["reference_default_run"]["s_projected"] /
                                                          c["reference_default_run_64x64x200"]["wall_s"])
            sw = c["thermal_sweep_26_points_64x64x200"]
            rd = cw["reference_default_run"]
            proj = sum(s * rd["s_per_sweep"] + o * rd["s_per_march"]
                       for s, o in zip(sw["sweeps_total"], sw["outer_iterations"]))
            cw["thermal_sweep_26_points"] = {
                "s_projected": proj, "gpu_speedup_batch": proj / sw["batch_s"],
                "method": "each point's sweep and outer-iteration counts times the default run's timed "
                          "sweep and march"}
            out["cpu_baseline"]["other_workloads"] = cw
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
