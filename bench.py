#!/usr/bin/env python3
"""Pressure-solve benchmark on one MI355X.

Metric: ms per pressure solve to 1e-10 relative residual (and HBM GB/s),
FP32-V-cycle / FP64-refinement form versus the FP64 GPU multigrid.

Workload (default): 256^3 box, periodic in x, walls in y and z, 64 droplets
(centres uniform, radii uniform in [0.03, 0.08], seed 1), density ratio 1000,
1.5-cell tanh interface, right-hand side = minus the divergence of iid
U(-1,1) face velocities, mean removed.  A step = one complete solve from
p = 0 to ||b - A p||_2 <= 1e-10 ||b||_2 on the device-resident 8-slot system.

Also reported on the same line:
  * roofline: live HIP-event timing (the dispatch's own start and end) of
    one in 7 launches of the finest-level smoother kernel during the timed
    region (algorithmic bytes / kernel time vs 8 TB/s);
  * cpu_baseline: the reference algorithm (NewSolverMG restated in C, on
    all OpenMP threads and on one) timed on this host for a bounded number
    of V-cycles, scaled to a whole solve with the iteration count of the
    same algorithm (counted by its GPU restatement, whose residual history
    is the reference's).
"""
import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "ms per pressure solve to 1e-10 rel. residual + HBM GB/s; mixed vs FP64 GPU MG"


def level_dims(n3, periodic=(0, 0, 0), lengths=(1.0, 1.0, 1.0)):
    """Level sizes of the fast hierarchy (the library's plan: every axis is
    halved, ceil(n / 2), unless its cells are more than 1.5 times as long as
    the shortest, it is periodic with an odd size, or it has one cell; a
    level with an odd periodic axis is the coarsest; stop at 64 cells)."""
    n = list(n3)
    h = [lengths[a] / n[a] for a in range(3)]
    dims = [tuple(n)]
    while n[0] * n[1] * n[2] > 64:
        if any(periodic[a] and n[a] % 2 for a in range(3)):
            break
        hmin = min(h)
        f = [2 if (n[a] > 1 and h[a] <= 1.5 * hmin) else 1 for a in range(3)]
        if max(f) == 1:
            break
        for a in range(3):
            if f[a] == 2:
                n[a] = (n[a] + 1) // 2
                h[a] *= 2.0
        dims.append(tuple(n))
    return dims


def solve_bytes(n3, iters, S, nu_pre=2, nu_post=2, tail_cells=4096):
    """Algorithmic bytes of one fast solve of a singular system whose A7 is
    the coupling sum (every array element counted once per launch that must
    touch it; S = bytes of the V-cycle type)."""
    dims = level_dims(n3)
    n0 = n3[0] * n3[1] * n3[2]
    v = 0.0
    for li, (a, b, c) in enumerate(dims[:-1]):
        n = a * b * c
        if li > 0 and n <= tail_cells:
            v += sum(x * y * z for (x, y, z) in dims[li:]) * S * 5    # LDS tail: staged once
            break
        v += n * S * (5 + 6 * (nu_pre - 1))                 # pre-smoothing (first: no x read)
        v += n * S * ((6 + 1 / 8) + 6 * (nu_post - 1))      # post (first adds the correction)
        v += n * S * (5 + 1 / 8)                            # residual + restriction
    else:
        nb = dims[-1][0] * dims[-1][1] * dims[-1][2]
        v += nb * nb * S + 2 * nb * S                       # dense bottom solve
    mixed = S == 4
    # the PCG step: (r, z) comes with the last sweep; d & A d reads z, d_old,
    # the FP64 couplings and x (the previous x += alpha d), writes d (type
    # S), q and x; the update reads r, q and writes r (+ its FP32 copy)
    outer = n0 * ((48 + 3 * S) + (24 + (4 if mixed else 0)))
    setup = n0 * (72 + 56 + (20 if mixed else 0))           # A, p in; op, b, x, r (+ FP32) out
    for (a, b, c) in dims[1:]:
        setup += a * b * c * S * (8 * 4 + 4)                # coarsening
    final = n0 * (16 + S)                                   # the last x += alpha d
    final += n0 * (40 + 8)                                  # true residual + write-back of p
    return setup + iters * (v + outer) + final


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_rb_sweep_drops256_traffic.md")


def measured_traffic():
    """HBM bytes per launch of the profiled sweep kernel, from the committed
    counter summary (FETCH_SIZE + WRITE_SIZE, calibrated against a copy of
    known size); None when no summary is present."""
    try:
        with open(TRAFFIC_FILE) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def gpu_run(pps, torch, grid, A, method, tol, steps, warmup, maxit=200):
    s = pps.Solver(grid, method=method)
    p = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
    for _ in range(warmup):
        p.zero_()
        s.solve(A, p, tol, 1.2, maxit, allow_maxit=False)
    # one launch in 7 of the timed sweep kernel carries the event pair (7 is
    # prime to the iteration count, so the samples rotate through the cycles)
    s.profile(True, every=7)
    s.profile_read()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = []
    for _ in range(steps):
        p.zero_()
        it, t2 = s.solve(A, p, tol, 1.2, maxit, allow_maxit=False)
        its.append(it)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kms, launches, kbytes = s.profile_read()
    s.profile(False)
    solve_ms_ev, setup_ms_ev = s.timing()
    s.close()
    return dict(ms=(t1 - t0) * 1e3 / steps, it=its[-1], tres2=t2, kernel_ms=kms,
                launches=launches, kernel_bytes=kbytes, event_solve_ms=solve_ms_ev,
                event_setup_ms=setup_ms_ev)


def reference_iterations(pps, torch, grid, A, tol, cap):
    """Iterations the reference algorithm needs (its GPU restatement), run
    for at most `cap` V-cycles; returns (count, extrapolated, history,
    wall ms of that capped run)."""
    s = pps.Solver(grid, method="reference")
    p = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
    t0 = time.perf_counter()
    it, t2 = s.solve(A, p, tol, 1.2, cap)
    gpu_ms = (time.perf_counter() - t0) * 1e3
    hist = s.history()
    s.close()
    if it <= cap:
        return it, False, hist, gpu_ms
    # not converged within the cap: extrapolate the geometric rate of the
    # last quarter of the history
    q = max(len(hist) // 4, 2)
    rate = (hist[-1] / hist[-q]) ** (1.0 / (q - 1))
    extra = math.log(tol / hist[-1]) / math.log(rate) if rate < 1 else float("inf")
    return cap + extra, True, hist, gpu_ms


def cpu_baseline(A_host, grid, cycles, n_solve_cycles, extrapolated):
    """The reference algorithm (the oracle's NewSolverMG restatement) on this
    host: `cycles` V-cycles on all cores the OpenMP runtime offers
    (OMP_NUM_THREADS) and on one core, scaled to a whole solve."""
    import numpy as np
    import __graft_entry__ as ge
    orc = ge.load_oracle()
    og = orc.OrcGrid.make(grid.nx, grid.ny, grid.nz, bc=tuple(grid.bdry_cond))
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass

    def per_cycle(threads):
        orc.set_threads(threads)
        p = np.zeros(og.shape_p)
        t0 = time.perf_counter()
        orc.newsolver_mg(og, A_host, p, 1e-300, 1.2, cycles)
        dt = time.perf_counter() - t0
        orc.set_threads(1)
        return dt / cycles

    nt = orc.max_threads()
    t_all = per_cycle(nt)
    t_one = per_cycle(1)
    how = "extrapolated" if extrapolated else "counted"
    return {
        "value": t_all * n_solve_cycles * 1e3,
        "unit": "ms",
        "cores": nt,
        "kind": "port",
        "sample": (f"{cycles} V-cycles of the reference NewSolverMG algorithm (oracle C port, "
                   f"OpenMP, {nt} threads, {cpu}) on the same system: {t_all * 1e3:.1f} ms/V-cycle x "
                   f"{n_solve_cycles:.0f} V-cycles to reach 1e-10 ({how} by the GPU restatement "
                   f"of the same algorithm); one thread: {t_one * 1e3:.1f} ms/V-cycle"),
        "ms_per_vcycle": t_all * 1e3,
        "vcycles_to_solve": n_solve_cycles,
        "one_core": {"value": t_one * n_solve_cycles * 1e3, "ms_per_vcycle": t_one * 1e3, "cores": 1},
    }


def free_run(pps, torch, args):
    """FreeSolver (solver_flag 1): cells with density above the mean are
    gas cells (pcmask 1, p = P_gas = 0), the rest liquid; a fixed number of
    iterations (maxError 0), timed with HIP events inside the library.
    Algorithmic bytes per iteration: the two colour passes read the 8 A
    planes, pcmask and p once and write p (84 B/cell), the residual pass
    reads the 8 planes, pcmask and p again (76 B/cell)."""
    import numpy as np
    print(f"[bench] building {args.workload}", file=sys.stderr, flush=True)
    grid, A, f = pps.make_problem(args.workload)
    rho = f["rho"]
    pcmask = (rho > rho.mean()).to(torch.int32).contiguous()
    pgas = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
    ms = []
    with pps.Solver(grid, method="sor") as s:
        for step in range(args.warmup + args.steps):
            p = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
            it, t2 = s.free_solve(A, p, pcmask, pgas, 1, 0.0, 1.2, args.free)
            if step >= args.warmup:
                ms.append(s.timing()[0])
    per_it = float(np.median(ms)) / args.free
    gbps = grid.cells * (84 + 76) / (per_it * 1e-3) / 1e9
    print(json.dumps({"mode": "free_solver", "workload": args.workload, "cells": grid.cells,
                      "iterations": args.free, "ms_per_iteration": round(per_it, 4),
                      "algorithmic_GBps": round(gbps, 1),
                      "gas_cells": int(pps.Grid.interior(grid, pcmask).sum().item()),
                      "steps": args.steps, "warmup": args.warmup, "dtype": "f64"}), flush=True)


def project_run(pps, torch, args):
    """Mixed solve followed by project_velocity: the fused entry point (the
    correction inside the last residual pass) against the solve alone and
    against the solve plus the standalone projection call (which also
    uploads the grid spacing on every call).  Also one time step's system
    from the fields: SetupPoisson on the device then the solve, against the
    solve that assembles its rows inside its setup pass (A never formed).
    CUDA events on the current stream; medians over the timed steps."""
    import numpy as np
    print(f"[bench] building {args.workload}", file=sys.stderr, flush=True)
    grid, A, f = pps.make_problem(args.workload)
    brms = float(torch.sqrt((A[7] ** 2).mean()))
    tol = 1e-10 * brms
    dt = 1.0
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    res = {"solve": [], "solve_then_project": [], "fused": [], "assemble_then_solve": [], "fields": []}
    with pps.Solver(grid, method="mixed") as s:
        p = torch.zeros(grid.shape_p, dtype=torch.float64, device="cuda")
        for step in range(args.warmup + args.steps):
            for mode in res:
                p.zero_()
                u = f["u"].clone()
                torch.cuda.synchronize()
                ev[0].record()
                if mode == "fused":
                    s.solve_project(A, p, f["u"], f["v"], f["w"], f["um"], f["vm"], f["wm"], f["rho"],
                                    dt, tol, 1.2, 100)
                elif mode == "assemble_then_solve":
                    A2 = pps.setup_poisson(grid, f["u"], f["v"], f["w"], f["um"], f["vm"], f["wm"], f["rho"], dt)
                    s.solve(A2, p, tol, 1.2, 100)
                    del A2
                elif mode == "fields":
                    s.solve_fields(f, p, tol, dt=dt, maxit=100)
                else:
                    s.solve(A, p, tol, 1.2, 100)
                    if mode == "solve_then_project":
                        pps.project_velocity(grid, u, f["v"].clone(), f["w"].clone(), f["um"],
                                             f["vm"], f["wm"], f["rho"], p, dt)
                ev[1].record()
                torch.cuda.synchronize()
                if step >= args.warmup:
                    res[mode].append(ev[0].elapsed_time(ev[1]))
    out = {k: round(float(np.median(v)), 3) for k, v in res.items()}
    print(json.dumps({"mode": "solve_project", "workload": args.workload, "cells": grid.cells,
                      "ms": out, "steps": args.steps, "warmup": args.warmup,
                      "dtype": "f32 V-cycle / f64 refinement"}), flush=True)


def autograd_run(pps, torch, args):
    """Forward (mixed solve from face values) + backward (adjoint solve and
    face kernel) of PressureSolve, J = 0.5 ||p - mean(p) - p_target||^2."""
    import numpy as np
    print(f"[bench] building {args.workload}", file=sys.stderr, flush=True)
    grid, A, f = pps.make_problem(args.workload)
    ng = grid.ng
    R = f["rho"][ng - 1:ng + grid.nz + 1, ng - 1:ng + grid.ny + 1, ng - 1:ng + grid.nx + 1]
    c = R[1:-1, 1:-1, 1:-1]
    cf = torch.stack([2.0 / (c + R[1:-1, 1:-1, 2:]), 2.0 / (c + R[1:-1, 2:, 1:-1]),
                      2.0 / (c + R[2:, 1:-1, 1:-1])]).contiguous()
    per = grid.periodic()
    if not per[0]:
        cf[0, :, :, -1] = 0.0
    if not per[1]:
        cf[1, :, -1, :] = 0.0
    if not per[2]:
        cf[2, -1, :, :] = 0.0
    b = A[7].clone()
    w = pps.face_weights(grid)
    rng = np.random.default_rng(21)
    ax = [(torch.arange(m, dtype=torch.float64, device="cuda") + 0.5) / m for m in (grid.nz, grid.ny, grid.nx)]
    pt = torch.zeros((grid.nz, grid.ny, grid.nx), dtype=torch.float64, device="cuda")
    for _ in range(8):
        k = rng.integers(1, 5, size=3)
        pt += rng.uniform(-1, 1) * torch.sin(2 * np.pi * (int(k[0]) * ax[2].view(1, 1, -1) + int(k[1]) * ax[1].view(1, -1, 1)
                                                          + int(k[2]) * ax[0].view(-1, 1, 1)) + rng.uniform(0, 2 * np.pi))
    PS = pps.PressureSolve
    fwd, bwd = [], []
    with pps.Solver(grid, method="mixed") as s:
        for step in range(args.warmup + args.steps):
            cc = cf.clone().requires_grad_(True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            p = PS.apply(cc, b, s, w, 1e-10, 200)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            it_f = len(s.history())
            J = 0.5 * (((p - p.mean()) - pt) ** 2).sum()
            J.backward()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            it_b = len(s.history())
            if step >= args.warmup:
                fwd.append((t1 - t0) * 1e3)
                bwd.append((t2 - t1) * 1e3)
    print(json.dumps({"mode": "autograd", "workload": args.workload, "cells": grid.cells,
                      "forward_ms": round(float(np.median(fwd)), 3), "backward_ms": round(float(np.median(bwd)), 3),
                      "iterations": {"forward": it_f, "adjoint": it_b}, "steps": args.steps,
                      "warmup": args.warmup, "dtype": "f32 V-cycle / f64 refinement"}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="drops256")
    ap.add_argument("--cpu-cycles", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ref-cap", type=int, default=6000,
                    help="V-cycles the reference method may run to count its iterations")
    ap.add_argument("--only", choices=("both", "mixed", "fp64"), default="both",
                    help="time one form only (for counter runs)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="run one device copy of MIB MiB (counter calibration) and exit")
    ap.add_argument("--free", type=int, default=0, metavar="ITERS",
                    help="time ITERS FreeSolver iterations (masked RB-SOR, solver_flag 1, the "
                         "droplets of --workload as gas cells) instead of the solve benchmark")
    ap.add_argument("--project", action="store_true",
                    help="time solve + project_velocity, fused against separate")
    ap.add_argument("--autograd", action="store_true",
                    help="time forward + backward of the differentiable solve on --workload "
                         "(mixed form) instead of the solve benchmark")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the solver runs on a single MI355X "
              "(multi-GPU sharding is designed, not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import __graft_entry__ as ge
    if not torch.cuda.is_available():
        print("no GPU", file=sys.stderr)
        sys.exit(1)
    if args.copy_calib:
        # a plain copy of known size: FETCH_SIZE / WRITE_SIZE of this kernel
        # calibrate the counters (bytes = MiB * 2^20 each way)
        n = args.copy_calib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 1.0, out=b)     # one elementwise kernel: read n, write n
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": n * 8}))
        return
    pps = ge.load_package()
    if args.autograd:
        autograd_run(pps, torch, args)
        return
    if args.project:
        project_run(pps, torch, args)
        return
    if args.free:
        free_run(pps, torch, args)
        return

    print(f"[bench] building {args.workload}", file=sys.stderr, flush=True)
    grid, A, _ = pps.make_problem(args.workload)
    n3 = (grid.nx, grid.ny, grid.nz)
    brms = float(torch.sqrt((A[7] ** 2).mean()))
    tol = 1e-10 * brms

    res = {}
    methods = ("fp64", "mixed") if args.only == "both" else (args.only,)
    for method in methods:
        print(f"[bench] {method}: {args.warmup} warmup + {args.steps} timed solves",
              file=sys.stderr, flush=True)
        res[method] = gpu_run(pps, torch, grid, A, method, tol, args.steps, args.warmup)
        r = res[method]
        S = 8 if method == "fp64" else 4
        r["solve_bytes"] = solve_bytes(n3, r["it"], S)
        r["solve_GBps"] = r["solve_bytes"] / (r["ms"] * 1e-3) / 1e9
        r["smoother_GBps"] = r["kernel_bytes"] / (r["kernel_ms"] * 1e-3) / 1e9
        print(f"[bench] {method}: {r['ms']:.3f} ms/solve, {r['it']} its, "
              f"rel.res {r['tres2'] / brms:.2e}, smoother {r['smoother_GBps']:.0f} GB/s",
              file=sys.stderr, flush=True)

    if args.only != "both":
        r = res[args.only]
        print(json.dumps({"form": args.only, "ms": r["ms"], "it": r["it"],
                          "smoother_GBps": r["smoother_GBps"]}), flush=True)
        return
    m = res["mixed"]
    f64 = res["fp64"]
    n_cells = grid.cells
    roof = {
        "kernel": "k_rb_sweep<float, false, 2, true, true, false, true, false, true> (finest-level "
                  "fused red-black Gauss-Seidel sweep, FP32, couplings staged through LDS)",
        "bound": "hbm",
        "achieved": round(m["smoother_GBps"], 1),
        "peak": HBM_PEAK_GBPS,
        "unit": "GB/s",
        "frac": round(m["smoother_GBps"] / HBM_PEAK_GBPS, 4),
        "traffic": measured_traffic(),
        "traffic_source": ("profiles/k_rb_sweep_drops256_traffic.md: rocprofv3 --pmc FETCH_SIZE and "
                           "WRITE_SIZE of this kernel in 'bench.py --only mixed --steps 1', calibrated "
                           "by 'bench.py --copy-calib 1024' (not measured in this run)"),
        "launches": m["launches"],
        "avg_launch_us": round(m["kernel_ms"] / max(m["launches"], 1) * 1e3, 2),
        "bytes_per_launch": round(m["kernel_bytes"] / max(m["launches"], 1)),
    }

    cpu = None
    ref_info = None
    if not args.no_cpu:
        print(f"[bench] reference algorithm iteration count on the GPU (cap {args.ref_cap})",
              file=sys.stderr, flush=True)
        nref, extrap, hist, ref_ms = reference_iterations(pps, torch, grid, A, tol, args.ref_cap)
        ref_info = {"vcycles_to_solve": nref, "extrapolated": extrap, "cap": args.ref_cap,
                    "vcycles_run": len(hist),
                    "rel_residual_at_end": (hist[-1] / brms) if hist else None,
                    "run_wall_ms": ref_ms,
                    "run": ("the reference method on the GPU, stopped at 1e-10 or at the cap; "
                            "run_wall_ms is that run's wall time (the whole solve when not "
                            "extrapolated)")}
        print(f"[bench] reference algorithm: {nref:.0f} V-cycles; CPU sample of "
              f"{args.cpu_cycles} cycles", file=sys.stderr, flush=True)
        A_host = np.ascontiguousarray(A.cpu().numpy())
        cpu = cpu_baseline(A_host, grid, args.cpu_cycles, nref, extrap)
        del A_host

    line = {
        "metric": METRIC,
        "value": round(m["ms"], 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(m["ms"], 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycle / f64 refinement",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {grid.nx}x{grid.ny}x{grid.nz}, bdry_cond "
                        f"{list(grid.bdry_cond)}, droplet field, density ratio 1000, "
                        f"rel. residual 1e-10",
            "cells": n_cells,
            "parallelism": "single GPU",
        },
        "fp64_ms": round(f64["ms"], 4),
        "speedup_vs_fp64": round(f64["ms"] / m["ms"], 3),
        "iterations": {"mixed": m["it"], "fp64": f64["it"]},
        "rel_residual": {"mixed": m["tres2"] / brms, "fp64": f64["tres2"] / brms},
        "solve_GBps": {"mixed": round(m["solve_GBps"], 1), "fp64": round(f64["solve_GBps"], 1)},
        "smoother_GBps": {"mixed": round(m["smoother_GBps"], 1), "fp64": round(f64["smoother_GBps"], 1)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "reference_algorithm_gpu": ref_info,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
