"""Benchmark of the per-cell HDG hot path on one MI355X.

Workload (the configuration the metric is quoted on): 2D silicon p-n diode,
256 x 256 quadrilateral cells, HDG degree 2, V/n/p, FP64 (81 cell dofs,
36 trace dofs per cell), state = thermal-equilibrium first guess under a
0.3 V bias boundary condition.  One step = pass 1 of a Newton step over all
65,536 cells: assembly, static condensation and the deterministic scatter
into the global trace system (hdg_condense), inputs resident in HBM.

Prints one JSON line.  ``value`` = cells assembled+condensed per second;
``roofline`` = the condense kernel against MI355X's FP64 peak with the
dense algorithmic FLOP count per cell; ``cpu_baseline`` = the oracle (the
reference-faithful C++ restatement: Gauss-Jordan inverse + mmult, WorkStream
threading, serial in-order scatter) timed on the host on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
# HBM bytes one pass-1 launch (32,768 cells of the 256x256 degree-2 workload)
# moves, from rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (factors from a
# plain device copy of known size): profiles/condense_traffic_2d_k2_256x256.md
TRAFFIC_PER_LAUNCH = {(256, 2): (96310.27 * 2.0 + 422006.80 * 1.0) * 1024}


_T0 = time.perf_counter()


def progress(msg):
    """one line on stderr as each part of the run ends (stdout keeps the
    single JSON line)"""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def flops_per_cell(N, T):
    """dense algorithmic count for pass 1: LU of the N x N cell block,
    solves for the T trace columns and the rhs, Schur product."""
    return 2 * N ** 3 / 3 + 2 * N ** 2 * (T + 1) + 2 * T * N * (T + 1)


def bytes_per_cell(N, T):
    """pass 1: read the cell state (N) and the trace state (T), write the
    condensed block (T*T) and right-hand side (T)."""
    return 8 * (N + T + T * T + T)


def host_cpu():
    """the host the CPU baseline runs on: logical CPUs, the ones this
    process may use, the model name"""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"nproc": os.cpu_count(), "usable": usable, "model": model}


def native_oracle():
    """builds the oracle for this host (-march=native, outside the tree); the
    shipped build targets x86-64-v3 so that it runs on any host.  Returns
    the flags of the build in use."""
    import subprocess
    import tempfile

    from oracle import oracle as O
    out = os.path.join(tempfile.gettempdir(), f"liboracle_native_{os.getpid()}.so")
    src = os.path.join(ROOT, "oracle", "hdg_oracle.cpp")
    r = subprocess.run(["g++", "-O3", "-march=native", "-std=c++17", "-fPIC", "-shared", "-o",
                        out, src, "-lpthread"], capture_output=True, text=True)
    if r.returncode == 0:
        O.use_library(out)
        return "-O3 -march=native (built on this host)"
    return "-O3 -march=x86-64-v3 (shipped build; the native build failed)"


def cpu_baseline(n_sample, degree, threads, bias, reps=5):
    """oracle pass 1 on an n_sample^2 diode of the same per-cell shape: one
    warm-up, then the median of reps timed runs"""
    import numpy as np

    from ddhdg_amd import configs
    from oracle import oracle as O
    from tests import helpers as H

    prob, params, adim = configs.diode(2, n_sample, degree, bias=bias)
    o = H.oracle_for(prob, params, adim)
    o.set_state(*H.equilibrium_guess(prob, adim, degree))
    o.pass1(False, threads)  # warm-up
    times = [o.pass1(False, threads) for _ in range(reps)]
    t = float(np.median(times))
    return o.ncells / t, o.ncells, t


def batched_sample(args, torch, stream):
    """pass 1 of many devices of uneven sizes and degrees: the degree-1 and
    degree-2 devices in one persistent launch per colour, the degree-3
    devices (pair kernel) in another (autotuned grids); against all
    devices condensed one after the other (two launches each).  FP64 and
    FP32 + refinement."""
    import numpy as np

    from ddhdg_amd import configs
    from ddhdg_amd.hip_lib import HdgBatch
    from ddhdg_amd.np_solver import NPSolver

    sv = []
    for prob, params, adim in configs.batched_devices(args.batch, 20261203):
        s = NPSolver(prob, params, adim, device="cuda")
        s.compute_local_charge_neutrality()
        s.project_densities_from_potential()
        sv.append(s)
    bufs = [(torch.empty(s.ctx.nnz, dtype=torch.float64, device="cuda"),
             torch.empty(s.ctx.ntrace, dtype=torch.float64, device="cuda"),
             torch.zeros(s.ctx.ntrace, dtype=torch.float64, device="cuda")) for s in sv]
    inb = [i for i, s in enumerate(sv) if s.degree <= 2]
    rest = [i for i, s in enumerate(sv) if s.degree > 2]
    groups = []
    for ids in (inb, rest):
        if ids:
            groups.append((HdgBatch([sv[i].ctx for i in ids]),
                           [sv[i].x_cell for i in ids], [sv[i].x_trace for i in ids],
                           *([bufs[i][j] for i in ids] for j in range(3))))

    def batched():
        for b, xc, xt, v, r, d in groups:
            b.condense(False, xc, xt, v, r, d)

    def sequential():
        for s, (vv, rr, dd) in zip(sv, bufs):
            s.ctx.condense(s.x_cell, s.x_trace, False, vv, rr, dd)

    def timed(fn, reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, e0.elapsed_time(e1) / reps

    reps = max(3, min(args.steps, 20))
    cells = sum(s.ctx.ncells for s in sv)
    out = {"devices": len(sv), "cells": int(cells),
           "degrees": {str(k): sum(1 for s in sv if s.degree == k) for k in (1, 2, 3)}}
    for prec in (0, 1, 2):
        for s in sv:
            s.ctx.set_precision(prec)
        batched()                      # first call autotunes the grid
        torch.cuda.synchronize()
        wb, gb = timed(batched, reps)
        ws, gs = timed(sequential, reps)
        # precision 1: FP32 + refinement in the one-wave kernels (degree 1-2),
        # degree 3 in FP64; precision 2: in the degree-3 pair kernel too
        key = ("fp64", "fp32_refine", "fp32_refine_every_kernel")[prec]
        out[key] = {"value": cells / wb, "unit": "cells/s", "gpu_ms": gb,
                    "sequential_value": cells / ws, "sequential_gpu_ms": gs}
    for s in sv:
        s.ctx.set_precision(0)
    out["value"] = out["fp64"]["value"]
    out["unit"] = "cells/s"
    out["autotune"] = groups[0][0].info()
    if len(groups) > 1:
        out["autotune_degree3"] = groups[1][0].info()
    if args.batch_newton:
        del groups, bufs
        # the trace factors are allocated by the library, outside torch's cache
        torch.cuda.empty_cache()
        progress("batched pass 1")
        out["newton"] = batched_newton(torch, sv)
    out["workload"] = (f"{len(sv)} 2D diodes from seed 20261203: sides uniform in 16..192 cells, "
                       "degree uniform in {1, 2, 3}, doping log-uniform 1e21-1e24, junction in "
                       "[0.3, 0.7], contact bias uniform in [0, 0.6] V; pass 1; degrees 1-2 in one "
                       "persistent launch per colour, degree 3 (pair kernel) in another")
    return out


def batched_newton(torch, sv):
    """the devices' Newton workflow (thermodynamic equilibrium, n and p,
    the coupled system) with FP32 + refinement condensation: all devices
    advanced together (BatchedNewton, one hipGraph per stage) against the
    devices solved one after the other (each with its own captured Newton
    iteration), from the same start; the final states are compared"""
    from ddhdg_amd.batch import BatchedNewton, batched_run

    for s in sv:
        s.ctx.set_precision(1)
    init = [(s.x_cell.clone(), s.x_trace.clone(), tuple(s.enabled)) for s in sv]

    def workflow(batched):
        rec = {}
        stages = [("equilibrium", None), ("np", (False, True, True)),
                  ("coupled", (True, True, True))]
        for name, en in stages:
            for s in sv:
                if en is not None:
                    s.set_enabled_components(*en)
                if name == "equilibrium":
                    s.set_enabled_components(True, False, False)
                # the factors of one component set per device at a time
                s.release_trace_solvers()
            torch.cuda.empty_cache()
            # the trace solvers of the stage (host plan, factor storage),
            # built before the clock starts in both modes
            # (and the charge-neutrality first guess, a host projection)
            ts0 = time.perf_counter()
            for s in sv:
                s.trace_solver()
                if name == "equilibrium":
                    s.compute_local_charge_neutrality()
            torch.cuda.synchronize()
            setup = time.perf_counter() - ts0
            t0 = time.perf_counter()
            if name == "equilibrium":
                nl = [s.parameters.nonlinear_parameters for s in sv]
                r = (BatchedNewton(sv, True).run() if batched else
                     [s._private_run(p.absolute_tolerance, p.relative_tolerance,
                                     p.max_number_of_iterations, True) for s, p in zip(sv, nl)])
            else:
                r = batched_run(sv) if batched else [s.run() for s in sv]
            torch.cuda.synchronize()
            elapsed = time.perf_counter() - t0
            if name == "equilibrium":
                for s in sv:
                    s.project_densities_from_potential(0.0, 0.0)
            rec[name] = {"s": elapsed, "setup_s": setup,
                         "iterations": [x.iterations for x in r],
                         "converged": sum(1 for x in r if x.converged)}
        return rec

    batched = workflow(True)
    progress("batched Newton workflow")
    final = [(s.x_cell.clone(), s.x_trace.clone()) for s in sv]
    for s, (c, tr, en) in zip(sv, init):
        s.set_enabled_components(*en)
        s.x_cell.copy_(c)
        s.x_trace.copy_(tr)
        s.parameters.graph_newton = True
    sequential = workflow(False)
    same = all(torch.equal(s.x_cell, c) and torch.equal(s.x_trace, tr)
               for s, (c, tr) in zip(sv, final))
    for s in sv:
        s.parameters.graph_newton = False
        s.ctx.set_precision(0)
    out = {"precision": "fp32_refine, precision 1 (the one-wave register kernels; degree 3 "
                        "and the generic kernel's stages in FP64)",
           "states_bitwise_equal": bool(same)}
    for name in batched:
        b, q = batched[name], sequential[name]
        out[name] = {"batched_s": b["s"], "sequential_s": q["s"],
                     "trace_solver_setup_s": b["setup_s"],
                     "speedup": q["s"] / b["s"] if b["s"] > 0 else None,
                     "iterations_max": max(b["iterations"]),
                     "iterations_total": sum(b["iterations"]),
                     "iterations_equal": b["iterations"] == q["iterations"],
                     "converged": b["converged"], "devices": len(sv)}
    out["workflow_batched_s"] = sum(batched[n]["s"] for n in batched)
    out["workflow_sequential_s"] = sum(sequential[n]["s"] for n in sequential)
    return out


def sample_3d(args, torch, stream):
    """the 3D workload: a 1 um cube diode, 48^3 hexahedra of degree 1, a
    Gaussian acceptor profile (peak 1e24 m^-3, sigma 0.1 um, centred in the
    cube) over uniform donors 1e22 m^-3; pass 1 in FP64 and in FP32 +
    refinement, the error of the latter against FP64 on a sample of cells"""
    from ddhdg_amd import configs
    from ddhdg_amd.np_solver import NPSolver

    n = args.cells3d
    prob, params, adim = configs.diode(3, n, 1, doping=configs.gaussian_doping())
    s = NPSolver(prob, params, adim, device="cuda")
    s.compute_local_charge_neutrality()
    s.project_densities_from_potential()
    ctx = s.ctx
    vals = torch.empty(ctx.nnz, dtype=torch.float64, device="cuda")
    rhs = torch.empty(ctx.ntrace, dtype=torch.float64, device="cuda")
    dirv = torch.zeros(ctx.ntrace, dtype=torch.float64, device="cuda")
    reps = max(3, min(args.steps, 10))
    out = {"workload": f"3D diode {n}^3 hexahedra, degree 1, V/n/p (96 cell dofs, 72 trace dofs), "
                       "Gaussian N_A peak 1e24 m^-3 sigma 0.1 um centred, N_D 1e22 m^-3; pass 1",
           "cells": int(ctx.ncells)}
    T = ctx.T
    sample = torch.arange(0, ctx.ncells, max(1, ctx.ncells // 4096), dtype=torch.int32,
                          device="cuda")
    blocks = {}
    for prec in (0, 1):
        ctx.set_precision(prec)
        for _ in range(2):
            ctx.condense(s.x_cell, s.x_trace, False, vals, rhs, dirv)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            ctx.condense(s.x_cell, s.x_trace, False, vals, rhs, dirv)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        S_ = torch.empty(sample.numel() * T * T, dtype=torch.float64, device="cuda")
        g_ = torch.empty(sample.numel() * T, dtype=torch.float64, device="cuda")
        ctx.cell_systems(s.x_cell, s.x_trace, False, sample, S_, g_)
        blocks[prec] = (S_.view(-1, T * T), g_.view(-1, T))
        key = "fp64" if prec == 0 else "fp32_refine"
        out[key] = {"value": ctx.ncells / (ms * 1e-3), "unit": "cells/s", "ms_per_pass": ms}
    ctx.set_precision(0)
    torch.cuda.synchronize()
    (Sd, gd), (Sm, gm) = blocks[0], blocks[1]
    out["fp32_refine"]["max_rel_err_S_vs_fp64"] = float(((Sm - Sd).norm(dim=1) /
                                                         Sd.norm(dim=1)).max())
    out["fp32_refine"]["max_rel_err_g_vs_fp64"] = float(((gm - gd).norm(dim=1) /
                                                         gd.norm(dim=1).clamp_min(1e-300)).max())
    out["fp32_refine"]["vs_fp64"] = out["fp32_refine"]["value"] / out["fp64"]["value"]
    # pass 2 at the same size (the generic kernel's reconstruct)
    dxt = torch.full((ctx.nfaces * 3 * ctx.nf,), 1e-3, dtype=torch.float64, device="cuda")
    dxc = torch.empty_like(s.x_cell)
    out["pass2_ms"] = event_ms(torch, stream, lambda: ctx.reconstruct(s.x_cell, s.x_trace, dxt,
                                                                     False, dxc), reps)
    # the first Newton stage at full size: the thermodynamic equilibrium (V
    # only; its trace factors take 63 GB, a coupled step's would take 567 GB)
    progress("3D degree-1 pass 1 and pass 2")
    s.set_enabled_components(True, False, False)
    out["newton_equilibrium"] = newton_phases(torch, stream, s, True, 2)
    progress("3D degree-1 equilibrium Newton step")
    s.set_enabled_components(True, True, True)
    del s
    # a coupled Newton step of the same device at 24^3
    prob, params, adim = configs.diode(3, 24, 1, doping=configs.gaussian_doping())
    s = NPSolver(prob, params, adim, device="cuda")
    s.compute_local_charge_neutrality()
    s.project_densities_from_potential()
    out["newton_step_24cube"] = newton_phases(torch, stream, s, False, 2)
    return out


def event_ms(torch, stream, fn, reps):
    """average device time of fn over reps calls (HIP events on the stream)"""
    fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def newton_phases(torch, stream, s, thermodyn, reps):
    """per-phase device times of a whole Newton step of solver s (its
    current enabled components): pass 1, trace factorisation, solve,
    pass 2; the step itself is not applied"""
    ctx = s.ctx
    ts = s.trace_solver()
    rp, ci, dirs, r2f = s._pattern()
    vals = torch.empty(ctx.nnz, dtype=torch.float64, device="cuda")
    rhs = torch.empty(ctx.ntrace, dtype=torch.float64, device="cuda")
    dirv = torch.zeros(ctx.ntrace, dtype=torch.float64, device="cuda")
    sol = torch.empty_like(rhs)
    dxt = torch.zeros(ctx.nfaces * 3 * ctx.nf, dtype=torch.float64, device="cuda")
    dxc = torch.empty_like(s.x_cell)
    ph = {"condense": event_ms(torch, stream, lambda: ctx.condense(s.x_cell, s.x_trace, thermodyn,
                                                                  vals, rhs, dirv), reps),
          "factor": event_ms(torch, stream, lambda: ts.factor(vals), reps),
          "solve": event_ms(torch, stream, lambda: ts.solve(rhs, sol), reps)}
    dxt[r2f] = sol
    ph["reconstruct"] = event_ms(torch, stream, lambda: ctx.reconstruct(
        s.x_cell, s.x_trace, dxt, thermodyn, dxc), reps)
    out = {"ms": sum(ph.values()), "phase_ms": ph, "trace_dofs": ctx.ntrace,
           "max_front": ts.max_front, "front_GB": ts.bytes / 1e9, "factor_GFLOP": ts.flops / 1e9,
           "factor_TFLOPs": ts.flops / (ph["factor"] * 1e-3) / 1e12}
    s._solver_cache.clear()
    return out


def sample_3d_k2(args, torch, stream):
    """3D at degree 2 on the generic kernels: a cube diode of n^3 hexahedra
    (324 cell dofs, 162 trace dofs per cell) with a step junction; pass 1,
    pass 2 and the trace factorisation and solve, FP64"""
    from ddhdg_amd import configs
    from ddhdg_amd.np_solver import NPSolver

    n = args.cells3d_k2
    prob, params, adim = configs.diode(3, n, 2, doping=configs.step_doping(1e22, 1e22))
    s = NPSolver(prob, params, adim, device="cuda")
    s.compute_local_charge_neutrality()
    s.project_densities_from_potential()
    ctx = s.ctx
    vals = torch.empty(ctx.nnz, dtype=torch.float64, device="cuda")
    rhs = torch.empty(ctx.ntrace, dtype=torch.float64, device="cuda")
    dirv = torch.zeros(ctx.ntrace, dtype=torch.float64, device="cuda")
    dxt = torch.zeros_like(s.x_trace)
    dxc = torch.empty_like(s.x_cell)
    ts = s.trace_solver()
    sol = torch.empty(ctx.ntrace, dtype=torch.float64, device="cuda")
    reps = 3

    def timed(fn):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    p1 = timed(lambda: ctx.condense(s.x_cell, s.x_trace, False, vals, rhs, dirv))
    p2 = timed(lambda: ctx.reconstruct(s.x_cell, s.x_trace, dxt, False, dxc))
    fm = timed(lambda: ts.factor(vals))
    sm = timed(lambda: ts.solve(rhs, sol))
    # the same passes with the generic kernel's FP32 elimination and FP64
    # refinement (precision 2), and its error against FP64 on a sample of
    # cells
    sample = torch.arange(0, ctx.ncells, max(1, ctx.ncells // 256), dtype=torch.int32,
                          device="cuda")
    T = ctx.T
    blocks = {}
    for prec in (0, 2):
        ctx.set_precision(prec)
        S_ = torch.empty(sample.numel() * T * T, dtype=torch.float64, device="cuda")
        g_ = torch.empty(sample.numel() * T, dtype=torch.float64, device="cuda")
        ctx.cell_systems(s.x_cell, s.x_trace, False, sample, S_, g_)
        blocks[prec] = (S_.view(-1, T * T), g_.view(-1, T))
    p1m = timed(lambda: ctx.condense(s.x_cell, s.x_trace, False, vals, rhs, dirv))
    p2m = timed(lambda: ctx.reconstruct(s.x_cell, s.x_trace, dxt, False, dxc))
    ctx.set_precision(0)
    (Sd, gd), (Sm, gm) = blocks[0], blocks[2]
    err_S = float(((Sm - Sd).norm(dim=1) / Sd.norm(dim=1)).max())
    return {"workload": f"3D diode {n}^3 hexahedra, degree 2, V/n/p (324 cell dofs, 162 trace "
                        "dofs), step junction N_D = N_A = 1e22 m^-3; generic kernels, FP64",
            "cells": int(ctx.ncells),
            "pass1": {"value": ctx.ncells / (p1 * 1e-3), "unit": "cells/s", "ms_per_pass": p1},
            "pass2_ms": p2, "factor_ms": fm, "solve_ms": sm,
            "fp32_refine": {"precision": 2, "pass1_ms": p1m, "pass2_ms": p2m,
                            "vs_fp64": p1 / p1m,
                            "max_rel_err_S_vs_fp64": err_S,
                            "error_sample_cells": int(sample.numel())},
            "factor_GFLOP": ts.flops / 1e9, "factor_TFLOPs": ts.flops / (fm * 1e-3) / 1e12,
            "max_front": int(ts.max_front)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=int, default=256)
    ap.add_argument("--degree", type=int, default=2)
    ap.add_argument("--bias", type=float, default=0.3)
    ap.add_argument("--cpu-sample", type=int, default=192,
                    help="cells per axis of the CPU baseline sample (about 10 s of CPU work)")
    ap.add_argument("--cpu-sample-1core", type=int, default=32,
                    help="cells per axis of the one-core CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--newton-steps", type=int, default=3,
                    help="whole Newton steps timed beside the metric")
    ap.add_argument("--cells3d", type=int, default=48,
                    help="cells per axis of the 3D sample (0: none)")
    ap.add_argument("--cells3d-k2", type=int, default=16,
                    help="cells per axis of the 3D degree-2 sample (0: none)")
    ap.add_argument("--batch", type=int, default=64,
                    help="devices condensed together beside the metric (0: none)")
    ap.add_argument("--no-batch-newton", dest="batch_newton", action="store_false",
                    help="skip the batched Newton workflow of the batched devices")
    ap.add_argument("--serve", type=int, default=16,
                    help="requests of the served stream served beside the metric (0: none)")
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-budget-gib", type=int, default=200)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device-to-device copy of this many MiB "
                         "(3 launches): the known-size kernel that calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the per-cell path runs on one GPU only (a mesh "
              "would be split into slabs of cells, one per GPU, with the trace rows of the "
              "cut faces exchanged); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8,
                          "launches": 3}), flush=True)
        return

    from ddhdg_amd import configs
    from ddhdg_amd.np_solver import NPSolver

    prob, params, adim = configs.diode(2, args.cells, args.degree, bias=args.bias)
    s = NPSolver(prob, params, adim, device="cuda")
    s.compute_local_charge_neutrality()
    s.project_densities_from_potential()
    ctx = s.ctx
    N, T = ctx.NC, ctx.T
    vals = torch.empty(ctx.nnz, dtype=torch.float64, device="cuda")
    rhs = torch.empty(ctx.ntrace, dtype=torch.float64, device="cuda")
    dirv = torch.zeros(ctx.ntrace, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()

    def step():
        ctx.condense(s.x_cell, s.x_trace, False, vals, rhs, dirv)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / args.steps       # two color launches per step
    ms_per_step = wall * 1e3 / args.steps
    cells = ctx.ncells
    progress(f"FP64 pass 1: {ms_per_step:.3f} ms per step")
    value = cells * args.steps / wall
    dxt = torch.zeros_like(s.x_trace)
    dxc = torch.empty_like(s.x_cell)

    # the same pass 1 with the local elimination in FP32 + one FP64
    # refinement step; its error against FP64 on a sample of cells
    ctx.set_precision(1)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0m, e1m = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    tm = time.perf_counter()
    e0m.record(stream)
    for _ in range(args.steps):
        step()
    e1m.record(stream)
    torch.cuda.synchronize()
    wall_m = time.perf_counter() - tm
    sample = torch.arange(0, ctx.ncells, max(1, ctx.ncells // 4096), dtype=torch.int32,
                          device="cuda")
    Sm = torch.empty(sample.numel() * T * T, dtype=torch.float64, device="cuda")
    gm = torch.empty(sample.numel() * T, dtype=torch.float64, device="cuda")
    ctx.cell_systems(s.x_cell, s.x_trace, False, sample, Sm, gm)
    ctx.set_precision(0)
    Sd, gd = torch.empty_like(Sm), torch.empty_like(gm)
    ctx.cell_systems(s.x_cell, s.x_trace, False, sample, Sd, gd)
    torch.cuda.synchronize()
    Sm, Sd = Sm.view(-1, T * T), Sd.view(-1, T * T)
    gm, gd = gm.view(-1, T), gd.view(-1, T)
    err_S = float(((Sm - Sd).norm(dim=1) / Sd.norm(dim=1)).max())
    err_g = float(((gm - gd).norm(dim=1) / gd.norm(dim=1).clamp_min(1e-300)).max())
    m_ms = e0m.elapsed_time(e1m) / args.steps / 2.0
    progress(f"FP32+refine pass 1: {wall_m * 1e3 / args.steps:.3f} ms per step")
    # the dense count split by precision: LU and solves in FP32; the
    # refinement residual and the Schur product in FP64
    f32 = 2 * N ** 3 / 3 + 2 * N ** 2 * (T + 1)
    f64 = 2 * N ** 2 * (T + 1) + 2 * T * N * (T + 1)
    per_s = (cells / 2) / (m_ms * 1e-3) / 1e12
    fp32_refine = {"value": cells * args.steps / wall_m, "unit": "cells/s",
                   "kernel_ms_per_launch": m_ms,
                   "roofline": {"fp32_flop_per_cell": f32, "fp64_flop_per_cell": f64,
                                "achieved_fp32_TFLOPs": f32 * per_s, "peak_fp32_TFLOPs": 157.3,
                                "achieved_fp64_TFLOPs": f64 * per_s,
                                "peak_fp64_TFLOPs": FP64_PEAK_TFLOPS},
                   "vs_fp64": (cells * args.steps / wall_m) / value,
                   "max_rel_err_S_vs_fp64": err_S, "max_rel_err_g_vs_fp64": err_g,
                   "error_sample_cells": int(sample.numel())}

    # a whole Newton step on the device: pass 1, the trace factorisation and
    # solve (nested-dissection multifrontal LU), pass 2
    ts = s.trace_solver() if args.newton_steps > 0 else None
    sol = torch.empty(ctx.ntrace, dtype=torch.float64, device="cuda")
    _, _, _, r2f = s._pattern()
    phases = {"condense": [], "factor": [], "solve": [], "reconstruct": []}

    def timed(name, fn):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        phases[name].append((e0, e1))

    def newton_step():
        timed("condense", step)
        timed("factor", lambda: ts.factor(vals))
        timed("solve", lambda: ts.solve(rhs, sol))
        dxt.zero_()
        dxt[r2f] = sol
        timed("reconstruct", lambda: s.ctx.reconstruct(s.x_cell, s.x_trace, dxt, False, dxc))

    n_newton = min(args.steps, args.newton_steps)
    newton_ms, phase_ms = None, None
    if n_newton > 0:
        newton_step()
        torch.cuda.synchronize()
        for v in phases.values():
            v.clear()
        w0 = time.perf_counter()
        for _ in range(n_newton):
            newton_step()
        torch.cuda.synchronize()
        newton_ms = (time.perf_counter() - w0) * 1e3 / n_newton
        phase_ms = {k: sum(a.elapsed_time(b) for a, b in v) / len(v) for k, v in phases.items()}
        progress(f"Newton step: {newton_ms:.2f} ms")

    graph_step_ms = None
    if n_newton > 0:
        # the same Newton iteration replayed from a captured hipGraph
        from ddhdg_amd.np_solver import NewtonIteration
        x0 = (s.x_cell.clone(), s.x_trace.clone())
        it = NewtonIteration(s, False, 1e-10, 1e-10, 100, graph=True)
        for _ in range(2):                      # eager, then capture + replay
            it.launch()
        it.sync()
        gw = time.perf_counter()
        for _ in range(n_newton):
            it.launch()
        it.sync()
        graph_step_ms = (time.perf_counter() - gw) * 1e3 / n_newton
        s.x_cell.copy_(x0[0])
        s.x_trace.copy_(x0[1])
        del it

    launch_ms = kernel_ms / 2.0
    cells_per_launch = cells / 2.0
    achieved_tflops = flops_per_cell(N, T) * cells_per_launch / (launch_ms * 1e-3) / 1e12
    achieved_gbs = bytes_per_cell(N, T) * cells_per_launch / (launch_ms * 1e-3) / 1e9
    line = {
        "metric": "cells assembled+condensed/s (FP64)",
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2D p-n diode {args.cells}x{args.cells} quads, HDG degree "
                               f"{args.degree}, V/n/p, FP64, pass 1 (assembly+condensation+scatter)",
                   "cells": cells, "cell_dofs": N, "trace_dofs_per_cell": T,
                   "global_trace_dofs": ctx.ntrace, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": achieved_tflops, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tflops / FP64_PEAK_TFLOPS,
                     "traffic": TRAFFIC_PER_LAUNCH.get((args.cells, args.degree))},
        "hbm": {"achieved_GBs": achieved_gbs, "peak_GBs": HBM_PEAK_GBS,
                "frac": achieved_gbs / HBM_PEAK_GBS, "bytes_per_cell": bytes_per_cell(N, T)},
        "kernel_ms_per_launch": launch_ms,
        "fp32_refine": fp32_refine,
    }
    if n_newton > 0:
        line["newton_step"] = {
            "ms": newton_ms, "steps": n_newton, "phase_ms": phase_ms,
            "graph_replay_ms": graph_step_ms,
            "trace_solver": {"levels": ts.levels, "max_front": ts.max_front,
                             "front_GB": ts.bytes / 1e9, "factor_GFLOP": ts.flops / 1e9,
                             "factor_TFLOPs": ts.flops / (phase_ms["factor"] * 1e-3) / 1e12}}
    if args.cells3d > 0:
        line["sample_3d"] = sample_3d(args, torch, stream)
        progress("3D degree-1 sample")
    if args.cells3d_k2 > 0:
        line["sample_3d_k2"] = sample_3d_k2(args, torch, stream)
        progress("3D degree-2 sample")
        torch.cuda.empty_cache()
    if args.batch > 0:
        line["batched"] = batched_sample(args, torch, stream)
        progress("batched devices")
    if args.serve > 0:
        # the served workload: the first requests of the seeded stream
        # (Poisson arrivals, lognormal 2D meshes 32^2-256^2, 20 % 3D cubes
        # 16^3-40^3, degree 1-2, 13-point sweeps), continuous batching, state
        # kept on the device across the bias points, admission by HBM budget
        # (requests larger than the whole budget are refused)
        from ddhdg_amd.serving import IVServer, memory_estimate, served_requests, summarize
        budget = int(args.serve_budget_gib) << 30
        reqs = served_requests(args.serve, seed=20261204, rate=args.serve_rate)
        runs = {}
        for prec in (0, 1):
            srv = IVServer(hbm_budget_bytes=budget, max_slots=args.serve_slots, graph=True,
                           precision=prec)
            res, wall = srv.run(reqs)
            runs[prec] = (srv, res, wall)
            progress(f"served stream, precision {prec}: {wall:.1f} s")
        keys = ("solves_per_s", "p50_latency_s", "p99_latency_s", "converged", "requests",
                "wall_s", "newton_iterations")

        def stats(srv, res, wall):
            admitted = [r for r in res if not (r.error and "budget" in r.error)]
            out = {k: v for k, v in summarize(admitted, wall).items() if k in keys}
            out["refused_over_budget"] = len(res) - len(admitted)
            out["max_concurrent"] = srv.max_active
            return out
        s64, r64, w64 = runs[0]
        s32, r32, w32 = runs[1]
        # FP32+refine against FP64: the largest difference of a sweep's
        # terminal currents relative to the sweep's largest current
        dI = max((max(abs(a - b) for a, b in zip(rm.currents, rf.currents)) /
                  max(max(abs(b) for b in rf.currents), 1e-300))
                 for rf, rm in zip(r64, r32) if rf.converged and rm.converged)
        same_its = all(rf.iterations == rm.iterations for rf, rm in zip(r64, r32))
        line["serving"] = {
            **stats(s64, r64, w64), "precision": "fp64", "graph_newton": True,
            "fp32_refine": {**stats(s32, r32, w32),
                            "max_current_diff_vs_fp64_rel_to_sweep_max": dI,
                            "same_iteration_counts_as_fp64": same_its},
            "requests_detail": [{"dim": q.dim, "side": q.n_cells[0], "degree": q.degree,
                                 "memory_GiB": round(memory_estimate(q) / 2**30, 2),
                                 "converged": r.converged, "latency_s": round(r.latency, 3),
                                 "error": r.error}
                                for q, r in zip(reqs, sorted(r64, key=lambda r: r.rid))],
            "workload": f"the first {args.serve} requests of the served stream (seed 20261204): "
                        f"Poisson arrivals {args.serve_rate}/s, 2D sides lognormal in 32..256, "
                        "20 % 3D cubes with sides lognormal in 16..40, degree 1-2, doping "
                        "log-uniform 1e21-1e24, 13-point sweeps 0-0.6 V; HBM budget "
                        f"{args.serve_budget_gib} GiB, {args.serve_slots} slots",
            "unit": "I-V solves/s"}
    if not args.no_cpu_baseline:
        host = host_cpu()
        build = native_oracle()
        # all the cores this process may use (the box grants 16 per GPU), and one
        threads = max(1, min(host["usable"], 16))
        v, nc, t = cpu_baseline(args.cpu_sample, args.degree, threads, args.bias)
        progress(f"CPU baseline, {threads} threads: {t:.2f} s per run")
        v1, nc1, t1 = cpu_baseline(args.cpu_sample_1core, args.degree, 1, args.bias)
        progress(f"CPU baseline, one core: {t1:.2f} s per run")
        line["cpu_baseline"] = {
            "value": v, "unit": "cells/s", "cores": threads, "kind": "port",
            "sample": f"oracle pass 1 (full-inverse Gauss-Jordan + mmult, WorkStream threads, "
                      f"in-order scatter) over a {args.cpu_sample}x{args.cpu_sample} diode of "
                      f"the same per-cell shape ({nc} cells), median of 5 runs after 1 warm-up, "
                      f"{t:.2f} s each",
            "one_core": {"value": v1, "unit": "cells/s", "cores": 1,
                         "sample": f"{args.cpu_sample_1core}x{args.cpu_sample_1core} diode "
                                   f"({nc1} cells), median of 5 runs, {t1:.2f} s each"},
            "host": host, "build": build}
        line["speedup_vs_cpu"] = value / v
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
