"""Timing of the GPU trace solver on a diode mesh.

    python -m ddhdg_amd.trace_bench --dim 2 --cells 256 --degree 2 --reps 3

Builds the condensed trace system of the charge-neutral first guess (pass 1
on the device), then times the nested-dissection factorisation and the
solve with HIP events on the launch stream, after two untimed rounds.
Prints one JSON line: the times, the plan's size (groups, largest front,
front storage) and the factorisation's FLOP count and rate."""
from __future__ import annotations

import argparse
import json


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--dim", type=int, default=2)
    ap.add_argument("--cells", type=int, default=256, help="cells per axis")
    ap.add_argument("--degree", type=int, default=2)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args(argv)

    import torch
    from . import configs
    from .np_solver import NPSolver

    prob, params, adim = configs.diode(args.dim, args.cells, args.degree)
    s = NPSolver(prob, params, adim, device="cuda")
    s.compute_local_charge_neutrality()
    s.project_densities_from_potential()
    ctx = s.ctx
    vals = torch.empty(ctx.nnz, dtype=torch.float64, device="cuda")
    rhs = torch.empty(ctx.ntrace, dtype=torch.float64, device="cuda")
    dirv = torch.zeros(ctx.ntrace, dtype=torch.float64, device="cuda")
    ctx.condense(s.x_cell, s.x_trace, False, vals, rhs, dirv)
    import time
    t0 = time.perf_counter()
    ts = s.trace_solver()
    create_s = time.perf_counter() - t0
    sol = torch.empty(ctx.ntrace, dtype=torch.float64, device="cuda")
    for _ in range(2):
        ts.factor(vals)
        ts.solve(rhs, sol)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(args.reps):
            fn()
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.reps

    f_ms = timed(lambda: ts.factor(vals))
    s_ms = timed(lambda: ts.solve(rhs, sol))
    print(json.dumps({"dim": args.dim, "cells_per_axis": args.cells, "degree": args.degree,
                      "trace_dofs": int(ctx.ntrace), "factor_ms": f_ms, "solve_ms": s_ms,
                      "create_s": create_s,
                      "factor_GFLOP": ts.flops / 1e9,
                      "factor_TFLOPs": ts.flops / (f_ms * 1e-3) / 1e12,
                      "groups": int(ts.levels), "max_front": int(ts.max_front),
                      "front_GB": ts.bytes / 1e9,
                      "reps": args.reps}), flush=True)


if __name__ == "__main__":
    main()
