"""Newton solves of many devices advanced together on one GPU.

The reference solves one device at a time (NPSolver::private_run,
src/np_solver.cpp:1452-1561).  BatchedNewton runs the same iteration for a
list of solvers at once:

  * pass 1 of every batchable device in persistent launches (hdg_batch: the
    2D devices of degree 1 and 2 with V/n/p or V alone in one launch per
    colour, the degree-3 devices with V/n/p in another); devices no batch
    kernel covers (3D, other component sets, operator variants) are
    condensed one by one in the same sequence;
  * every device's trace factorisation and solve, Dirichlet overwrite,
    pass 2, norms and update, the devices spread over a few streams forked
    from the iteration's stream (parallel branches of the graph: small
    devices overlap large ones);
  * after an eager first iteration the whole sequence is captured into one
    hipGraph and replayed, so an iteration of all devices costs one graph
    launch and one host synchronisation (the norms of every device arrive
    in one pinned array).

Each device keeps its own L-infinity criterion (src/np_solver.cpp:1521-1560).
A device that has converged (or failed) leaves the iteration: the
per-device sequence runs for the devices still iterating only, the next
iteration runs eagerly and the one after it is captured again (pass 1
batches keep their finished devices, whose outputs go to their own buffers
and whose state no longer changes).  A device's state, iteration count and
last update norm are therefore those of the same solve run alone (pass 1
of a batch is bitwise equal to each device's own, the rest is the same
per-device sequence).
"""
from __future__ import annotations

import numpy as np

from .hip_lib import HdgBatch, max_abs4
from .np_solver import NonlinearIterationResults
from .problem import CellFaceTauNPSolverParameters


def _batch_groups(solvers, thermodyn):
    """indices of the devices pass 1 can run batched, as lists (one per
    persistent launch), and the rest"""
    low, deg3, alone = [], [], []
    for i, s in enumerate(solvers):
        en = tuple(bool(e) for e in s.enabled)
        full = en == (True, True, True)
        vonly = en == (True, False, False)
        default = (s.parameters.dd_flux_type == "use_cell" and not s.parameters.phi_linearize
                   and not isinstance(s.parameters, CellFaceTauNPSolverParameters))
        if s.mesh.dim != 2 or not default:
            alone.append(i)
        elif s.degree <= 2 and ((thermodyn and vonly) or (not thermodyn and full)):
            low.append(i)
        elif s.degree == 3 and not thermodyn and full:
            deg3.append(i)
        else:
            alone.append(i)
    groups = [g for g in (low, deg3) if g]
    # one precision per batch
    out = []
    for g in groups:
        by_prec = {}
        for i in g:
            by_prec.setdefault(solvers[i].ctx.precision, []).append(i)
        out.extend(by_prec.values())
    return out, alone


class BatchedNewton:
    """Newton solves of `solvers` (NPSolver objects on one device, each with
    its own state, enabled components and boundary data) advanced together;
    `thermodyn` as NPSolver._private_run.  run() returns one
    NonlinearIterationResults per solver."""

    def __init__(self, solvers, thermodyn=False, abs_tol=None, rel_tol=None, max_it=None,
                 graph=True, streams=4):
        if not solvers:
            raise ValueError("BatchedNewton needs at least one solver")
        self.s = list(solvers)
        t = self.t = self.s[0].torch
        dev = self.s[0].device
        for s in self.s:
            if s.parameters.trace_solver != "gpu" or s.parameters.iterative_linear_solver:
                raise ValueError("BatchedNewton needs the device trace solver of every device")
            if thermodyn and tuple(s.enabled) != (True, False, False):
                raise ValueError("thermodynamic equilibrium runs with V only")
        nl = [s.parameters.nonlinear_parameters for s in self.s]
        self.abs_tol = [n.absolute_tolerance if abs_tol is None else abs_tol for n in nl]
        self.rel_tol = [n.relative_tolerance if rel_tol is None else rel_tol for n in nl]
        mi = [n.max_number_of_iterations if max_it is None else max_it for n in nl]
        self.max_it = [m if m >= 0 else 1 << 30 for m in mi]
        self.thermodyn = thermodyn
        self.use_graph = graph
        self.graph = None
        cur = t.cuda.current_stream()
        self.stream = t.cuda.Stream() if graph and cur == t.cuda.default_stream() else cur
        n = len(self.s)
        self.bufs = []
        for s in self.s:
            ctx = s.ctx
            # the factors of one component set per device at a time
            s.release_trace_solvers()
            s._sync_tau()
            _, _, dirs, r2f = s._pattern()
            self.bufs.append({
                "vals": t.empty(ctx.nnz, dtype=t.float64, device=dev),
                "rhs": t.empty(ctx.ntrace, dtype=t.float64, device=dev),
                "sol": t.empty(ctx.ntrace, dtype=t.float64, device=dev),
                "dirv": t.zeros(ctx.ntrace, dtype=t.float64, device=dev),
                "upd_trace": t.zeros_like(s.x_trace),
                "upd_cell": t.zeros_like(s.x_cell),
                "dirs": t.as_tensor(dirs, device=dev),
                "r2f": r2f,
                "nscratch": t.empty(1024, dtype=t.float64, device=dev),
            })
        self.dnorms = t.zeros(4 * n, dtype=t.float64, device=dev)
        self.norms = t.zeros(4 * n, dtype=t.float64).pin_memory()
        self.alpha = [s.parameters.nonlinear_parameters.alpha for s in self.s]
        self.step = 0
        self.done = [False] * n
        self.converged = [False] * n
        self.iterations = [0] * n
        self.last_norm = [0.0] * n
        self.error = [None] * n
        # the per-device part of an iteration runs on `streams` streams
        # forked from the iteration's stream (parallel branches of the
        # graph), so that small devices overlap large ones
        self.fork = [t.cuda.Stream() for _ in range(max(1, streams))]
        groups, self.alone = _batch_groups(self.s, thermodyn)
        self.batches = [(HdgBatch([self.s[i].ctx for i in g]), g) for g in groups]
        self.captures = 0
        self.active = list(range(n))
        self.graph = None
        self.eager_next = True

    def _regroup(self, active):
        """the devices still iterating: the next iteration runs eagerly, the
        one after it is captured again.  Pass 1 batches keep their finished
        devices (their outputs go to their own buffers; their state no
        longer changes), so no batch is rebuilt or retuned."""
        self.active = active
        self.graph = None
        self.eager_next = True

    def close(self):
        for b, _ in self.batches:
            b.close()
        self.batches = []

    def _iteration(self):
        th = self.thermodyn
        for b, g in self.batches:
            b.condense(th, [self.s[i].x_cell for i in g], [self.s[i].x_trace for i in g],
                       [self.bufs[i]["vals"] for i in g], [self.bufs[i]["rhs"] for i in g],
                       [self.bufs[i]["dirv"] for i in g])
        t = self.t
        main = t.cuda.current_stream()
        for st in self.fork:
            st.wait_stream(main)
        alone = set(self.alone)
        for k, i in enumerate(self.active):
            with t.cuda.stream(self.fork[k % len(self.fork)]):
                self._device_step(i, i in alone)
        for st in self.fork:
            main.wait_stream(st)
        self.norms.copy_(self.dnorms, non_blocking=True)

    def _device_step(self, i, condense):
        th = self.thermodyn
        s, b = self.s[i], self.bufs[i]
        if condense:
            s.ctx.condense(s.x_cell, s.x_trace, th, b["vals"], b["rhs"], b["dirv"])
        s._solve_trace(b["vals"], b["rhs"], b["sol"])
        b["sol"][b["dirs"]] = b["dirv"][b["dirs"]]
        b["upd_trace"].zero_()
        b["upd_trace"][b["r2f"]] = b["sol"]
        s.ctx.reconstruct(s.x_cell, s.x_trace, b["upd_trace"], th, b["upd_cell"])
        s.update_cell = b["upd_cell"]
        max_abs4([b["upd_cell"], b["upd_trace"], s.x_cell, s.x_trace], b["nscratch"],
                 self.dnorms[4 * i:4 * i + 4])
        s.x_cell.add_(b["upd_cell"], alpha=self.alpha[i])
        s.x_trace.add_(b["upd_trace"], alpha=self.alpha[i])

    def launch(self):
        t = self.t
        if self.stream != t.cuda.current_stream():
            self.stream.wait_stream(t.cuda.current_stream())
        with t.cuda.stream(self.stream):
            if not self.use_graph or self.eager_next:
                self._iteration()
                self.eager_next = False
            else:
                if self.graph is None:
                    self.graph = t.cuda.CUDAGraph()
                    with t.cuda.graph(self.graph, stream=self.stream):
                        self._iteration()
                    self.captures += 1
                self.graph.replay()
        self.step += 1

    def finish(self):
        """the criterion of every device still iterating; True when all are
        done"""
        nrm = self.norms.numpy().reshape(-1, 4)
        stop = []
        for i in self.active:
            upd = max(float(nrm[i, 0]), float(nrm[i, 1]))
            cur = max(float(nrm[i, 2]), float(nrm[i, 3]))
            self.iterations[i] = self.step
            self.last_norm[i] = float(nrm[i, 0])
            if upd < self.abs_tol[i] or upd < self.rel_tol[i] * cur:
                self.converged[i] = True
            elif np.isfinite(upd) and self.step < self.max_it[i]:
                continue
            elif not np.isfinite(upd):
                self.error[i] = "the Newton update is not finite"
            self.done[i] = True
            stop.append(i)
        if all(self.done):
            return True
        if stop:
            self._regroup([i for i in self.active if not self.done[i]])
        return False

    def run(self):
        while True:
            self.launch()
            self.stream.synchronize()
            if self.finish():
                break
        self.stream.synchronize()
        return [NonlinearIterationResults(c, it, nrm)
                for c, it, nrm in zip(self.converged, self.iterations, self.last_norm)]


def batched_run(solvers, graph=True, **kw):
    """NPSolver.run of every solver (its current components), batched"""
    bn = BatchedNewton(solvers, False, graph=graph, **kw)
    try:
        return bn.run()
    finally:
        bn.close()


def batched_equilibrium(solvers, graph=True, **kw):
    """NPSolver.compute_thermodynamic_equilibrium of every solver, batched
    (src/thermodynamic_equilibrium.cpp:365-405): V alone from the local
    charge-neutrality guess, then the densities from the potential; each
    solver's enabled components are restored"""
    saved = [tuple(s.enabled) for s in solvers]
    for s in solvers:
        s.set_enabled_components(True, False, False)
        s.compute_local_charge_neutrality()
    bn = BatchedNewton(solvers, True, graph=graph, **kw)
    try:
        res = bn.run()
    finally:
        bn.close()
    for s, en in zip(solvers, saved):
        s.project_densities_from_potential(0.0, 0.0)
        s.set_enabled_components(*en)
    return res
