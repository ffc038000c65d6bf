"""Continuous-batching server for I-V sweep requests on one GPU.

A request is a diode I-V sweep: a mesh, a degree, a doping profile and a list
of bias points.  It runs the reference's diode workflow (thermodynamic
equilibrium, densities from the potential, n and p with V frozen, the coupled
system; examples/residual.ipynb cells 4-17) and then one coupled Newton solve
per bias point, the Newton state kept on the device from one point to the
next; each converged point records the terminal current at both contacts.

Every Newton iteration after a request's first is replayed from a captured
hipGraph (one launch instead of hundreds); the graph is kept across the bias
points, which change only boundary data in place.

Scheduling: requests wait in arrival order and are admitted while a slot is
free and their device-memory estimate (context tables, trace-solver fronts,
Newton buffers) fits the HBM budget, so requests arriving while others run
join the running set as soon as slots and memory free up.  Every active
request keeps one Newton iteration in flight on the stream of its slot
(default priority), with an event recorded behind it; the loop polls the
events and advances each request whose iteration has finished (its state
machine reads the update norms from pinned host memory) and launches its next
one, so a slow request (a large 3D cube) does not hold back the others.
Requests share no buffers and every kernel is deterministic, so a request's
results are the same alone and in any mix.
"""
from __future__ import annotations

import functools
import time
from dataclasses import dataclass, field

import numpy as np

from . import configs
from .np_solver import NewtonIteration, NPSolver


@dataclass
class IVRequest:
    rid: int
    dim: int
    n_cells: tuple
    degree: int
    biases: tuple
    nd: float = 1e23
    na: float = 1e23
    junction: float = 0.5
    arrival: float = 0.0          # seconds after the server starts
    max_it: int = 0               # Newton iteration limit per stage (0: the solver's)


@dataclass
class IVResult:
    rid: int
    biases: list
    currents: list                # A (per metre of depth in 2D), contact at x = 1
    iterations: list              # Newton iterations per stage
    converged: bool
    arrival: float
    admitted: float
    finished: float
    state: tuple = field(default=None, repr=False)
    error: str = None             # why a failed request failed (None when converged)

    @property
    def latency(self):
        return self.finished - self.arrival


@functools.lru_cache(maxsize=256)
def _trace_solver_bytes(dim, n_cells, tdf):
    """device bytes of the trace solver of a mesh: fronts, work vectors,
    pivots and the groups' index lists, from the plan's exact sizes (built
    on the host without its maps); the front-assembly maps are bounded by
    two ints per front index (8 bytes per work entry) and 24 bytes per
    node"""
    from .hip_lib import HdgError, trace_plan_size
    try:
        z = trace_plan_size(dim, n_cells, tdf)
    except HdgError:
        return 1 << 62          # no plan (a front beyond the solver's limits)
    return (8 * z["front_entries"] + 32 * z["work_entries"] + 4 * z["pivot_entries"] +
            24 * z["nodes"])


def memory_estimate(req: IVRequest) -> int:
    """device bytes a request holds while it runs: cell and trace state and
    updates, the trace matrix, the trace solver of the full V/n/p system (the
    largest of its stages; it holds one at a time), the generic kernel's
    scratch where it needs one, tables"""
    k, d = req.degree, req.dim
    n = tuple(list(req.n_cells) + [1] * (3 - len(req.n_cells)))
    ncells = int(np.prod(n[:d]))
    nb, nf = (k + 1) ** d, (k + 1) ** (d - 1)
    nfaces = sum(int(np.prod([n[b] + (1 if a == b else 0) for b in range(d)])) for a in range(d))
    NC = 3 * (d + 1) * nb
    tdf = 3 * nf
    nnz = nfaces * (4 * d - 1) * tdf * tdf
    # the matrix-entry map: front offset (8 bytes) and entry index (4)
    solver = _trace_solver_bytes(d, n[:d], tdf) + 12 * nnz
    # the generic kernel's global scratch where its reduced systems do not
    # fit LDS (3D degree 2): 512 slices of R
    maxr = 3 * nb + 2 * d * 3 * nf
    scratch = 512 * maxr * (maxr + 2) * 8 if (d == 3 and k >= 2) else 0
    return int(8 * (4 * ncells * NC + 6 * nfaces * 3 * nf + nnz) + solver + scratch + (64 << 20))


class _Slot:
    """one admitted request: its solver, stream and stage machine"""

    def __init__(self, req: IVRequest, t0: float, now: float, torch, graph=True, precision=0):
        self.req, self.torch, self.graph = req, torch, graph
        self.stream = torch.cuda.Stream()
        with torch.cuda.stream(self.stream):
            prob, params, adim = configs.diode(
                req.dim, req.n_cells, req.degree,
                doping=configs.step_doping(req.nd, req.na, req.junction))
            self.solver = NPSolver(prob, params, adim)
            s = self.solver
            s.ctx.set_precision(precision)
            s.set_enabled_components(True, False, False)
            s.compute_local_charge_neutrality()
        self.max_it = int(req.max_it) or None
        self.stage = 0                        # 0 eq, 1 n/p, 2 coupled, 3.. biases
        self.iterations, self.currents = [], []
        self.result = None
        self.admitted = now
        self.t0 = t0
        self.ok = True
        self.error = None
        self.done = torch.cuda.Event()
        self.inflight = False
        self._start_stage()

    def _start_stage(self):
        s, nl = self.solver, self.solver.parameters.nonlinear_parameters
        with self.torch.cuda.stream(self.stream):
            if self.stage == 1:
                s.project_densities_from_potential()
                s.set_enabled_components(False, True, True)
            elif self.stage == 2:
                s.set_enabled_components(True, True, True)
            elif self.stage >= 3:
                configs.set_bias(s, self.req.biases[self.stage - 3])
            # a request holds the trace solver of its current stage only
            s.release_trace_solvers()
            if self.stage >= 3:
                # bias points: same operator and buffers, boundary data
                # changed in place; the captured iteration is replayed
                self.it.reset()
            else:
                self.it = NewtonIteration(s, self.stage == 0, nl.absolute_tolerance,
                                          nl.relative_tolerance,
                                          self.max_it or nl.max_number_of_iterations,
                                          graph=self.graph)

    def close(self):
        """release the request's device memory: trace solvers, graph,
        context"""
        s = getattr(self, "solver", None)
        if s is None:
            return
        s.release_trace_solvers(keep_current=False)
        self.it = None
        s.ctx.close()
        self.solver = None

    def launch(self):
        with self.torch.cuda.stream(self.stream):
            self.it.launch()
            self.done.record(self.stream)
        self.inflight = True

    def poll(self):
        """True once the iteration in flight has finished on the device"""
        if self.inflight and self.done.query():
            self.inflight = False
            return True
        return False

    def advance(self, now):
        """after the iteration in flight has finished: True when the request
        is done"""
        status = self.it.finish()
        if status == "continue":
            return False
        self.iterations.append(self.it.step)
        if status == "failed":
            self.ok = False
            self.error = f"Newton did not converge in stage {self.stage}"
        elif self.stage >= 3:
            with self.torch.cuda.stream(self.stream):
                self.currents.append(self.solver.terminal_current(1))
        self.stage += 1
        if status == "failed" or self.stage >= 3 + len(self.req.biases):
            return True
        self._start_stage()
        return False


class IVServer:
    """Admits I-V requests by HBM budget into at most max_slots concurrent
    slots and runs them with continuous batching (see the module doc)."""

    def __init__(self, hbm_budget_bytes=64 << 30, max_slots=8, keep_state=False, graph=True,
                 precision=0):
        """precision: 0 FP64, 1 FP32 local elimination + FP64 refinement
        (hdg_set_precision) in every request's pass 1"""
        import torch
        self.torch = torch
        self.budget, self.max_slots = int(hbm_budget_bytes), int(max_slots)
        self.keep_state, self.graph, self.precision = keep_state, graph, int(precision)

    def run(self, requests, clock=None):
        """serve requests (arrival times in seconds from the start); returns
        the results in request order and the wall time.

        A request that raises (its set-up, a launch or a stage change) or
        whose Newton solve fails is reported failed, with the reason, and
        evicted: its device memory and slot are released and the others go
        on unchanged (they share no buffers with it).  A request whose memory
        estimate exceeds the whole budget is rejected at admission."""
        torch = self.torch
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        active, done, used = [], {}, 0
        # ticks: Newton iterations launched
        self.max_active, self.ticks = 0, 0
        t0 = time.perf_counter() if clock is None else clock()
        now = lambda: (time.perf_counter() if clock is None else clock()) - t0

        def failed(req, admitted, err):
            done[req.rid] = IVResult(req.rid, list(req.biases), [], [], False, req.arrival,
                                     admitted, now(), None, err)

        while pending or active:
            # admission: arrived, fits the budget, a slot is free
            while pending and len(active) < self.max_slots and pending[0].arrival <= now():
                need = memory_estimate(pending[0])
                if need > self.budget:
                    req = pending.pop(0)
                    failed(req, now(), f"needs {need / 2**30:.1f} GiB, budget "
                                       f"{self.budget / 2**30:.1f} GiB")
                    continue
                if used + need > self.budget:
                    break
                req = pending.pop(0)
                try:
                    slot = _Slot(req, t0, now(), torch, self.graph, self.precision)
                except Exception as e:          # noqa: BLE001 - isolate the request
                    failed(req, now(), f"{type(e).__name__}: {e}")
                    continue
                used += need
                active.append((slot, need))
            if not active:
                if pending:
                    wait = pending[0].arrival - now()
                    if wait > 0:
                        time.sleep(min(wait, 0.05))
                continue
            self.max_active = max(self.max_active, len(active))
            ready = []
            for slot, _ in active:
                try:
                    if not slot.inflight:
                        slot.launch()
                        self.ticks += 1
                    ready.append(slot.poll())
                except Exception as e:          # noqa: BLE001
                    slot.ok, slot.error = False, f"{type(e).__name__}: {e}"
                    ready.append(None)
            if not any(r is not False for r in ready):
                time.sleep(5e-5)
                continue
            t = now()
            keep = []
            for (slot, need), fin in zip(active, ready):
                finished = fin is None
                if fin:
                    try:
                        finished = slot.advance(t)
                    except Exception as e:      # noqa: BLE001
                        slot.ok, slot.error = False, f"{type(e).__name__}: {e}"
                        finished = True
                if finished:
                    r = slot.req
                    state = None
                    if self.keep_state and slot.ok:
                        state = slot.solver.get_state()
                    done[r.rid] = IVResult(r.rid, list(r.biases), slot.currents, slot.iterations,
                                           slot.ok, r.arrival, slot.admitted, t, state,
                                           slot.error)
                    slot.stream.synchronize()
                    slot.close()
                    used -= need
                else:
                    keep.append((slot, need))
            active = keep
        wall = now()
        return [done[r.rid] for r in sorted(requests, key=lambda r: r.rid)], wall


def summarize(results, wall):
    lat = np.array([r.latency for r in results])
    ok = sum(r.converged for r in results)
    return {"requests": len(results), "converged": int(ok), "wall_s": wall,
            "solves_per_s": len(results) / wall if wall > 0 else 0.0,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "newton_iterations": int(sum(sum(r.iterations) for r in results))}


def lognormal_side(rng, lo, hi):
    """a mesh side lognormal between lo and hi: log-side normal with its mean
    at the middle of [log lo, log hi] and its standard deviation a quarter of
    that interval (so 95 % of the draws fall inside), redrawn outside it"""
    mu, sd = 0.5 * (np.log(lo) + np.log(hi)), 0.25 * (np.log(hi) - np.log(lo))
    while True:
        side = int(round(np.exp(rng.normal(mu, sd))))
        if lo <= side <= hi:
            return side


def served_requests(n=2000, seed=20261204, rate=4.0, frac3d=0.2, sides2d=(32, 256),
                    sides3d=(16, 40), degrees=(1, 2), points=13, vmax=0.6):
    """the served workload: n I-V sweep requests with Poisson arrivals at
    `rate` per second; 2D meshes with sides lognormal in sides2d (32^2 to
    256^2 cells) or, for a fraction frac3d, 3D cubes with sides lognormal in
    sides3d (16^3 to 40^3); degree uniform in `degrees`; donor and acceptor
    densities log-uniform in [1e21, 1e24] m^-3 and a junction uniform in
    [0.3, 0.7] um; each a `points`-point continuation over [0, vmax] V (the
    coupled solve at 0 V, then points - 1 bias points)"""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    biases = tuple(float(b) for b in np.linspace(0.0, vmax, points)[1:])
    for i in range(n):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        dim = 3 if rng.uniform() < frac3d else 2
        side = lognormal_side(rng, *(sides3d if dim == 3 else sides2d))
        degree = int(rng.choice(degrees))
        out.append(IVRequest(
            rid=i, dim=dim, n_cells=(side,) * dim, degree=degree, biases=biases,
            nd=float(10 ** rng.uniform(21, 24)), na=float(10 ** rng.uniform(21, 24)),
            junction=float(rng.uniform(0.3, 0.7)), arrival=t))
    return out

