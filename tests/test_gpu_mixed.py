"""FP32 local elimination with FP64 refinement (GPU tests): the per-cell
condensed blocks and pass 2 against the FP64 oracle, with the error stated
per case, and Newton on the diode converging to the FP64 solution.
Precision 1 runs FP32 in the one-wave register kernels only; precision 2 in
every kernel (the per-cell tests use it, so that the degree-3 pair kernel
and the generic kernel are covered too)."""
import numpy as np
import pytest

from ddhdg_amd import configs
from tests import helpers as H

pytestmark = pytest.mark.gpu

# normwise relative bound on S and g per cell after one refinement step: the
# FP32 elimination leaves about kappa * 2^-24 per column, one step squares it
MIXED_BAR = 1e-9


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU test on a machine without a GPU"
    return torch


# (2, 3, 3): the pair kernel (V/n/p only: its equilibrium runs the generic
# kernel in FP64)
MIXED_CASES = [(2, 4, 1), (2, 4, 2), (2, 3, 3), (3, 3, 1)]


@pytest.mark.parametrize("dim,n_cells,k", MIXED_CASES)
@pytest.mark.parametrize("thermodyn", [False, True])
def test_mixed_cell_systems_error(dim, n_cells, k, thermodyn):
    t = _torch()
    print(f"case mixed cell systems dim={dim} n={n_cells} k={k} thermo={thermodyn}", flush=True)
    from ddhdg_amd.np_solver import NPSolver
    prob, params, adim = configs.diode(dim, n_cells, k)
    s = NPSolver(prob, params, adim)
    enabled = (True, False, False) if thermodyn else (True, True, True)
    s.set_enabled_components(*enabled)
    s.ctx.set_precision(2)
    o = H.oracle_for(prob, params, adim, enabled)
    cell, tr = H.perturbed(*H.equilibrium_guess(prob, adim, k), seed=31 + k)
    s.set_state(cell, tr)
    o.set_state(cell, tr)
    ctx = s.ctx
    cells = t.arange(ctx.ncells, dtype=t.int32, device="cuda")
    S = t.empty(ctx.ncells * ctx.T * ctx.T, dtype=t.float64, device="cuda")
    g = t.empty(ctx.ncells * ctx.T, dtype=t.float64, device="cuda")
    ctx.cell_systems(s.x_cell, s.x_trace, thermodyn, cells, S, g)
    t.cuda.synchronize()
    S = S.cpu().numpy().reshape(ctx.ncells, ctx.T, ctx.T)
    g = g.cpu().numpy().reshape(ctx.ncells, ctx.T)
    eS = eg = 0.0
    for c in range(ctx.ncells):
        So, go, *_ = o.condensed_cell(c, thermodyn)
        eS, eg = max(eS, H.rel_fro(S[c], So)), max(eg, H.rel_fro(g[c], go))
    print(f"  FP32+refine vs FP64 oracle: worst rel S {eS:.2e} g {eg:.2e}")
    assert eS < MIXED_BAR and eg < MIXED_BAR


@pytest.mark.parametrize("dim,n_cells,k", MIXED_CASES)
@pytest.mark.parametrize("thermodyn", [False, True])
def test_mixed_reconstruct_error(dim, n_cells, k, thermodyn):
    """pass 2 with the FP32 elimination and one FP64 refinement step against
    the FP64 oracle's pass 2 (error stated per case)"""
    t = _torch()
    print(f"case mixed reconstruct dim={dim} n={n_cells} k={k} thermo={thermodyn}", flush=True)
    from ddhdg_amd.np_solver import NPSolver
    prob, params, adim = configs.diode(dim, n_cells, k)
    s = NPSolver(prob, params, adim)
    enabled = (True, False, False) if thermodyn else (True, True, True)
    s.set_enabled_components(*enabled)
    s.ctx.set_precision(2)
    o = H.oracle_for(prob, params, adim, enabled)
    cell, tr = H.perturbed(*H.equilibrium_guess(prob, adim, k), seed=41 + k)
    s.set_state(cell, tr)
    o.set_state(cell, tr)
    rng = np.random.default_rng(6)
    upd = 1e-2 * rng.standard_normal(s.x_trace.numel())
    if thermodyn:
        upd.reshape(-1, 3, s.ctx.nf)[:, 1:] = 0
    dxt = t.as_tensor(upd, device="cuda")
    dxc = t.empty_like(s.x_cell)
    s.ctx.reconstruct(s.x_cell, s.x_trace, dxt, thermodyn, dxc)
    t.cuda.synchronize()
    err = H.rel_fro(dxc.cpu().numpy(), o.pass2(upd, thermodyn))
    print(f"  FP32+refine pass 2 vs FP64 oracle: rel err {err:.2e}")
    assert err < MIXED_BAR


@pytest.mark.parametrize("n_cells,k,prec32", [(8, 2, 1), (6, 3, 2)])
def test_mixed_newton_reaches_fp64_solution(n_cells, k, prec32):
    """the diode workflow (equilibrium, n/p, coupled, 0.1 V, 0.3 V) with the mixed
    elimination (both passes) converges to the FP64 run's solution (relative
    L-infinity 1e-10)"""
    _torch()
    print(f"case mixed Newton diode 2D n={n_cells} k={k}", flush=True)
    from ddhdg_amd.np_solver import NPSolver
    states, its = [], []
    for prec in (0, prec32):
        prob, params, adim = configs.diode(2, n_cells, k)
        s = NPSolver(prob, params, adim)
        s.ctx.set_precision(prec)
        s.set_enabled_components(True, False, False)
        s.compute_local_charge_neutrality()
        res = configs.solve_equilibrium_and_bias(s, (0.1, 0.3))
        print(f"  precision {prec}: {[(r.converged, r.iterations) for r in res]}")
        assert all(r.converged for r in res)
        its.append([r.iterations for r in res])
        states.append(s.get_state())
    print(f"  iterations FP64 {its[0]} FP32+refine {its[1]}")
    for a, b in zip(states[0], states[1]):
        assert np.abs(a - b).max() <= 1e-10 * np.abs(a).max()


# the generic kernel (forced; it also runs 3D degree 2 and the non-default
# operator variants) at precision 2 (precision 1 leaves it in FP64): (3, 2,
# 2) keeps its reduced systems in global scratch
GENERIC_CASES = [(2, 4, 1), (2, 4, 2), (2, 3, 3), (3, 2, 1), (3, 2, 2)]


@pytest.mark.parametrize("dim,n_cells,k", GENERIC_CASES)
@pytest.mark.parametrize("thermodyn", [False, True])
def test_mixed_generic_kernel_error(dim, n_cells, k, thermodyn):
    """the generic kernel with the FP32 elimination and one FP64 refinement
    step: per-cell S, g and pass 2 against the FP64 oracle"""
    t = _torch()
    print(f"case mixed generic dim={dim} n={n_cells} k={k} thermo={thermodyn}", flush=True)
    from ddhdg_amd.np_solver import NPSolver
    prob, params, adim = configs.diode(dim, n_cells, k)
    s = NPSolver(prob, params, adim)
    enabled = (True, False, False) if thermodyn else (True, True, True)
    s.set_enabled_components(*enabled)
    s.ctx.set_kernel_variant(1)
    s.ctx.set_precision(2)
    o = H.oracle_for(prob, params, adim, enabled)
    cell, tr = H.perturbed(*H.equilibrium_guess(prob, adim, k), seed=51 + k)
    s.set_state(cell, tr)
    o.set_state(cell, tr)
    ctx = s.ctx
    cells = t.arange(ctx.ncells, dtype=t.int32, device="cuda")
    S = t.empty(ctx.ncells * ctx.T * ctx.T, dtype=t.float64, device="cuda")
    g = t.empty(ctx.ncells * ctx.T, dtype=t.float64, device="cuda")
    ctx.cell_systems(s.x_cell, s.x_trace, thermodyn, cells, S, g)
    rng = np.random.default_rng(7)
    upd = 1e-2 * rng.standard_normal(s.x_trace.numel())
    if thermodyn:
        upd.reshape(-1, 3, s.ctx.nf)[:, 1:] = 0
    dxt = t.as_tensor(upd, device="cuda")
    dxc = t.empty_like(s.x_cell)
    ctx.reconstruct(s.x_cell, s.x_trace, dxt, thermodyn, dxc)
    t.cuda.synchronize()
    S = S.cpu().numpy().reshape(ctx.ncells, ctx.T, ctx.T)
    g = g.cpu().numpy().reshape(ctx.ncells, ctx.T)
    eS = eg = 0.0
    for c in range(ctx.ncells):
        So, go, *_ = o.condensed_cell(c, thermodyn)
        eS, eg = max(eS, H.rel_fro(S[c], So)), max(eg, H.rel_fro(g[c], go))
    e2 = H.rel_fro(dxc.cpu().numpy(), o.pass2(upd, thermodyn))
    print(f"  FP32+refine vs FP64 oracle: worst rel S {eS:.2e} g {eg:.2e}, pass 2 {e2:.2e}")
    assert eS < MIXED_BAR and eg < MIXED_BAR and e2 < MIXED_BAR


@pytest.mark.parametrize("k,variant", [(2, 1), (3, 0)])
def test_precision_1_leaves_other_kernels_fp64(k, variant):
    """precision 1 runs the generic kernel (variant 1) and the degree-3 pair
    kernel in FP64: their blocks equal the FP64 ones bitwise"""
    t = _torch()
    print(f"case precision 1 k={k} variant={variant}", flush=True)
    from ddhdg_amd.np_solver import NPSolver
    prob, params, adim = configs.diode(2, 4, k)
    s = NPSolver(prob, params, adim)
    s.ctx.set_kernel_variant(variant)
    s.set_state(*H.perturbed(*H.equilibrium_guess(prob, adim, k), seed=71))
    ctx = s.ctx
    cells = t.arange(ctx.ncells, dtype=t.int32, device="cuda")
    out = []
    for prec in (0, 1):
        ctx.set_precision(prec)
        S = t.empty(ctx.ncells * ctx.T * ctx.T, dtype=t.float64, device="cuda")
        g = t.empty(ctx.ncells * ctx.T, dtype=t.float64, device="cuda")
        ctx.cell_systems(s.x_cell, s.x_trace, False, cells, S, g)
        out.append((S, g))
    t.cuda.synchronize()
    assert t.equal(out[0][0], out[1][0]) and t.equal(out[0][1], out[1][1])


def test_mixed_generic_variant_matches_fp64():
    """a non-default operator variant (use_trace flux, only on the generic
    kernel) with the FP32 elimination (precision 2): the global system
    against the same kernel in FP64"""
    t = _torch()
    print("case mixed generic variant use_trace 2D k=2", flush=True)
    from ddhdg_amd.np_solver import NPSolver
    prob, params, adim = configs.diode(2, 5, 2, dd_flux_type="use_trace")
    out = []
    for prec in (0, 2):
        s = NPSolver(prob, params, adim)
        s.ctx.set_precision(prec)
        s.set_state(*H.perturbed(*H.equilibrium_guess(prob, adim, 2), seed=61))
        A, b, _, _ = s.assemble_system()
        out.append((A.toarray(), b))
    eA = H.rel_fro(out[1][0], out[0][0])
    eb = H.rel_fro(out[1][1], out[0][1])
    print(f"  FP32+refine vs FP64 global system: rel {eA:.2e}, rhs {eb:.2e}")
    assert eA < MIXED_BAR and eb < MIXED_BAR
