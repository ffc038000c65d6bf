"""Parity of the HIP path with the CPU oracle (GPU tests).

Bars: per-cell condensed blocks S, g and the global trace system to 1e-12
(normwise relative, FP64); reconstructed cell updates to 1e-11; Newton
iteration counts equal and final states within 1e-9 (relative L-infinity).
"""
import numpy as np
import pytest

from ddhdg_amd import configs
from tests import helpers as H

pytestmark = pytest.mark.gpu

CASES = [(2, 4, 1), (2, 4, 2), (2, 3, 3), (3, 2, 1)]


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU test on a machine without a GPU"
    return torch


def _solver(dim, n_cells, k, **kw):
    from ddhdg_amd.np_solver import NPSolver
    prob, params, adim = configs.diode(dim, n_cells, k, **kw)
    return NPSolver(prob, params, adim), prob, params, adim


def _state(prob, adim, k, seed):
    return H.perturbed(*H.equilibrium_guess(prob, adim, k), seed=seed)


@pytest.mark.parametrize("dim,n_cells,k", CASES)
@pytest.mark.parametrize("thermodyn", [False, True])
def test_cell_systems_match_oracle(dim, n_cells, k, thermodyn):
    t = _torch()
    print(f"case cell_systems dim={dim} n={n_cells} k={k} thermo={thermodyn}", flush=True)
    s, prob, params, adim = _solver(dim, n_cells, k)
    enabled = (True, False, False) if thermodyn else (True, True, True)
    s.set_enabled_components(*enabled)
    o = H.oracle_for(prob, params, adim, enabled)
    cell, tr = _state(prob, adim, k, seed=11 + dim + k)
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
    worst_S = worst_g = 0.0
    for c in range(ctx.ncells):
        So, go, _, _, _ = o.condensed_cell(c, thermodyn)
        worst_S = max(worst_S, H.rel_fro(S[c], So))
        worst_g = max(worst_g, H.rel_fro(g[c], go))
    print(f"  worst rel S {worst_S:.3e} g {worst_g:.3e}")
    assert worst_S < 1e-12 and worst_g < 1e-12


@pytest.mark.parametrize("dim,n_cells,k", CASES)
def test_global_system_matches_oracle(dim, n_cells, k):
    t = _torch()
    print(f"case global_system dim={dim} n={n_cells} k={k}", flush=True)
    s, prob, params, adim = _solver(dim, n_cells, k)
    o = H.oracle_for(prob, params, adim)
    cell, tr = _state(prob, adim, k, seed=7)
    s.set_state(cell, tr)
    o.set_state(cell, tr)
    A, b, dirs, dv = s.assemble_system()
    o.pass1()
    Ao, bo, dio, dvo = o.system()
    assert np.array_equal(A.indptr, Ao.indptr) and np.array_equal(A.indices, Ao.indices)
    assert np.array_equal(np.sort(dirs), np.sort(dio))
    keep = np.ones(A.shape[0], dtype=bool)
    keep[dirs] = False
    Ak, Aok = A[keep][:, keep], Ao[keep][:, keep]
    assert H.rel_fro(Ak.toarray(), Aok.toarray()) < 1e-12
    assert H.rel_fro(b[keep], bo[keep]) < 1e-12
    order = np.argsort(dio)
    assert H.rel_fro(dv[np.argsort(dirs)], dvo[order]) < 1e-13
    # Dirichlet rows: unit diagonal, zero elsewhere, zero rhs
    Ad = A[dirs].toarray()
    assert np.all(Ad[np.arange(len(dirs)), dirs] == 1.0)
    assert np.count_nonzero(Ad) == len(dirs) and np.all(b[dirs] == 0)


@pytest.mark.parametrize("dim,n_cells,k", CASES)
@pytest.mark.parametrize("thermodyn", [False, True])
def test_reconstruct_matches_oracle(dim, n_cells, k, thermodyn):
    t = _torch()
    print(f"case reconstruct dim={dim} n={n_cells} k={k} thermo={thermodyn}", flush=True)
    s, prob, params, adim = _solver(dim, n_cells, k)
    enabled = (True, False, False) if thermodyn else (True, True, True)
    s.set_enabled_components(*enabled)
    o = H.oracle_for(prob, params, adim, enabled)
    cell, tr = _state(prob, adim, k, seed=21)
    s.set_state(cell, tr)
    o.set_state(cell, tr)
    rng = np.random.default_rng(5)
    upd = 1e-2 * rng.standard_normal(s.x_trace.numel())
    if thermodyn:
        upd.reshape(-1, 3, s.ctx.nf)[:, 1:] = 0
    dxt = t.as_tensor(upd, device="cuda")
    dxc = t.empty_like(s.x_cell)
    s.ctx.reconstruct(s.x_cell, s.x_trace, dxt, thermodyn, dxc)
    t.cuda.synchronize()
    ref = o.pass2(upd, thermodyn)
    err = H.rel_fro(dxc.cpu().numpy(), ref)
    print(f"  rel err {err:.3e}")
    assert err < 1e-11


def test_condense_is_deterministic():
    t = _torch()
    print("case determinism 2D k=2 n=8", flush=True)
    s, prob, params, adim = _solver(2, 8, 2)
    s.set_state(*_state(prob, adim, 2, seed=3))
    A1, b1, _, _ = s.assemble_system()
    A2, b2, _, _ = s.assemble_system()
    assert np.array_equal(A1.data, A2.data) and np.array_equal(b1, b2)


@pytest.mark.parametrize("dim,n_cells,k", [(2, 12, 1), (2, 8, 2)])
def test_newton_diode_matches_oracle(dim, n_cells, k):
    """equilibrium (V-only Boltzmann Newton) then 0.1 V forward bias: equal
    Newton iteration counts and final states."""
    _torch()
    print(f"case newton dim={dim} n={n_cells} k={k}", flush=True)
    s, prob, params, adim = _solver(dim, n_cells, k)
    o = H.oracle_for(prob, params, adim)
    s.set_enabled_components(True, False, False)
    s.compute_local_charge_neutrality()
    o.set_state(*s.get_state())
    r_gpu = s._private_run(1e-10, 1e-10, 50, True)
    r_orc = o.run_newton(enabled=(True, False, False), thermodyn=True, max_it=50)
    print(f"  equilibrium iterations gpu={r_gpu.iterations} oracle={r_orc.iterations}")
    assert r_gpu.converged and r_orc.converged and r_gpu.iterations == r_orc.iterations
    cg, tg = s.get_state()
    co, to = o.get_state()
    assert np.abs(cg - co).max() <= 1e-9 * np.abs(co).max()
    s.project_densities_from_potential()
    s.set_enabled_components(True, True, True)
    o.set_state(*s.get_state())
    configs.set_bias(s, 0.1)
    prob.boundary_handler  # same handler object: refresh the oracle's values
    o2 = H.oracle_for(prob, params, adim)
    o2.set_state(*s.get_state())
    r_gpu = s.run()
    r_orc = o2.run_newton(max_it=100)
    print(f"  0.1 V iterations gpu={r_gpu.iterations} oracle={r_orc.iterations}")
    assert r_gpu.converged and r_orc.converged and r_gpu.iterations == r_orc.iterations
    cg, tg = s.get_state()
    co, to = o2.get_state()
    assert np.abs(cg - co).max() <= 1e-9 * np.abs(co).max()
    assert np.abs(tg - to).max() <= 1e-9 * np.abs(to).max()
