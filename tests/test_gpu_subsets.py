"""Every subset of {V, n, p} on the GPU (the mask walk of reference
include/templatized_parameters.h:44-101; tests/disabling_components.cpp:
148-276), on the manufactured solution that pins the oracle
(tests/test_oracle_subsets.py): per-cell condensed blocks for both kernels,
and Newton solves with equal iteration counts, states within 1e-9 of the
oracle's and the disabled components left bitwise unchanged."""
import numpy as np
import pytest

from tests import helpers as H
from tests.test_oracle_subsets import SUBSETS, start_state, subset_problem

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU test on a machine without a GPU"
    return torch


def _solver(dim, n_cells, k, enabled):
    from ddhdg_amd.np_solver import NPSolver
    prob, params, adim = subset_problem(n_cells, k, dim)
    s = NPSolver(prob, params, adim)
    s.set_enabled_components(*enabled)
    o = H.oracle_for(prob, params, adim, enabled)
    return s, o, prob


@pytest.mark.parametrize("enabled", SUBSETS[1:])
@pytest.mark.parametrize("dim,n_cells,k", [(2, 4, 1), (2, 4, 2), (2, 3, 3), (3, 2, 1), (3, 2, 2)])
@pytest.mark.parametrize("variant", [1, 0])
def test_subset_cell_systems_match_oracle(enabled, dim, n_cells, k, variant):
    t = _torch()
    print(f"case subset cell_systems en={enabled} dim={dim} n={n_cells} k={k} variant={variant}",
          flush=True)
    s, o, prob = _solver(dim, n_cells, k, enabled)
    s.ctx.set_kernel_variant(variant)
    rng = np.random.default_rng(7 + dim + k)
    cell, tr = start_state(prob.mesh, k, enabled)
    cell = cell * (1 + 0.05 * rng.standard_normal(cell.shape))
    s.set_state(cell, tr)
    o.set_state(cell, tr)
    ctx = s.ctx
    cells = t.arange(ctx.ncells, dtype=t.int32, device="cuda")
    S = t.empty(ctx.ncells * ctx.T * ctx.T, dtype=t.float64, device="cuda")
    g = t.empty(ctx.ncells * ctx.T, dtype=t.float64, device="cuda")
    ctx.cell_systems(s.x_cell, s.x_trace, False, cells, S, g)
    t.cuda.synchronize()
    S = S.cpu().numpy().reshape(ctx.ncells, ctx.T, ctx.T)
    g = g.cpu().numpy().reshape(ctx.ncells, ctx.T)
    worst_S = worst_g = 0.0
    for c in range(ctx.ncells):
        So, go, *_ = o.condensed_cell(c, False)
        worst_S = max(worst_S, H.rel_fro(S[c], So))
        worst_g = max(worst_g, H.rel_fro(g[c], go))
    print(f"  worst rel S {worst_S:.3e} g {worst_g:.3e}", flush=True)
    assert worst_S < 1e-12 and worst_g < 1e-12


@pytest.mark.parametrize("enabled", SUBSETS)
@pytest.mark.parametrize("dim,n_cells,k", [(2, 8, 1), (2, 8, 2), (2, 4, 3), (3, 3, 1)])
def test_subset_newton_matches_oracle(enabled, dim, n_cells, k):
    _torch()
    print(f"case subset newton en={enabled} dim={dim} n={n_cells} k={k}", flush=True)
    s, o, prob = _solver(dim, n_cells, k, enabled)
    c0, t0 = start_state(prob.mesh, k, enabled)
    s.set_state(c0, t0)
    o.set_state(c0, t0)
    r = s.run()
    ro = o.run_newton(enabled=enabled, max_it=30)
    print(f"  iterations gpu={r.iterations} oracle={ro.iterations}", flush=True)
    assert r.converged and ro.converged and r.iterations == ro.iterations
    cg, tg = s.get_state()
    co, to = o.get_state()
    assert np.abs(cg - co).max() <= 1e-9 * np.abs(co).max()
    assert np.abs(tg - to).max() <= 1e-9 * np.abs(to).max()
    m = prob.mesh
    cc, cs = cg.reshape(m.ncells, 3, -1), c0.reshape(m.ncells, 3, -1)
    tt, ts = tg.reshape(m.nfaces, 3, -1), t0.reshape(m.nfaces, 3, -1)
    for comp in range(3):
        if not enabled[comp]:
            assert np.array_equal(cc[:, comp], cs[:, comp])
            assert np.array_equal(tt[:, comp], ts[:, comp])
