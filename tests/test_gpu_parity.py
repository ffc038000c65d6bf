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

# (3, 2, 2): the generic kernel with its reduced systems in global scratch
CASES = [(2, 4, 1), (2, 4, 2), (2, 3, 3), (3, 2, 1), (3, 2, 2)]


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU test on a machine without a GPU"
    return torch


def _solver(dim, n_cells, k, variant=0, **kw):
    from ddhdg_amd.np_solver import NPSolver
    prob, params, adim = configs.diode(dim, n_cells, k, **kw)
    s = NPSolver(prob, params, adim)
    s.ctx.set_kernel_variant(variant)
    return s, prob, params, adim


def _state(prob, adim, k, seed):
    return H.perturbed(*H.equilibrium_guess(prob, adim, k), seed=seed)


@pytest.mark.parametrize("dim,n_cells,k", CASES)
@pytest.mark.parametrize("thermodyn", [False, True])
@pytest.mark.parametrize("variant", [1, 0])
@pytest.mark.parametrize("consistent", [False, True])
def test_cell_systems_match_oracle(dim, n_cells, k, thermodyn, variant, consistent):
    t = _torch()
    print(f"case cell_systems dim={dim} n={n_cells} k={k} thermo={thermodyn} "
          f"variant={variant} consistent={consistent}", flush=True)
    s, prob, params, adim = _solver(dim, n_cells, k, variant,
                                    consistent_drift_jacobian=consistent)
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


@pytest.mark.parametrize("dim,n_cells,k", [(2, 4, 2), (3, 2, 1), (3, 2, 2)])
@pytest.mark.parametrize("variant", [1, 0])
def test_cell_systems_carriers_only_match_oracle(dim, n_cells, k, variant):
    """n and p enabled with V frozen (the second stage of the diode
    workflow): two components, so 2 (k+1)^d scalar rows per cell"""
    t = _torch()
    print(f"case cell_systems n,p dim={dim} n={n_cells} k={k} variant={variant}", flush=True)
    s, prob, params, adim = _solver(dim, n_cells, k, variant)
    enabled = (False, True, True)
    s.set_enabled_components(*enabled)
    o = H.oracle_for(prob, params, adim, enabled)
    cell, tr = _state(prob, adim, k, seed=5 + dim + k)
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
        So, go, _, _, _ = o.condensed_cell(c, False)
        worst_S = max(worst_S, H.rel_fro(S[c], So))
        worst_g = max(worst_g, H.rel_fro(g[c], go))
    print(f"  worst rel S {worst_S:.3e} g {worst_g:.3e}")
    assert worst_S < 1e-12 and worst_g < 1e-12


@pytest.mark.parametrize("dim,n_cells,k", CASES)
@pytest.mark.parametrize("variant", [1, 0])
def test_global_system_matches_oracle(dim, n_cells, k, variant):
    t = _torch()
    print(f"case global_system dim={dim} n={n_cells} k={k} variant={variant}", flush=True)
    s, prob, params, adim = _solver(dim, n_cells, k, variant)
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
@pytest.mark.parametrize("variant", [1, 0])
def test_reconstruct_matches_oracle(dim, n_cells, k, thermodyn, variant):
    t = _torch()
    print(f"case reconstruct dim={dim} n={n_cells} k={k} thermo={thermodyn} variant={variant}",
          flush=True)
    s, prob, params, adim = _solver(dim, n_cells, k, variant)
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


def _oracle_stage(o, enabled, thermodyn=False):
    return o.run_newton(enabled=enabled, thermodyn=thermodyn, max_it=100)


def _close(s, o, tol=1e-9):
    cg, tg = s.get_state()
    co, to = o.get_state()
    return (np.abs(cg - co).max() <= tol * np.abs(co).max() and
            np.abs(tg - to).max() <= tol * np.abs(to).max())


@pytest.mark.parametrize("dim,n_cells,k", [(2, 12, 1), (2, 8, 2), (2, 6, 3), (3, 2, 2)])
def test_newton_diode_matches_oracle(dim, n_cells, k):
    """the reference's diode workflow (equilibrium, n and p with V frozen,
    coupled system, 0.1 V and 0.3 V forward bias): equal Newton iteration
    counts at every stage and equal final states.  Both sides start every
    stage from the same state."""
    _torch()
    print(f"case newton dim={dim} n={n_cells} k={k}", flush=True)
    s, prob, params, adim = _solver(dim, n_cells, k)
    o = H.oracle_for(prob, params, adim)
    s.set_enabled_components(True, False, False)
    s.compute_local_charge_neutrality()
    o.set_state(*s.get_state())
    r_gpu = s._private_run(1e-10, 1e-10, 100, True)
    r_orc = _oracle_stage(o, (True, False, False), True)
    print(f"  equilibrium iterations gpu={r_gpu.iterations} oracle={r_orc.iterations}")
    assert r_gpu.converged and r_orc.converged and r_gpu.iterations == r_orc.iterations
    assert _close(s, o)
    s.project_densities_from_potential()
    stages = [("n,p", (False, True, True), None), ("coupled 0 V", (True, True, True), None),
              ("0.1 V", (True, True, True), 0.1), ("0.3 V", (True, True, True), 0.3)]
    for name, en, bias in stages:
        s.set_enabled_components(*en)
        if bias is not None:
            configs.set_bias(s, bias)
        o = H.oracle_for(prob, params, adim)
        o.set_state(*s.get_state())
        r_gpu = s.run()
        r_orc = _oracle_stage(o, en)
        print(f"  {name}: iterations gpu={r_gpu.iterations} oracle={r_orc.iterations}")
        assert r_gpu.converged and r_orc.converged, name
        assert r_gpu.iterations == r_orc.iterations, name
        assert _close(s, o), name


def test_newton_reference_linearisation_matches_oracle():
    """the reference's own drift linearisation (an inexact Newton that only
    converges linearly) on a lightly doped diode: the same iterates, step by
    step, for a fixed number of iterations"""
    _torch()
    print("case newton reference linearisation 2D n=8 k=1 doping 1e21", flush=True)
    s, prob, params, adim = _solver(2, 8, 1, doping=configs.step_doping(1e21, 1e21),
                                    consistent_drift_jacobian=False)
    s.set_enabled_components(True, False, False)
    s.compute_local_charge_neutrality()
    s._private_run(1e-10, 1e-10, 100, True)
    s.project_densities_from_potential()
    s.set_enabled_components(True, True, True)
    configs.set_bias(s, 0.1)
    o = H.oracle_for(prob, params, adim)
    o.set_state(*s.get_state())
    ng, no = [], []
    r_gpu = s.run(max_number_of_iterations=15, log=lambda st, u, c: ng.append(u))
    r_orc = o.run_newton(max_it=15, log=lambda st, u, c: no.append(u))
    print(f"  update norms gpu={ng[-1]:.3e} oracle={no[-1]:.3e} after {len(ng)} steps")
    assert r_gpu.iterations == r_orc.iterations == 15
    assert np.allclose(ng, no, rtol=1e-8, atol=0)
    assert _close(s, o)


@pytest.mark.parametrize("dim,n_cells,k", [(2, (6, 5), 2), (2, (4, 4), 1), (3, (2, 3, 2), 1)])
def test_terminal_current_matches_oracle(dim, n_cells, k):
    """integrals of the numerical fluxes of n and p over every boundary id,
    GPU against the oracle on the same perturbed state"""
    t = _torch()
    print(f"case terminal current dim={dim} n={n_cells} k={k}", flush=True)
    from ddhdg_amd.np_solver import NPSolver
    prob, params, adim = configs.diode(dim, n_cells, k)
    s = NPSolver(prob, params, adim)
    o = H.oracle_for(prob, params, adim)
    cell, tr = H.perturbed(*H.equilibrium_guess(prob, adim, k), seed=13)
    s.set_state(cell, tr)
    o.set_state(cell, tr)
    out = t.empty(2, dtype=t.float64, device="cuda")
    for bid in range(2 * dim):
        s.ctx.terminal_current(s.x_cell, s.x_trace, bid, out)
        g = out.cpu().numpy()
        r = o.terminal_current(bid)
        print(f"  bid {bid}: gpu {g} oracle {r}")
        assert np.allclose(g, r, rtol=1e-12, atol=1e-14 * np.abs(r).max())


def test_iv_points_match_oracle():
    """forward-bias I-V points of a resolved diode strip (128 x 1 cells,
    degree 2): every bias solve starts both sides from the same state; equal
    Newton iteration counts, equal states, equal terminal currents, and the
    current conserved between the two contacts"""
    t = _torch()
    print("case I-V strip 128x1 k=2", flush=True)
    s, prob, params, adim = _solver(2, (128, 1), 2)
    s.set_enabled_components(True, False, False)
    s.compute_local_charge_neutrality()
    assert s._private_run(1e-10, 1e-10, 100, True).converged
    s.project_densities_from_potential()
    for en in ((False, True, True), (True, True, True)):
        s.set_enabled_components(*en)
        assert s.run().converged
    for bias in (0.3, 0.4):
        configs.set_bias(s, bias)
        o = H.oracle_for(prob, params, adim)
        o.set_state(*s.get_state())
        r_gpu = s.run()
        r_orc = o.run_newton()
        I_gpu = [s.terminal_current(b) for b in (0, 1)]
        I_orc = [(o.terminal_current(b)[0] - o.terminal_current(b)[1]) * s.current_scale()
                 for b in (0, 1)]
        print(f"  {bias} V: iterations gpu={r_gpu.iterations} oracle={r_orc.iterations} "
              f"I gpu={I_gpu} oracle={I_orc} A/m")
        assert r_gpu.converged and r_orc.converged
        assert r_gpu.iterations == r_orc.iterations
        assert _close(s, o)
        # the contact current is a small difference of O(1) drift and
        # diffusion fluxes (dimensionless): it agrees to their rounding
        assert np.allclose(I_gpu, I_orc, rtol=1e-8, atol=1e-13 * s.current_scale())
        assert abs(I_gpu[0] + I_gpu[1]) < 1e-5 * abs(I_gpu[1]) + 1e-14 * s.current_scale()


@pytest.mark.parametrize("n_cells", [(3, 2, 4), (5, 4, 3)])
@pytest.mark.parametrize("thermodyn", [False, True])
def test_3d_register_kernel_matches_lds_kernel(n_cells, thermodyn):
    """3D degree 1: the register kernel (two columns per lane) against the
    LDS kernel and the oracle on the same state, per-cell blocks and the
    global system"""
    t = _torch()
    print(f"case 3D register kernel n={n_cells} thermo={thermodyn}", flush=True)
    from ddhdg_amd.np_solver import NPSolver
    prob, params, adim = configs.diode(3, n_cells, 1)
    s = NPSolver(prob, params, adim)
    enabled = (True, False, False) if thermodyn else (True, True, True)
    s.set_enabled_components(*enabled)
    o = H.oracle_for(prob, params, adim, enabled)
    cell, tr = _state(prob, adim, 1, seed=19)
    s.set_state(cell, tr)
    o.set_state(cell, tr)
    ctx = s.ctx
    cells = t.arange(ctx.ncells, dtype=t.int32, device="cuda")
    out = {}
    for variant in (2, 1):
        ctx.set_kernel_variant(variant)
        S = t.empty(ctx.ncells * ctx.T * ctx.T, dtype=t.float64, device="cuda")
        g = t.empty(ctx.ncells * ctx.T, dtype=t.float64, device="cuda")
        ctx.cell_systems(s.x_cell, s.x_trace, thermodyn, cells, S, g)
        A, b, dirs, dv = s.assemble_system(thermodyn)
        t.cuda.synchronize()
        out[variant] = (S.cpu().numpy().reshape(ctx.ncells, ctx.T, ctx.T),
                        g.cpu().numpy().reshape(ctx.ncells, ctx.T), A, b)
    worst = 0.0
    for c in range(ctx.ncells):
        So, go, _, _, _ = o.condensed_cell(c, thermodyn)
        worst = max(worst, H.rel_fro(out[2][0][c], So), H.rel_fro(out[2][1][c], go))
    print(f"  register kernel vs oracle: worst rel {worst:.3e}")
    assert worst < 1e-12
    assert H.rel_fro(out[2][2].toarray(), out[1][2].toarray()) < 1e-12
    assert H.rel_fro(out[2][3], out[1][3]) < 1e-12


@pytest.mark.parametrize("n_cells", [(5, 4), (7, 3), (1, 1)])
def test_2d_k3_register_kernel_matches_lds_kernel(n_cells):
    """2D degree 3 with V/n/p: the register kernel (a pair of waves per
    cell) against the LDS kernel and the oracle on the same state, per-cell
    blocks and the global system.  Odd cell counts per colour leave a pair
    of the last workgroup without a cell of its own."""
    t = _torch()
    print(f"case 2D k=3 register kernel n={n_cells}", flush=True)
    from ddhdg_amd.np_solver import NPSolver
    prob, params, adim = configs.diode(2, n_cells, 3)
    s = NPSolver(prob, params, adim)
    o = H.oracle_for(prob, params, adim)
    cell, tr = _state(prob, adim, 3, seed=23)
    s.set_state(cell, tr)
    o.set_state(cell, tr)
    ctx = s.ctx
    cells = t.arange(ctx.ncells, dtype=t.int32, device="cuda")
    out = {}
    for variant in (2, 1):
        ctx.set_kernel_variant(variant)
        S = t.empty(ctx.ncells * ctx.T * ctx.T, dtype=t.float64, device="cuda")
        g = t.empty(ctx.ncells * ctx.T, dtype=t.float64, device="cuda")
        ctx.cell_systems(s.x_cell, s.x_trace, False, cells, S, g)
        A, b, dirs, dv = s.assemble_system(False)
        t.cuda.synchronize()
        out[variant] = (S.cpu().numpy().reshape(ctx.ncells, ctx.T, ctx.T),
                        g.cpu().numpy().reshape(ctx.ncells, ctx.T), A, b, dv)
    worst = 0.0
    for c in range(ctx.ncells):
        So, go, _, _, _ = o.condensed_cell(c, False)
        worst = max(worst, H.rel_fro(out[2][0][c], So), H.rel_fro(out[2][1][c], go))
    print(f"  register kernel vs oracle: worst rel {worst:.3e}")
    assert worst < 1e-12
    assert H.rel_fro(out[2][2].toarray(), out[1][2].toarray()) < 1e-12
    assert H.rel_fro(out[2][3], out[1][3]) < 1e-12
    assert np.array_equal(out[2][4], out[1][4])


def test_newton_3d_register_kernel_matches_oracle():
    """the 3D diode workflow on the register-resident 3D kernel (two columns
    per lane): equal Newton iteration counts and states at every stage"""
    _torch()
    print("case newton 3D n=(8,2,2) k=1 doping 1e22", flush=True)
    s, prob, params, adim = _solver(3, (8, 2, 2), 1, doping=configs.step_doping(1e22, 1e22))
    o = H.oracle_for(prob, params, adim)
    s.set_enabled_components(True, False, False)
    s.compute_local_charge_neutrality()
    o.set_state(*s.get_state())
    r_gpu = s._private_run(1e-10, 1e-10, 100, True)
    r_orc = _oracle_stage(o, (True, False, False), True)
    assert r_gpu.converged and r_orc.converged and r_gpu.iterations == r_orc.iterations
    assert _close(s, o)
    s.project_densities_from_potential()
    for name, en, bias in [("n,p", (False, True, True), None), ("coupled", (True, True, True), None),
                           ("0.1 V", (True, True, True), 0.1)]:
        s.set_enabled_components(*en)
        if bias is not None:
            configs.set_bias(s, bias)
        o = H.oracle_for(prob, params, adim)
        o.set_state(*s.get_state())
        r_gpu = s.run()
        r_orc = _oracle_stage(o, en)
        print(f"  {name}: iterations gpu={r_gpu.iterations} oracle={r_orc.iterations}")
        assert r_gpu.converged and r_orc.converged, name
        assert r_gpu.iterations == r_orc.iterations, name
        assert _close(s, o), name


def test_3d_k2_generic_kernel_multiple_rounds():
    """3D degree 2 on the generic kernel, whose reduced systems live in a
    global scratch slice per workgroup of a persistent grid of 512: an
    11x11x10 mesh has about 605 cells per colour, so part of the grid runs a
    second round on its reused slice.  Cells of the second round against the
    oracle, the global system against the oracle's, and pass 2 against the
    oracle's on every cell."""
    t = _torch()
    print("case 3D k=2 generic kernel, 11x11x10 (two rounds per colour)", flush=True)
    s, prob, params, adim = _solver(3, (11, 11, 10), 2)
    o = H.oracle_for(prob, params, adim)
    cell, tr = _state(prob, adim, 2, seed=31)
    s.set_state(cell, tr)
    o.set_state(cell, tr)
    ctx = s.ctx
    # all 1210 cells in one call: the grid of 512 takes list entries 512 and
    # beyond in its second and third rounds
    idx = t.arange(ctx.ncells, dtype=t.int32, device="cuda")
    S = t.empty(ctx.ncells * ctx.T * ctx.T, dtype=t.float64, device="cuda")
    g = t.empty(ctx.ncells * ctx.T, dtype=t.float64, device="cuda")
    ctx.cell_systems(s.x_cell, s.x_trace, False, idx, S, g)
    S = S.cpu().numpy().reshape(-1, ctx.T, ctx.T)
    g = g.cpu().numpy().reshape(-1, ctx.T)
    sample = np.concatenate([np.arange(0, 512, 97), np.arange(512, ctx.ncells, 23)])
    worst = 0.0
    for c in sample:
        So, go, *_ = o.condensed_cell(int(c), False)
        worst = max(worst, H.rel_fro(S[c], So), H.rel_fro(g[c], go))
    print(f"  {len(sample)} cells, most of later rounds: worst rel {worst:.3e}", flush=True)
    assert worst < 1e-12
    A, b, dirs, dv = s.assemble_system()
    o.pass1(False, 16)
    Ao, bo, dio, dvo = o.system()
    keep = np.ones(A.shape[0], dtype=bool)
    keep[dirs] = False
    errA = H.rel_fro(A[keep][:, keep].data, Ao[keep][:, keep].data)
    errb = H.rel_fro(b[keep], bo[keep])
    print(f"  global system: values {errA:.3e} rhs {errb:.3e}", flush=True)
    assert errA < 1e-12 and errb < 1e-12
    rng = np.random.default_rng(6)
    upd = 1e-2 * rng.standard_normal(s.x_trace.numel())
    dxc = t.empty_like(s.x_cell)
    ctx.reconstruct(s.x_cell, s.x_trace, t.as_tensor(upd, device="cuda"), False, dxc)
    err = H.rel_fro(dxc.cpu().numpy(), o.pass2(upd, False, 16))
    print(f"  pass 2 rel err {err:.3e}", flush=True)
    assert err < 1e-11
