"""GPU parity of the non-default operator variants with the oracle: the
use_trace and qiu_shi_stabilization fluxes, per-cell-face tau
(CellFaceTauNPSolverParameters) and phi_linearize.

Bars as for the default operator: per-cell S, g and the global system to
1e-12, reconstructed cell updates to 1e-11 (normwise relative, FP64), equal
Newton iteration counts and final states within 1e-9, terminal currents to
their rounding.  These variants run on the LDS kernel (every shape)."""
import numpy as np
import pytest

from ddhdg_amd import configs, fe
from ddhdg_amd.problem import V, n, p
from tests import helpers as H

pytestmark = pytest.mark.gpu

CASES = [(2, 4, 1), (2, 4, 2), (2, 3, 3), (3, 2, 1)]
VARIANTS = [("use_trace", "fixed", False), ("qiu_shi_stabilization", "fixed", False),
            ("use_cell", "cellface", False), ("use_trace", "cellface", False),
            ("qiu_shi_stabilization", "cellface", False), ("use_cell", "fixed", True),
            ("use_trace", "cellface", True)]


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU test on a machine without a GPU"
    return torch


def _tau_table(prob, adim, seed, lo, hi):
    rng = np.random.default_rng(seed)
    rf = np.array([adim.get_tau_rescaling_factor(c) for c in range(3)])
    return rng.uniform(lo, hi, (prob.mesh.ncells, 2 * prob.mesh.dim, 3)) * rf


def _problem(dim, n_cells, k, flux, tau_kind, phi, lo=0.5, hi=2.0, **kw):
    prob, params, adim = configs.diode(dim, n_cells, k, dd_flux_type=flux,
                                       phi_linearize=phi, **kw)
    if tau_kind == "cellface":
        tab = _tau_table(prob, adim, 17 + dim + k, lo, hi)
        prob, params, adim = configs.diode(dim, n_cells, k, dd_flux_type=flux,
                                           phi_linearize=phi, cell_face_tau=tab, **kw)
    return prob, params, adim


def _solver(prob, params, adim):
    from ddhdg_amd.np_solver import NPSolver
    return NPSolver(prob, params, adim)


def _smooth_state(prob, k, seed):
    """projection of smooth fields with positive densities at every
    quadrature point (phi_linearize takes their logarithm) and an electric
    field whose normal component changes sign across the mesh (both
    branches of the qiu_shi switch).  The potential stays within 0.2 V
    (about 8 U_T): the equilibrium operator's Boltzmann terms grow as
    exp(V/U_T) and a larger swing only measures their conditioning."""
    mesh = prob.mesh
    ref = fe.RefCell(mesh, k)
    rng = np.random.default_rng(seed)
    a = rng.uniform(1, 3, 3)
    cq, fq = H._cell_qpoints(mesh, ref), H._face_qpoints(mesh, ref)
    cfq = fq[mesh.cell_faces]
    funcs = {V: lambda x: 0.2 * np.sin(a[0] * x[..., 0] + a[1] * x[..., -1]),
             n: lambda x: 0.6 * np.exp(0.7 * np.cos(a[2] * x[..., 0] - x[..., 1])),
             p: lambda x: 0.3 * np.exp(0.5 * np.sin(a[1] * x[..., 1] + x[..., 0]))}
    d, nb = mesh.dim, ref.nb
    cell = np.zeros((mesh.ncells, 3, d + 1, nb))
    trace = np.zeros((mesh.nfaces, 3, ref.nf))
    for c, f in funcs.items():
        u, W = fe.project_cells(mesh, ref, f(cq), f(cfq))
        cell[:, c, d], cell[:, c, :d] = u, W
        trace[:, c] = fe.trace_project(mesh, ref, f(fq))
    for c in (n, p):
        assert (cell[:, c, d] @ ref.phic.T).min() > 0
    return cell.ravel(), trace.ravel()


@pytest.mark.parametrize("dim,n_cells,k", CASES)
@pytest.mark.parametrize("flux,tau_kind,phi", VARIANTS)
@pytest.mark.parametrize("thermodyn", [False, True])
def test_variant_cell_systems_match_oracle(dim, n_cells, k, flux, tau_kind, phi, thermodyn):
    t = _torch()
    print(f"case variant cell_systems dim={dim} n={n_cells} k={k} flux={flux} "
          f"tau={tau_kind} phi={phi} thermo={thermodyn}", flush=True)
    prob, params, adim = _problem(dim, n_cells, k, flux, tau_kind, phi)
    s = _solver(prob, params, adim)
    enabled = (True, False, False) if thermodyn else (True, True, True)
    s.set_enabled_components(*enabled)
    o = H.oracle_for(prob, params, adim, enabled)
    cell, tr = _smooth_state(prob, k, seed=dim + 3 * k)
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
@pytest.mark.parametrize("flux,tau_kind,phi", VARIANTS)
def test_variant_reconstruct_matches_oracle(dim, n_cells, k, flux, tau_kind, phi):
    """pass 2, including the reconstruct pass's omission of the upwind trace
    term for use_trace / qiu_shi (template.h:2122-2160)"""
    t = _torch()
    print(f"case variant reconstruct dim={dim} n={n_cells} k={k} flux={flux} "
          f"tau={tau_kind} phi={phi}", flush=True)
    prob, params, adim = _problem(dim, n_cells, k, flux, tau_kind, phi)
    s = _solver(prob, params, adim)
    o = H.oracle_for(prob, params, adim)
    cell, tr = _smooth_state(prob, k, seed=5 + dim + k)
    s.set_state(cell, tr)
    o.set_state(cell, tr)
    rng = np.random.default_rng(7)
    upd = 1e-2 * rng.standard_normal(s.x_trace.numel())
    dxt = t.as_tensor(upd, device="cuda")
    dxc = t.empty_like(s.x_cell)
    s.ctx.reconstruct(s.x_cell, s.x_trace, dxt, False, dxc)
    t.cuda.synchronize()
    err = H.rel_fro(dxc.cpu().numpy(), o.pass2(upd))
    print(f"  rel err {err:.3e}")
    assert err < 1e-11


@pytest.mark.parametrize("flux", ["use_trace", "qiu_shi_stabilization"])
def test_variant_global_system_matches_oracle(flux):
    t = _torch()
    print(f"case variant global system 2D n=5 k=2 flux={flux} cellface", flush=True)
    prob, params, adim = _problem(2, 5, 2, flux, "cellface", False)
    s = _solver(prob, params, adim)
    o = H.oracle_for(prob, params, adim)
    cell, tr = H.perturbed(*H.equilibrium_guess(prob, adim, 2), seed=3)
    s.set_state(cell, tr)
    o.set_state(cell, tr)
    A, b, dirs, dv = s.assemble_system()
    o.pass1()
    Ao, bo, dio, dvo = o.system()
    keep = np.ones(A.shape[0], dtype=bool)
    keep[dirs] = False
    assert H.rel_fro(A[keep][:, keep].toarray(), Ao[keep][:, keep].toarray()) < 1e-12
    assert H.rel_fro(b[keep], bo[keep]) < 1e-12


def _close(s, o, tol=1e-9):
    cg, tg = s.get_state()
    co, to = o.get_state()
    return (np.abs(cg - co).max() <= tol * np.abs(co).max() and
            np.abs(tg - to).max() <= tol * np.abs(to).max())


@pytest.mark.parametrize("n_cells,k,flux,tau_kind", [
    (8, 2, "use_trace", "fixed"), (12, 1, "qiu_shi_stabilization", "fixed"),
    (8, 2, "use_trace", "cellface"), (8, 1, "use_cell", "cellface")])
def test_variant_newton_diode_matches_oracle(n_cells, k, flux, tau_kind):
    """the diode workflow with the variant operator (per-face tau within
    +-10 % of the default): equal iteration counts and states at every
    stage, both sides starting each stage from the same state"""
    _torch()
    print(f"case variant newton 2D n={n_cells} k={k} flux={flux} tau={tau_kind}", flush=True)
    prob, params, adim = _problem(2, n_cells, k, flux, tau_kind, False, lo=0.9, hi=1.1)
    s = _solver(prob, params, adim)
    o = H.oracle_for(prob, params, adim)
    s.set_enabled_components(True, False, False)
    s.compute_local_charge_neutrality()
    o.set_state(*s.get_state())
    r_gpu = s._private_run(1e-10, 1e-10, 100, True)
    r_orc = o.run_newton(enabled=(True, False, False), thermodyn=True, max_it=100)
    assert r_gpu.converged and r_orc.converged and r_gpu.iterations == r_orc.iterations
    assert _close(s, o)
    s.project_densities_from_potential()
    for name, en, bias in [("n,p", (False, True, True), None), ("coupled", (True, True, True), None),
                           ("0.1 V", (True, True, True), 0.1), ("0.3 V", (True, True, True), 0.3)]:
        s.set_enabled_components(*en)
        if bias is not None:
            configs.set_bias(s, bias)
        o = H.oracle_for(prob, params, adim)
        o.set_state(*s.get_state())
        r_gpu = s.run()
        r_orc = o.run_newton(enabled=en, max_it=100)
        print(f"  {name}: iterations gpu={r_gpu.iterations} oracle={r_orc.iterations}")
        assert r_gpu.converged and r_orc.converged, name
        assert r_gpu.iterations == r_orc.iterations, name
        assert _close(s, o), name


@pytest.mark.parametrize("n_cells,k", [((32, 2), 1), ((16, 2), 2)])
def test_phi_linearize_newton_matches_oracle(n_cells, k):
    """phi_linearize bias solves (frozen quasi-Fermi potentials in the V
    rows: an inexact Newton) on a resolved lightly doped diode, from the
    converged default-operator state: equal iteration counts and states"""
    _torch()
    print(f"case phi_linearize newton 2D n={n_cells} k={k}", flush=True)
    dop = configs.step_doping(1e21, 1e21)
    prob, params, adim = configs.diode(2, n_cells, k, doping=dop)
    s = _solver(prob, params, adim)
    s.set_enabled_components(True, False, False)
    s.compute_local_charge_neutrality()
    assert s._private_run(1e-10, 1e-10, 100, True).converged
    s.project_densities_from_potential()
    for en in ((False, True, True), (True, True, True)):
        s.set_enabled_components(*en)
        assert s.run().converged
    probp, paramsp, _ = configs.diode(2, n_cells, k, doping=dop, phi_linearize=True)
    sp = _solver(probp, paramsp, adim)
    sp.set_state(*s.get_state())
    for bias in (0.1, 0.3):
        configs.set_bias(sp, bias)
        o = H.oracle_for(probp, paramsp, adim)
        o.set_state(*sp.get_state())
        r_gpu = sp.run()
        r_orc = o.run_newton(max_it=100)
        print(f"  {bias} V: iterations gpu={r_gpu.iterations} oracle={r_orc.iterations}")
        assert r_gpu.converged and r_orc.converged
        assert r_gpu.iterations == r_orc.iterations
        assert _close(sp, o)


@pytest.mark.parametrize("flux,tau_kind", [("use_trace", "fixed"),
                                           ("qiu_shi_stabilization", "cellface"),
                                           ("use_cell", "cellface")])
def test_variant_terminal_current_matches_oracle(flux, tau_kind):
    t = _torch()
    print(f"case variant terminal current 2D (6,5) k=2 flux={flux} tau={tau_kind}", flush=True)
    prob, params, adim = _problem(2, (6, 5), 2, flux, tau_kind, False)
    s = _solver(prob, params, adim)
    o = H.oracle_for(prob, params, adim)
    cell, tr = _smooth_state(prob, 2, seed=23)
    s.set_state(cell, tr)
    o.set_state(cell, tr)
    out = t.empty(2, dtype=t.float64, device="cuda")
    for bid in range(4):
        s.ctx.terminal_current(s.x_cell, s.x_trace, bid, out)
        g = out.cpu().numpy()
        r = o.terminal_current(bid)
        print(f"  bid {bid}: gpu {g} oracle {r}")
        assert np.allclose(g, r, rtol=1e-12, atol=1e-14 * np.abs(r).max())


def test_variant_graph_newton_equals_eager_with_tau_update():
    """graph-captured Newton with per-face tau and use_trace equals eager
    execution bitwise, also after the tau values change in place between
    two bias points"""
    _torch()
    print("case variant graph 2D n=8 k=2 use_trace cellface", flush=True)

    def workflow(graph):
        prob, params, adim = _problem(2, 8, 2, "use_trace", "cellface", False, lo=0.9, hi=1.1)
        params.graph_newton = graph
        s = _solver(prob, params, adim)
        s.set_enabled_components(True, False, False)
        s.compute_local_charge_neutrality()
        res = configs.solve_equilibrium_and_bias(s, (0.1,))
        tab = _tau_table(prob, adim, 99, 0.95, 1.05)
        for c in range(prob.mesh.ncells):
            for f in range(4):
                params.set_face(c, f, *tab[c, f])
        configs.set_bias(s, 0.2)
        res.append(s.run())
        return [(r.converged, r.iterations) for r in res], s.get_state()

    its_e, st_e = workflow(False)
    its_g, st_g = workflow(True)
    print(f"  eager {its_e} graph {its_g}")
    assert its_e == its_g and all(c for c, _ in its_e)
    assert all(np.array_equal(a, b) for a, b in zip(st_e, st_g))


def test_variants_are_refused_by_the_register_kernel_paths():
    """the register kernel and the batched launch run the default operator
    only: forcing them on a variant context is an error, not a silent
    fallback"""
    from ddhdg_amd.hip_lib import HdgBatch, HdgError
    _torch()
    prob, params, adim = _problem(2, 4, 2, "use_cell", "cellface", False)
    s = _solver(prob, params, adim)
    with pytest.raises(HdgError):
        s.ctx.set_kernel_variant(2)
    t = _torch()
    b = HdgBatch([s.ctx])
    z = lambda m: t.zeros(m, dtype=t.float64, device="cuda")
    with pytest.raises(HdgError):
        b.condense(False, [s.x_cell], [s.x_trace], [z(s.ctx.nnz)], [z(s.ctx.ntrace)],
                   [z(s.ctx.ntrace)])
