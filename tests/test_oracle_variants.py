"""Pins of the oracle's non-default operator variants (no GPU needed): the
drift-diffusion flux types use_trace and qiu_shi_stabilization, per-cell-face
tau (CellFaceTauComputer) and phi_linearize.

What fixes them:
- an exact discrete solution with a drift current (V linear, n and p
  constant): its residual is zero for every flux type, tau and
  phi_linearize;
- closed forms of the face residual for a single trace jump, which separate
  the flux types and the (cell, face, component) indexing of tau;
- the local matrix against central differences of the local residual;
- the reference's own CellFace-vs-Fixed tau test
  (tests/cell_face_tau_computer.cpp:144-209);
- for phi_linearize, the Boltzmann identity of the residual and the closed
  form of its V-row derivative at frozen quasi-Fermi potentials;
- the reconstruct pass of use_trace / qiu_shi, which omits the upwind trace
  term (include/assembling_methods.template.h:2122-2160), against the local
  blocks of the use_cell operator.
"""
import numpy as np
import pytest

from ddhdg_amd import configs, fe
from ddhdg_amd.problem import (DIRICHLET, KB, Q, LinearRecombinationTerm, V, n, p)
from tests import helpers as H

FLUX = ["use_cell", "use_trace", "qiu_shi_stabilization"]


def _random_tau(prob, adim, seed, lo=0.5, hi=2.0):
    """physical per-(cell, face, component) tau whose rescaled values lie in
    [lo, hi]"""
    mesh = prob.mesh
    rng = np.random.default_rng(seed)
    rf = np.array([adim.get_tau_rescaling_factor(c) for c in range(3)])
    return rng.uniform(lo, hi, (mesh.ncells, 2 * mesh.dim, 3)) * rf


def _zero_doping(x):
    return 0.0 * x[..., 0]


# ------------------------------------------------------------ exact state
def _drift_state(prob, adim, k, e, n0, p0):
    """V = -e x (so E = e along x), n = n0, p = p0 (dimensionless), cells
    and traces, by projection (exact: all are in the discrete spaces)"""
    mesh = prob.mesh
    ref = fe.RefCell(mesh, k)
    cq, fq = H._cell_qpoints(mesh, ref), H._face_qpoints(mesh, ref)
    cfq = fq[mesh.cell_faces]
    d, nb = mesh.dim, ref.nb
    cell = np.zeros((mesh.ncells, 3, d + 1, nb))
    trace = np.zeros((mesh.nfaces, 3, ref.nf))
    funcs = {V: lambda x: -e * x[..., 0], n: lambda x: n0 + 0 * x[..., 0],
             p: lambda x: p0 + 0 * x[..., 0]}
    for c, f in funcs.items():
        u, W = fe.project_cells(mesh, ref, f(cq), f(cfq))
        cell[:, c, d], cell[:, c, :d] = u, W
        trace[:, c] = fe.trace_project(mesh, ref, f(fq))
    return cell.ravel(), trace.ravel()


def _drift_problem(dim, ncells, k, e, n0, p0, **kw):
    """the diode materials with no recombination, constant net doping
    n0 - p0 and Dirichlet data of the exact state on ids 0 and 1"""
    nrf = 1e23
    prob, params, adim = configs.diode(dim, ncells, k,
                                       doping=lambda x: (n0 - p0) * nrf + 0 * x[..., 0], **kw)
    prob.recombination_term = LinearRecombinationTerm(0, 0, 0)
    Vrf = adim.get_component_rescaling_factor(V)
    bh = prob.boundary_handler
    for bid in (0, 1):
        bh.replace_boundary_condition(bid, DIRICHLET, V, lambda x: -e * x[..., 0] * Vrf)
        bh.replace_boundary_condition(bid, DIRICHLET, n, lambda x: n0 * nrf + 0 * x[..., 0])
        bh.replace_boundary_condition(bid, DIRICHLET, p, lambda x: p0 * nrf + 0 * x[..., 0])
    return prob, params, adim


@pytest.mark.parametrize("dim,ncells,k", [(2, 3, 1), (2, 3, 2), (3, 2, 1)])
@pytest.mark.parametrize("flux", FLUX)
@pytest.mark.parametrize("tau_kind", ["fixed", "cellface"])
@pytest.mark.parametrize("phi_lin", [False, True])
def test_exact_drift_state_has_zero_residual(dim, ncells, k, flux, tau_kind, phi_lin):
    """V linear and n, p constant with n - p = doping and no recombination
    solve the drift-diffusion system exactly, with a nonzero drift current
    n mu_n E; every HDG variant reproduces such a solution, so the assembled
    residual vanishes and one Newton step leaves the state"""
    e, n0, p0 = 0.5, 0.8, 0.3
    kw = dict(dd_flux_type=flux, phi_linearize=phi_lin)
    prob, params, adim = _drift_problem(dim, ncells, k, e, n0, p0, **kw)
    if tau_kind == "cellface":
        tab = _random_tau(prob, adim, seed=dim * 10 + k)
        prob, params, adim = _drift_problem(dim, ncells, k, e, n0, p0, cell_face_tau=tab, **kw)
    o = H.oracle_for(prob, params, adim)
    cell, tr = _drift_state(prob, adim, k, e, n0, p0)
    o.set_state(cell, tr)
    o.pass1()
    A, b, di, dv = o.system()
    assert np.abs(b).max() < 1e-12, np.abs(b).max()
    assert np.abs(dv).max() < 1e-12
    res = o.run_newton(max_it=1)
    c2, t2 = o.get_state()
    assert np.abs(c2 - cell).max() < 1e-11 and np.abs(t2 - tr).max() < 1e-11


# --------------------------------------------------- closed face residuals
def _jump_case(flux, comp, e, delta, lf, tau_tab=None):
    """2D 3x3 k=2, zero doping, no recombination; state zero except the E
    flux (= e along x, all cells) and the trace of component comp on local
    face lf of the middle cell (= delta).  Returns (sum of that face's trace
    rows of the residual t, sum of the cell's comp scalar rows of c, tau_r,
    mu_r, E.nu, face measure)."""
    kw = dict(dd_flux_type=flux, doping=_zero_doping)
    prob, params, adim = configs.diode(2, 3, 2, **kw)
    if tau_tab is not None:
        prob, params, adim = configs.diode(2, 3, 2, cell_face_tau=tau_tab, **kw)
    prob.recombination_term = LinearRecombinationTerm(0, 0, 0)
    o = H.oracle_for(prob, params, adim)
    mesh = prob.mesh
    ref = fe.RefCell(mesh, 2)
    d, nb, nf = 2, ref.nb, ref.nf
    cell = np.zeros((mesh.ncells, 3, d + 1, nb))
    trace = np.zeros((mesh.nfaces, 3, nf))
    cell[:, V, 0, :] = e
    mid = 4
    trace[mesh.cell_faces[mid, lf], comp, :] = delta
    o.set_state(cell.ravel(), trace.ravel())
    A, CT, TC, TT, c, t = o.local_system(mid)
    tdf = 3 * nf
    tsum = t[lf * tdf + comp * nf: lf * tdf + (comp + 1) * nf].sum()
    base = comp * (d + 1) * nb + d * nb
    csum = c[base: base + nb].sum()
    if tau_tab is not None:
        tau_r = tau_tab[mid, lf, comp] / adim.get_tau_rescaling_factor(comp)
    else:
        tau_r = (params.V_tau, params.n_tau, params.p_tau)[comp] / adim.get_tau_rescaling_factor(comp)
    mu = (None, prob.n_mobility.mu, prob.p_mobility.mu)[comp]
    mu_r = None if mu is None else mu / adim.get_mobility_rescaling_factor()
    a, side = divmod(lf, 2)
    Enu = e * (1.0 if side else -1.0) if a == 0 else 0.0
    area = mesh.h[1 - a]
    return tsum, csum, tau_r, mu_r, Enu, area


@pytest.mark.parametrize("flux", FLUX)
@pytest.mark.parametrize("lf", [0, 1, 2])
@pytest.mark.parametrize("comp", [n, p])
def test_face_residual_closed_form(flux, lf, comp):
    """single trace jump delta of a carrier on one face, constant field e:
    the face's trace residual and the cell's scalar residual both sum to
    (-/+ c^ mu E.nu - tau delta)|F| with c^ the flux's drift value: the cell
    value 0 (use_cell), the trace value delta (use_trace), or the cell value
    where E.nu > 0 and the trace value elsewhere (qiu_shi_stabilization)"""
    e, delta = 0.7, 0.3
    tsum, csum, tau_r, mu_r, Enu, area = _jump_case(flux, comp, e, delta, lf)
    if flux == "use_cell":
        chat = 0.0
    elif flux == "use_trace":
        chat = delta
    else:
        chat = 0.0 if Enu > 0 else delta
    sgn = -1.0 if comp == n else 1.0
    expected = (sgn * chat * mu_r * Enu - tau_r * delta) * area
    assert tsum == pytest.approx(expected, rel=1e-12, abs=1e-14)
    assert csum == pytest.approx(expected, rel=1e-12, abs=1e-14)


@pytest.mark.parametrize("comp", [V, n, p])
@pytest.mark.parametrize("lf", [0, 1, 2, 3])
def test_cell_face_tau_indexing(comp, lf):
    """with per-(cell, face, component) tau the jump residual picks the
    value of the middle cell, that face and that component"""
    prob, params, adim = configs.diode(2, 3, 2)
    tab = _random_tau(prob, adim, seed=99)
    e = 0.0 if comp == V else 0.4
    tsum, csum, tau_r, mu_r, Enu, area = _jump_case("use_cell", comp, e, 0.25, lf, tab)
    expected = (tau_r if comp == V else -tau_r) * 0.25 * area
    assert tau_r == pytest.approx(tab[4, lf, comp] / adim.get_tau_rescaling_factor(comp))
    assert tsum == pytest.approx(expected, rel=1e-12, abs=1e-14)
    assert csum == pytest.approx(expected, rel=1e-12, abs=1e-14)


# ------------------------------------------------ reference CellFace test
@pytest.mark.parametrize("dim,ncells,k", [(2, 4, 1), (3, 2, 1)])
def test_cell_face_tau_equals_fixed_tau(dim, ncells, k):
    """reference tests/cell_face_tau_computer.cpp:144-209: the same tau set
    on every (cell, face) as CellFaceTau gives the FixedTau results (there to
    1e-9 after full solves; here the condensed systems are identical)"""
    prob, params, adim = configs.diode(dim, ncells, k)
    taus = np.array([params.V_tau, params.n_tau, params.p_tau]) * np.array([1.0, 100.0, 1e4])
    prob, params, adim = configs.diode(dim, ncells, k, tau=tuple(taus))
    tab = np.broadcast_to(taus, (prob.mesh.ncells, 2 * dim, 3))
    probc, paramsc, _ = configs.diode(dim, ncells, k, cell_face_tau=tab)
    o1 = H.oracle_for(prob, params, adim)
    o2 = H.oracle_for(probc, paramsc, adim)
    state = H.perturbed(*H.equilibrium_guess(prob, adim, k), seed=2)
    o1.set_state(*state)
    o2.set_state(*state)
    o1.pass1()
    o2.pass1()
    A1, b1, _, _ = o1.system()
    A2, b2, _, _ = o2.system()
    assert np.array_equal(A1.data, A2.data) and np.array_equal(b1, b2)


# ------------------------------------------------------- Jacobian checks
@pytest.mark.parametrize("flux", ["use_trace", "qiu_shi_stabilization"])
@pytest.mark.parametrize("tau_kind", ["fixed", "cellface"])
def test_flux_variant_jacobian_consistent_with_residual(flux, tau_kind):
    """the local matrix of the use_trace / qiu_shi operators is the
    derivative of their local residual (consistent volume drift, R = 0),
    central differences on an n-side and a p-side cell.  A uniform field
    (0.3, 0.2) is added to E so that no face has E.nu near 0, where the
    qiu_shi switch makes the residual non-differentiable."""
    prob, params, adim = configs.diode(2, 4, 1, dd_flux_type=flux)
    if tau_kind == "cellface":
        prob, params, adim = configs.diode(2, 4, 1, dd_flux_type=flux,
                                           cell_face_tau=_random_tau(prob, adim, 5))
    prob.recombination_term = LinearRecombinationTerm(0, 0, 0)
    o = H.oracle_for(prob, params, adim)
    c0, t0 = H.perturbed(*H.equilibrium_guess(prob, adim, 1), seed=4)
    cs = c0.reshape(prob.mesh.ncells, 3, 3, 4)
    cs[:, V, 0] += 0.3
    cs[:, V, 1] += 0.2
    NC, h = o.NC, 1e-6
    for cell in (5, 6):
        o.set_state(c0, t0)
        A, CT, TC, TT, cc, t = o.local_system(cell)
        _, _, dofs, _, _ = o.condensed_cell(cell)
        rng = np.random.default_rng(cell)
        vu, vl = rng.standard_normal(A.shape[0]), rng.standard_normal(TT.shape[0])

        def resid(s):
            ce, te = c0.copy(), t0.copy()
            ce[cell * NC:(cell + 1) * NC] += s * vu
            te[dofs] += s * vl
            o.set_state(ce, te)
            out = o.local_system(cell)
            return out[4], out[5]
        (cp, tp), (cm, tm) = resid(h), resid(-h)
        dc, dt = (cp - cm) / (2 * h), (tp - tm) / (2 * h)
        assert H.rel_fro(dc, -(A @ vu + CT @ vl)) < 1e-8, cell
        assert H.rel_fro(dt, -(-TC @ vu + TT @ vl)) < 1e-8, cell


# ------------------------------------------------------- phi_linearize
def test_phi_linearize_residual_and_v_rows():
    """phi_linearize: the Boltzmann densities at the quasi-Fermi potentials
    of n and p are n and p again, so the residual equals the plain one; the
    V scalar rows lose their n and p columns and gain the derivative at
    frozen potentials, (n + p)/U_T times the mass matrix"""
    k = 2
    prob, params, adim = configs.diode(2, 3, k)
    probp, paramsp, _ = configs.diode(2, 3, k, phi_linearize=True)
    o1 = H.oracle_for(prob, params, adim)
    o2 = H.oracle_for(probp, paramsp, adim)
    mesh = prob.mesh
    ref = fe.RefCell(mesh, k)
    d, nb = 2, ref.nb
    # multiplicative perturbation only: the quasi-Fermi potentials need
    # positive densities at the quadrature points (the reference takes
    # their logarithm, include/solver.h:343-375)
    cell, tr = H.equilibrium_guess(prob, adim, k)
    rng = np.random.default_rng(6)
    cell = cell * (1 + 1e-2 * rng.standard_normal(cell.shape))
    tr = tr * (1 + 1e-2 * rng.standard_normal(tr.shape))
    o1.set_state(cell, tr)
    o2.set_state(cell, tr)
    kT = KB * float(prob.temperature) / (Q * adim.get_component_rescaling_factor(V))
    cs = cell.reshape(mesh.ncells, 3, d + 1, nb)
    for c in (0, 3, 2, 8):    # away from the junction at x = 0.5
        assert (ref.phic @ cs[c, n, d]).min() > 0 and (ref.phic @ cs[c, p, d]).min() > 0
        A1, CT1, TC1, TT1, c1, t1 = o1.local_system(c)
        A2, CT2, TC2, TT2, c2, t2 = o2.local_system(c)
        assert H.rel_fro(c2, c1) < 1e-12 and H.rel_fro(t2, t1) < 1e-12
        assert np.array_equal(CT2, CT1) and np.array_equal(TC2, TC1) and np.array_equal(TT2, TT1)
        vs = V * (d + 1) * nb + d * nb
        rows = slice(vs, vs + nb)
        nq = ref.phic @ cs[c, n, d]
        pq = ref.phic @ cs[c, p, d]
        Mw = np.einsum("q,qi,qj->ij", ref.JxWc * (nq + pq) / kT, ref.phic, ref.phic)
        dA = A2 - A1
        assert H.rel_fro(dA[rows, vs:vs + nb], Mw) < 1e-12
        for comp, sgn in ((n, 1.0), (p, -1.0)):
            cs_ = comp * (d + 1) * nb + d * nb
            assert np.abs(A2[rows, cs_:cs_ + nb]).max() == 0.0
            assert H.rel_fro(-dA[rows, cs_:cs_ + nb], sgn * ref.M) < 1e-12
        other = np.ones(A1.shape[0], dtype=bool)
        other[rows] = False
        assert np.array_equal(A2[other], A1[other])


# ------------------------------------------------- reconstruct quirk (Q5)
@pytest.mark.parametrize("flux", ["use_trace", "qiu_shi_stabilization"])
def test_reconstruct_omits_upwind_trace_term(flux):
    """the reconstruct pass of the upwinded fluxes adds only the use_cell
    part of CT lambda to the cell rhs (template.h:2122-2160 vs :983-1009):
    on interior cells u = A^-1 (c - CT_use_cell lambda), with A, c of the
    variant and CT of use_cell at the same state"""
    prob, params, adim = configs.diode(2, 4, 1, dd_flux_type=flux)
    prob0, params0, _ = configs.diode(2, 4, 1)
    o = H.oracle_for(prob, params, adim)
    o0 = H.oracle_for(prob0, params0, adim)
    state = H.perturbed(*H.equilibrium_guess(prob, adim, 1), seed=9)
    o.set_state(*state)
    o0.set_state(*state)
    rng = np.random.default_rng(1)
    upd = 1e-2 * rng.standard_normal(o.nfaces * 3 * o.nf)
    u = o.pass2(upd).reshape(o.ncells, o.NC)
    r2f = o.restricted_to_full((1, 1, 1))
    worst, moved = 0.0, 0.0
    for c in (5, 6, 9, 10):   # no face on the x = 0 or x = 1 contacts
        A, CT, TC, TT, cc, t = o.local_system(c)
        _, CT0, _, _, _, _ = o0.local_system(c)
        _, _, dofs, _, _ = o.condensed_cell(c)
        lam = upd[r2f[dofs]]
        ref_u = np.linalg.solve(A, cc - CT0 @ lam)
        worst = max(worst, H.rel_fro(u[c], ref_u))
        moved = max(moved, H.rel_fro(np.linalg.solve(A, cc - CT @ lam), ref_u))
    assert worst < 1e-12
    assert moved > 1e-8     # the upwind term is not zero on this state
