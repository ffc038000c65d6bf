"""The continuous-batching I-V server (GPU tests): each request's results are
the same alone and in a mix of concurrent requests (bitwise: requests share
no buffers and every kernel is deterministic), admission honours the memory
budget, and the currents are those of a plain sequential solve."""
import numpy as np
import pytest

from ddhdg_amd import configs
from ddhdg_amd.serving import IVRequest, IVServer, memory_estimate

pytestmark = pytest.mark.gpu

REQS = [IVRequest(0, 2, (12, 12), 1, (0.1, 0.2), arrival=0.0),
        IVRequest(1, 2, (8, 16), 2, (0.1, 0.2, 0.3), nd=3e22, na=2e23, junction=0.4, arrival=0.0),
        IVRequest(2, 2, (16, 16), 2, (0.15,), arrival=0.05),
        IVRequest(3, 2, (10, 6), 1, (0.1, 0.3), nd=1e22, na=1e22, junction=0.6, arrival=0.1)]


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU test on a machine without a GPU"
    return torch


def _same(a, b):
    assert a.converged and b.converged
    assert a.iterations == b.iterations
    assert np.array_equal(np.asarray(a.currents), np.asarray(b.currents))
    assert all(np.array_equal(x, y) for x, y in zip(a.state, b.state))


def test_results_same_alone_and_in_mix():
    _torch()
    print("case serving mix of 4 requests vs each alone", flush=True)
    mix, wall = IVServer(max_slots=4, keep_state=True).run(REQS)
    for r in mix:
        print(f"  request {r.rid}: iterations {r.iterations} currents {r.currents} "
              f"latency {r.latency:.3f} s")
    for req in REQS:
        alone, _ = IVServer(max_slots=1, keep_state=True).run(
            [IVRequest(**{**req.__dict__, "arrival": 0.0})])
        _same(mix[req.rid], alone[0])


def test_admission_honours_budget():
    _torch()
    print("case serving budget of one request at a time", flush=True)
    budget = max(memory_estimate(r) for r in REQS[:2])
    srv = IVServer(hbm_budget_bytes=budget, max_slots=4)
    res, _ = srv.run(REQS[:2])
    assert srv.max_active == 1 and all(r.converged for r in res)
    srv = IVServer(hbm_budget_bytes=1 << 40, max_slots=4)
    srv.run(REQS[:2])
    assert srv.max_active == 2


def test_served_currents_match_sequential_solve():
    """a request's currents equal those of the plain NPSolver workflow"""
    _torch()
    print("case serving vs sequential NPSolver", flush=True)
    from ddhdg_amd.np_solver import NPSolver
    req = REQS[1]
    res, _ = IVServer(max_slots=1).run([req])
    prob, params, adim = configs.diode(req.dim, req.n_cells, req.degree,
                                       doping=configs.step_doping(req.nd, req.na, req.junction))
    s = NPSolver(prob, params, adim)
    s.set_enabled_components(True, False, False)
    s.compute_local_charge_neutrality()
    configs.solve_equilibrium_and_bias(s)
    cur = []
    for b in req.biases:
        configs.set_bias(s, b)
        assert s.run().converged
        cur.append(s.terminal_current(1))
    assert np.array_equal(np.asarray(cur), np.asarray(res[0].currents))


def test_3d_requests_same_alone_and_in_mix():
    """3D requests (the register kernel at degree 1, the generic kernel at
    degree 2) served beside 2D ones: each result bitwise the same as alone"""
    _torch()
    print("case serving mix with 3D requests vs each alone", flush=True)
    reqs = [REQS[0], REQS[2],
            IVRequest(4, 3, (3, 3, 3), 1, (0.1, 0.2), arrival=0.0),
            IVRequest(5, 3, (2, 2, 2), 2, (0.1,), nd=3e22, na=2e23, junction=0.4, arrival=0.02)]
    mix, _ = IVServer(max_slots=4, keep_state=True).run(reqs)
    by_id = {r.rid: r for r in mix}
    for req in reqs:
        r = by_id[req.rid]
        print(f"  request {r.rid}: iterations {r.iterations} currents {r.currents}")
        alone, _ = IVServer(max_slots=1, keep_state=True).run(
            [IVRequest(**{**req.__dict__, "arrival": 0.0})])
        _same(r, alone[0])


def test_failed_requests_are_evicted_and_isolated():
    """one request stopped by an iteration limit of 1, one whose set-up
    raises (degree 4 is not supported), one larger than the whole budget:
    each is reported failed with its reason, and the others finish bitwise
    the same as alone"""
    _torch()
    print("case serving failure isolation", flush=True)
    bad = [IVRequest(10, 2, (12, 12), 1, (0.1,), arrival=0.0, max_it=1),
           IVRequest(11, 2, (8, 8), 4, (0.1,), arrival=0.01),
           IVRequest(12, 2, (4096, 4096), 2, (0.1,), arrival=0.02)]
    good = [REQS[0], REQS[1]]
    budget = 8 << 30
    assert memory_estimate(bad[2]) > budget
    mix, _ = IVServer(hbm_budget_bytes=budget, max_slots=4, keep_state=True).run(good + bad)
    by_id = {r.rid: r for r in mix}
    for rid in (10, 11, 12):
        r = by_id[rid]
        print(f"  request {rid}: failed with {r.error!r}")
        assert not r.converged and r.error
    assert "did not converge" in by_id[10].error
    assert "not supported" in by_id[11].error
    assert "budget" in by_id[12].error
    for req in good:
        alone, _ = IVServer(max_slots=1, keep_state=True).run(
            [IVRequest(**{**req.__dict__, "arrival": 0.0})])
        _same(by_id[req.rid], alone[0])
