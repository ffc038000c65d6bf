"""Batched Newton (BatchedNewton: the iteration of many devices in one
launch sequence, replayed from one hipGraph) against each device solved
alone: the reference's diode workflow (equilibrium, n/p, coupled, one bias
step) on devices of uneven sizes and degrees 1-3, in FP64 and with FP32
elimination + FP64 refinement, gives every device the same iteration count
and a bitwise equal state; a device that stops early (its own iteration
limit) leaves the others unchanged."""
import numpy as np
import pytest

from ddhdg_amd import configs

pytestmark = pytest.mark.gpu

# (side, degree, N_D, N_A, junction, bias)
DEVICES = [(12, 1, 1e23, 1e23, 0.5, 0.05), (12, 2, 3e22, 3e23, 0.4, 0.1),
           (7, 3, 1e21, 5e22, 0.6, 0.05), (16, 1, 2e23, 1e22, 0.35, 0.15),
           (6, 3, 1e24, 1e23, 0.5, 0.1), (10, 2, 5e21, 5e21, 0.65, 0.05)]


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU test on a machine without a GPU"
    return torch


def _solvers(precision=0, max_it=None):
    from ddhdg_amd.np_solver import NPSolver
    out = []
    for side, k, nd, na, j, _ in DEVICES:
        prob, params, adim = configs.diode(2, side, k, doping=configs.step_doping(nd, na, j))
        if max_it is not None:
            params.nonlinear_parameters.max_number_of_iterations = max_it
        s = NPSolver(prob, params, adim)
        s.ctx.set_precision(precision)
        out.append(s)
    return out


def _same(a, b):
    ca, ta = a.get_state()
    cb, tb = b.get_state()
    return np.array_equal(ca, cb) and np.array_equal(ta, tb)


@pytest.mark.parametrize("precision", [0, 1])
def test_batched_workflow_equals_alone(precision):
    _torch()
    print(f"case batched Newton workflow, {len(DEVICES)} devices, precision {precision}",
          flush=True)
    from ddhdg_amd.batch import batched_equilibrium, batched_run
    sb, sa = _solvers(precision), _solvers(precision)
    stages = [("equilibrium", None), ("n,p", (False, True, True)),
              ("coupled", (True, True, True)), ("bias", (True, True, True))]
    for name, en in stages:
        if name == "equilibrium":
            rb = batched_equilibrium(sb)
            ra = [s.compute_thermodynamic_equilibrium() for s in sa]
        else:
            for (side, k, nd, na, j, bias), x, y in zip(DEVICES, sb, sa):
                x.set_enabled_components(*en)
                y.set_enabled_components(*en)
                if name == "bias":
                    configs.set_bias(x, bias)
                    configs.set_bias(y, bias)
            rb = batched_run(sb)
            ra = [s.run() for s in sa]
        print(f"  {name}: batched {[r.iterations for r in rb]} alone {[r.iterations for r in ra]}",
              flush=True)
        for i, (x, y) in enumerate(zip(rb, ra)):
            assert x.converged and y.converged, (name, i)
            assert x.iterations == y.iterations, (name, i)
            assert x.last_update_norm == y.last_update_norm, (name, i)
        for i, (x, y) in enumerate(zip(sb, sa)):
            assert _same(x, y), (name, i)


def test_batched_device_limit_leaves_others_unchanged():
    """one device with an iteration limit of 1 stops after one step (not
    converged) and keeps that state; the others finish as alone"""
    _torch()
    print("case batched Newton with one device stopped at its limit", flush=True)
    from ddhdg_amd.batch import BatchedNewton, batched_equilibrium
    sb, sa = _solvers(), _solvers()
    batched_equilibrium(sb)
    for s in sa:
        s.compute_thermodynamic_equilibrium()
    for (side, k, nd, na, j, bias), x, y in zip(DEVICES, sb, sa):
        configs.set_bias(x, bias)
        configs.set_bias(y, bias)
    sb[2].parameters.nonlinear_parameters.max_number_of_iterations = 1
    sa[2].parameters.nonlinear_parameters.max_number_of_iterations = 1
    bn = BatchedNewton(sb)
    rb = bn.run()
    bn.close()
    ra = [s.run() for s in sa]
    print(f"  batched {[(r.converged, r.iterations) for r in rb]}", flush=True)
    assert not rb[2].converged and rb[2].iterations == 1
    for i, (x, y) in enumerate(zip(rb, ra)):
        assert (x.converged, x.iterations) == (y.converged, y.iterations), i
    for i, (x, y) in enumerate(zip(sb, sa)):
        assert _same(x, y), i


def test_batched_graph_equals_eager():
    """the captured iteration replays bitwise like the eager one"""
    _torch()
    print("case batched Newton graph vs eager", flush=True)
    from ddhdg_amd.batch import batched_equilibrium
    sg, se = _solvers(), _solvers()
    rg = batched_equilibrium(sg, graph=True)
    re_ = batched_equilibrium(se, graph=False)
    assert [r.iterations for r in rg] == [r.iterations for r in re_]
    for x, y in zip(sg, se):
        assert _same(x, y)
