"""hipGraph-captured Newton iterations (GPU tests): the diode workflow with
every iteration after the first replayed from a captured graph takes the same
iterations and reaches bitwise the same states as eager execution, with eager
work and synchronisations between the replays."""
import numpy as np
import pytest

from ddhdg_amd import configs

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU test on a machine without a GPU"
    return torch


def _workflow(n, k, graph):
    from ddhdg_amd.np_solver import NPSolver
    prob, params, adim = configs.diode(2, n, k)
    params.graph_newton = graph
    s = NPSolver(prob, params, adim)
    s.set_enabled_components(True, False, False)
    s.compute_local_charge_neutrality()
    res = configs.solve_equilibrium_and_bias(s, (0.1, 0.3))
    return [(r.converged, r.iterations) for r in res], s.get_state(), s.terminal_current(1)


@pytest.mark.parametrize("n,k", [(8, 2), (12, 1), (48, 2)])
def test_graph_newton_equals_eager(n, k):
    _torch()
    print(f"case graph Newton 2D n={n} k={k}", flush=True)
    its_e, st_e, I_e = _workflow(n, k, False)
    its_g, st_g, I_g = _workflow(n, k, True)
    print(f"  eager {its_e} graph {its_g} current {I_e} {I_g}")
    assert its_e == its_g and all(c for c, _ in its_e)
    assert all(np.array_equal(a, b) for a, b in zip(st_e, st_g))
    assert I_e == I_g
