"""Batched pass 1 (GPU tests): several meshes of uneven sizes and degrees in
one persistent launch per colour give, for every mesh, exactly the outputs of
its own hdg_condense (bitwise: same kernels, disjoint writes per colour)."""
import numpy as np
import pytest

from ddhdg_amd import configs
from tests import helpers as H

pytestmark = pytest.mark.gpu

MESHES = [((12, 10), 1), ((9, 9), 2), ((16, 5), 2), ((7, 13), 1)]
# degree 3 (the pair kernel): odd cell counts per colour pad the chunks
MESHES3 = [((7, 5), 3), ((4, 9), 3), ((1, 1), 3), ((10, 6), 3)]


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU test on a machine without a GPU"
    return torch


def _solvers(thermodyn, precision=0, meshes=MESHES):
    from ddhdg_amd.np_solver import NPSolver
    out = []
    for i, (n, k) in enumerate(meshes):
        prob, params, adim = configs.diode(2, n, k)
        s = NPSolver(prob, params, adim)
        if thermodyn:
            s.set_enabled_components(True, False, False)
        s.ctx.set_precision(precision)
        s.set_state(*H.perturbed(*H.equilibrium_guess(prob, adim, k), seed=40 + i))
        out.append(s)
    return out


@pytest.mark.parametrize("thermodyn", [False, True])
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_batch_equals_per_mesh(thermodyn, precision):
    t = _torch()
    print(f"case batch of {len(MESHES)} meshes thermo={thermodyn} precision={precision}",
          flush=True)
    from ddhdg_amd.hip_lib import HdgBatch
    sv = _solvers(thermodyn, precision)
    bufs = []
    for s in sv:
        c = s.ctx
        bufs.append((t.full((c.nnz,), 7.0, dtype=t.float64, device="cuda"),
                     t.full((c.ntrace,), 7.0, dtype=t.float64, device="cuda"),
                     t.zeros(c.ntrace, dtype=t.float64, device="cuda")))
    b = HdgBatch([s.ctx for s in sv])
    for _ in range(2):   # the first call also autotunes the grid
        b.condense(thermodyn, [s.x_cell for s in sv], [s.x_trace for s in sv],
                   [x[0] for x in bufs], [x[1] for x in bufs], [x[2] for x in bufs])
    t.cuda.synchronize()
    print(f"  {b.info()}")
    for s, (v, r, d) in zip(sv, bufs):
        c = s.ctx
        v1 = t.empty_like(v)
        r1 = t.empty_like(r)
        d1 = t.zeros_like(d)
        c.condense(s.x_cell, s.x_trace, thermodyn, v1, r1, d1)
        t.cuda.synchronize()
        assert t.equal(v, v1) and t.equal(r, r1) and t.equal(d, d1)


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_degree3_batch_equals_per_mesh(precision):
    """degree-3 meshes in one persistent launch per colour of the pair kernel
    (FP64 and FP32 + refinement): every mesh's outputs bitwise equal to its
    own hdg_condense"""
    t = _torch()
    print(f"case degree-3 batch of {len(MESHES3)} meshes precision={precision}", flush=True)
    from ddhdg_amd.hip_lib import HdgBatch
    sv = _solvers(False, precision, MESHES3)
    bufs = []
    for s in sv:
        c = s.ctx
        bufs.append((t.full((c.nnz,), 7.0, dtype=t.float64, device="cuda"),
                     t.full((c.ntrace,), 7.0, dtype=t.float64, device="cuda"),
                     t.zeros(c.ntrace, dtype=t.float64, device="cuda")))
    b = HdgBatch([s.ctx for s in sv])
    for _ in range(2):   # the first call also autotunes the grid
        b.condense(False, [s.x_cell for s in sv], [s.x_trace for s in sv],
                   [x[0] for x in bufs], [x[1] for x in bufs], [x[2] for x in bufs])
    t.cuda.synchronize()
    print(f"  {b.info()}")
    for s, (v, r, d) in zip(sv, bufs):
        c = s.ctx
        v1 = t.empty_like(v)
        r1 = t.empty_like(r)
        d1 = t.zeros_like(d)
        c.condense(s.x_cell, s.x_trace, False, v1, r1, d1)
        t.cuda.synchronize()
        assert t.equal(v, v1) and t.equal(r, r1) and t.equal(d, d1)


def test_batch_rejects_mixed_degree3():
    """a batch holds degree-1/2 meshes or degree-3 meshes, not both"""
    _torch()
    from ddhdg_amd.hip_lib import HdgBatch, HdgError
    sv = _solvers(False, 0, [((4, 4), 2), ((4, 4), 3)])
    with pytest.raises(HdgError, match="degree 3"):
        HdgBatch([s.ctx for s in sv])
