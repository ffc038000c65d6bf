"""A C++ host program (tests/cpp/hdg_newton_step.cpp) drives one Newton step
through the C ABI of libddhdg_hip.so with hipMalloc buffers, the call
sequence of a C++ NPSolver binding; its state after the step equals the
Python binding's: bitwise when both upload the same problem data, and to
rounding when the C++ host evaluates the doping and contact data itself
(its libm exp may differ from numpy's in the last bit)."""
import os
import struct
import subprocess

import numpy as np
import pytest

from ddhdg_amd import configs
from tests import helpers as H

pytestmark = pytest.mark.gpu

PROG = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "hdg_newton_step")


@pytest.mark.parametrize("dim,n_cells,k", [(2, (12, 10), 2), (2, (9, 9), 1), (3, (4, 3, 3), 1)])
@pytest.mark.parametrize("from_file", [1, 0])
def test_cpp_host_newton_step_matches_python(dim, n_cells, k, from_file, tmp_path):
    import torch
    assert torch.cuda.is_available()
    assert os.path.exists(PROG), "build the C++ test program (tests/cpp/Makefile) first"
    print(f"case C++ host newton step dim={dim} n={n_cells} k={k} from_file={from_file}",
          flush=True)
    from ddhdg_amd.np_solver import NPSolver
    nd, na, junction, bias = 3e22, 2e23, 0.45, 0.25
    prob, params, adim = configs.diode(dim, n_cells, k, bias=bias,
                                       doping=configs.step_doping(nd, na, junction))
    s = NPSolver(prob, params, adim)
    s.compute_local_charge_neutrality()
    s.project_densities_from_potential()
    cell, tr = H.perturbed(*s.get_state(), seed=7, rel=1e-3)
    s.set_state(cell, tr)
    n3 = list(n_cells) + [1] * (3 - len(n_cells))
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(struct.pack("<6i", dim, k, *n3, from_file))
        f.write(struct.pack("<4d", bias, nd, na, junction))
        f.write(struct.pack("<2q", cell.size, tr.size))
        f.write(np.ascontiguousarray(cell, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(tr, dtype="<f8").tobytes())
        if from_file:
            dirv, neuv = s.bc_values()
            for a in (s.doping_values(), dirv, neuv):
                f.write(np.ascontiguousarray(a, dtype="<f8").ravel().tobytes())
    r = subprocess.run([PROG, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(r.stdout.strip(), r.stderr.strip(), flush=True)
    assert r.returncode == 0
    out = np.fromfile(fout, dtype="<f8")
    cc, tc = out[:cell.size], out[cell.size:cell.size + tr.size]
    res = s.run(max_number_of_iterations=1)
    assert res.iterations == 1
    cg, tg = s.get_state()
    if from_file:
        assert np.array_equal(cc, cg) and np.array_equal(tc, tg)
    else:
        err = max(np.abs(cc - cg).max() / np.abs(cg).max(), np.abs(tc - tg).max() / np.abs(tg).max())
        print(f"  C++-evaluated data: state difference {err:.2e}", flush=True)
        assert err < 1e-13
