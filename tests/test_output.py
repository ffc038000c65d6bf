"""VTK output (CPU): the written .vtu holds the cell and trace fields at
their nodes with the reference's names and rescalings."""
import base64
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from ddhdg_amd import fe, output
from ddhdg_amd.problem import Adimensionalizer, V, n, p


def _read(path):
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    arrays = {}
    for da in root.iter("DataArray"):
        raw = base64.b64decode(da.text.strip())
        nbytes = int(np.frombuffer(raw[:4], dtype="<u4")[0])
        dt = {"Float64": "<f8", "Int64": "<i8", "UInt8": "<u1"}[da.get("type")]
        a = np.frombuffer(raw[4:4 + nbytes], dtype=dt)
        arrays[da.get("Name")] = a.reshape(-1, int(da.get("NumberOfComponents")))
    return int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells")), arrays


@pytest.mark.parametrize("dim,k", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_cell_and_trace_output(dim, k, tmp_path):
    mesh = fe.BoxMesh(dim, [3, 2, 2][:dim], [0.0] * dim, [1.0, 2.0, 1.5][:dim])
    adim = Adimensionalizer(1e-6, 11604.5, 1e23, 1.0)
    nb, nf = (k + 1) ** dim, (k + 1) ** (dim - 1)
    xs = output.cell_nodes(mesh, k)
    # a state whose dofs are functions of the node coordinates
    cell = np.zeros((mesh.ncells, 3, dim + 1, nb))
    cell[:, V, dim] = xs[..., 0] + 2 * xs[..., 1]
    cell[:, n, dim] = 1 + xs[..., 0] ** 2
    cell[:, p, dim] = 3.0
    for a in range(dim):
        cell[:, V, a] = -(1.0 if a == 0 else 2.0 if a == 1 else 0.0)
    trace = np.arange(mesh.nfaces * 3 * nf, dtype=float)
    fc, ft = tmp_path / "c.vtu", tmp_path / "t.vtu"
    output.write_cells(fc, mesh, k, cell.ravel(), adim, update=0.5 * cell.ravel(),
                       doping=lambda x: 1e22 * x[..., 0])
    npts, ncells, arr = _read(fc)
    assert npts == mesh.ncells * nb and ncells == mesh.ncells * k ** dim
    assert np.allclose(arr["Points"].reshape(mesh.ncells, nb, 3), xs)
    Vrf = adim.get_component_rescaling_factor(V)
    pot = arr["electric_potential"].reshape(mesh.ncells, nb)
    assert np.allclose(pot, (xs[..., 0] + 2 * xs[..., 1]) * Vrf)
    assert np.allclose(arr["electron_density"].reshape(mesh.ncells, nb), (1 + xs[..., 0] ** 2) * 1e23)
    assert np.allclose(arr["hole_density_updates"].reshape(mesh.ncells, nb), 1.5e23)
    E = arr["electric_field"].reshape(mesh.ncells, nb, 3)
    assert np.allclose(E[..., 0], -Vrf / 1e-6) and np.allclose(E[..., 2], 0.0 if dim == 2 else E[..., 2])
    assert np.allclose(arr["doping"].ravel(), 1e22 * xs[..., 0].ravel())
    # every sub-cell's vertices are nodes of one cell, corners of a box
    conn = arr["connectivity"].ravel().reshape(ncells, -1)
    assert np.all(conn // nb == conn[:, :1] // nb)
    output.write_trace(ft, mesh, k, trace, adim)
    npts, ncells, arr = _read(ft)
    assert npts == mesh.nfaces * nf
    assert np.allclose(arr["electron_density"].ravel(),
                       trace.reshape(mesh.nfaces, 3, nf)[:, n].ravel() * 1e23)


@pytest.mark.gpu
def test_solver_output_results(tmp_path):
    """NPSolver.output_results after an equilibrium solve: the written
    potential and the saved update are the solver's state and last update"""
    import torch
    assert torch.cuda.is_available()
    from ddhdg_amd import configs
    from ddhdg_amd.np_solver import NPSolver
    print("case output_results 2D 8x6 k=2", flush=True)
    prob, params, adim = configs.diode(2, (8, 6), 2)
    s = NPSolver(prob, params, adim)
    assert s.compute_thermodynamic_equilibrium().converged
    f, ft = tmp_path / "sol.vtu", tmp_path / "trace.vtu"
    s.output_results(str(f), str(ft), save_update=True)
    _, _, arr = _read(f)
    cell, tr = s.get_state()
    c = cell.reshape(s.ctx.ncells, 3, 3, 9)
    assert np.allclose(arr["electric_potential"].ravel(),
                       c[:, V, 2].ravel() * adim.get_component_rescaling_factor(V), rtol=1e-15)
    u = s.update_cell.cpu().numpy().reshape(s.ctx.ncells, 3, 3, 9)
    assert np.allclose(arr["electric_potential_updates"].ravel(),
                       u[:, V, 2].ravel() * adim.get_component_rescaling_factor(V), rtol=1e-15)
    _, _, tarr = _read(ft)
    assert np.allclose(tarr["hole_density"].ravel(),
                       tr.reshape(-1, 3, 3)[:, p].ravel() * adim.doping_magnitude, rtol=1e-15)
