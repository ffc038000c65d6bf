"""VTK output of the cell and trace fields, for visual checks.

The counterpart of NPSolver::output_results (reference src/np_solver.cpp:
1737-2033): the cell solution (electric field, potential, electron and hole
displacements and densities, with the reference's field names), optionally
the last Newton update, and the doping, written as a VTK XML unstructured
grid (.vtu).  Each cell's scalar basis is Lagrange on (k+1)^d equidistant
nodes, so the nodal values are the dofs; a cell is written as k^d linear
sub-cells over its own nodes (the field is discontinuous across cells).
The trace output writes the trace dofs at their Gauss-Lobatto face nodes
as a point cloud.  Quantities are redimensionalised as the reference's
Adimensionalizer::redimensionalize_dof_vector does (src/adimensionalizer.cpp:
367-420): potentials times the potential scale, densities times the doping
magnitude, displacements also divided by the scale length.
"""
from __future__ import annotations

import base64

import numpy as np

from . import fe
from .problem import V, evaluate, n, p

# include/components.h: component and displacement names
COMPONENT_NAMES = ("electric_potential", "electron_density", "hole_density")
DISPLACEMENT_NAMES = ("electric_field", "electron_displacement", "hole_displacement")

_VTK_QUAD, _VTK_HEX, _VTK_VERTEX = 9, 12, 1


def _b64(a, dtype):
    """binary inline data: a UInt32 byte count, then the array"""
    raw = np.ascontiguousarray(a, dtype=dtype).tobytes()
    return base64.b64encode(np.uint32(len(raw)).tobytes() + raw).decode()


def _data_array(name, arr, ncomp=1, dtype="<f8", vtype="Float64"):
    return (f'<DataArray type="{vtype}" Name="{name}" NumberOfComponents="{ncomp}" '
            f'format="binary">{_b64(arr, dtype)}</DataArray>')


def _write(filename, points, conn, offsets, types, point_data):
    npts, ncell = len(points), len(offsets)
    parts = ['<?xml version="1.0"?>',
             '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian" '
             'header_type="UInt32">',
             '<UnstructuredGrid>',
             f'<Piece NumberOfPoints="{npts}" NumberOfCells="{ncell}">',
             '<PointData>']
    for name, (arr, ncomp) in point_data.items():
        parts.append(_data_array(name, arr, ncomp))
    parts += ['</PointData>', '<Points>', _data_array("Points", points, 3), '</Points>',
              '<Cells>', _data_array("connectivity", conn, 1, "<i8", "Int64"),
              _data_array("offsets", offsets, 1, "<i8", "Int64"),
              _data_array("types", types, 1, "<u1", "UInt8"), '</Cells>',
              '</Piece>', '</UnstructuredGrid>', '</VTKFile>']
    with open(filename, "w") as f:
        f.write("\n".join(parts) + "\n")


def cell_nodes(mesh, degree):
    """coordinates (ncells, nb, 3) of the equidistant scalar nodes of every
    cell, in the cell basis order (first axis fastest)"""
    d, k = mesh.dim, degree
    nb = (k + 1) ** d
    idx = np.array(np.unravel_index(np.arange(nb), [k + 1] * d, order="F")).T
    x = np.zeros((mesh.ncells, nb, 3))
    for a in range(d):
        x[..., a] = mesh.lower[a] + (mesh.ic[:, a][:, None] + idx[None, :, a] / k) * mesh.h[a]
    return x


def _subcells(d, k):
    """local node indices of the k^d linear sub-cells of a cell, in VTK
    quad / hexahedron vertex order"""
    def node(i, j, l=0):
        return i + (k + 1) * (j + (k + 1) * l)
    out = []
    if d == 2:
        for j in range(k):
            for i in range(k):
                out.append([node(i, j), node(i + 1, j), node(i + 1, j + 1), node(i, j + 1)])
    else:
        for l in range(k):
            for j in range(k):
                for i in range(k):
                    out.append([node(i, j, l), node(i + 1, j, l), node(i + 1, j + 1, l),
                                node(i, j + 1, l), node(i, j, l + 1), node(i + 1, j, l + 1),
                                node(i + 1, j + 1, l + 1), node(i, j + 1, l + 1)])
    return np.array(out)


def write_cells(filename, mesh, degree, cell_state, adim, update=None,
                redimensionalize=True, doping=None):
    """the cell fields of a state vector (dimensionless, the library's cell
    layout) as a .vtu file; `doping` is a function of the points (physical)"""
    d, k = mesh.dim, degree
    nb = (k + 1) ** d
    pts = cell_nodes(mesh, k)
    sub = _subcells(d, k)
    conn = (np.arange(mesh.ncells)[:, None, None] * nb + sub[None]).reshape(-1)
    nv = sub.shape[1]
    nsub = mesh.ncells * len(sub)
    offsets = nv * np.arange(1, nsub + 1)
    types = np.full(nsub, _VTK_QUAD if d == 2 else _VTK_HEX)
    data = {}
    for tag, vec in (("", cell_state), ("_updates", update)):
        if vec is None:
            continue
        c = np.asarray(vec, dtype=float).reshape(mesh.ncells, 3, d + 1, nb)
        for comp in (V, n, p):
            rf = adim.get_component_rescaling_factor(comp) if redimensionalize else 1.0
            drf = rf / adim.scale_length if redimensionalize else 1.0
            disp = np.zeros((mesh.ncells, nb, 3))
            disp[..., :d] = c[:, comp, :d, :].transpose(0, 2, 1) * drf
            data[DISPLACEMENT_NAMES[comp] + tag] = (disp.reshape(-1, 3), 3)
            data[COMPONENT_NAMES[comp] + tag] = ((c[:, comp, d, :] * rf).reshape(-1), 1)
    if doping is not None:
        scale = 1.0 if redimensionalize else 1.0 / adim.doping_magnitude
        data["doping"] = (evaluate(doping, pts[..., :d]).reshape(-1) * scale, 1)
    _write(filename, pts.reshape(-1, 3), conn, offsets, types, data)


def write_trace(filename, mesh, degree, trace_state, adim, redimensionalize=True):
    """the trace dofs of a state vector (full trace layout) at their face
    nodes (Gauss-Lobatto points of each face) as a .vtu point cloud"""
    d, k = mesh.dim, degree
    nf = (k + 1) ** (d - 1)
    xi = fe.lobatto01(k + 1)
    pts = np.zeros((mesh.nfaces, nf, 3))
    fidx = np.array(np.unravel_index(np.arange(nf), [k + 1] * (d - 1), order="F")).T
    for a in range(d):
        m = mesh.face_dims[a]
        cnt = mesh.face_off[a + 1] - mesh.face_off[a]
        fi = np.arange(cnt)
        pos = np.stack([fi % m[0], (fi // m[0]) % m[1], fi // (m[0] * m[1])], axis=1)
        sl = slice(mesh.face_off[a], mesh.face_off[a + 1])
        for b in range(d):
            pts[sl, :, b] = (mesh.lower[b] + pos[:, b] * mesh.h[b])[:, None]
        for mm in range(d - 1):
            b = fe.TANGENTS[d][a][mm]
            pts[sl, :, b] += xi[fidx[:, mm]][None, :] * mesh.h[b]
    t = np.asarray(trace_state, dtype=float).reshape(mesh.nfaces, 3, nf)
    data = {}
    for comp in (V, n, p):
        rf = adim.get_component_rescaling_factor(comp) if redimensionalize else 1.0
        data[COMPONENT_NAMES[comp]] = ((t[:, comp, :] * rf).reshape(-1), 1)
    npts = mesh.nfaces * nf
    _write(filename, pts.reshape(-1, 3), np.arange(npts), np.arange(1, npts + 1),
           np.full(npts, _VTK_VERTEX), data)
