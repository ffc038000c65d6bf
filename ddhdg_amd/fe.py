"""Reference-cell tables and L2 projections on the structured box mesh (host).

Set-up utilities around the hot path: first guesses and boundary data are
computed here on the host, the Newton steps run in the HIP kernels.  The
element choices follow the reference (src/np_solver.cpp:292-325):
FE_DGQ(k) on equidistant nodes for cells, FE_FaceQ(k) on Gauss-Lobatto
nodes for the trace, QGauss(k+1) rules; numbering as include/ddhdg_hip.h.
"""
from __future__ import annotations

import numpy as np


def gauss01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1), 0.5 * w


def lobatto01(n):
    if n == 2:
        return np.array([0.0, 1.0])
    c = np.zeros(n)
    c[-1] = 1
    inner = np.polynomial.legendre.Legendre(c).deriv().roots()
    return np.concatenate([[0.0], np.sort(0.5 * (inner.real + 1)), [1.0]])


def lagrange(nodes, x):
    """values (len(x), len(nodes)) and derivatives of the Lagrange basis."""
    x = np.asarray(x, dtype=float)
    n = len(nodes)
    V = np.ones((x.size, n))
    D = np.zeros((x.size, n))
    for i in range(n):
        for m in range(n):
            if m != i:
                V[:, i] *= (x - nodes[m]) / (nodes[i] - nodes[m])
        for l in range(n):
            if l == i:
                continue
            t = np.full(x.size, 1.0 / (nodes[i] - nodes[l]))
            for m in range(n):
                if m != i and m != l:
                    t *= (x - nodes[m]) / (nodes[i] - nodes[m])
            D[:, i] += t
    return V, D


TANGENTS = {2: [[1], [0]], 3: [[1, 2], [2, 0], [0, 1]]}


class BoxMesh:
    """Structured box mesh with colorized boundary ids (2a+s), the mesh
    GridGenerator::subdivided_hyper_rectangle(..., colorize=true) builds."""

    def __init__(self, dim, n_cells, lower, upper):
        self.dim = dim
        self.n = [int(n_cells[a]) if a < dim else 1 for a in range(3)]
        self.lower = [float(lower[a]) if a < dim else 0.0 for a in range(3)]
        self.upper = [float(upper[a]) if a < dim else 1.0 for a in range(3)]
        self.h = [(self.upper[a] - self.lower[a]) / self.n[a] for a in range(3)]
        self.ncells = int(np.prod(self.n))
        self.face_dims = [[self.n[b] + (1 if a == b else 0) for b in range(3)] for a in range(3)]
        self.face_off = [0]
        for a in range(3):
            cnt = int(np.prod(self.face_dims[a])) if a < dim else 0
            self.face_off.append(self.face_off[-1] + cnt)
        self.nfaces = self.face_off[dim]
        c = np.arange(self.ncells)
        self.ic = np.stack([c % self.n[0], (c // self.n[0]) % self.n[1],
                            c // (self.n[0] * self.n[1])], axis=1)
        nfc = 2 * dim
        self.cell_faces = np.zeros((self.ncells, nfc), dtype=np.int64)
        self.cell_bid = np.full((self.ncells, nfc), -1, dtype=np.int64)
        for lf in range(nfc):
            a, s = divmod(lf, 2)
            pos = self.ic.copy()
            pos[:, a] += s
            m = self.face_dims[a]
            self.cell_faces[:, lf] = self.face_off[a] + pos[:, 0] + m[0] * (pos[:, 1] + m[1] * pos[:, 2])
            on_b = (self.ic[:, a] == 0) if s == 0 else (self.ic[:, a] == self.n[a] - 1)
            self.cell_bid[on_b, lf] = lf

    @classmethod
    def hyper_rectangle(cls, dim, p1, p2, n_cells):
        return cls(dim, n_cells, p1, p2)


class RefCell:
    """Shape tables of one box cell of the mesh (all cells share them)."""

    def __init__(self, mesh: BoxMesh, k: int):
        d = mesh.dim
        self.dim, self.k = d, k
        self.nb, self.nf = (k + 1) ** d, (k + 1) ** (d - 1)
        self.Qc, self.Qf, self.nfc = self.nb, self.nf, 2 * d
        qx, qw = gauss01(k + 1)
        self.qx = qx
        cn = np.arange(k + 1) / k
        tn = lobatto01(k + 1)
        Vc, Dc = lagrange(cn, qx)
        h = mesh.h
        # cell tables, lexicographic (first axis fastest)
        idx = np.array(np.unravel_index(np.arange(self.nb), [k + 1] * d, order="F")).T
        self.phic = np.ones((self.Qc, self.nb))
        self.gradc = np.ones((self.Qc, self.nb, d))
        self.JxWc = np.ones(self.Qc) * np.prod(h[:d])
        for q in range(self.Qc):
            qi = idx[q]
            for a in range(d):
                self.JxWc[q] *= qw[qi[a]]
            for j in range(self.nb):
                ji = idx[j]
                self.phic[q, j] = np.prod([Vc[qi[a], ji[a]] for a in range(d)])
                for a in range(d):
                    self.gradc[q, j, a] = np.prod(
                        [(Dc if b == a else Vc)[qi[b], ji[b]] for b in range(d)]) / h[a]
        # faces
        fidx = np.array(np.unravel_index(np.arange(self.nf), [k + 1] * (d - 1), order="F")).T \
            if d > 1 else np.zeros((1, 0), dtype=int)
        Vt, _ = lagrange(tn, qx)
        self.xi = np.ones((self.Qf, self.nf))
        for s in range(self.Qf):
            for j in range(self.nf):
                self.xi[s, j] = np.prod([Vt[fidx[s][m], fidx[j][m]] for m in range(d - 1)])
        self.phif = np.zeros((self.nfc, self.Qf, self.nb))
        self.JxWf = np.zeros((self.nfc, self.Qf))
        for g in range(self.nfc):
            a, side = divmod(g, 2)
            area = np.prod([h[b] for b in range(d) if b != a])
            for s in range(self.Qf):
                pt = np.zeros(d)
                pt[a] = side
                w = area
                for m in range(d - 1):
                    pt[TANGENTS[d][a][m]] = qx[fidx[s][m]]
                    w *= qw[fidx[s][m]]
                self.JxWf[g, s] = w
                Vp, _ = lagrange(cn, pt)
                for j in range(self.nb):
                    self.phif[g, s, j] = np.prod([Vp[b, idx[j][b]] for b in range(d)])
        self.M = np.einsum("q,qi,qj->ij", self.JxWc, self.phic, self.phic)
        self.Minv = np.linalg.inv(self.M)
        self.Mf = {a: np.einsum("s,si,sj->ij", self.JxWf[2 * a], self.xi, self.xi) for a in range(d)}


def project_cells(mesh: BoxMesh, ref: RefCell, fc, ff):
    """L2 projection of a scalar function and of minus its gradient (as the
    reference project_component_one_cell, src/project_component.cpp:
    282-452).  fc: values at cell qps (ncells, Qc); ff: values at the
    cells' face qps (ncells, nfc, Qf).  Returns scalar dofs (ncells, nb) and
    flux dofs (ncells, dim, nb)."""
    d = mesh.dim
    u = np.einsum("ij,cj->ci", ref.Minv, np.einsum("q,qi,cq->ci", ref.JxWc, ref.phic, fc))
    W = np.zeros((mesh.ncells, d, ref.nb))
    for a in range(d):
        r = np.einsum("q,qi,cq->ci", ref.JxWc, ref.gradc[:, :, a], fc)
        for s_ in range(2):
            g = 2 * a + s_
            nu = 1.0 if s_ else -1.0
            r -= nu * np.einsum("s,si,cs->ci", ref.JxWf[g], ref.phif[g], ff[:, g, :])
        W[:, a, :] = np.einsum("ij,cj->ci", ref.Minv, r)
    return u, W


def face_values(mesh: BoxMesh, ref: RefCell, u):
    """values of cell scalar dofs u (ncells, nb) at the cells' face qps."""
    return np.einsum("gsj,cj->cgs", ref.phif, u)


def trace_project(mesh: BoxMesh, ref: RefCell, fvals):
    """face L2 projection of values at face qps (nfaces, Qf) -> (nfaces, nf)."""
    out = np.zeros((mesh.nfaces, ref.nf))
    for a in range(mesh.dim):
        sl = slice(mesh.face_off[a], mesh.face_off[a + 1])
        rhs = np.einsum("s,sj,fs->fj", ref.JxWf[2 * a], ref.xi, fvals[sl])
        out[sl] = np.linalg.solve(ref.Mf[a], rhs.T).T
    return out


def trace_l2_average(mesh: BoxMesh, ref: RefCell, cell_face_vals):
    """trace from cell values: L2 projection of the average of the two sides
    (the single side on the boundary), the reference's l2_average strategy
    (src/copy_on_trace.cpp:629-700, 949-993)."""
    acc = np.zeros((mesh.nfaces, ref.Qf))
    cnt = np.zeros(mesh.nfaces)
    for lf in range(2 * mesh.dim):
        np.add.at(acc, mesh.cell_faces[:, lf], cell_face_vals[:, lf, :])
        np.add.at(cnt, mesh.cell_faces[:, lf], 1)
    return trace_project(mesh, ref, acc / cnt[:, None])


def cell_face_points(mesh: BoxMesh, face_qpoints):
    """(ncells, nfc, Qf, dim) coordinates of each cell's face qps."""
    return face_qpoints[mesh.cell_faces]
