"""ctypes binding of ``libddhdg_hip.so`` (C ABI in include/ddhdg_hip.h).

The native library is required: importing the context raises if it is not
built, and every call that returns a negative status raises ``HdgError`` with
the library's message.  Device buffers are torch tensors on the current HIP
device; their raw pointers cross the C ABI.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libddhdg_hip.so")
_LIB = None

(HDG_NCELLS, HDG_NFACES, HDG_NC, HDG_QC, HDG_QF, HDG_NB, HDG_NF, HDG_NE, HDG_NS, HDG_T,
 HDG_NTRACE, HDG_NNZ, HDG_NSIZES) = range(13)

# every symbol include/ddhdg_hip.h declares
EXPORTS = [
    "hdg_last_error", "hdg_ctx_create", "hdg_ctx_destroy", "hdg_sizes", "hdg_set_enabled",
    "hdg_csr_pattern", "hdg_dirichlet_dofs", "hdg_cell_qpoints", "hdg_face_qpoints",
    "hdg_set_doping", "hdg_set_bc_values", "hdg_condense", "hdg_reconstruct",
    "hdg_cell_systems", "hdg_set_kernel_variant",
    "hdg_trace_solver_create", "hdg_trace_solver_destroy", "hdg_trace_solver_info",
    "hdg_trace_factor", "hdg_trace_factor_status", "hdg_trace_solve",
    "hdg_trace_plan_selftest", "hdg_terminal_current", "hdg_set_precision",
    "hdg_batch_create", "hdg_batch_destroy", "hdg_batch_condense", "hdg_batch_info",
    "hdg_max_abs4", "hdg_set_cell_face_tau", "hdg_max_abs4_noalloc", "hdg_trace_plan_size",
    "hdg_local_system", "hdg_gmres_solve", "hdg_batch_tuning",
]

(HDG_TS_FLOPS, HDG_TS_BYTES, HDG_TS_LEVELS, HDG_TS_MAX_FRONT, HDG_TS_NODES,
 HDG_TS_NINFO) = range(6)


class HdgError(RuntimeError):
    pass


class MeshDesc(ct.Structure):
    _fields_ = [("dim", ct.c_int), ("degree", ct.c_int), ("n_cells", ct.c_int * 3),
                ("lower", ct.c_double * 3), ("upper", ct.c_double * 3)]


class Phys(ct.Structure):
    _fields_ = [("eps", ct.c_double), ("mu_n", ct.c_double), ("mu_p", ct.c_double),
                ("U_T", ct.c_double), ("tau", ct.c_double * 3), ("recombination", ct.c_int),
                ("rec", ct.c_double * 3), ("n_rescale", ct.c_double),
                ("rec_rescale", ct.c_double), ("kT_over_q", ct.c_double),
                ("ec_r", ct.c_double), ("ev_r", ct.c_double), ("nc", ct.c_double),
                ("nv", ct.c_double), ("flux_type", ct.c_int), ("phi_linearize", ct.c_int),
                ("bc_type", (ct.c_int * 3) * 6), ("drift_jacobian", ct.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise HdgError(f"{LIB_PATH} is not built (python -m ddhdg_amd.build)")
        L = ct.CDLL(LIB_PATH)
        P = ct.POINTER
        i, v, dp, lp, ip = ct.c_int, ct.c_void_p, P(ct.c_double), P(ct.c_longlong), P(ct.c_int)
        sigs = {
            "hdg_last_error": (ct.c_char_p, []),
            "hdg_ctx_create": (i, [P(MeshDesc), P(Phys), P(v)]),
            "hdg_ctx_destroy": (None, [v]),
            "hdg_sizes": (i, [v, lp]),
            "hdg_set_enabled": (i, [v, i, i, i]),
            "hdg_set_kernel_variant": (i, [v, i]),
            "hdg_csr_pattern": (i, [v, lp, ip]),
            "hdg_dirichlet_dofs": (i, [v, lp]),
            "hdg_cell_qpoints": (i, [v, dp]),
            "hdg_face_qpoints": (i, [v, dp]),
            "hdg_set_doping": (i, [v, dp]),
            "hdg_set_bc_values": (i, [v, dp, dp]),
            "hdg_condense": (i, [v, v, v, i, v, v, v, v]),
            "hdg_reconstruct": (i, [v, v, v, v, i, v, v]),
            "hdg_cell_systems": (i, [v, v, v, i, v, i, v, v, v]),
            "hdg_local_system": (i, [v, v, v, i, v, i, v, v, v, v, v, v, v]),
            "hdg_trace_solver_create": (i, [v, P(v)]),
            "hdg_trace_solver_destroy": (None, [v]),
            "hdg_trace_solver_info": (i, [v, dp]),
            "hdg_trace_factor": (i, [v, v, v]),
            "hdg_trace_factor_status": (i, [v, v]),
            "hdg_trace_solve": (i, [v, v, v, v]),
            "hdg_trace_plan_selftest": (i, [i, ip, i, ct.c_uint, dp, dp]),
            "hdg_terminal_current": (i, [v, v, v, i, v, v]),
            "hdg_set_precision": (i, [v, i]),
            "hdg_batch_create": (i, [P(v), i, P(v)]),
            "hdg_batch_destroy": (None, [v]),
            "hdg_batch_condense": (i, [v, i, P(v), P(v), P(v), P(v), P(v), v]),
            "hdg_batch_info": (i, [v, dp]),
            "hdg_batch_tuning": (i, [v, dp, i]),
            "hdg_max_abs4": (i, [P(v), lp, v, v, v]),
            "hdg_max_abs4_noalloc": (i, [P(v), lp, v, v]),
            "hdg_trace_plan_size": (i, [i, ip, i, dp]),
            "hdg_set_cell_face_tau": (i, [v, dp]),
            "hdg_gmres_solve": (i, [v, v, v, ct.c_longlong, v, v, i, ct.c_double,
                                    ct.c_longlong, lp, dp, v]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _check(rc):
    if rc < 0:
        raise HdgError(lib().hdg_last_error().decode())
    return rc


def _dp(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def _ptr(t):
    return ct.c_void_p(t.data_ptr())


def _stream():
    import torch
    return ct.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(t, name, numel, dtype="float64"):
    """host-side check of a device buffer before its pointer crosses the C
    ABI: a CUDA tensor, contiguous, of the kernel's element type and at
    least numel elements (the kernels index it without bounds)"""
    import torch
    want = getattr(torch, dtype)
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise HdgError(f"{name}: expected a device tensor")
    if t.dtype != want:
        raise HdgError(f"{name}: expected {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise HdgError(f"{name}: expected a contiguous tensor")
    if t.numel() < numel:
        raise HdgError(f"{name}: {t.numel()} elements, {numel} needed")
    return ct.c_void_p(t.data_ptr())


class HdgContext:
    """One mesh on the GPU: owns the device tables of the C ABI context."""

    def __init__(self, mesh: MeshDesc, phys: Phys):
        self.L = lib()
        self.ctx = ct.c_void_p()
        _check(self.L.hdg_ctx_create(ct.byref(mesh), ct.byref(phys), ct.byref(self.ctx)))
        self.mesh, self.phys = mesh, phys
        self.dim = mesh.dim
        self.precision = 0
        self._refresh()

    def close(self):
        if self.ctx:
            self.L.hdg_ctx_destroy(self.ctx)
            self.ctx = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _refresh(self):
        s = np.zeros(HDG_NSIZES, dtype=np.int64)
        _check(self.L.hdg_sizes(self.ctx, s.ctypes.data_as(ct.POINTER(ct.c_longlong))))
        self.sizes = s
        (self.ncells, self.nfaces, self.NC, self.Qc, self.Qf, self.nb, self.nf, self.ne,
         self.Ns, self.T, self.ntrace, self.nnz) = (int(x) for x in s)

    def set_enabled(self, V, n, p):
        _check(self.L.hdg_set_enabled(self.ctx, int(V), int(n), int(p)))
        self._refresh()

    def set_precision(self, precision):
        """0: FP64; 1: FP32 elimination with FP64 refinement in the one-wave
        register kernels (2D degree 1-2, 3D degree 1; both passes), FP64
        elsewhere; 2: FP32 with FP64 refinement in every kernel"""
        _check(self.L.hdg_set_precision(self.ctx, int(precision)))
        self.precision = int(precision)

    def set_kernel_variant(self, variant):
        """0 automatic, 1 LDS workgroup-per-cell kernel, 2 register-resident."""
        _check(self.L.hdg_set_kernel_variant(self.ctx, int(variant)))

    def csr_pattern(self):
        rp = np.zeros(self.ntrace + 1, dtype=np.int64)
        ci = np.zeros(self.nnz, dtype=np.int32)
        _check(self.L.hdg_csr_pattern(self.ctx, rp.ctypes.data_as(ct.POINTER(ct.c_longlong)),
                                      ci.ctypes.data_as(ct.POINTER(ct.c_int))))
        return rp, ci

    def dirichlet_dofs(self):
        n = _check(self.L.hdg_dirichlet_dofs(self.ctx, None))
        idx = np.zeros(max(n, 1), dtype=np.int64)
        _check(self.L.hdg_dirichlet_dofs(self.ctx, idx.ctypes.data_as(ct.POINTER(ct.c_longlong))))
        return idx[:n]

    def cell_qpoints(self):
        x = np.zeros((self.ncells, self.Qc, self.dim))
        _check(self.L.hdg_cell_qpoints(self.ctx, _dp(x)))
        return x

    def face_qpoints(self):
        x = np.zeros((self.nfaces, self.Qf, self.dim))
        _check(self.L.hdg_face_qpoints(self.ctx, _dp(x)))
        return x

    def set_doping(self, rescaled_doping_qp):
        a = np.ascontiguousarray(rescaled_doping_qp, dtype=np.float64)
        assert a.size == self.ncells * self.Qc
        _check(self.L.hdg_set_doping(self.ctx, _dp(a)))

    def set_cell_face_tau(self, rescaled_tau):
        """per-(cell, face, component) tau, rescaled, shape (ncells, 2*dim,
        3); None returns to the fixed tau"""
        if rescaled_tau is None:
            _check(self.L.hdg_set_cell_face_tau(self.ctx, None))
            return
        a = np.ascontiguousarray(rescaled_tau, dtype=np.float64)
        assert a.size == self.ncells * 2 * self.dim * 3
        _check(self.L.hdg_set_cell_face_tau(self.ctx, _dp(a)))

    def set_bc_values(self, dirichlet=None, neumann=None):
        n = self.nfaces * 3 * self.Qf
        d = None if dirichlet is None else np.ascontiguousarray(dirichlet, dtype=np.float64)
        m = None if neumann is None else np.ascontiguousarray(neumann, dtype=np.float64)
        for a in (d, m):
            assert a is None or a.size == n
        _check(self.L.hdg_set_bc_values(self.ctx, None if d is None else _dp(d),
                                        None if m is None else _dp(m)))

    # -- device-pointer calls (torch tensors on the current device)
    def _state(self, x_cell, x_trace):
        return (_dev(x_cell, "x_cell", self.ncells * self.NC),
                _dev(x_trace, "x_trace", self.nfaces * 3 * self.nf))

    def condense(self, x_cell, x_trace, thermodyn, vals, rhs, dir_vals):
        xc, xt = self._state(x_cell, x_trace)
        _check(self.L.hdg_condense(self.ctx, xc, xt, int(thermodyn),
                                   _dev(vals, "vals", self.nnz), _dev(rhs, "rhs", self.ntrace),
                                   _dev(dir_vals, "dir_vals", self.ntrace), _stream()))

    def reconstruct(self, x_cell, x_trace, dx_trace, thermodyn, dx_cell):
        xc, xt = self._state(x_cell, x_trace)
        _check(self.L.hdg_reconstruct(self.ctx, xc, xt,
                                      _dev(dx_trace, "dx_trace", self.nfaces * 3 * self.nf),
                                      int(thermodyn), _dev(dx_cell, "dx_cell", self.ncells * self.NC),
                                      _stream()))

    def cell_systems(self, x_cell, x_trace, thermodyn, cells, S, g):
        """condensed blocks of the listed cells (int32 indices in [0, ncells)),
        S cell-major (T x T each), g (T each)"""
        xc, xt = self._state(x_cell, x_trace)
        m = int(cells.numel())
        cp = _dev(cells, "cells", m, "int32")
        if m and (int(cells.min()) < 0 or int(cells.max()) >= self.ncells):
            raise HdgError(f"cells: indices outside [0, {self.ncells})")
        _check(self.L.hdg_cell_systems(self.ctx, xc, xt, int(thermodyn), cp, m,
                                       _dev(S, "S", m * self.T * self.T), _dev(g, "g", m * self.T),
                                       _stream()))

    def local_system(self, x_cell, x_trace, thermodyn, cells, A, CT, TC, TT, c, t):
        """uncondensed local blocks of the listed cells (N = NE*(dim+1)*NB):
        A (N x N), CT (N x T), TC (T x N, negated), TT (T x T), c, t per cell"""
        xc, xt = self._state(x_cell, x_trace)
        m = int(cells.numel())
        cp = _dev(cells, "cells", m, "int32")
        if m and (int(cells.min()) < 0 or int(cells.max()) >= self.ncells):
            raise HdgError(f"cells: indices outside [0, {self.ncells})")
        N, T = self.ne * (self.dim + 1) * self.nb, self.T
        _check(self.L.hdg_local_system(
            self.ctx, xc, xt, int(thermodyn), cp, m, _dev(A, "A", m * N * N),
            _dev(CT, "CT", m * N * T), _dev(TC, "TC", m * T * N), _dev(TT, "TT", m * T * T),
            _dev(c, "c", m * N), _dev(t, "t", m * T), _stream()))

    def terminal_current(self, x_cell, x_trace, boundary_id, out):
        """out (device, 2): integrals of the numerical fluxes of n and p over
        the faces of boundary_id (dimensionless)."""
        xc, xt = self._state(x_cell, x_trace)
        _check(self.L.hdg_terminal_current(self.ctx, xc, xt, int(boundary_id),
                                           _dev(out, "out", 2), _stream()))


class TraceSolver:
    """GPU nested-dissection multifrontal LU of a context's restricted trace
    system (the global solve; SparseDirectUMFPACK in the reference).  Built
    for the context's current enabled components."""

    def __init__(self, ctx: HdgContext):
        self.L = lib()
        self.h = ct.c_void_p()
        _check(self.L.hdg_trace_solver_create(ctx.ctx, ct.byref(self.h)))
        info = np.zeros(HDG_TS_NINFO)
        _check(self.L.hdg_trace_solver_info(self.h, _dp(info)))
        self.flops, self.bytes = float(info[HDG_TS_FLOPS]), float(info[HDG_TS_BYTES])
        self.levels, self.max_front = int(info[HDG_TS_LEVELS]), int(info[HDG_TS_MAX_FRONT])
        self.nodes = int(info[HDG_TS_NODES])
        self._ctx = ctx
        self.nnz, self.ntrace = ctx.nnz, ctx.ntrace

    def close(self):
        if self.h:
            self.L.hdg_trace_solver_destroy(self.h)
            self.h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def factor(self, vals, check=False):
        _check(self.L.hdg_trace_factor(self.h, _dev(vals, "vals", self.nnz), _stream()))
        if check:
            _check(self.L.hdg_trace_factor_status(self.h, _stream()))

    def solve(self, rhs, x):
        _check(self.L.hdg_trace_solve(self.h, _dev(rhs, "rhs", self.ntrace),
                                      _dev(x, "x", self.ntrace), _stream()))


def trace_plan_selftest(dim, n_cells, dofs_per_face, seed=1, csr_vals=None):
    """host-only check of the trace solver's plan (no GPU): returns the
    relative error of the plan's multifrontal solve against dense
    elimination (on csr_vals, a host array in the trace CSR order, or a
    random matrix), the number of tree nodes and of levels."""
    n = (ct.c_int * 3)(*(list(n_cells) + [1] * (3 - len(n_cells))))
    out = np.zeros(3)
    vals = None
    if csr_vals is not None:
        vals = np.ascontiguousarray(csr_vals, dtype=np.float64)
    _check(lib().hdg_trace_plan_selftest(int(dim), n, int(dofs_per_face), int(seed),
                                         _dp(vals) if vals is not None else None, _dp(out)))
    return float(out[0]), int(out[1]), int(out[2])


def trace_plan_size(dim, n_cells, dofs_per_face):
    """host-only sizes of the trace solver's plan (no GPU): front, work and
    pivot entries, factorisation FLOPs, largest front, tree nodes"""
    n = (ct.c_int * 3)(*(list(n_cells) + [1] * (3 - len(n_cells))))
    out = np.zeros(6)
    _check(lib().hdg_trace_plan_size(int(dim), n, int(dofs_per_face), _dp(out)))
    return {"front_entries": int(out[0]), "work_entries": int(out[1]),
            "pivot_entries": int(out[2]), "flops": float(out[3]), "max_front": int(out[4]),
            "nodes": int(out[5])}


def gmres_solve(rowptr, colind, vals, rhs, x, restart=28, tol=0.0, max_it=None):
    """x = A^-1 rhs by restarted GMRES without preconditioner from x = 0
    (hdg_gmres_solve), for the device CSR matrix (rowptr int64 n+1, colind
    int32, vals float64) of order n = rhs.numel(); the column indices must lie
    in [0, n) (the product reads x[colind] unchecked: pass a pattern that
    hdg_csr_pattern wrote).  Returns (iterations, residual estimate); raises
    HdgError when it does not converge within max_it (default 10 n)"""
    n = rhs.numel()
    nnz = vals.numel()
    max_it = 10 * n if max_it is None else int(max_it)
    it, res = ct.c_longlong(0), ct.c_double(0.0)
    _check(lib().hdg_gmres_solve(_dev(rowptr, "rowptr", n + 1, "int64"),
                                 _dev(colind, "colind", nnz, "int32"),
                                 _dev(vals, "vals", nnz), ct.c_longlong(n),
                                 _dev(rhs, "rhs", n), _dev(x, "x", n), int(restart),
                                 float(tol), ct.c_longlong(max_it), ct.byref(it),
                                 ct.byref(res), _stream()))
    return int(it.value), float(res.value)


class HdgBatch:
    """pass 1 of several contexts (2D, same components; degree 1/2, or all
    of degree 3 with V/n/p) in one persistent launch per colour; the grid is
    autotuned on the first call."""

    def __init__(self, contexts):
        self.L = lib()
        self.ctxs = list(contexts)
        arr = (ct.c_void_p * len(self.ctxs))(*[c.ctx.value for c in self.ctxs])
        self.h = ct.c_void_p()
        _check(self.L.hdg_batch_create(arr, len(self.ctxs), ct.byref(self.h)))

    def close(self):
        if self.h:
            self.L.hdg_batch_destroy(self.h)
            self.h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def condense(self, thermodyn, x_cells, x_traces, vals, rhs, dirv):
        n = len(self.ctxs)
        for lst in (x_cells, x_traces, vals, rhs, dirv):
            if len(lst) != n:
                raise HdgError(f"batch of {n} contexts, {len(lst)} buffers")

        def arr(ts, name, size):
            return (ct.c_void_p * n)(*[_dev(t, f"{name}[{i}]", size(c)).value
                                       for i, (t, c) in enumerate(zip(ts, self.ctxs))])
        _check(self.L.hdg_batch_condense(
            self.h, int(thermodyn), arr(x_cells, "x_cell", lambda c: c.ncells * c.NC),
            arr(x_traces, "x_trace", lambda c: c.nfaces * 3 * c.nf),
            arr(vals, "vals", lambda c: c.nnz), arr(rhs, "rhs", lambda c: c.ntrace),
            arr(dirv, "dir_vals", lambda c: c.ntrace), _stream()))

    def info(self):
        """the autotuner's choice (after the first condense): cells, grid
        (workgroups), CUs, tile (waves per workgroup, or pairs for degree 3)
        and every candidate (tile, grid, ms of a two-colour pass)"""
        out = np.zeros(6)
        _check(self.L.hdg_batch_info(self.h, _dp(out)))
        cand = np.zeros(3 * 8)
        n = _check(self.L.hdg_batch_tuning(self.h, _dp(cand), 8))
        return {"cells": int(out[0]), "grid": int(out[1]), "cus": int(out[2]),
                "tile": int(out[3]), "ms": float(out[5]),
                "candidates": [{"tile": int(cand[3 * k]), "grid": int(cand[3 * k + 1]),
                                "ms": float(cand[3 * k + 2])} for k in range(min(n, 8))]}


def max_abs4(vectors, scratch, out):
    """out[v] = max |vectors[v]| (four device tensors; scratch: 1024 doubles)"""
    if len(vectors) != 4:
        raise HdgError("max_abs4 takes four vectors")
    arr = (ct.c_void_p * 4)(*[_dev(t, f"vectors[{i}]", t.numel()).value
                               for i, t in enumerate(vectors)])
    lens = (ct.c_longlong * 4)(*[t.numel() for t in vectors])
    _check(lib().hdg_max_abs4(arr, lens, _dev(scratch, "scratch", 1024), _dev(out, "out", 4),
                              _stream()))


def max_abs4_noalloc(vectors, out):
    """out[v] = max |vectors[v]| with the library's static scratch (no
    allocation; concurrent calls on one GPU race)"""
    if len(vectors) != 4:
        raise HdgError("max_abs4 takes four vectors")
    arr = (ct.c_void_p * 4)(*[_dev(t, f"vectors[{i}]", t.numel()).value
                               for i, t in enumerate(vectors)])
    lens = (ct.c_longlong * 4)(*[t.numel() for t in vectors])
    _check(lib().hdg_max_abs4_noalloc(arr, lens, _dev(out, "out", 4), _stream()))

