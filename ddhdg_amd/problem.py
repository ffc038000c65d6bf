"""Problem description mirroring the reference's host API.

Names, argument meaning and defaults follow the reference classes:
  Adimensionalizer              include/adimensionalizer.h:15-262
  HomogeneousPermittivity       include/permittivity.h:10-88
  HomogeneousMobility           include/mobility.h:10-87
  ShockleyReadHallFixedTemperature, LinearRecombinationTerm
                                include/recombination_term.h, src/recombination_term.cpp
  BoundaryConditionHandler      include/boundary_conditions.h:125-200
  HomogeneousProblem            include/problem.h:13-52
  NonlinearSolverParameters, FixedTauNPSolverParameters
                                include/np_solver_parameters.h:48-64, 299-333
Functions are Python callables evaluated on arrays of points (..., dim),
or plain numbers for constants.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# include/constants.h:12-15
PI = math.pi
EPSILON0 = 8.85418781762e-12
Q = 1.602176634e-19
KB = 1.380649e-23
EV = 1.602176565e-19

V, n, p = 0, 1, 2
COMPONENTS = (V, n, p)
DIRICHLET, NEUMANN = "dirichlet", "neumann"


def evaluate(f, pts):
    """evaluate a function spec at points (..., dim)."""
    pts = np.asarray(pts, dtype=float)
    if callable(f):
        return np.broadcast_to(np.asarray(f(pts), dtype=float), pts.shape[:-1]).copy()
    return np.full(pts.shape[:-1], float(f))


class Adimensionalizer:
    def __init__(self, scale_length=1.0, temperature_magnitude=Q / KB, doping_magnitude=1.0,
                 mobility_magnitude=1.0):
        self.scale_length = scale_length
        self.temperature_magnitude = temperature_magnitude
        self.doping_magnitude = doping_magnitude
        self.mobility_magnitude = mobility_magnitude

    def get_component_rescaling_factor(self, c):
        if c == V:
            return self.temperature_magnitude * KB / Q
        return self.doping_magnitude

    def get_permittivity_rescaling_factor(self):
        return (self.scale_length ** 2 * Q * Q * self.doping_magnitude /
                (self.temperature_magnitude * KB))

    def get_thermal_voltage_rescaling_factor(self):
        return self.temperature_magnitude * KB / Q

    def get_mobility_rescaling_factor(self):
        return self.mobility_magnitude

    def get_recombination_rescaling_factor(self):
        return (KB * self.doping_magnitude * self.temperature_magnitude * self.mobility_magnitude /
                (self.scale_length ** 2 * Q))

    def get_tau_rescaling_factor(self, c):
        Tv = self.temperature_magnitude * KB / Q
        if c == V:
            return Q * self.doping_magnitude / Tv
        return Tv * self.mobility_magnitude / self.scale_length ** 2

    def get_neumann_boundary_condition_rescaling_factor(self, c):
        if c == V:
            return Q
        return (KB * self.doping_magnitude * self.temperature_magnitude * self.mobility_magnitude /
                self.scale_length ** 2)


@dataclass
class HomogeneousPermittivity:
    epsilon: float


@dataclass
class HomogeneousMobility:
    mu: float
    component: int = n


@dataclass
class ShockleyReadHallFixedTemperature:
    intrinsic_carrier_concentration: float
    electron_life_time: float
    hole_life_time: float

    @classmethod
    def from_band_parameters(cls, nc, nv, ec, ev, temperature, tau_n, tau_p):
        # reference quirk kept: the constructor stores Nc*Nv*exp(..), i.e.
        # n_i^2, as the intrinsic concentration (include/recombination_term.h:219-229)
        ni = nc * nv * math.exp((ev - ec) / (KB * temperature))
        return cls(ni, tau_n, tau_p)

    def compute_recombination_term(self, n_, p_, rescaling_factor):
        ni = self.intrinsic_carrier_concentration / rescaling_factor
        num = n_ * p_ - ni * ni
        den = self.hole_life_time * (n_ + ni) + self.electron_life_time * (p_ + ni)
        return num * rescaling_factor / den


@dataclass
class LinearRecombinationTerm:
    constant_term: float = 0.0
    n_linear_coefficient: float = 0.0
    p_linear_coefficient: float = 0.0


class BoundaryConditionHandler:
    def __init__(self):
        self.bcs = {}

    def add_boundary_condition(self, boundary_id, bc_type, component, f):
        self.bcs[(int(boundary_id), int(component))] = (bc_type, f)

    def replace_boundary_condition(self, boundary_id, bc_type, component, f):
        self.add_boundary_condition(boundary_id, bc_type, component, f)

    def get(self, boundary_id, component):
        return self.bcs.get((int(boundary_id), int(component)))

    def bc_type_table(self):
        t = [[0, 0, 0] for _ in range(6)]
        for (bid, c), (kind, _) in self.bcs.items():
            if 0 <= bid < 6:
                t[bid][c] = 1 if kind == DIRICHLET else 2
        return t


@dataclass
class HomogeneousProblem:
    mesh: object                      # fe.BoxMesh
    permittivity: HomogeneousPermittivity
    n_mobility: HomogeneousMobility
    p_mobility: HomogeneousMobility
    recombination_term: object
    temperature: float
    doping: object
    boundary_handler: BoundaryConditionHandler
    conduction_band_density: float
    valence_band_density: float
    conduction_band_edge_energy: float = 0.0
    valence_band_edge_energy: float = 0.0


@dataclass
class NonlinearSolverParameters:
    absolute_tolerance: float = 1e-10
    relative_tolerance: float = 1e-10
    max_number_of_iterations: int = 100
    alpha: float = 1.0


@dataclass
class FixedTauNPSolverParameters:
    """FixedTauNPSolverParameters (include/np_solver_parameters.h:295-330):
    degree 1 and tau 1 for every component by default, as there"""
    V_degree: int = 1
    n_degree: int = 1
    p_degree: int = 1
    nonlinear_parameters: NonlinearSolverParameters = field(default_factory=NonlinearSolverParameters)
    V_tau: float = 1.0
    n_tau: float = 1.0
    p_tau: float = 1.0
    iterative_linear_solver: bool = False
    multithreading: bool = True
    dd_flux_type: str = "use_cell"
    phi_linearize: bool = False
    # global trace solve: "gpu" = nested-dissection multifrontal LU on the
    # device (the default), "host" = SciPy's SuperLU on the host (the
    # reference's plain sparse direct form, UMFPACK, src/np_solver.cpp:1347-1351)
    trace_solver: str = "gpu"
    # replay each Newton iteration from a captured hipGraph (one launch per
    # iteration instead of hundreds); needs trace_solver "gpu"
    graph_newton: bool = False
    # False: the reference's linearisation of the volume drift term
    # (include/assembling_methods.template.h:500, 521-522), whose sign is
    # opposite to the derivative of its own residual (:656, :673), so that
    # Newton is inexact; True: the consistent linearisation
    consistent_drift_jacobian: bool = False


# DDFluxType (include/np_solver_parameters.h:34-39)
DD_FLUX_TYPES = {"use_cell": 0, "use_trace": 1, "qiu_shi_stabilization": 2}


def flux_type_code(name):
    if name not in DD_FLUX_TYPES:
        raise ValueError(f"dd_flux_type must be one of {sorted(DD_FLUX_TYPES)}, not {name!r}")
    return DD_FLUX_TYPES[name]


@dataclass
class CellFaceTauNPSolverParameters:
    """CellFaceTauNPSolverParameters (include/np_solver_parameters.h:335-404,
    src/np_solver_parameters.cpp:239-356; Python binding
    src/pyddhdg/pyddhdg_common.cpp:149-250): tau given per (cell, face) and
    component by set_face / set_multiple_faces, rescaled by the same factors
    as the fixed tau (CellFaceTauComputer, include/np_solver_parameters.h:
    139-259).  A cell is its index on the structured mesh (level 0, x
    fastest, the numbering of GridGenerator::subdivided_hyper_rectangle), or
    a (level, index) pair, or an object with level() and index().  As in the
    reference, a cell that was never set is an error when the solver
    assembles, and unset faces of a set cell hold tau = 0."""
    V_degree: int = 1
    n_degree: int = 1
    p_degree: int = 1
    nonlinear_parameters: NonlinearSolverParameters = field(default_factory=NonlinearSolverParameters)
    iterative_linear_solver: bool = False
    multithreading: bool = True
    dd_flux_type: str = "use_cell"
    phi_linearize: bool = False
    trace_solver: str = "gpu"
    graph_newton: bool = False
    consistent_drift_jacobian: bool = False
    version: int = field(default=0, init=False)

    def __post_init__(self):
        self._tau = {}

    @staticmethod
    def _cell_id(cell):
        if hasattr(cell, "level") and hasattr(cell, "index"):
            lv, ix = (cell.level() if callable(cell.level) else cell.level,
                      cell.index() if callable(cell.index) else cell.index)
        elif isinstance(cell, tuple):
            lv, ix = cell
        else:
            lv, ix = 0, cell
        if int(lv) != 0:
            raise ValueError("the structured meshes have level-0 cells only")
        return int(ix)

    def set_face(self, cell, face, V_tau_value, n_tau_value, p_tau_value, faces_per_cell=None):
        cid = self._cell_id(cell)
        fpc = faces_per_cell or 6
        if not 0 <= int(face) < fpc:
            raise ValueError("face index out of range")
        row = self._tau.setdefault(cid, {})
        row[int(face)] = (float(V_tau_value), float(n_tau_value), float(p_tau_value))
        self.version += 1

    def set_multiple_faces(self, dim, cell_levels, cell_indices, faces, V_tau_value,
                           n_tau_value, p_tau_value):
        m = len(cell_levels)
        for lst in (cell_indices, faces, V_tau_value, n_tau_value, p_tau_value):
            if len(lst) != m:
                raise ValueError("The list passed to this function must have the same size")
        for i in range(m):
            self.set_face((cell_levels[i], cell_indices[i]), faces[i], V_tau_value[i],
                          n_tau_value[i], p_tau_value[i], 2 * dim)

    def clear(self):
        self._tau.clear()
        self.version += 1

    def tau_table(self, ncells, faces_per_cell):
        """physical tau, shape (ncells, faces_per_cell, 3)"""
        out = np.zeros((ncells, faces_per_cell, 3))
        for c in range(ncells):
            row = self._tau.get(c)
            if row is None:
                raise KeyError(f"CellFaceTauNPSolverParameters: no tau set for cell {c}")
            for f, v in row.items():
                if f >= faces_per_cell:
                    raise ValueError(f"cell {c}: face {f} out of range")
                out[c, f] = v
        return out


def compute_density(problem, c, qf_potential, potential, temperature):
    """Solver::compute_density (include/solver.h:378-412), physical units."""
    kbt = KB * temperature
    if c == n:
        e = (Q * (potential - qf_potential) - problem.conduction_band_edge_energy * EV) / kbt
        return problem.conduction_band_density * np.exp(e)
    e = (Q * (qf_potential - potential) + problem.valence_band_edge_energy * EV) / kbt
    return problem.valence_band_density * np.exp(e)


def local_charge_neutrality_potential(problem, doping_values, temperature):
    """compute_local_charge_neutrality_on_a_point
    (src/thermodynamic_equilibrium.cpp:9-43), physical volts."""
    nc, nv = problem.conduction_band_density, problem.valence_band_density
    ec = problem.conduction_band_edge_energy * EV
    ev = problem.valence_band_edge_energy * EV
    ut = temperature * KB / Q
    ni2 = nc * nv * np.exp((ev - ec) / (temperature * KB))
    return ((ec + ev) / (2 * Q) + ut / 2 * np.log(nv / nc) +
            ut * np.arcsinh(np.asarray(doping_values) / (2 * np.sqrt(ni2))))
