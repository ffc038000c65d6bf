/*
 * hdg_oracle.h -- C interface of the CPU restatement of ddhdg's per-cell
 * HDG path (assembly, LocalCondenser, Gauss-Jordan Schur complement,
 * serial in-order scatter, reconstruction).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path
 * and the timed CPU baseline.  Product code never links or calls it; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do.
 *
 * Everything here follows the reference loops one-for-one (file:line
 * citations are in hdg_oracle.cpp), including the quirks the reference
 * has (SRH derivative scaling, Gauss-Jordan full inverse, two assembly
 * passes per Newton step).
 *
 * Mesh: structured box, n[0] x n[1] (x n[2]) cells, colorized boundary ids
 * (id 2a+s for the face with normal axis a, side s), the same meshes
 * GridGenerator::subdivided_hyper_rectangle(..., colorize=true) makes.
 *
 * Numbering shared with the HIP path (so vectors compare 1:1):
 *   cell c = ix + nx*(iy + ny*iz)
 *   cell dofs (all three components, dimensionless):
 *       c*NC + comp*(d+1)*nb + a*nb + j   (flux, axis a, basis j)
 *       c*NC + comp*(d+1)*nb + d*nb + j   (scalar, basis j)
 *   faces: axis-major, then lexicographic position on the face grid
 *   full trace vector:       f*3*nf + comp*nf + j
 *   restricted trace vector: f*ne*nf + ci*nf + j  (ci = index among the
 *                            enabled components, in V, n, p order)
 */
#ifndef HDG_ORACLE_H
#define HDG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

typedef struct
{
  int    dim;          /* 2 or 3 */
  int    degree;       /* k >= 1, same for V, n, p */
  int    n_cells[3];
  double lower[3];
  double upper[3];
  /* physics, physical units (reference include/problem.h:13-52) */
  double epsilon;      /* absolute permittivity, F/m */
  double mu_n, mu_p;   /* m^2/(V s) */
  double temperature;  /* K, constant */
  double nc, nv;       /* band densities */
  double ec, ev;       /* band edge energies, eV */
  int    recombination;  /* 0: none, 1: SRH fixed T, 2: linear */
  double rec_params[3];  /* SRH: ni, tau_n, tau_p; linear: a, b, c */
  double tau[3];         /* V, n, p (FixedTauComputer, physical) */
  int    flux_type;      /* 0 use_cell, 1 use_trace, 2 qiu_shi */
  int    phi_linearize;
  double adim[4];        /* scale_length, temperature_magnitude,
                            doping_magnitude, mobility_magnitude */
  int    bc_type[6][3];  /* per boundary id, per component:
                            0 none, 1 Dirichlet, 2 Neumann */
  int    drift_jacobian; /* 0: the reference's linearisation of the volume
                            drift term (template.h:500, 521-522, sign opposite
                            to the residual's J at :656, :673); 1: the
                            consistent one (Newton Jacobian exact) */
} orc_problem;

int  orc_create(const orc_problem *prob, orc_ctx **out);
void orc_destroy(orc_ctx *ctx);
const char *orc_last_error(void);

/* sizes: [0]=ncells [1]=nfaces [2]=NC (all-component cell dofs per cell)
 * [3]=Qc [4]=Qf [5]=nb [6]=nf [7]=ne (enabled) [8]=N (enabled cell dofs)
 * [9]=T (enabled trace dofs per cell) [10]=n restricted trace dofs
 * [11]=csr nnz */
void orc_sizes(const orc_ctx *ctx, long long out[12]);

int orc_set_enabled(orc_ctx *ctx, int V, int n, int p);
/* cell quadrature points: ncells*Qc*dim; face quadrature points:
 * nfaces*Qf*dim (face-local order, first tangent axis fastest) */
void orc_cell_qpoints(const orc_ctx *ctx, double *xyz);
void orc_face_qpoints(const orc_ctx *ctx, double *xyz);
/* doping at cell quadrature points, physical (ncells*Qc) */
void orc_set_doping(orc_ctx *ctx, const double *doping_qp);
/* boundary values at face quadrature points, physical, nfaces*3*Qf
 * (entries of interior faces are ignored); NULL leaves zeros */
void orc_set_bc_values(orc_ctx *ctx, const double *dirichlet,
                       const double *neumann);
/* CellFaceTauComputer: tau per (cell, face, component), physical,
 * tau[(cell*2d + face)*3 + comp]; NULL returns to the problem's fixed tau */
void orc_set_cell_face_tau(orc_ctx *ctx, const double *tau);
void orc_set_state(orc_ctx *ctx, const double *cell, const double *trace);
void orc_get_state(const orc_ctx *ctx, double *cell, double *trace);

/* get_local_cell_system equivalent (after the LocalCondenser, before
 * gauss_jordan): A (N*N), CT (N*T), TC (T*N, stored negated as in the
 * reference), TT (T*T), c (N), t (T).  Row-major. */
int orc_local_system(orc_ctx *ctx, int cell, int thermodyn, double *A,
                     double *CT, double *TC, double *TT, double *c,
                     double *t);
/* condensed cell: S (T*T), g (T), restricted global trace indices (T),
 * Dirichlet list (restricted global index, value) with count. */
int orc_condensed_cell(orc_ctx *ctx, int cell, int thermodyn, double *S,
                       double *g, long long *dofs, int *n_dir,
                       long long *dir_idx, double *dir_val);

/* restricted trace CSR pattern (rowptr n+1, colind nnz) */
void orc_csr_pattern(const orc_ctx *ctx, long long *rowptr, int *colind);
/* pass 1 over cells [0, ncells_limit) (all if <0): zeroes the global matrix
 * and rhs, then WorkStream-style pipeline (nthreads workers, serial in-order
 * copier).  Returns the wall seconds of the cell pipeline (zeroing not
 * included), negative on error. */
double orc_pass1(orc_ctx *ctx, int thermodyn, int nthreads,
                 long long ncells_limit);
void   orc_get_system(const orc_ctx *ctx, double *vals, double *rhs);
/* Dirichlet list from the last pass 1 (restricted indices, values) */
int  orc_n_dirichlet(const orc_ctx *ctx);
void orc_get_dirichlet(const orc_ctx *ctx, long long *idx, double *val);
/* pass 2: reconstruct with the full trace update (nfaces*3*nf), write the
 * cell update (ncells*NC).  Returns seconds. */
double orc_pass2(orc_ctx *ctx, int thermodyn, int nthreads,
                 const double *update_trace_full, double *update_cell);

/* recombination model of prob at one point: out = {R, dR/dn, dR/dp} with the
 * reference's compute_recombination_term / compute_derivative_of_
 * recombination_term semantics (densities rescaled by rf) */
/* integrals of the numerical fluxes of n and p over the faces of one
 * boundary id (outward normal, dimensionless): out[0] (n), out[1] (p) */
int orc_terminal_current(orc_ctx *ctx, int boundary_id, double out[2]);

void orc_recombination(const orc_problem *prob, double n, double p, double rf,
                       double out[3]);

#ifdef __cplusplus
}
#endif
#endif
