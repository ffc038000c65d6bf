/*
 * ddhdg_hip.h -- C ABI of the MI355X (gfx950) per-cell HDG hot path.
 *
 * Replaces, at whole-mesh granularity, the reference operator
 *   NPSolver<dim,Problem>::assemble_system(trace_reconstruct,
 *                                          thermodyn_eq, multithreading)
 *   (reference include/np_solver.h:627-630,
 *    include/assembling_methods.template.h:10-89)
 * together with the per-cell operator it dispatches
 *   assemble_system_one_cell_internal<prm, hasBC, Tau, schur>
 *   (include/assembling_methods.template.h:2290-2698)
 * and its copier copy_local_to_global (src/np_solver.cpp:1293-1322).
 *
 * Plain C: pointers and sizes only.  Every function returns 0 on success or
 * a negative code; hdg_last_error() (thread-local) holds the message.  A
 * context owns its device tables; it is used from one host thread and one
 * HIP stream at a time.
 *
 * State vectors live in device memory (HBM) and are passed as device
 * pointers; the numbering is
 *   cell c = ix + nx*(iy + ny*iz)
 *   cell vector (all three components, dimensionless), per cell NC doubles:
 *     comp*(dim+1)*nb + a*nb + j   flux (E, Wn, Wp) axis a, basis j
 *     comp*(dim+1)*nb + dim*nb + j scalar (V, n, p), basis j
 *   faces: axis-major, lexicographic on each face grid
 *   full trace vector:       f*3*nf + comp*nf + j
 *   restricted trace vector: f*ne*nf + ci*nf + j  (enabled components only)
 * as the reference's deal.II FESystem/FE_FaceQ ordering within a cell.
 */
#ifndef DDHDG_HIP_H
#define DDHDG_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hdg_ctx hdg_ctx;

/* Structured box mesh with colorized boundary ids (id 2a+s for the faces
 * with normal axis a on side s), as GridGenerator::subdivided_hyper_
 * rectangle(..., colorize=true) builds it.  Coordinates in scale_length
 * units, as the reference's triangulation. */
typedef struct
{
  int    dim;        /* 2 or 3 */
  int    degree;     /* k: 2D 1..3, 3D 1..2 */
  int    n_cells[3];
  double lower[3];
  double upper[3];
} hdg_mesh_desc;

/* Dimensionless coefficients, folded by the host layer from Problem and
 * Adimensionalizer exactly as the reference ScratchData computes them
 * (include/permittivity.h:67-83, include/mobility.h:66-82,
 *  src/np_solver_parameters.cpp:13-28, include/adimensionalizer.h). */
typedef struct
{
  double eps;            /* epsilon / permittivity rescaling factor */
  double mu_n, mu_p;     /* mobility / mobility rescaling factor */
  double U_T;            /* k_B T / q / thermal voltage rescaling */
  double tau[3];         /* FixedTauComputer, rescaled, V n p */
  int    recombination;  /* 0 none, 1 SRH fixed T, 2 linear */
  double rec[3];         /* SRH: n_i (physical), tau_n, tau_p;
                            linear: a, b, c */
  double n_rescale;      /* doping magnitude (n, p rescaling factor) */
  double rec_rescale;    /* recombination rescaling factor */
  /* Boltzmann terms of the thermodynamic-equilibrium operator
   * (include/assembling_methods.template.h:442-459, 677-693) */
  double kT_over_q;      /* k_B T / (q V_rescale) */
  double ec_r, ev_r;     /* band edges / (q V_rescale) */
  double nc, nv;         /* band densities, physical */
  int    flux_type;      /* DDFluxType (include/np_solver_parameters.h:
                            34-39): 0 use_cell, 1 use_trace,
                            2 qiu_shi_stabilization */
  int    phi_linearize;  /* NPSolverParameters::phi_linearize: the V rows
                            use Boltzmann densities at the quasi-Fermi
                            potentials of the current n and p
                            (include/assembling_methods.template.h:229-264,
                            442-459, 677-693) */
  int    bc_type[6][3];  /* per boundary id and component:
                            0 none, 1 Dirichlet, 2 Neumann */
  int    drift_jacobian; /* 0: the reference's linearisation of the volume
                            drift term in E (template.h:500, 521-522), whose
                            sign is opposite to the derivative of its residual
                            (:656, :673); 1: the consistent (exact) one */
} hdg_phys;

/* sizes returned by hdg_sizes */
enum
{
  HDG_NCELLS = 0,
  HDG_NFACES,
  HDG_NC,        /* cell dofs per cell, all components */
  HDG_QC,        /* cell quadrature points */
  HDG_QF,        /* face quadrature points */
  HDG_NB,        /* scalar basis per cell */
  HDG_NF,        /* trace basis per face */
  HDG_NE,        /* enabled components */
  HDG_NS,        /* enabled scalar dofs per cell (reduced cell block) */
  HDG_T,         /* enabled trace dofs per cell */
  HDG_NTRACE,    /* restricted trace dofs (global) */
  HDG_NNZ,       /* restricted trace matrix nonzeros */
  HDG_NSIZES
};

const char *hdg_last_error(void);

/* Context over one mesh on the current HIP device.  Allocates the device
 * tables (geometry, lifting operators, scatter map) and, for 3D degree 2,
 * the generic kernel's working storage for its reduced systems (512
 * slices of 476 KB, 244 MB).  All components enabled on creation. */
int  hdg_ctx_create(const hdg_mesh_desc *mesh, const hdg_phys *phys,
                    hdg_ctx **out);
void hdg_ctx_destroy(hdg_ctx *ctx);
int  hdg_sizes(const hdg_ctx *ctx, long long out[HDG_NSIZES]);

/* NPSolver::set_enabled_components (reference include/np_solver.h:172-175);
 * rebuilds the restricted trace pattern. */
int hdg_set_enabled(hdg_ctx *ctx, int V, int n, int p);

/* Kernel choice for pass 1: 0 = automatic (a register-resident kernel
 * where one applies: 2D degree 1-2 and 3D degree 1, all components or the
 * V-only equilibrium operator; 2D degree 3 with all components; otherwise
 * the generic workgroup-per-cell kernel), 1 = always the generic kernel,
 * 2 = register-resident (error if none applies).  No counterpart in the
 * reference (one CPU code path). */
int hdg_set_kernel_variant(hdg_ctx *ctx, int variant);

/* Precision of the local elimination of both passes: 0 = FP64 (default);
 * 1 = the scalar block eliminated in FP32 with FP64 iterative refinement
 * (assembly, the residual, the refinement residuals and the Schur product
 * in FP64) in the one-wave register kernels (2D degree 1-2, 3D degree 1),
 * FP64 elsewhere; 2 = the same in every kernel, also the degree-3 pair
 * kernel and the generic kernel, whose FP32 forms run slower than their
 * FP64 ones.  No counterpart in the reference (FP64 throughout). */
int hdg_set_precision(hdg_ctx *ctx, int precision);

/* Restricted trace matrix pattern, host arrays: rowptr (NTRACE+1),
 * colind (NNZ); rows are sorted. */
int hdg_csr_pattern(const hdg_ctx *ctx, long long *rowptr, int *colind);
/* Restricted indices of the Dirichlet-constrained trace dofs (host array,
 * returns the count when idx is NULL). */
int hdg_dirichlet_dofs(const hdg_ctx *ctx, long long *idx);
/* Quadrature points (host arrays): cells ncells*QC*dim, faces
 * nfaces*QF*dim (face-local order, first tangent axis fastest). */
int hdg_cell_qpoints(const hdg_ctx *ctx, double *xyz);
int hdg_face_qpoints(const hdg_ctx *ctx, double *xyz);

/* Rescaled doping at cell quadrature points (host array, ncells*QC),
 * i.e. the reference's rescaled_doping evaluated at the quadrature points
 * (src/np_solver.cpp:341-343). */
int hdg_set_doping(hdg_ctx *ctx, const double *doping_qp);
/* CellFaceTauComputer (reference include/np_solver_parameters.h:139-259,
 * CellFaceTauNPSolverParameters::set_face, src/np_solver_parameters.cpp:
 * 328-356): tau per cell, face and component, rescaled like hdg_phys.tau,
 * host array tau[(cell*2*dim + face)*3 + comp].  NULL returns to the fixed
 * tau of hdg_phys.  Values must be finite and >= 0.  Later calls overwrite
 * the values in place (synchronous copy), so launches recorded in a HIP
 * graph see them; a NULL call frees the device copy and must not be made
 * while such a graph may still run. */
int hdg_set_cell_face_tau(hdg_ctx *ctx, const double *tau);
/* Boundary data at face quadrature points (host arrays, nfaces*3*QF):
 * Dirichlet values divided by the component rescaling factor and Neumann
 * values divided by the Neumann rescaling factor; entries of interior
 * faces are ignored.  NULL leaves the current values. */
int hdg_set_bc_values(hdg_ctx *ctx, const double *dirichlet,
                      const double *neumann);

/* Pass 1 of a Newton step: assembly + LocalCondenser + static condensation
 * + deterministic scatter.  Replaces assemble_system(false, thermodyn, *)
 * (template.h:10-89 with :2290-2687) and copy_local_to_global.
 *   x_cell, x_trace   device, current state (ncells*NC, nfaces*3*nf)
 *   csr_vals          device, NNZ values in the hdg_csr_pattern order
 *   rhs               device, NTRACE (= system_rhs, minus the residual)
 *   dir_vals          device, NTRACE; on Dirichlet dofs receives the
 *                     projected Dirichlet update (template.h:1592-1691),
 *                     other entries untouched
 * Dirichlet rows carry a unit diagonal and a zero right-hand side.
 * stream: hipStream_t (NULL for the default stream). */
int hdg_condense(hdg_ctx *ctx, const double *x_cell, const double *x_trace,
                 int thermodyn, double *csr_vals, double *rhs,
                 double *dir_vals, void *stream);

/* Pass 2: reconstruction of the cell update from the full trace update
 * (nfaces*3*nf, Dirichlet values included).  Replaces
 * assemble_system(true, thermodyn, *) (template.h:2357-2366, 2642-2643,
 * 2688-2697).  dx_cell: device, ncells*NC, disabled components zero. */
int hdg_reconstruct(hdg_ctx *ctx, const double *x_cell,
                    const double *x_trace, const double *dx_trace,
                    int thermodyn, double *dx_cell, void *stream);

/* Condensed blocks of selected cells, the condensed counterpart of the
 * reference's get_local_cell_system (src/np_solver.cpp:2337-2445):
 * cells (device, ncell ints); S (device, ncell*T*T, row-major, local trace
 * order face-major then component then basis) and g (device, ncell*T).
 * Dirichlet rows and columns are zero, as in the reference. */
int hdg_cell_systems(hdg_ctx *ctx, const double *x_cell,
                     const double *x_trace, int thermodyn, const int *cells,
                     int ncell, double *S, double *g, void *stream);

/* The uncondensed local system of selected cells: the counterpart of the
 * reference's get_local_cell_system (src/np_solver.cpp:2337-2445), which runs
 * the per-cell operator without its Schur complement.  Per listed cell, in
 * the reference's local order (N = NE*(dim+1)*NB cell dofs, flux axes then
 * scalar per enabled component; T trace dofs as in hdg_cell_systems):
 *   A  (N x N), CT (N x T), TC (T x N, stored negated as the reference does),
 *   TT (T x T), c (N), t (T),
 * row-major, device arrays of ncell blocks each, after the LocalCondenser
 * has moved the Dirichlet columns of CT into c.  Synchronises the stream
 * (a temporary workspace is allocated and freed). */
int hdg_local_system(hdg_ctx *ctx, const double *x_cell, const double *x_trace,
                     int thermodyn, const int *cells, int ncell, double *A,
                     double *CT, double *TC, double *TT, double *c, double *t,
                     void *stream);

/* Terminal current through the faces of one boundary id (outward normal):
 * out (device, 2 doubles) receives the integrals of the HDG numerical fluxes
 *   F_n =  n mu_n E.nu - U_T mu_n W_n.nu - tau_n (n - lambda_n)
 *   F_p = -p mu_p E.nu - U_T mu_p W_p.nu - tau_p (p - lambda_p)
 * in dimensionless units; the physical current leaving through the boundary
 * is mu_mag N k_B T_mag l^(dim-2) (out[0] - out[1]) (A per metre of depth in
 * 2D, A in 3D).  The reference has no contact-current integral (it projects
 * Jn, Jp fields, src/derivative_quantities.cpp:1080-1115); the fluxes are the
 * ones of its trace equations (include/assembling_methods.template.h:
 * 1311-1447). */
int hdg_terminal_current(hdg_ctx *ctx, const double *x_cell,
                         const double *x_trace, int boundary_id, double *out,
                         void *stream);

/* Newton norms: out[v] = max_k |vectors[v][k]| for four device vectors of
 * the given lengths (out and scratch in device memory, scratch 1024
 * doubles); two kernels, no memset, deterministic.  The L-infinity norms of
 * the reference's convergence test (src/np_solver.cpp:1521-1560). */
int hdg_max_abs4(const double *const *vectors, const long long *lengths,
                 double *scratch, double *out, void *stream);

/* hdg_max_abs4 without the scratch argument: the partial maxima live in a
 * static device array of the library (4 x 256 doubles, 8 KiB, one copy per
 * GPU, held until the process ends), so the call allocates nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops).  The scratch is fixed
 * in size, so vectors of any length are accepted; a negative length is an
 * error.  Every call on one GPU shares that GPU's array: calls that overlap
 * on one GPU (different streams or threads) race. */
int hdg_max_abs4_noalloc(const double *const *vectors, const long long *lengths,
                         double *out, void *stream);

/* ---- Batched pass 1 over several meshes --------------------------------
 * Pass 1 (hdg_condense) of several contexts (2D with the same enabled
 * components: degree 1 and 2 mixed, or all of degree 3 with V, n and p) in
 * one persistent launch per colour; each context's outputs are exactly
 * those of its own hdg_condense, in the contexts' precision (they must
 * share it).  A degree-3 batch runs the full operator only.  The autotuner picks the tile (waves per workgroup
 * of the degree 1-2 kernel: 4 or 8, one cell per wave) and the persistent
 * grid (workgroups per CU) by timing each candidate on the first call
 * (hdg_batch_info reports the choice, hdg_batch_tuning every candidate).
 * Pointer arguments are host arrays of n device pointers, in the order of
 * ctxs.  A call with the
 * buffers and data of the previous call reuses its uploaded parameter
 * blocks, so it can be captured into a HIP graph (after an eager first
 * call); a captured call with other buffers returns -2.  No counterpart in
 * the reference (one mesh per NPSolver). */
typedef struct hdg_batch hdg_batch;

int  hdg_batch_create(hdg_ctx *const *ctxs, int n, hdg_batch **out);
void hdg_batch_destroy(hdg_batch *batch);
int  hdg_batch_condense(hdg_batch *batch, int thermodyn,
                        const double *const *x_cell,
                        const double *const *x_trace, double *const *csr_vals,
                        double *const *rhs, double *const *dir_vals,
                        void *stream);
/* out = {cells, persistent workgroups chosen, CUs, tile chosen (waves per
 * workgroup; pairs of waves for degree 3), candidates timed, ms of a
 * two-colour pass with the choice} */
int  hdg_batch_info(const hdg_batch *batch, double out[6]);
/* The autotuner's candidates: out[3k..3k+2] = {tile, grid, ms} for k <
 * min(cap, candidates); returns the number of candidates (0 before the
 * first hdg_batch_condense). */
int  hdg_batch_tuning(const hdg_batch *batch, double *out, int cap);

/* ---- Global trace solve ------------------------------------------------
 * Sparse direct LU of the restricted trace system, on the GPU: a nested-
 * dissection multifrontal factorisation over the structured cell grid
 * (dense fronts in HBM, eliminated per tree level in panels of whole faces
 * by batched HIP kernels and rocBLAS gemm).  Replaces
 * the reference's SparseDirectUMFPACK factorize + solve of system_matrix
 * (src/np_solver.cpp:1326-1354).  A solver is built for the context's mesh
 * and its enabled components at creation; rebuild it after
 * hdg_set_enabled. */
typedef struct hdg_trace_solver hdg_trace_solver;

enum
{
  HDG_TS_FLOPS = 0,  /* factorisation flops (dense fronts, padded) */
  HDG_TS_BYTES,      /* device bytes held by the fronts (= the factors) */
  HDG_TS_LEVELS,     /* tree levels (batched launches per sweep) */
  HDG_TS_MAX_FRONT,  /* largest front dimension */
  HDG_TS_NODES,      /* tree nodes */
  HDG_TS_NINFO
};

/* Builds the plan and its index maps on the device, then times rocBLAS's
 * GEMM solutions for the largest Schur updates (at most about 0.3 s of wall
 * time per process for shapes not met before; the fastest is kept). */
int  hdg_trace_solver_create(const hdg_ctx *ctx, hdg_trace_solver **out);
void hdg_trace_solver_destroy(hdg_trace_solver *solver);
int  hdg_trace_solver_info(const hdg_trace_solver *solver, double out[HDG_TS_NINFO]);
/* Factorise the matrix whose values (device, NNZ, hdg_csr_pattern order)
 * hdg_condense wrote; asynchronous on stream. */
int  hdg_trace_factor(hdg_trace_solver *solver, const double *csr_vals,
                      void *stream);
/* Synchronises stream; -5 if a front met an exactly zero pivot. */
int  hdg_trace_factor_status(hdg_trace_solver *solver, void *stream);
/* x = A^-1 rhs for the last factorised A (device vectors, NTRACE). */
int  hdg_trace_solve(hdg_trace_solver *solver, const double *rhs, double *x,
                     void *stream);
/* Host-only check of the solver's plan for a dim-D n_cells grid with
 * dofs_per_face trace dofs per face: every index against its buffer, then the
 * plan's multifrontal arithmetic in host loops against dense elimination, on
 * csr_vals (host, NNZ, hdg_csr_pattern order) or, when csr_vals is NULL, on a
 * random diagonally dominant matrix in the trace pattern.  out = {relative
 * error, tree nodes, levels}.  Small grids only (about 3000 trace dofs at
 * most).  No counterpart in the reference (UMFPACK is a library call
 * there). */
int  hdg_trace_plan_selftest(int dim, const int *n_cells, int dofs_per_face,
                             unsigned seed, const double *csr_vals, double out[3]);

/* Host-only sizes of the trace solver's plan for a dim-D n_cells grid with
 * dofs_per_face trace dofs per face, without building its index maps (cheap
 * at any size): out = {front entries (doubles of factor storage), work-vector
 * entries, pivot entries, factorisation FLOPs, largest front, tree nodes}.
 * The device memory hdg_trace_solver_create would take is about
 * 8*out[0] + 8*out[1] + 4*out[2] + 8*NNZ bytes.  For admission by memory
 * budget; no counterpart in the reference. */
int  hdg_trace_plan_size(int dim, const int *n_cells, int dofs_per_face, double out[6]);

/* The iterative branch of the reference's solve_linear_problem
 * (src/np_solver.cpp:1326-1354, parameters->iterative_linear_solver): GMRES
 * restarted every `restart` iterations (deal.II's default keeps 30 vectors,
 * i.e. 28), no preconditioner, zero initial guess, stopping when the
 * residual estimate is at most tol (the reference: 1e-10 |rhs|_2) or with
 * -6 after max_it iterations (the reference: 10 n).  rowptr (n+1), colind,
 * vals: a device CSR matrix of order n (hdg_csr_pattern layout); rhs, x:
 * device vectors of n entries.  iterations and residual (host) receive the
 * iteration count and the last residual estimate.  Synchronises stream;
 * allocates a workspace of (restart + 2) n doubles for the call.  -5 if the
 * residual is not a number. */
int  hdg_gmres_solve(const long long *rowptr, const int *colind, const double *vals,
                     long long n, const double *rhs, double *x, int restart,
                     double tol, long long max_it, long long *iterations,
                     double *residual, void *stream);

#ifdef __cplusplus
}
#endif
#endif
