// hdg_kernels.h -- kernel parameter block shared by the host layer and the
// gfx950 kernels of the per-cell HDG path.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace hdg
{
  struct DevTables
  {
    const double *JxWc, *phic, *gradc, *phif, *JxWf, *xi, *Bz, *Bt;
    const double *LDc, *LDf, *LCc, *LCf, *PhiMinvC, *PhiMinvF;
    const double *Minv, *MinvD, *MinvC, *Mfinv;
  };

  struct KParams
  {
    // mesh
    int       n[3];
    long long face_off[4];
    long long face_dims[3][3];
    // enabled components
    int en[3];
    int ne;
    int encomp[3];
    int thermo;
    // physics (dimensionless)
    double eps, mun, mup, UT, tau[3];
    int    rec;
    double rec0, rec1, rec2, n_rf, rec_rf;
    double kTq, ec_r, ev_r, nc, nv;
    int    bc[6][3];
    double sdrift; // +1 reference linearisation, -1 consistent
    int    mixed;  // 1: FP32 elimination + FP64 refinement in the register kernel's pass 1
    int    deg;    // polynomial degree (batched launches mix degrees)
    int    flux_type; // 0 use_cell, 1 use_trace, 2 qiu_shi_stabilization
    int    phi_lin;   // phi_linearize: Boltzmann V rows at frozen quasi-Fermi potentials
    DevTables tb;
    // per-mesh data
    const double *doping;   // ncells*QC
    const double *dirv;     // nfaces*3*QF
    const double *neuv;     // nfaces*3*QF
    const double *tau_cf;   // ncells*2d*3 per-(cell, face, comp) tau, or null: tau[]
    // scatter map
    const long long     *blkptr;   // nfaces
    const int           *nbk;      // nfaces
    const unsigned char *cslot;    // ncells*NFC*NFC
    // state
    const double *xc, *xt, *dxt;
    // outputs
    double *vals, *rhs, *dirout, *dxc, *Sout, *gout;
    // global scratch of the generic kernel where its reduced system does not
    // fit LDS (generic_scratch_bytes)
    double *rgscr;
  };

  // launchers (hdg_kernels.hip)
  hipError_t launch_condense(int dim, int deg, const KParams &p, const int *cells,
                             int ncell, int color, int mode, hipStream_t s);
  hipError_t launch_reconstruct(int dim, int deg, const KParams &p, int ncells,
                                hipStream_t s);
  bool supported(int dim, int deg);
  // device bytes of the generic kernel's global scratch for a shape (0 when
  // its reduced system fits LDS)
  size_t generic_scratch_bytes(int dim, int deg);
  // register-resident pass 1 (hdg_fast2d.hip): 2D, degree 1-2, all three
  // components or the V-only thermodynamic-equilibrium operator
  bool       fast_supported(int dim, int deg, int ne, int thermo, const int en[3]);
  hipError_t launch_fast_condense(int deg, const KParams &p, const int *cells, int ncell,
                                  int color, int mode, hipStream_t s);
  // out[v] = max |a[v]_k| (device); scratch: 4*256 doubles
  hipError_t launch_norms(const double *const a[4], const long long n[4], double *scratch,
                          double *out, hipStream_t s);
  // the same with the partials in a static device array (allocates nothing;
  // concurrent calls on one GPU race)
  hipError_t launch_norms_static(const double *const a[4], const long long n[4], double *out,
                                 hipStream_t s);
  // terminal current through boundary id bid: out[0] = integral of the
  // numerical flux of n, out[1] of p (device, dimensionless)
  hipError_t launch_current(int dim, int nb, int nf, int qf, const KParams &p, int bid,
                            double *out, hipStream_t s);
  // pass 1 of several meshes in one persistent launch per colour (register
  // kernel, degree 1 or 2): plist (device) holds one parameter block per
  // mesh, chunk_dev / items (device) the work list of nchunks chunks of
  // batch_waves_per_group() cells; tile: waves per workgroup (4 or 8)
  int        batch_waves_per_group();
  hipError_t launch_batch_condense(int ne, int thermo, int mixed, const KParams *plist,
                                   const int *chunk_dev, const int *items, int nchunks,
                                   int color, int grid, int tile, hipStream_t s);
  // register-resident pass 1 for 3D degree 1 (hdg_fast3d.hip): two columns
  // per lane, all three components or the V-only equilibrium operator
  bool       fast3_supported(int dim, int deg, int ne, int thermo, const int en[3]);
  hipError_t launch_fast3_condense(const KParams &p, const int *cells, int ncell, int color,
                                   int mode, hipStream_t s);
  // register-resident pass 1 for 2D degree 3 with V/n/p (hdg_fast2dp.hip):
  // a pair of waves per cell
  bool       fast2p_supported(int dim, int deg, int ne, int thermo, const int en[3]);
  hipError_t launch_fast2p_condense(const KParams &p, const int *cells, int ncell, int color,
                                    int mode, hipStream_t s);
  // its pass 2 (all cells in order, dx_cell out)
  hipError_t launch_fast2p_reconstruct(const KParams &p, int ncells, hipStream_t s);
  // the same pass 1 over several degree-3 meshes in one persistent launch
  // per colour (work list in chunks of batch_pairs_per_group() cells), FP64
  // or, at precision 2 (mixed == 2), FP32 + refinement
  int        batch_pairs_per_group();
  hipError_t launch_batch_pair(const KParams *plist, const int *chunk_dev, const int *items,
                               int nchunks, int color, int grid, int mixed, hipStream_t s);
  // register-resident pass 2 (same shapes as launch_fast_condense)
  hipError_t launch_fast_reconstruct(int deg, const KParams &p, int ncells, hipStream_t s);
  // the uncondensed local system of listed cells (hdg_local.hip); G: a
  // workspace of local_system_workspace(dim, deg, ne) bytes per cell
  size_t     local_system_workspace(int dim, int deg, int ne);
  hipError_t launch_local_system(int dim, int deg, const KParams &p, const int *cells, int ncell,
                                 double *G, double *A, double *CT, double *TC, double *TT,
                                 double *c, double *t, hipStream_t s);
  // pass 2 on the 3D degree-1 register kernel (same operators as its pass 1)
  hipError_t launch_fast3_reconstruct(const KParams &p, int ncells, hipStream_t s);
  // restarted GMRES, no preconditioner, zero initial guess (hdg_gmres.hip):
  // x = A^-1 b for the device CSR matrix (rp, ci, vals) of order n; stops
  // when the Arnoldi residual estimate is at most tol.  Synchronises s.
  // 0, or -3 (device failure), -5 (residual not a number), -6 (max_it
  // iterations without convergence), with err set
  int        gmres_solve(const long long *rp, const int *ci, const double *vals, long long n,
                         const double *b, double *x, int restart, double tol, long long max_it,
                         long long *iterations, double *residual, hipStream_t s,
                         std::string &err);
} // namespace hdg
