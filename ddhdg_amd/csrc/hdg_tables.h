// hdg_tables.h -- host-built reference-cell tables for the HDG kernels.
//
// One set of tables per mesh: every cell of a structured box mesh has the
// same shape, so the quadrature weights, shape values and the flux-lifting
// operators below are shared by all cells and live in device memory once.
//
// Flux elimination.  In every cell the rows of the flux test functions
// (q1, q2, q3 in include/assembling_methods.template.h:269-532, 885-1019)
// read, for component c and axis a,
//     M W_{c,a} - D_a u_c + sum_f C_{a,f} lambda_{c,f} = rhs_{q,c,a}
// with M the scalar DG mass matrix, D_a(i,j) = (phi_j, d_a phi_i) and
// C_{a,f}(i,m) = <xi_m, phi_i nu_a>_f: geometry only, independent of the
// Newton state.  So W = M^{-1}(rhs_q + D_a u - C_a lambda) exactly, and the
// lifting operators M^{-1}D_a, M^{-1}C_{a,f}, M^{-1} evaluated at the
// quadrature points are precomputed here.  The kernels then condense the
// reduced system on (scalar dofs, trace dofs) only; its Schur complement is
// the same matrix S = TT - TC A^{-1} CT as the reference's.
#pragma once
#include <vector>

namespace hdg
{
  struct HostTables
  {
    int dim, deg, nb, nf, qc, qf, nfc, sz;
    std::vector<double> JxWc;     // qc
    std::vector<double> phic;     // qc*nb
    std::vector<double> gradc;    // qc*nb*dim (physical)
    std::vector<double> phif;     // nfc*qf*nb  cell basis at face qps
    std::vector<double> JxWf;     // nfc*qf
    std::vector<double> xi;       // qf*nf      trace basis at face qps
    std::vector<double> Bz;       // sz*nb      z-row test table (x JxW)
    std::vector<double> Bt;       // nfc*qf*nf  trace-row test table (x JxW)
    std::vector<double> LDc;      // qc*dim*nb  lifted flux of scalar dof j
    std::vector<double> LDf;      // nfc*qf*nb  its normal component on faces
    std::vector<double> LCc;      // nfc*qc*nf  lifted flux (axis of f) of trace dof (f,m)
    std::vector<double> LCf;      // nfc*nfc*qf*nf  normal component on face g
    std::vector<double> PhiMinvC; // qc*nb      phi(q) M^{-1}
    std::vector<double> PhiMinvF; // nfc*qf*nb  phi(g,s) M^{-1} (normal sign folded in)
    std::vector<double> Minv;     // nb*nb
    std::vector<double> MinvD;    // dim*nb*nb
    std::vector<double> MinvC;    // nfc*nb*nf
    std::vector<double> Mfinv;    // dim*nf*nf  face mass inverse per normal axis
    std::vector<double> qx1;      // 1D gauss points on [0,1]
    int                 tang[3][2];
  };

  // h: cell size per axis
  bool build_tables(int dim, int deg, const double h[3], HostTables &T);
} // namespace hdg
