// hdg_kernels.hip -- gfx950 kernels of the per-cell HDG path (generic form).
//
// One workgroup (256 threads = 4 wave64) per cell.  The cell's reduced
// bordered system
//        [ A'  CT' | c' ]      rows: scalar test functions (ne*nb)
//        [ TC' TT' | t' ]            trace test functions (2d*ne*nf)
// is built in LDS after the flux unknowns have been eliminated with the
// precomputed lifting operators (hdg_tables.h), then ne*nb steps of
// Gaussian elimination with row-scaled partial pivoting reduce the trace
// rows to [ S | g ] = [TT - TC A^-1 CT | t - TC A^-1 c], the same blocks the
// reference builds with gauss_jordan + mmult
// (include/assembling_methods.template.h:2654-2668).
//
// The reference's terms map onto three per-quadrature-point forms applied
// to a column's "bundle" (its scalar values u, lifted fluxes W and trace
// values lambda at the quadrature points):
//   z rows, cell qp:  phi_i * alpha_c + grad phi_i . beta_c
//        add_cell_products_to_cc_matrix/rhs (template.h:269-760)
//   z rows, face qp:  phi_i * F_c      add_border_products_to_cc_matrix/rhs,
//        assemble_ct_matrix, add_ct_matrix_terms_to_cc_rhs (:885-1141,
//        :1737-2118)
//   trace rows:       xi_i * F_c       assemble_tc_matrix, assemble_tt_matrix,
//        add_tc/tt_matrix_terms_to_tt_rhs (:1145-1588), apply_nbc_on_face
// where F_c is the HDG numerical flux of component c: its drift part uses
// the cell value (use_cell), the trace value (use_trace) or, for
// qiu_shi_stabilization, the cell value where E.nu > 0 and the trace value
// elsewhere (template.h:983-1009, 1249-1263, 1417-1435, 1523-1536,
// 1865-1903, 2036-2071).  tau is per face: FixedTauComputer or
// CellFaceTauComputer (include/np_solver_parameters.h:92-115, 157-236).
#include <mutex>
#include <algorithm>
#include <utility>

#include "hdg_device.h"
#include "hdg_kernels.h"

namespace hdg
{
  namespace
  {
    constexpr int NT = 256;

    template <int DIM, int DEG>
    struct Sh
    {
      static constexpr int K1    = DEG + 1;
      static constexpr int NB    = DIM == 2 ? K1 * K1 : K1 * K1 * K1;
      static constexpr int NF    = DIM == 2 ? K1 : K1 * K1;
      static constexpr int QC    = NB;
      static constexpr int QF    = NF;
      static constexpr int NFC   = 2 * DIM;
      static constexpr int NCD   = 3 * (DIM + 1) * NB;
      static constexpr int MAXNS = 3 * NB;
      static constexpr int MAXT  = NFC * 3 * NF;
      static constexpr int MAXR  = MAXNS + MAXT;
      static constexpr int LD    = MAXR + 2;
      static constexpr int SZ    = QC * (1 + DIM) + NFC * QF;
      // the reduced system R and the column data G (doubles, from the base
      // rg_base returns): in LDS with the rest, or, when they do not fit
      // (3D degree 2: 794 KB), in a global scratch slice per workgroup
      static constexpr bool BIG   = DIM == 3 && DEG >= 2;
      static constexpr int oR     = 0;
      static constexpr int oG     = oR + MAXR * LD; // unused when BIG
      static constexpr int RGD    = (MAXR + (BIG ? 0 : SZ)) * LD;
      // LDS layout (doubles)
      static constexpr int oXS    = BIG ? 0 : RGD;
      static constexpr int oXT    = oXS + NCD;
      static constexpr int oU0    = oXT + NFC * 3 * NF;
      static constexpr int oW0    = oU0 + 3 * QC;
      static constexpr int oU0F   = oW0 + 3 * DIM * QC;
      static constexpr int oW0F   = oU0F + 3 * NFC * QF;
      static constexpr int oTR0   = oW0F + 3 * NFC * QF;
      static constexpr int oQD    = oTR0 + 3 * NFC * QF; // 6*QC point data
      static constexpr int oRHSQ  = oQD + 6 * QC;
      static constexpr int oWRC   = oRHSQ + 3 * DIM * NB;
      static constexpr int oWRF   = oWRC + 3 * DIM * QC;
      static constexpr int oLAMF  = oWRF + 3 * NFC * QF;
      static constexpr int oDEL   = oLAMF + 3 * NFC * QF;
      static constexpr int oLUPD  = oDEL + NFC * 3 * NF;
      static constexpr int oRS    = oLUPD + NFC * 3 * NF;
      static constexpr int oMUL   = oRS + MAXR;
      static constexpr int oSOL   = oMUL + MAXR;
      static constexpr int oTAU   = oSOL + MAXNS;   // tau per (face, comp)
      // BIG: the scalar block (MAXNS x AL) factorised in LDS, later the
      // staging tiles of the Schur product
      static constexpr int AL     = MAXNS + 1;
      static constexpr int oAW    = oTAU + NFC * 3;
      static constexpr int NDBL   = oAW + (BIG ? MAXNS * AL : 0);
      // ints after the doubles
      static constexpr int iUSED  = 0;
      static constexpr int iPIV   = iUSED + MAXR;
      static constexpr int iDIR   = iPIV + MAXNS;
      static constexpr int iNEU   = iDIR + NFC * 3;
      static constexpr int iFACE  = iNEU + NFC * 3;
      static constexpr int iBID   = iFACE + 2 * NFC; // face index as 2 ints
      static constexpr int NINT   = iBID + NFC;
      static constexpr size_t bytes = NDBL * sizeof(double) + NINT * sizeof(int);
      static_assert(bytes <= 160 * 1024, "LDS part of the layout");
      static_assert(!BIG || 2 * bytes <= 160 * 1024, "two workgroups per CU");
    };

    typedef const __attribute__((address_space(4))) double ctabg;

    // f(std::integral_constant<int, I>) for I = 0, 1, ...: loops whose
    // index selects a register of a per-thread array stay expanded
    template <class F, int... I>
    __device__ __forceinline__ void static_for(F &&f, std::integer_sequence<int, I...>)
    {
      (f(std::integral_constant<int, I>{}), ...);
    }

    // workgroups of a generic launch whose R and G live in global scratch
    constexpr int GRID_BIG = 512;
    // the thread index behind an empty asm: what is derived from it is
    // formed where it is used instead of being hoisted out of the
    // persistent cell loop and kept live (spilled) across it
    __device__ __forceinline__ int opaque_tid()
    {
      int t = threadIdx.x;
      asm volatile("" : "+v"(t));
      return t;
    }

    // base of R and G for this workgroup
    template <int DIM, int DEG>
    __device__ __forceinline__ double *rg_base(const KParams &p, double *sm)
    {
      using S = Sh<DIM, DEG>;
      if constexpr (S::BIG)
        return p.rgscr + (size_t)blockIdx.x * S::RGD;
      else
        return sm;
    }

    // ------------------------------------------------------------------
    // Shared first half of both passes: load, state at quadrature points,
    // flux-row residual and its lifting, Dirichlet projection (pass 1) or
    // lifted trace update (pass 2), and the reduced system in LDS.
    // pass2: rows = z rows only, columns = scalar dofs + one rhs column.
    template <int DIM, int DEG>
    __device__ void build_reduced(const KParams &p, long long cell, bool pass2,
                                  double *sm, int *si, int &ns_out, int &T_out)
    {
      using S           = Sh<DIM, DEG>;
      constexpr int NB  = S::NB, NF = S::NF, QC = S::QC, QF = S::QF;
      constexpr int NFC = S::NFC, LD = S::LD, SZ = S::SZ;
      const int     tid = opaque_tid();
      double *rg = rg_base<DIM, DEG>(p, sm);
      double *R = rg + S::oR, *G = rg + S::oG, *xs = sm + S::oXS, *xt = sm + S::oXT;
      double *u0 = sm + S::oU0, *W0 = sm + S::oW0, *u0f = sm + S::oU0F;
      double *W0f = sm + S::oW0F, *tr0 = sm + S::oTR0, *qd = sm + S::oQD;
      double *rhsq = sm + S::oRHSQ, *Wrc = sm + S::oWRC, *Wrf = sm + S::oWRF;
      double *lamf = sm + S::oLAMF, *del = sm + S::oDEL, *lupd = sm + S::oLUPD;
      double *tauv = sm + S::oTAU;
      int    *isdir = si + S::iDIR, *isneu = si + S::iNEU, *bidv = si + S::iBID;
      long long *facev = reinterpret_cast<long long *>(si + S::iFACE);
      // point data: r, drn, drp, phi'_V, phi_rhs, doping
      double *qr = qd, *qdrn = qd + QC, *qdrp = qd + 2 * QC, *qphd = qd + 3 * QC;
      double *qphr = qd + 4 * QC, *qdop = qd + 5 * QC;
      const DevTables &tb = p.tb;

      const int ne   = p.ne;
      const int ns   = ne * NB;
      const int tdf  = ne * NF;
      const int T    = NFC * tdf;
      const int nrow = pass2 ? ns : ns + T;
      const int ncol = pass2 ? ns + 1 : ns + T + 1;
      ns_out         = ns;
      T_out          = T;

      int ic[3];
      ic[0] = (int)(cell % p.n[0]);
      ic[1] = (int)((cell / p.n[0]) % p.n[1]);
      ic[2] = (int)(cell / ((long long)p.n[0] * p.n[1]));

      // ---- A: load state and boundary flags
      if (tid < NFC)
        {
          const int lf = tid, a = lf >> 1, sd = lf & 1;
          facev[lf]    = face_of(p, ic[0], ic[1], ic[2], lf);
          const bool bnd = sd ? (ic[a] == p.n[a] - 1) : (ic[a] == 0);
          const int  bid = bnd ? lf : -1;
          bidv[lf]       = bid;
          for (int c = 0; c < 3; ++c)
            {
              const int bt         = (bid >= 0 && p.en[c]) ? p.bc[bid][c] : 0;
              isdir[lf * 3 + c]    = bt == 1;
              isneu[lf * 3 + c]    = bt == 2 && !p.thermo;  // zero Neumann data at equilibrium
            }
        }
      for (int i = tid; i < S::NCD; i += NT)
        xs[i] = p.xc[cell * S::NCD + i];
      for (int i = tid; i < QC; i += NT)
        qdop[i] = p.doping[cell * QC + i];
      if (tid < NFC * 3)
        tauv[tid] = p.tau_cf ? p.tau_cf[cell * NFC * 3 + tid] : p.tau[tid % 3];
      __syncthreads();
      for (int i = tid; i < NFC * 3 * NF; i += NT)
        {
          const int lf = i / (3 * NF), r = i % (3 * NF);
          xt[i]        = p.xt[facev[lf] * 3 * NF + r];
          if (pass2)
            lupd[i] = p.dxt[facev[lf] * 3 * NF + r];
        }
      __syncthreads();

      // ---- B: state at quadrature points
      // cell: u0[c][q], W0[c][a][q]
      for (int e = tid; e < 3 * (DIM + 1) * QC; e += NT)
        {
          const int c = e / ((DIM + 1) * QC), r = e % ((DIM + 1) * QC);
          const int blk = r / QC, q = r % QC; // blk < DIM: flux axis, == DIM scalar
          const double *x = xs + c * (DIM + 1) * NB + blk * NB;
          double        v = 0;
          for (int j = 0; j < NB; ++j)
            v += x[j] * tb.phic[q * NB + j];
          if (blk == DIM)
            u0[c * QC + q] = v;
          else
            W0[(c * DIM + blk) * QC + q] = v;
        }
      // faces: u0f, W0f (normal comp), tr0; [c][g][s]
      for (int e = tid; e < 3 * NFC * QF; e += NT)
        {
          const int c = e / (NFC * QF), g = (e / QF) % NFC, s = e % QF;
          const int a = g >> 1;
          const double nu = (g & 1) ? 1.0 : -1.0;
          const double *xsc = xs + c * (DIM + 1) * NB;
          const double *ph  = tb.phif + (g * QF + s) * NB;
          double        u = 0, w = 0, t = 0;
          for (int j = 0; j < NB; ++j)
            {
              u += xsc[DIM * NB + j] * ph[j];
              w += xsc[a * NB + j] * ph[j];
            }
          for (int m = 0; m < NF; ++m)
            t += xt[(g * 3 + c) * NF + m] * tb.xi[s * NF + m];
          u0f[e] = u;
          W0f[e] = nu * w;
          tr0[e] = t;
          if (pass2)
            {
              double l = 0;
              for (int m = 0; m < NF; ++m)
                l += lupd[(g * 3 + c) * NF + m] * tb.xi[s * NF + m];
              lamf[e] = l;
            }
          else
            lamf[e] = 0;
        }
      __syncthreads();

      // ---- C: point data and flux-row residual
      for (int q = tid; q < QC; q += NT)
        {
          double r = 0, a = 0, b = 0;
          if (p.en[1] || p.en[2])
            recomb(p, u0[QC + q], u0[2 * QC + q], r, a, b);
          qr[q]   = r;
          qdrn[q] = a;
          qdrp[q] = b;
          if (p.thermo || p.phi_lin)
            {
              // template.h:442-459 and 677-693: phi_n = phi_p = 0 in
              // thermodynamic equilibrium; with phi_linearize the
              // quasi-Fermi potentials of the current n and p
              // (template.h:229-264, include/solver.h:343-375), zero for a
              // disabled carrier (the scratch's zero-initialised vectors)
              const double V0 = u0[q], kT = p.kTq;
              double       pn = 0, pq = 0;
              if (!p.thermo)
                {
                  if (p.en[1])
                    pn = V0 - p.ec_r - kT * log(u0[QC + q] * p.n_rf / p.nc);
                  if (p.en[2])
                    pq = V0 - p.ev_r + kT * log(u0[2 * QC + q] * p.n_rf / p.nv);
                }
              const double e1 = exp((p.ev_r - V0 + pq) / kT);
              const double e2 = exp((V0 - pn - p.ec_r) / kT);
              qphd[q] = -1 / p.n_rf * (p.nv / kT * e1 + p.nc / kT * e2);
              qphr[q] = 1 / p.n_rf * (p.nv * e1 - p.nc * e2);
            }
          else
            qphd[q] = qphr[q] = 0;
        }
      // rhs_q[c][a][i] = sum_q (u0 d_a phi_i - W0_a phi_i) JxW
      //                 - sum_{faces of axis a} tr0 phi_i nu JxW
      for (int e = tid; e < 3 * DIM * NB; e += NT)
        {
          const int c = e / (DIM * NB), a = (e / NB) % DIM, i = e % NB;
          double    v = 0;
          if (p.en[c])
            {
              for (int q = 0; q < QC; ++q)
                v += (u0[c * QC + q] * tb.gradc[(q * NB + i) * DIM + a] -
                      W0[(c * DIM + a) * QC + q] * tb.phic[q * NB + i]) *
                     tb.JxWc[q];
              for (int sd = 0; sd < 2; ++sd)
                {
                  const int    g  = 2 * a + sd;
                  const double nu = sd ? 1.0 : -1.0;
                  for (int s = 0; s < QF; ++s)
                    v -= tr0[(c * NFC + g) * QF + s] * tb.phif[(g * QF + s) * NB + i] *
                         nu * tb.JxWf[g * QF + s];
                }
            }
          rhsq[e] = v;
        }
      // Dirichlet update: L2 projection of g/rf - lambda0 on the face
      // (apply_dbc_on_face, template.h:1592-1691); zero in thermodynamic
      // equilibrium (template.h:1662-1665)
      if (!pass2)
        for (int e = tid; e < NFC * 3 * NF; e += NT)
          {
            const int g = e / (3 * NF), c = (e / NF) % 3, m = e % NF;
            double    v = 0;
            if (isdir[g * 3 + c] && !p.thermo)
              {
                const int a = g >> 1;
                for (int n2 = 0; n2 < NF; ++n2)
                  {
                    double b = 0;
                    for (int s = 0; s < QF; ++s)
                      b += tb.xi[s * NF + n2] * tb.JxWf[g * QF + s] *
                           (p.dirv[(facev[g] * 3 + c) * QF + s] -
                            tr0[(c * NFC + g) * QF + s]);
                    v += tb.Mfinv[(a * NF + m) * NF + n2] * b;
                  }
              }
            del[e] = v;
          }
      __syncthreads();

      // ---- D: lifted residual flux W^r (plus lifted trace update, pass 2)
      for (int e = tid; e < 3 * DIM * QC; e += NT)
        {
          const int c = e / (DIM * QC), a = (e / QC) % DIM, q = e % QC;
          double    v = 0;
          if (p.en[c])
            {
              for (int i = 0; i < NB; ++i)
                v += tb.PhiMinvC[q * NB + i] * rhsq[(c * DIM + a) * NB + i];
              if (pass2)
                for (int sd = 0; sd < 2; ++sd)
                  {
                    const int f = 2 * a + sd;
                    for (int m = 0; m < NF; ++m)
                      v += tb.LCc[(f * QC + q) * NF + m] * lupd[(f * 3 + c) * NF + m];
                  }
            }
          Wrc[e] = v;
        }
      for (int e = tid; e < 3 * NFC * QF; e += NT)
        {
          const int c = e / (NFC * QF), g = (e / QF) % NFC, s = e % QF;
          const int a = g >> 1;
          double    v = 0;
          if (p.en[c])
            {
              for (int i = 0; i < NB; ++i)
                v += tb.PhiMinvF[(g * QF + s) * NB + i] * rhsq[(c * DIM + a) * NB + i];
              if (pass2)
                for (int sd = 0; sd < 2; ++sd)
                  {
                    const int f = 2 * a + sd;
                    for (int m = 0; m < NF; ++m)
                      v += tb.LCf[((f * NFC + g) * QF + s) * NF + m] *
                           lupd[(f * 3 + c) * NF + m];
                  }
            }
          Wrf[e] = v;
        }
      __syncthreads();

      // ---- E: reduced system, one row block (component) at a time
      const double eps = p.eps, mun = p.mun, mup = p.mup, UT = p.UT;
      const bool   phim = p.thermo || p.phi_lin;
      const int    ft   = p.flux_type;
      for (int ci = 0; ci < ne; ++ci)
        {
          const int c = p.encomp[ci];
          // column m: kind 0 = scalar dof j of component cc, 1 = trace dof j
          // of component cc on face f, 2 = the rhs
          auto decode = [&](int m, int &kind, int &cc, int &j, int &f) {
              cc = 0;
              j  = 0;
              f  = 0;
              if (m < ns)
                {
                  kind = 0;
                  cc   = p.encomp[m / NB];
                  j    = m % NB;
                }
              else if (!pass2 && m < ns + T)
                {
                  kind         = 1;
                  const int lt = m - ns;
                  f            = lt / tdf;
                  cc           = p.encomp[(lt % tdf) / NF];
                  j            = lt % NF;
                }
              else
                kind = 2;
          };
          // the bundle entry of column m at quadrature slot `slot` (cell
          // points: value and flux axes; then face points)
          auto colval = [&](int slot, int kind, int cc, int j, int f) -> double {
              double val = 0;
              if (slot < QC * (1 + DIM))
                {
                  const int q = slot / (1 + DIM), cs = slot % (1 + DIM);
                  double    uu[3] = {0, 0, 0};
                  if (kind == 0)
                    uu[cc] = tb.phic[q * NB + j];
                  if (cs == 0)
                    {
                      if (c == 0)
                        val = phim ? -qphd[q] * uu[0] : uu[1] - uu[2];
                      else
                        val = -(qdrn[q] * uu[1] + qdrp[q] * uu[2]);
                      if (kind == 2)
                        {
                          if (c == 0)
                            val += qdop[q] +
                                   (phim ? qphr[q] : u0[2 * QC + q] - u0[QC + q]);
                          else
                            val += qr[q];
                        }
                    }
                  else
                    {
                      const int a  = cs - 1;
                      double    WW[3] = {0, 0, 0};
                      if (kind == 0)
                        WW[cc] = tb.LDc[(q * DIM + a) * NB + j];
                      else if (kind == 1)
                        WW[cc] = ((f >> 1) == a) ? tb.LCc[(f * QC + q) * NF + j] : 0.0;
                      else
                        for (int c2 = 0; c2 < 3; ++c2)
                          WW[c2] = -Wrc[(c2 * DIM + a) * QC + q];
                      const double E0a = W0[(0 * DIM + a) * QC + q];
                      if (c == 0)
                        {
                          val = -eps * WW[0];
                          if (kind == 2)
                            val += eps * E0a;
                        }
                      else if (c == 1)
                        {
                          const double n0 = u0[QC + q];
                          val = -mun * E0a * uu[1] + p.sdrift * n0 * mun * WW[0] + UT * mun * WW[1];
                          if (kind == 2)
                            val += n0 * mun * E0a - UT * mun * W0[(1 * DIM + a) * QC + q];
                        }
                      else
                        {
                          const double p0 = u0[2 * QC + q];
                          val = mup * E0a * uu[2] - p.sdrift * p0 * mup * WW[0] + UT * mup * WW[2];
                          if (kind == 2)
                            val += -p0 * mup * E0a - UT * mup * W0[(2 * DIM + a) * QC + q];
                        }
                    }
                }
              else
                {
                  const int fs = slot - QC * (1 + DIM);
                  const int g = fs / QF, s = fs % QF;
                  double    uu[3] = {0, 0, 0}, WN[3] = {0, 0, 0}, LL[3] = {0, 0, 0};
                  if (kind == 0)
                    {
                      uu[cc] = tb.phif[(g * QF + s) * NB + j];
                      WN[cc] = tb.LDf[(g * QF + s) * NB + j];
                    }
                  else if (kind == 1)
                    {
                      WN[cc] = tb.LCf[((f * NFC + g) * QF + s) * NF + j];
                      if (f == g)
                        LL[cc] = tb.xi[s * NF + j];
                    }
                  else
                    for (int c2 = 0; c2 < 3; ++c2)
                      {
                        WN[c2] = -Wrf[(c2 * NFC + g) * QF + s];
                        LL[c2] = -lamf[(c2 * NFC + g) * QF + s];
                      }
                  const int    fi  = g * QF + s;
                  const double E0n = W0f[(0 * NFC) * QF + fi];
                  const double tc  = tauv[g * 3 + c];
                  // drift of the numerical flux on the trace value
                  const bool up = ft == 1 || (ft == 2 && E0n <= 0);
                  if (c == 0)
                    {
                      val = eps * WN[0] + tc * uu[0] - tc * LL[0];
                      if (kind == 2)
                        val += -eps * E0n - tc * u0f[fi] + tc * tr0[fi];
                    }
                  else if (c == 1)
                    {
                      const double n0 = u0f[NFC * QF + fi], l0 = tr0[NFC * QF + fi];
                      // the rhs column carries no upwind trace term: zero
                      // trace update in pass 1, and the reconstruct pass
                      // omits it (template.h:2122-2160)
                      const double dr = up ? (kind == 2 ? 0.0 : LL[1] * mun * E0n) + l0 * mun * WN[0]
                                           : uu[1] * mun * E0n + n0 * mun * WN[0];
                      val = dr - UT * mun * WN[1] - tc * uu[1] + tc * LL[1];
                      if (kind == 2)
                        val += -((up ? l0 : n0) * mun * E0n - UT * mun * W0f[NFC * QF + fi]) +
                               tc * n0 - tc * l0;
                    }
                  else
                    {
                      const double p0 = u0f[2 * NFC * QF + fi], l0 = tr0[2 * NFC * QF + fi];
                      const double dr = up ? (kind == 2 ? 0.0 : LL[2] * mup * E0n) + l0 * mup * WN[0]
                                           : uu[2] * mup * E0n + p0 * mup * WN[0];
                      val = -dr - UT * mup * WN[2] - tc * uu[2] + tc * LL[2];
                      if (kind == 2)
                        val += -(-(up ? l0 : p0) * mup * E0n - UT * mup * W0f[2 * NFC * QF + fi]) +
                               tc * p0 - tc * l0;
                    }
                }
              return val;
          };
          if constexpr (!S::BIG)
            {
              for (int e = tid; e < SZ * ncol; e += NT)
                {
                  const int slot = e / ncol, m = e % ncol;
                  int       kind, cc, j, f;
                  decode(m, kind, cc, j, f);
                  G[slot * LD + m] = colval(slot, kind, cc, j, f);
                }
              __syncthreads();
              // rows of this component: z rows (NB) then trace rows (NFC*NF)
              const int nzr = NB + (pass2 ? 0 : NFC * NF);
              for (int e = tid; e < nzr * ncol; e += NT)
                {
                  const int rr = e / ncol, m = e % ncol;
                  double    v  = 0;
                  int       row;
                  if (rr < NB)
                    {
                      row = ci * NB + rr;
                      for (int sl = 0; sl < SZ; ++sl)
                        v += tb.Bz[sl * NB + rr] * G[sl * LD + m];
                    }
                  else
                    {
                      const int tr = rr - NB, g = tr / NF, i = tr % NF;
                      row = ns + g * tdf + ci * NF + i;
                      if (!isdir[g * 3 + c])
                        {
                          for (int s = 0; s < QF; ++s)
                            v += tb.Bt[(g * QF + s) * NF + i] *
                                 G[(QC * (1 + DIM) + g * QF + s) * LD + m];
                          if (m == ncol - 1 && isneu[g * 3 + c])
                            for (int s = 0; s < QF; ++s)
                              v += tb.Bt[(g * QF + s) * NF + i] *
                                   p.neuv[(facev[g] * 3 + c) * QF + s];
                        }
                      else
                        v = 0;
                    }
                  R[row * LD + m] = v;
                }
              __syncthreads();
            }
          else
            {
              // R and G in global memory: one thread per column computes
              // its bundle entries slot by slot and folds each into the z
              // rows and the trace rows of the slot's face at once (the
              // same sums in the same order as the two-step form), so G is
              // never stored and R is written once
              ctabg *Bzc = (ctabg *)tb.Bz, *Btc = (ctabg *)tb.Bt;
              for (int m = tid; m < ncol; m += NT)
                {
                  int kind, cc, j, f;
                  decode(m, kind, cc, j, f);
                  double az[NB];
#pragma unroll
                  for (int rr = 0; rr < NB; ++rr)
                    az[rr] = 0;
                  // two (three) slots per iteration: their table loads are
                  // in flight together (measured against 1, 4 and 6)
#pragma unroll 2
                  for (int slot = 0; slot < QC * (1 + DIM); ++slot)
                    {
                      const double v = colval(slot, kind, cc, j, f);
#pragma unroll
                      for (int rr = 0; rr < NB; ++rr)
                        az[rr] += Bzc[slot * NB + rr] * v;
                    }
                  for (int g = 0; g < NFC; ++g)
                    {
                      double at[NF];
#pragma unroll
                      for (int i = 0; i < NF; ++i)
                        at[i] = 0;
#pragma unroll 3
                      for (int s = 0; s < QF; ++s)
                        {
                          const int    slot = QC * (1 + DIM) + g * QF + s;
                          const double v    = colval(slot, kind, cc, j, f);
#pragma unroll
                          for (int rr = 0; rr < NB; ++rr)
                            az[rr] += Bzc[slot * NB + rr] * v;
#pragma unroll
                          for (int i = 0; i < NF; ++i)
                            at[i] += Btc[(g * QF + s) * NF + i] * v;
                        }
                      if (!pass2)
                        {
                          const bool dir = isdir[g * 3 + c];
                          const bool neu = m == ncol - 1 && isneu[g * 3 + c];
#pragma unroll
                          for (int i = 0; i < NF; ++i)
                            {
                              double w = at[i];
                              if (neu)
                                for (int s = 0; s < QF; ++s)
                                  w += tb.Bt[(g * QF + s) * NF + i] *
                                       p.neuv[(facev[g] * 3 + c) * QF + s];
                              R[(ns + g * tdf + ci * NF + i) * LD + m] = dir ? 0.0 : w;
                            }
                        }
                    }
#pragma unroll
                  for (int rr = 0; rr < NB; ++rr)
                    R[(ci * NB + rr) * LD + m] = az[rr];
                }
            }
        }
      if constexpr (S::BIG)
        __syncthreads();
      (void)nrow;
    }

    // Gaussian elimination of the first ns columns with row-scaled partial
    // pivoting (rows stay in place; used[] marks pivot rows)
    template <int DIM, int DEG>
    __device__ void eliminate(double *rg, double *sm, int *si, int ns, int nrow, int ncol)
    {
      using S          = Sh<DIM, DEG>;
      constexpr int LD = S::LD;
      const int     tid = opaque_tid();
      double *R = rg + S::oR, *rs = sm + S::oRS, *mul = sm + S::oMUL;
      int    *used = si + S::iUSED, *piv = si + S::iPIV;
      for (int i = tid; i < nrow; i += NT)
        {
          double mx = 0;
          if (i < ns)
            for (int j = 0; j < ns; ++j)
              mx = fmax(mx, fabs(R[i * LD + j]));
          rs[i]   = mx > 0 ? 1.0 / mx : 1.0;
          used[i] = i < ns ? 0 : 1; // trace rows never pivot
        }
      __syncthreads();
      for (int k = 0; k < ns; ++k)
        {
          if (tid < 64)
            {
              double best = -1;
              int    bi   = -1;
              for (int i = tid; i < ns; i += 64)
                if (!used[i])
                  {
                    const double sc = fabs(R[i * LD + k]) * rs[i];
                    if (sc > best)
                      {
                        best = sc;
                        bi   = i;
                      }
                  }
              for (int off = 32; off > 0; off >>= 1)
                {
                  const double ob = __shfl_xor(best, off);
                  const int    oi = __shfl_xor(bi, off);
                  if (ob > best || (ob == best && oi >= 0 && (bi < 0 || oi < bi)))
                    {
                      best = ob;
                      bi   = oi;
                    }
                }
              if (tid == 0)
                {
                  piv[k]   = bi;
                  used[bi] = 2; // current pivot
                }
            }
          __syncthreads();
          const int    pr = piv[k];
          const double pv = R[pr * LD + k];
          for (int i = tid; i < nrow; i += NT)
            mul[i] = (i < ns ? used[i] == 0 : true) && i != pr ? R[i * LD + k] / pv : 0.0;
          __syncthreads();
          const int w = ncol - k - 1;
          for (int e = tid; e < nrow * w; e += NT)
            {
              const int i = e / w, j = k + 1 + e % w;
              const double m = mul[i];
              if (m != 0.0)
                R[i * LD + j] -= m * R[pr * LD + j];
            }
          __syncthreads();
          if (tid == 0)
            used[pr] = 1;
          __syncthreads();
        }
    }

    // The same elimination on an nr x nc block A in LDS (leading dimension
    // ld, nr scalar rows, the first nr columns eliminated), with each
    // row's multiplier of step k (rows not yet pivoted) stored in A[i][k]:
    // afterwards row piv[k] holds [L(k, 0..k-1) | U(k, k..)] in its own
    // place.  Pivot choice and update order are those of eliminate().
    template <int DIM, int DEG, typename Real = double>
    __device__ __forceinline__ void lu_rows(Real *A, int ld, double *sm, int *si, int nr, int nc)
    {
      using S   = Sh<DIM, DEG>;
      static_assert(2 * S::MAXNS <= NT, "one update lane pair per row");
      const int tid = opaque_tid();
      double *rs = sm + S::oRS;
      int    *used = si + S::iUSED, *piv = si + S::iPIV;
      for (int i = tid; i < nr; i += NT)
        {
          double mx = 0;
          for (int j = 0; j < nr; ++j)
            mx = fmax(mx, fabs((double)A[i * ld + j]));
          rs[i]   = mx > 0 ? 1.0 / mx : 1.0;
          used[i] = 0;
        }
      __syncthreads();
      // update lanes: a lane pair per row (row tid / 2, every other column
      // from tid % 2), in one wave, so both read the row's entry in column k
      // before the first of them overwrites it with the multiplier
      const int ui = tid >> 1, uh = tid & 1;
      for (int k = 0; k < nr; ++k)
        {
          if (tid < 64)
            {
              double best = -1;
              int    bi   = -1;
              for (int i = tid; i < nr; i += 64)
                if (!used[i])
                  {
                    const double sc = fabs((double)A[i * ld + k]) * rs[i];
                    if (sc > best || bi < 0) // a NaN column still picks a row
                      {
                        best = sc;
                        bi   = i;
                      }
                  }
              for (int off = 32; off > 0; off >>= 1)
                {
                  const double ob = __shfl_xor(best, off);
                  const int    oi = __shfl_xor(bi, off);
                  if ((bi < 0 && oi >= 0) || ob > best ||
                      (ob == best && oi >= 0 && (bi < 0 || oi < bi)))
                    {
                      best = ob;
                      bi   = oi;
                    }
                }
              if (tid == 0)
                {
                  piv[k]   = bi;
                  used[bi] = 2; // current pivot
                }
            }
          __syncthreads();
          // rows still to pivot (the pivot rows of this and earlier steps
          // keep used = 2)
          const int  pr = piv[k];
          const Real pv = A[pr * ld + k];
          if (ui < nr && used[ui] == 0)
            {
              const Real m = A[ui * ld + k] / pv;
              if (m != Real(0))
                for (int j = k + 1 + uh; j < nc; j += 2)
                  A[ui * ld + j] -= m * A[pr * ld + j];
              if (uh == 0)
                A[ui * ld + k] = m;
            }
          __syncthreads();
        }
    }

    // FP32 elimination with FP64 iterative refinement, one column (one
    // thread): x = A'^-1 b for b in col[i * LD], i < ns, overwritten by x.
    // A32 holds lu_rows' FP32 factors of the equilibrated block Dr A' Dc
    // (row piv[k]: [L(k, 0..k-1) | U(k, k..)], leading dimension al); rsc,
    // csc the diagonals of Dr, Dc (powers of two); Ad is A' in FP64 (leading
    // dimension LD), anorm its infinity norm.  Refinement as in LAPACK's
    // dsgesv:
    //   x = 0, r = b;  repeat:  d = Dc U^-1 L^-1 P Dr r  (FP32 factors),
    //   x += d,  r -= A' d  (FP64),  until ||r|| <= ||x|| ||A'|| eps sqrt(ns)
    //   and ||d|| <= 2^-44 ||x||
    // (at most MIXED_IT solves; three where the block is well conditioned).
    // The second condition bounds the forward error too: the local blocks of
    // the carriers are ill conditioned where the doping is high, and a
    // solution with an FP64-level residual alone (or one predicted from the
    // contraction of the corrections) differs from the FP64 one by enough
    // to add Newton steps (measured on the 12^3 Gaussian cube).
    // NSS >= ns bounds the register arrays (ns is uniform).
    constexpr int MIXED_IT = 8;
    template <int NSS, int LD>
    __device__ __forceinline__ void mixed_col_solve(double *col, const double *Ad,
                                                    const float *A32, int al, const int *piv,
                                                    const double *rsc, const double *csc,
                                                    double anorm, int ns)
    {
      auto lu_solve = [&](float (&y)[NSS]) __attribute__((always_inline)) {
        static_for(
          [&](auto kc) {
            constexpr int k = decltype(kc)::value;
            if (k < ns)
              {
                const float *a   = A32 + piv[k] * al;
                float        acc = (float)(col[piv[k] * LD] * rsc[piv[k]]);
#pragma unroll
                for (int l = 0; l < k; ++l)
                  acc -= a[l] * y[l];
                y[k] = acc;
              }
          },
          std::make_integer_sequence<int, NSS>{});
        static_for(
          [&](auto kc) {
            constexpr int k = NSS - 1 - decltype(kc)::value;
            if (k < ns)
              {
                const float *a   = A32 + piv[k] * al;
                float        acc = y[k];
#pragma unroll
                for (int l = k + 1; l < NSS; ++l)
                  if (l < ns)
                    acc -= a[l] * y[l];
                y[k] = acc / a[k];
              }
          },
          std::make_integer_sequence<int, NSS>{});
      };
      const double tol = anorm * 0x1p-52 * sqrt((double)ns);
      double       x[NSS];
#pragma unroll
      for (int k = 0; k < NSS; ++k)
        x[k] = 0.0;
      // one expansion of lu_solve (the loop is not unrolled)
#pragma unroll 1
      for (int it = 0; it < MIXED_IT; ++it)
        {
          float y[NSS];
          lu_solve(y);
          double xmax = 0.0, dmax = 0.0;
#pragma unroll
          for (int k = 0; k < NSS; ++k)
            if (k < ns)
              {
                const double d = (double)y[k] * csc[k];
                x[k] += d;
                xmax = fmax(xmax, fabs(x[k]));
                dmax = fmax(dmax, fabs(d));
              }
          double rmax = 0.0;
          for (int i = 0; i < ns; ++i)
            {
              const double *a   = Ad + i * LD;
              double        acc = col[i * LD];
#pragma unroll
              for (int k = 0; k < NSS; ++k)
                if (k < ns)
                  acc -= a[k] * ((double)y[k] * csc[k]);
              col[i * LD] = acc;
              rmax        = fmax(rmax, fabs(acc));
            }
          if (!(rmax > xmax * tol) && !(dmax > xmax * 0x1p-44))
            break;
        }
#pragma unroll
      for (int k = 0; k < NSS; ++k)
        if (k < ns)
          col[k * LD] = x[k];
    }

    // rsum: the row sums of |A'| (mixed_factor), for its infinity norm
    template <int DIM, int DEG>
    __device__ __forceinline__ void mixed_col(double *col, const double *Ad, const float *A32,
                                              int al, const int *piv, const double *rsc,
                                              const double *csc, const double *rsum, int ns)
    {
      using S      = Sh<DIM, DEG>;
      double anorm = 0.0;
      for (int i = 0; i < ns; ++i)
        anorm = fmax(anorm, rsum[i]);
      mixed_col_solve<S::MAXNS, S::LD>(col, Ad, A32, al, piv, rsc, csc, anorm, ns);
    }

    // the FP32 copy of the equilibrated scalar block: Dr A' Dc with Dr, Dc
    // powers of two (exact) that bring every row's, then every column's,
    // largest entry into [1, 2): the local blocks of the carriers mix
    // densities many orders of magnitude apart, beyond what one refinement
    // step of a plain FP32 factorisation recovers.  Then lu_rows in FP32.
    // rsc, csc, rsum (the row sums of |A'|): ns doubles each (LDS)
    template <int DIM, int DEG>
    __device__ __forceinline__ void mixed_factor(const double *R, float *A32, int al, double *sm,
                                                 int *si, double *rsc, double *csc,
                                                 double *rsum, int ns)
    {
      using S          = Sh<DIM, DEG>;
      constexpr int LD = S::LD;
      const int     tid = opaque_tid();
      for (int i = tid; i < ns; i += NT)
        {
          double m = 0, sum = 0;
          for (int j = 0; j < ns; ++j)
            {
              m = fmax(m, fabs(R[i * LD + j]));
              sum += fabs(R[i * LD + j]);
            }
          rsc[i]  = m > 0 ? ldexp(1.0, -ilogb(m)) : 1.0;
          rsum[i] = sum;
        }
      __syncthreads();
      for (int j = tid; j < ns; j += NT)
        {
          double m = 0;
          for (int i = 0; i < ns; ++i)
            m = fmax(m, fabs(R[i * LD + j] * rsc[i]));
          csc[j] = m > 0 ? ldexp(1.0, -ilogb(m)) : 1.0;
        }
      __syncthreads();
      for (int e = tid; e < ns * ns; e += NT)
        {
          const int i = e / ns, j = e % ns;
          A32[i * al + j] = (float)(R[i * LD + j] * rsc[i] * csc[j]);
        }
      __syncthreads();
      lu_rows<DIM, DEG, float>(A32, al, sm, si, ns, ns);
    }

    // where the FP32 factors live: the G region of an LDS layout (dead after
    // the reduced system is built), the scalar-block area of a global one
    template <int DIM, int DEG>
    __device__ __forceinline__ float *mixed_a32(double *rg, double *sm, int &al)
    {
      using S = Sh<DIM, DEG>;
      if constexpr (S::BIG)
        {
          al = S::AL;
          return reinterpret_cast<float *>(sm + S::oAW);
        }
      else
        {
          static_assert((S::MAXNS * (S::MAXNS + 1) + 1) / 2 <= S::SZ * S::LD,
                        "the FP32 factors fit the G region");
          al = S::MAXNS + 1;
          return reinterpret_cast<float *>(rg + S::oG);
        }
    }

    // pass 1 with the FP32 elimination and FP64 refinement:
    //   X = A'^-1 [CT' | c']   one thread per column (mixed_col)
    //   [S | g] = [TT' | t'] - TC' X   in FP64, written over the trace rows
    template <int DIM, int DEG>
    __device__ void mixed_schur(double *rg, double *sm, int *si, int ns, int T);
    template <int DIM, int DEG>
    __device__ __forceinline__ void schur_product(double *R, double *A, const int *rowof, int ns,
                                                  int T);

    // The two triangular solves of schur_big for NSS scalar rows (rows of A
    // in place, leading dimensions LD for R and AL for A):
    //   Y = L^-1 P [CT' | c']  thread j < T + 1 holds column NSS + j
    //   M = TC' U^-1           thread t < T holds trace row NSS + t
    template <int NSS, int LD, int AL>
    __device__ __forceinline__ void tri_solves(double *R, const double *A, const int *piv,
                                               int tid, int T)
    {
      if (tid < T + 1)
        {
          double *col = R + NSS + tid;
          double  y[NSS];
#pragma unroll
          for (int k = 0; k < NSS; ++k)
            y[k] = col[piv[k] * LD];
          static_for(
            [&](auto kc) {
              constexpr int k = decltype(kc)::value;
              if constexpr (k >= 1)
                {
                  const double *a   = A + piv[k] * AL;
                  double        acc = y[k];
#pragma unroll
                  for (int l = 0; l < k; ++l)
                    acc -= a[l] * y[l];
                  y[k] = acc;
                }
            },
            std::make_integer_sequence<int, NSS>{});
#pragma unroll
          for (int k = 0; k < NSS; ++k)
            col[piv[k] * LD] = y[k];
        }
      if (tid < T)
        {
          double *row = R + (NSS + tid) * LD;
          double  x[NSS];
#pragma unroll
          for (int k = 0; k < NSS; ++k)
            x[k] = row[k];
          static_for(
            [&](auto kc) {
              constexpr int k = decltype(kc)::value;
              const double *a = A + piv[k] * AL;
              const double  m = x[k] / a[k];
              row[k]          = m; // final: stored at once, not held
#pragma unroll
              for (int j = k + 1; j < NSS; ++j)
                x[j] -= m * a[j];
            },
            std::make_integer_sequence<int, NSS>{});
        }
    }

    // eliminate() for shapes whose R lives in global memory (ns scalar rows,
    // T trace rows, T + 1 right-hand columns), as a blocked Schur complement:
    //   A' = P L U            lu_rows on the scalar block, in LDS
    //   Y  = L^-1 P [CT' | c']  one thread per column, in place in R
    //   M  = TC' U^-1          one thread per trace row, in place in R
    //   [S | g] = [TT' | t'] - M Y   LDS-staged product
    // Every entry receives the same multiply-adds in the same order as in
    // eliminate(), which reads and writes the whole of R once per pivot.
    template <int DIM, int DEG>
    __device__ __forceinline__ void schur_big(double *R, double *sm, int *si, int ns, int T)
    {
      using S             = Sh<DIM, DEG>;
      constexpr int LD    = S::LD, AL = S::AL, MAXNS = S::MAXNS;
      constexpr int KC    = 9;                      // k per staged chunk
      constexpr int RA    = 6;                      // row slots per pass
      constexpr int CB    = (S::MAXT + 1 + 15) / 16; // column slots
      static_assert(S::NB % KC == 0, "ns is a multiple of the chunk");
      static_assert(S::MAXT + 1 <= NT, "one thread per right-hand column and trace row");
      static_assert(16 * RA * KC + KC * 16 * CB <= MAXNS * AL, "staging fits");
      double    *A   = sm + S::oAW;
      const int *piv = si + S::iPIV;
      const int  tid = opaque_tid();
      const int  W   = T + 1;
      for (int e = tid; e < ns * ns; e += NT)
        {
          const int i = e / ns, j = e % ns;
          A[i * AL + j] = R[i * LD + j];
        }
      __syncthreads();
      lu_rows<DIM, DEG>(A, AL, sm, si, ns, ns);
      if (ns == MAXNS)
        tri_solves<MAXNS, LD, AL>(R, A, piv, tid, T);
      else if (ns == 2 * S::NB)
        tri_solves<2 * S::NB, LD, AL>(R, A, piv, tid, T);
      else
        tri_solves<S::NB, LD, AL>(R, A, piv, tid, T);
      __syncthreads();
      schur_product<DIM, DEG>(R, A, piv, ns, T);
    }

    // [S | g] -= M Y over the trace rows of R (global or LDS), in passes of
    // 16 RA rows with KC-deep chunks of M and Y staged in LDS at A: M in
    // R's trace rows, columns 0..ns (its column k pairs with Y's row k);
    // Y's row k at R row rowof[k] (rowof null: row k), columns ns..ns+T
    template <int DIM, int DEG>
    __device__ __forceinline__ void schur_product(double *R, double *A, const int *rowof, int ns,
                                                  int T)
    {
      using S             = Sh<DIM, DEG>;
      constexpr int LD    = S::LD;
      constexpr int KC    = 9;                      // k per staged chunk
      constexpr int RA    = 6;                      // row slots per pass
      constexpr int CB    = (S::MAXT + 1 + 15) / 16; // column slots
      const int     tid   = opaque_tid();
      const int     W     = T + 1;
      constexpr int WP = 16 * CB;       // staged Y width (zero beyond W)
      double   *Ms = A;                // [16 RA][KC], rows of this pass
      double   *Ys = A + 16 * RA * KC; // [KC][WP]
      const int tx = tid & 15, ty = tid >> 4;
      for (int r0 = 0; r0 < T; r0 += 16 * RA)
        {
          double acc[RA][CB];
#pragma unroll
          for (int a = 0; a < RA; ++a)
#pragma unroll
            for (int b = 0; b < CB; ++b)
              {
                const int t = r0 + ty + 16 * a, c = tx + 16 * b;
                acc[a][b]   = (t < T && c < W) ? R[(ns + t) * LD + ns + c] : 0.0;
              }
          for (int k0 = 0; k0 < ns; k0 += KC)
            {
              __syncthreads();
              for (int e = tid; e < 16 * RA * KC; e += NT)
                {
                  const int t = r0 + e / KC;
                  Ms[e]       = t < T ? R[(ns + t) * LD + k0 + e % KC] : 0.0;
                }
              for (int e = tid; e < KC * WP; e += NT)
                {
                  const int kk = e / WP, c = e % WP;
                  const int yr = rowof ? rowof[k0 + kk] : k0 + kk;
                  Ys[e]        = c < W ? R[yr * LD + ns + c] : 0.0;
                }
              __syncthreads();
#pragma unroll
              for (int kk = 0; kk < KC; ++kk)
                {
                  double mv[RA], yv[CB];
#pragma unroll
                  for (int a = 0; a < RA; ++a)
                    mv[a] = Ms[(ty + 16 * a) * KC + kk];
#pragma unroll
                  for (int b = 0; b < CB; ++b)
                    yv[b] = Ys[kk * WP + tx + 16 * b];
#pragma unroll
                  for (int a = 0; a < RA; ++a)
#pragma unroll
                    for (int b = 0; b < CB; ++b)
                      acc[a][b] -= mv[a] * yv[b];
                }
            }
#pragma unroll
          for (int a = 0; a < RA; ++a)
#pragma unroll
            for (int b = 0; b < CB; ++b)
              {
                const int t = r0 + ty + 16 * a, c = tx + 16 * b;
                if (t < T && c < W)
                  R[(ns + t) * LD + ns + c] = acc[a][b];
              }
        }
      __syncthreads();
    }

    template <int DIM, int DEG>
    __device__ void mixed_schur(double *rg, double *sm, int *si, int ns, int T)
    {
      using S          = Sh<DIM, DEG>;
      constexpr int LD = S::LD;
      double       *R   = rg + S::oR;
      const int    *piv = si + S::iPIV;
      const int     tid = opaque_tid();
      int           al;
      float        *A32 = mixed_a32<DIM, DEG>(rg, sm, al);
      double       *rsc = sm + S::oMUL, *csc = rsc + S::MAXNS, *rsum = sm + S::oSOL;
      static_assert(2 * S::MAXNS <= S::MAXR, "both scales fit the multiplier area");
      mixed_factor<DIM, DEG>(R, A32, al, sm, si, rsc, csc, rsum, ns);
      if (tid < T + 1)
        mixed_col<DIM, DEG>(R + ns + tid, R, A32, al, piv, rsc, csc, rsum, ns);
      __syncthreads();
      if constexpr (S::BIG)
        schur_product<DIM, DEG>(R, sm + S::oAW, nullptr, ns, T);
      else
        {
          // every (trace row, column) entry: its dot product over the X
          // rows (rows below ns and columns from ns are written only here)
          const int W = T + 1;
          for (int e = tid; e < T * W; e += NT)
            {
              const int     t   = e / W, c = e % W;
              const double *m   = R + (ns + t) * LD;
              double        acc = m[ns + c];
              for (int k = 0; k < ns; ++k)
                acc -= m[k] * R[k * LD + ns + c];
              R[(ns + t) * LD + ns + c] = acc;
            }
          __syncthreads();
        }
    }

    // MIXED: the elimination in FP32 with one FP64 refinement step
    // (mixed_schur); a kernel of its own, so that the FP64 kernel's
    // registers are its own
    template <int DIM, int DEG, bool MIXED>
    __global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2)))
      condense_kernel(KParams p, const int *cells, int ncell, int color, int mode)
    {
      using S = Sh<DIM, DEG>;
      extern __shared__ double smem[];
      double *sm = smem;
      int    *si = reinterpret_cast<int *>(smem + S::NDBL);
      // one cell per workgroup and round (a single round unless R and G are
      // in global scratch, where the grid is capped)
      for (int b = blockIdx.x; b < ncell; b += gridDim.x)
      {
      const long long cell = cells[b];
      int             ns, T;
      build_reduced<DIM, DEG>(p, cell, false, sm, si, ns, T);
      constexpr int NF = S::NF, NFC = S::NFC, LD = S::LD;
      double       *rg  = rg_base<DIM, DEG>(p, sm);
      double       *R   = rg + S::oR;
      double       *del = sm + S::oDEL;
      int          *isdir = si + S::iDIR, *bidv = si + S::iBID;
      long long    *facev = reinterpret_cast<long long *>(si + S::iFACE);
      const int     tid   = opaque_tid();
      const int     tdf   = p.ne * NF;
      const int     nrow = ns + T, ncol = ns + T + 1;
      // LocalCondenser::condense_ct_matrix (src/local_condenser.cpp:141-160)
      // on the reduced columns: rhs -= delta_j * col_j, col_j = 0
      bool anydir = false;
      for (int g = 0; g < 3 * NFC; ++g)
        anydir = anydir || isdir[g];
      for (int r = tid; anydir && r < nrow; r += NT)
        {
          double acc = 0;
          for (int lt = 0; lt < T; ++lt)
            {
              const int g = lt / tdf, c = p.encomp[(lt % tdf) / NF], m = lt % NF;
              if (isdir[g * 3 + c])
                {
                  acc += del[(g * 3 + c) * NF + m] * R[r * LD + ns + lt];
                  R[r * LD + ns + lt] = 0;
                }
            }
          R[r * LD + ncol - 1] -= acc;
        }
      __syncthreads();
      if constexpr (MIXED)
        mixed_schur<DIM, DEG>(rg, sm, si, ns, T);
      else if constexpr (S::BIG)
        schur_big<DIM, DEG>(R, sm, si, ns, T);
      else
        eliminate<DIM, DEG>(rg, sm, si, ns, nrow, ncol);
      // output
      if (mode == 1)
        {
          for (int e = tid; e < T * T; e += NT)
            {
              const int i = e / T, j = e % T;
              p.Sout[(long long)b * T * T + e] = R[(ns + i) * LD + ns + j];
            }
          for (int i = tid; i < T; i += NT)
            p.gout[(long long)b * T + i] = R[(ns + i) * LD + ncol - 1];
          __syncthreads();
          continue;
        }
      const unsigned char *slot = p.cslot + cell * NFC * NFC;
      for (int e = tid; e < T * T; e += NT)
        {
          const int i = e / T, j = e % T;
          const int gi = i / tdf, gj = j / tdf, ri = i % tdf, rj = j % tdf;
          const int ci = p.encomp[ri / NF];
          const long long F = facev[gi];
          const bool dir = isdir[gi * 3 + ci];
          const double v = dir ? (i == j ? 1.0 : 0.0) : R[(ns + i) * LD + ns + j];
          const long long pos = p.blkptr[F] * tdf * tdf + (long long)ri * p.nbk[F] * tdf +
                                (long long)slot[gi * NFC + gj] * tdf + rj;
          const bool add = color == 1 && gi == gj && bidv[gi] < 0;
          if (add)
            p.vals[pos] += v;
          else
            p.vals[pos] = v;
        }
      for (int i = tid; i < T; i += NT)
        {
          const int gi = i / tdf, ri = i % tdf;
          const int ci = p.encomp[ri / NF];
          const long long F   = facev[gi];
          const long long row = F * tdf + ri;
          const bool dir = isdir[gi * 3 + ci];
          const double v = dir ? 0.0 : R[(ns + i) * LD + ncol - 1];
          if (color == 1 && bidv[gi] < 0)
            p.rhs[row] += v;
          else
            p.rhs[row] = v;
          if (dir)
            p.dirout[row] = del[(gi * 3 + ci) * NF + (ri % NF)];
        }
      // the next round reuses the workgroup's LDS and scratch
      __syncthreads();
      }
    }

    template <int DIM, int DEG, bool MIXED>
    __global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2))) reconstruct_kernel(KParams p, int ncells)
    {
      using S = Sh<DIM, DEG>;
      extern __shared__ double smem[];
      double *sm = smem;
      int    *si = reinterpret_cast<int *>(smem + S::NDBL);
      for (long long cell = blockIdx.x; cell < ncells; cell += gridDim.x)
      {
      int ns, T;
      build_reduced<DIM, DEG>(p, cell, true, sm, si, ns, T);
      constexpr int NB = S::NB, NF = S::NF, LD = S::LD, DIMc = DIM;
      double       *rg = rg_base<DIM, DEG>(p, sm);
      double       *R = rg + S::oR, *sol = sm + S::oSOL, *rhsq = sm + S::oRHSQ;
      double       *lupd = sm + S::oLUPD;
      int          *piv  = si + S::iPIV;
      const int     tid  = opaque_tid();
      if constexpr (MIXED)
        {
          // u = A'^-1 c' in FP32 with one FP64 refinement step
          int     al;
          float  *A32 = mixed_a32<DIM, DEG>(rg, sm, al);
          // the row sums go to sol, which receives u only afterwards
          double *rsc = sm + S::oMUL, *csc = rsc + S::MAXNS, *rsum = sol;
          mixed_factor<DIM, DEG>(R, A32, al, sm, si, rsc, csc, rsum, ns);
          if (tid == 0)
            mixed_col<DIM, DEG>(R + ns, R, A32, al, piv, rsc, csc, rsum, ns);
          __syncthreads();
          for (int k = tid; k < ns; k += NT)
            sol[k] = R[k * LD + ns];
          __syncthreads();
        }
      else
        {
          // U and the reduced rhs: R itself, or for R in global memory a
          // copy of the ns x (ns + 1) system factorised in LDS
          const double *U  = R;
          int           lu = LD;
          if constexpr (S::BIG)
            {
              double *A = sm + S::oAW;
              for (int e = tid; e < ns * (ns + 1); e += NT)
                {
                  const int i = e / (ns + 1), j = e % (ns + 1);
                  A[i * S::AL + j] = R[i * LD + j];
                }
              __syncthreads();
              lu_rows<DIM, DEG>(A, S::AL, sm, si, ns, ns + 1);
              U  = A;
              lu = S::AL;
            }
          else
            eliminate<DIM, DEG>(rg, sm, si, ns, ns, ns + 1);
          if (tid == 0)
            for (int k = ns - 1; k >= 0; --k)
              {
                const int pr = piv[k];
                double    v  = U[pr * lu + ns];
                for (int j = k + 1; j < ns; ++j)
                  v -= U[pr * lu + j] * sol[j];
                sol[k] = v / U[pr * lu + k];
              }
          __syncthreads();
        }
      // W = M^-1 (rhs_q + D_a u - sum_f C_{a,f} lambda_f), written with u
      double *out = p.dxc + cell * S::NCD;
      for (int e = tid; e < S::NCD; e += NT)
        {
          const int c = e / ((DIMc + 1) * NB), r = e % ((DIMc + 1) * NB);
          const int blk = r / NB, i = r % NB;
          double    v   = 0;
          if (p.en[c])
            {
              int ci = 0;
              for (int c2 = 0; c2 < c; ++c2)
                ci += p.en[c2];
              const double *u = sol + ci * NB;
              if (blk == DIMc)
                v = u[i];
              else
                {
                  const int a = blk;
                  for (int l = 0; l < NB; ++l)
                    v += p.tb.Minv[i * NB + l] * rhsq[(c * DIMc + a) * NB + l] +
                         p.tb.MinvD[(a * NB + i) * NB + l] * u[l];
                  for (int sd = 0; sd < 2; ++sd)
                    {
                      const int f = 2 * a + sd;
                      for (int m = 0; m < NF; ++m)
                        v -= p.tb.MinvC[(f * NB + i) * NF + m] * lupd[(f * 3 + c) * NF + m];
                    }
                }
            }
          out[e] = v;
        }
      __syncthreads();
      }
    }

    // the dynamic-LDS attribute of the four kernels of a shape, once
    template <int DIM, int DEG>
    hipError_t generic_attrs()
    {
      using S          = Sh<DIM, DEG>;
      static bool attr = false;
      if (attr)
        return hipSuccess;
      for (const void *k : {(const void *)condense_kernel<DIM, DEG, false>,
                            (const void *)condense_kernel<DIM, DEG, true>,
                            (const void *)reconstruct_kernel<DIM, DEG, false>,
                            (const void *)reconstruct_kernel<DIM, DEG, true>})
        {
          hipError_t e =
            hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)S::bytes);
          if (e != hipSuccess)
            return e;
        }
      attr = true;
      return hipSuccess;
    }

    template <int DIM, int DEG>
    hipError_t launch_c(const KParams &p, const int *cells, int ncell, int color, int mode,
                        hipStream_t s)
    {
      using S      = Sh<DIM, DEG>;
      hipError_t e = generic_attrs<DIM, DEG>();
      if (e != hipSuccess)
        return e;
      if (ncell <= 0)
        return hipSuccess;
      if (S::BIG && !p.rgscr)
        return hipErrorInvalidValue;
      const int grid = S::BIG ? std::min(ncell, GRID_BIG) : ncell;
      // precision 2 only: the generic kernel's FP32 form is slower than its
      // FP64 one (precision 1 keeps FP32 to the register kernels)
      if (p.mixed == 2)
        hipLaunchKernelGGL((condense_kernel<DIM, DEG, true>), dim3(grid), dim3(NT), S::bytes, s,
                           p, cells, ncell, color, mode);
      else
        hipLaunchKernelGGL((condense_kernel<DIM, DEG, false>), dim3(grid), dim3(NT), S::bytes, s,
                           p, cells, ncell, color, mode);
      return hipGetLastError();
    }

    template <int DIM, int DEG>
    hipError_t launch_r(const KParams &p, int ncells, hipStream_t s)
    {
      using S      = Sh<DIM, DEG>;
      hipError_t e = generic_attrs<DIM, DEG>();
      if (e != hipSuccess)
        return e;
      if (ncells <= 0)
        return hipSuccess;
      if (S::BIG && !p.rgscr)
        return hipErrorInvalidValue;
      const int grid = S::BIG ? std::min(ncells, GRID_BIG) : ncells;
      if (p.mixed == 2)
        hipLaunchKernelGGL((reconstruct_kernel<DIM, DEG, true>), dim3(grid), dim3(NT), S::bytes,
                           s, p, ncells);
      else
        hipLaunchKernelGGL((reconstruct_kernel<DIM, DEG, false>), dim3(grid), dim3(NT),
                           S::bytes, s, p, ncells);
      return hipGetLastError();
    }
  } // namespace

  namespace
  {
    // Terminal current: the HDG numerical fluxes of n and p integrated over
    // the faces of one boundary id (outward normal).  The fluxes are the ones
    // the trace equations make single-valued (add_tc_matrix_terms_to_tt_rhs,
    // include/assembling_methods.template.h:1311-1447):
    //   F_n =  n^ mu_n E.nu - U_T mu_n W_n.nu - tau_n (n - lambda_n)
    //   F_p = -p^ mu_p E.nu - U_T mu_p W_p.nu - tau_p (p - lambda_p)
    // n^ = n (use_cell), lambda_n (use_trace), or n where E.nu > 0 and
    // lambda_n elsewhere (qiu_shi_stabilization, template.h:1417, 1435).
    // One workgroup; fixed-order reduction (deterministic).
    constexpr int CUR_T = 256;

    __global__ void __launch_bounds__(CUR_T)
      current_kernel(KParams p, int dim, int nb, int nf, int qf, int bid, double *out)
    {
      __shared__ double red[2][CUR_T];
      const int a = bid >> 1, side = bid & 1;
      const int o1 = (a + 1) % 3, o2 = (a + 2) % 3;
      const long long cnt = (long long)p.n[o1] * p.n[o2];
      const double    nu  = side ? 1.0 : -1.0;
      const int       ncd = 3 * (dim + 1) * nb;
      const DevTables &tb = p.tb;
      double sn = 0, sp = 0;
      for (long long e = threadIdx.x; e < cnt; e += CUR_T)
        {
          int ic[3];
          ic[a]  = side ? p.n[a] - 1 : 0;
          ic[o1] = (int)(e % p.n[o1]);
          ic[o2] = (int)(e / p.n[o1]);
          const long long cell = ic[0] + (long long)p.n[0] * (ic[1] + (long long)p.n[1] * ic[2]);
          const long long F    = face_of(p, ic[0], ic[1], ic[2], bid);
          const double   *x    = p.xc + cell * ncd;
          const double   *xt   = p.xt + F * 3 * nf;
          for (int s = 0; s < qf; ++s)
            {
              const double *ph = tb.phif + (bid * qf + s) * nb;
              double        u[3] = {0, 0, 0}, w[3] = {0, 0, 0}, lam[3] = {0, 0, 0};
              for (int c = 0; c < 3; ++c)
                {
                  const double *xc = x + c * (dim + 1) * nb;
                  for (int j = 0; j < nb; ++j)
                    {
                      u[c] += xc[dim * nb + j] * ph[j];
                      w[c] += xc[a * nb + j] * ph[j];
                    }
                  w[c] *= nu;
                  for (int m = 0; m < nf; ++m)
                    lam[c] += xt[c * nf + m] * tb.xi[s * nf + m];
                }
              const double JW = tb.JxWf[bid * qf + s];
              const double *tf = p.tau_cf ? p.tau_cf + (cell * 2 * dim + bid) * 3 : p.tau;
              const bool    cl = p.flux_type == 0 || (p.flux_type == 2 && w[0] > 0);
              const double  nn = cl ? u[1] : lam[1], pp = cl ? u[2] : lam[2];
              sn += JW * (nn * p.mun * w[0] - p.UT * p.mun * w[1] - tf[1] * (u[1] - lam[1]));
              sp += JW * (-pp * p.mup * w[0] - p.UT * p.mup * w[2] - tf[2] * (u[2] - lam[2]));
            }
        }
      red[0][threadIdx.x] = sn;
      red[1][threadIdx.x] = sp;
      __syncthreads();
      for (int off = CUR_T / 2; off > 0; off >>= 1)
        {
          if ((int)threadIdx.x < off)
            {
              red[0][threadIdx.x] += red[0][threadIdx.x + off];
              red[1][threadIdx.x] += red[1][threadIdx.x + off];
            }
          __syncthreads();
        }
      if (threadIdx.x == 0)
        {
          out[0] = red[0][0];
          out[1] = red[1][0];
        }
    }
  } // namespace

  namespace
  {
    // max |a_k| of four device vectors: block partial maxima, then one
    // workgroup (max is exact in any order: deterministic)
    constexpr int NRM_T = 256, NRM_B = 256;

    // max that propagates NaN (fmax drops it): a NaN update must not read
    // as a converged Newton step
    __device__ __forceinline__ double nan_max(double a, double b)
    {
      return (a > b || a != a) ? a : b;
    }

    __global__ void __launch_bounds__(NRM_T)
      norms_partial(const double *a0, long long n0, const double *a1, long long n1,
                    const double *a2, long long n2, const double *a3, long long n3,
                    double *part)
    {
      __shared__ double red[NRM_T];
      const double   *a[4] = {a0, a1, a2, a3};
      const long long n[4] = {n0, n1, n2, n3};
      for (int v = 0; v < 4; ++v)
        {
          double m = 0;
          for (long long i = blockIdx.x * (long long)NRM_T + threadIdx.x; i < n[v];
               i += (long long)gridDim.x * NRM_T)
            m = nan_max(m, fabs(a[v][i]));
          red[threadIdx.x] = m;
          __syncthreads();
          for (int off = NRM_T / 2; off > 0; off >>= 1)
            {
              if ((int)threadIdx.x < off)
                red[threadIdx.x] = nan_max(red[threadIdx.x], red[threadIdx.x + off]);
              __syncthreads();
            }
          if (threadIdx.x == 0)
            part[v * NRM_B + blockIdx.x] = red[0];
          __syncthreads();
        }
    }

    __global__ void __launch_bounds__(NRM_T) norms_final(const double *part, double *out)
    {
      __shared__ double red[NRM_T];
      for (int v = 0; v < 4; ++v)
        {
          red[threadIdx.x] = threadIdx.x < NRM_B ? part[v * NRM_B + threadIdx.x] : 0.0;
          __syncthreads();
          for (int off = NRM_T / 2; off > 0; off >>= 1)
            {
              if ((int)threadIdx.x < off)
                red[threadIdx.x] = nan_max(red[threadIdx.x], red[threadIdx.x + off]);
              __syncthreads();
            }
          if (threadIdx.x == 0)
            out[v] = red[0];
          __syncthreads();
        }
    }
  } // namespace

  hipError_t launch_norms(const double *const a[4], const long long n[4], double *scratch,
                          double *out, hipStream_t s)
  {
    hipLaunchKernelGGL(norms_partial, dim3(NRM_B), dim3(NRM_T), 0, s, a[0], n[0], a[1], n[1],
                       a[2], n[2], a[3], n[3], scratch);
    hipLaunchKernelGGL(norms_final, dim3(1), dim3(NRM_T), 0, s, scratch, out);
    return hipGetLastError();
  }

  // Partial maxima of launch_norms_static: a file-scope device array, so the
  // entry point that uses it allocates nothing.  Every GPU holds its own copy
  // until the process ends; calls that overlap on one GPU share it and race.
  // Each call's first kernel overwrites every partial it later reads, so
  // nothing has to be cleared between calls.
  constexpr int NRM_SCRATCH = 4 * NRM_B;
  static __device__ __align__(256) double scratch[NRM_SCRATCH];

  namespace
  {
    constexpr int MAX_DEVICES = 64;
    std::mutex    scratch_mutex;
    double       *scratch_ptr[MAX_DEVICES] = {};
  } // namespace

  hipError_t norms_static_scratch(double **ptr)
  {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
      return e;
    if (dev < 0 || dev >= MAX_DEVICES)
      return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev])
      {
        // the device address of this device's copy: the array's name is a
        // host-side stand-in and must never reach a kernel
        void *p = nullptr;
        e       = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
          return e;
        scratch_ptr[dev] = static_cast<double *>(p);
      }
    *ptr = scratch_ptr[dev];
    return hipSuccess;
  }

  hipError_t launch_norms_static(const double *const a[4], const long long n[4], double *out,
                                 hipStream_t s)
  {
    double    *part = nullptr;
    hipError_t e    = norms_static_scratch(&part);
    if (e != hipSuccess)
      return e;
    static_assert(4 * NRM_B <= NRM_SCRATCH, "partials exceed the static scratch");
    return launch_norms(a, n, part, out, s);
  }

  hipError_t launch_current(int dim, int nb, int nf, int qf, const KParams &p, int bid,
                            double *out, hipStream_t s)
  {
    hipLaunchKernelGGL(current_kernel, dim3(1), dim3(CUR_T), 0, s, p, dim, nb, nf, qf, bid, out);
    return hipGetLastError();
  }

  bool supported(int dim, int deg)
  {
    return (dim == 2 && deg >= 1 && deg <= 3) || (dim == 3 && deg >= 1 && deg <= 2);
  }

  size_t generic_scratch_bytes(int dim, int deg)
  {
    if (dim == 3 && deg == 2)
      return (size_t)GRID_BIG * Sh<3, 2>::RGD * sizeof(double);
    return 0;
  }

  hipError_t launch_condense(int dim, int deg, const KParams &p, const int *cells, int ncell,
                             int color, int mode, hipStream_t s)
  {
    if (dim == 2 && deg == 1)
      return launch_c<2, 1>(p, cells, ncell, color, mode, s);
    if (dim == 2 && deg == 2)
      return launch_c<2, 2>(p, cells, ncell, color, mode, s);
    if (dim == 2 && deg == 3)
      return launch_c<2, 3>(p, cells, ncell, color, mode, s);
    if (dim == 3 && deg == 1)
      return launch_c<3, 1>(p, cells, ncell, color, mode, s);
    if (dim == 3 && deg == 2)
      return launch_c<3, 2>(p, cells, ncell, color, mode, s);
    return hipErrorInvalidValue;
  }

  hipError_t launch_reconstruct(int dim, int deg, const KParams &p, int ncells, hipStream_t s)
  {
    if (dim == 2 && deg == 1)
      return launch_r<2, 1>(p, ncells, s);
    if (dim == 2 && deg == 2)
      return launch_r<2, 2>(p, ncells, s);
    if (dim == 2 && deg == 3)
      return launch_r<2, 3>(p, ncells, s);
    if (dim == 3 && deg == 1)
      return launch_r<3, 1>(p, ncells, s);
    if (dim == 3 && deg == 2)
      return launch_r<3, 2>(p, ncells, s);
    return hipErrorInvalidValue;
  }
} // namespace hdg
