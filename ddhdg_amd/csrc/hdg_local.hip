// hdg_local.hip -- the uncondensed local system of listed cells
// (NPSolver::get_local_cell_system, reference src/np_solver.cpp:2337-2445).
//
// The reference runs the per-cell operator without its Schur complement
// (execute_schur_complement = false) and returns the blocks as assembled:
//   cc = A (N x N), ct = CT (N x T), tc = TC (T x N, stored negated,
//   include/assembling_methods.template.h:1227-1230), tt = TT (T x T),
//   c (N), t (T),
// after the LocalCondenser has moved the Dirichlet columns of CT into c
// (src/local_condenser.cpp:141-160).  N = ne (d+1) nb cell dofs (flux then
// scalar per enabled component), T = 2d ne nf trace dofs.
//
// One workgroup per cell.  The rows of the flux test functions are the
// geometric forms (mass, -grad, face coupling; template.h:269-532,
// 885-1019).  The rows of the scalar and trace test functions are the same
// three per-point forms the condensation kernels use (hdg_kernels.hip),
// applied to the unlifted bundle of every column: a flux column's basis
// function, a scalar column's, a trace column's, and the residual.  The
// column data at the quadrature slots go through a per-cell workspace in
// global memory.
#include "hdg_device.h"
#include "hdg_kernels.h"

namespace hdg
{
  namespace
  {
    constexpr int LT = 256;

    template <int DIM, int DEG>
    struct LShape
    {
      static constexpr int K1  = DEG + 1;
      static constexpr int NB  = DIM == 2 ? K1 * K1 : K1 * K1 * K1;
      static constexpr int NF  = DIM == 2 ? K1 : K1 * K1;
      static constexpr int QC  = NB, QF = NF, NFC = 2 * DIM;
      static constexpr int NCD = 3 * (DIM + 1) * NB;
      static constexpr int SZ  = QC * (1 + DIM) + NFC * QF;
      // LDS (doubles)
      static constexpr int oXS   = 0;
      static constexpr int oXT   = oXS + NCD;
      static constexpr int oU0   = oXT + NFC * 3 * NF;
      static constexpr int oW0   = oU0 + 3 * QC;
      static constexpr int oU0F  = oW0 + 3 * DIM * QC;
      static constexpr int oW0F  = oU0F + 3 * NFC * QF;
      static constexpr int oTR0  = oW0F + 3 * NFC * QF;
      static constexpr int oQD   = oTR0 + 3 * NFC * QF;
      static constexpr int oRHSQ = oQD + 6 * QC;
      static constexpr int oDEL  = oRHSQ + 3 * DIM * NB;
      static constexpr int oTAU  = oDEL + NFC * 3 * NF;
      static constexpr int NDBL  = oTAU + NFC * 3;
      static constexpr int iDIR = 0, iNEU = iDIR + NFC * 3, iFACE = iNEU + NFC * 3;
      static constexpr int NINT = iFACE + 2 * NFC;
    };

    struct LocalOut
    {
      double *A, *CT, *TC, *TT, *c, *t;
    };

    template <int DIM, int DEG>
    __global__ void __launch_bounds__(LT)
      local_system_kernel(KParams p, const int *__restrict__ cells, int ncell, double *G,
                          LocalOut o)
    {
      using S           = LShape<DIM, DEG>;
      constexpr int NB  = S::NB, NF = S::NF, QC = S::QC, QF = S::QF, NFC = S::NFC, SZ = S::SZ;
      __shared__ double sm[S::NDBL];
      __shared__ int    si[S::NINT];
      const int         tid = threadIdx.x;
      const int         k   = blockIdx.x;
      if (k >= ncell)
        return;
      const long long cell = cells[k];
      const DevTables &tb  = p.tb;
      double *xs = sm + S::oXS, *xt = sm + S::oXT, *u0 = sm + S::oU0, *W0 = sm + S::oW0;
      double *u0f = sm + S::oU0F, *W0f = sm + S::oW0F, *tr0 = sm + S::oTR0, *qd = sm + S::oQD;
      double *rhsq = sm + S::oRHSQ, *del = sm + S::oDEL, *tauv = sm + S::oTAU;
      double *qr = qd, *qdrn = qd + QC, *qdrp = qd + 2 * QC, *qphd = qd + 3 * QC;
      double *qphr = qd + 4 * QC, *qdop = qd + 5 * QC;
      int       *isdir = si + S::iDIR, *isneu = si + S::iNEU;
      long long *facev = reinterpret_cast<long long *>(si + S::iFACE);

      const int ne = p.ne, nb1 = (DIM + 1) * NB;
      const int N = ne * nb1, tdf = ne * NF, T = NFC * tdf, ncol = N + T + 1;
      int ic[3];
      ic[0] = (int)(cell % p.n[0]);
      ic[1] = (int)((cell / p.n[0]) % p.n[1]);
      ic[2] = (int)(cell / ((long long)p.n[0] * p.n[1]));

      // ---- state, boundary flags, tau
      if (tid < NFC)
        {
          const int lf = tid, a = lf >> 1, sd = lf & 1;
          facev[lf]      = face_of(p, ic[0], ic[1], ic[2], lf);
          const bool bnd = sd ? (ic[a] == p.n[a] - 1) : (ic[a] == 0);
          for (int c = 0; c < 3; ++c)
            {
              const int bt      = (bnd && p.en[c]) ? p.bc[lf][c] : 0;
              isdir[lf * 3 + c] = bt == 1;
              isneu[lf * 3 + c] = bt == 2 && !p.thermo;
            }
        }
      for (int i = tid; i < S::NCD; i += LT)
        xs[i] = p.xc[cell * S::NCD + i];
      for (int i = tid; i < QC; i += LT)
        qdop[i] = p.doping[cell * QC + i];
      if (tid < NFC * 3)
        tauv[tid] = p.tau_cf ? p.tau_cf[cell * NFC * 3 + tid] : p.tau[tid % 3];
      __syncthreads();
      for (int i = tid; i < NFC * 3 * NF; i += LT)
        xt[i] = p.xt[facev[i / (3 * NF)] * 3 * NF + i % (3 * NF)];
      __syncthreads();
      // ---- the state at the quadrature points
      for (int e = tid; e < 3 * (DIM + 1) * QC; e += LT)
        {
          const int c = e / ((DIM + 1) * QC), r = e % ((DIM + 1) * QC);
          const int blk = r / QC, q = r % QC;
          const double *x = xs + c * nb1 + blk * NB;
          double        v = 0;
          for (int j = 0; j < NB; ++j)
            v += x[j] * tb.phic[q * NB + j];
          if (blk == DIM)
            u0[c * QC + q] = v;
          else
            W0[(c * DIM + blk) * QC + q] = v;
        }
      for (int e = tid; e < 3 * NFC * QF; e += LT)
        {
          const int c = e / (NFC * QF), g = (e / QF) % NFC, s = e % QF;
          const double nu  = (g & 1) ? 1.0 : -1.0;
          const double *xc = xs + c * nb1;
          const double *ph = tb.phif + (g * QF + s) * NB;
          double        u = 0, w = 0, t = 0;
          for (int j = 0; j < NB; ++j)
            {
              u += xc[DIM * NB + j] * ph[j];
              w += xc[(g >> 1) * NB + j] * ph[j];
            }
          for (int m = 0; m < NF; ++m)
            t += xt[(g * 3 + c) * NF + m] * tb.xi[s * NF + m];
          u0f[e] = u;
          W0f[e] = nu * w;
          tr0[e] = t;
        }
      __syncthreads();
      // ---- point data (SRH and its derivatives; the Boltzmann terms of the
      // equilibrium and phi_linearize operators), the flux-row residual and
      // the projected Dirichlet update (apply_dbc_on_face, template.h:
      // 1592-1691; zero in thermodynamic equilibrium)
      for (int q = tid; q < QC; q += LT)
        {
          double r = 0, a = 0, b = 0;
          if (p.en[1] || p.en[2])
            recomb(p, u0[QC + q], u0[2 * QC + q], r, a, b);
          qr[q]   = r;
          qdrn[q] = a;
          qdrp[q] = b;
          if (p.thermo || p.phi_lin)
            {
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
      for (int e = tid; e < 3 * DIM * NB; e += LT)
        {
          const int c = e / (DIM * NB), a = (e / NB) % DIM, i = e % NB;
          double    v = 0;
          for (int q = 0; q < QC; ++q)
            v += (u0[c * QC + q] * tb.gradc[(q * NB + i) * DIM + a] -
                  W0[(c * DIM + a) * QC + q] * tb.phic[q * NB + i]) *
                 tb.JxWc[q];
          for (int sd = 0; sd < 2; ++sd)
            {
              const int    g  = 2 * a + sd;
              const double nu = sd ? 1.0 : -1.0;
              for (int s = 0; s < QF; ++s)
                v -= tr0[(c * NFC + g) * QF + s] * tb.phif[(g * QF + s) * NB + i] * nu *
                     tb.JxWf[g * QF + s];
            }
          rhsq[e] = v;
        }
      for (int e = tid; e < NFC * 3 * NF; e += LT)
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
                         (p.dirv[(facev[g] * 3 + c) * QF + s] - tr0[(c * NFC + g) * QF + s]);
                  v += tb.Mfinv[(a * NF + m) * NF + n2] * b;
                }
            }
          del[e] = v;
        }
      __syncthreads();

      // ---- the forms of the scalar and trace test rows for every column,
      // at every quadrature slot: G[ci][slot][m] of this cell
      const double eps = p.eps, mun = p.mun, mup = p.mup, UT = p.UT;
      const bool   phim = p.thermo || p.phi_lin;
      const int    ft   = p.flux_type;
      double      *Gc   = G + (size_t)k * ne * SZ * ncol;
      for (int e = tid; e < ne * SZ * ncol; e += LT)
        {
          const int ci = e / (SZ * ncol), slot = (e / ncol) % SZ, m = e % ncol;
          const int c  = p.encomp[ci];
          // column m: 0 scalar dof, 3 flux dof (axis ax), 1 trace dof on
          // face f, 2 the residual
          int kind, cc = 0, j = 0, ax = 0, f = 0;
          if (m < N)
            {
              const int r = m % nb1;
              cc   = p.encomp[m / nb1];
              j    = r % NB;
              ax   = r / NB;
              kind = ax == DIM ? 0 : 3;
            }
          else if (m < N + T)
            {
              const int lt = m - N;
              kind         = 1;
              f            = lt / tdf;
              cc           = p.encomp[(lt % tdf) / NF];
              j            = lt % NF;
            }
          else
            kind = 2;
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
                    val += c == 0 ? qdop[q] + (phim ? qphr[q] : u0[2 * QC + q] - u0[QC + q])
                                  : qr[q];
                }
              else
                {
                  const int a     = cs - 1;
                  double    WW[3] = {0, 0, 0};
                  if (kind == 3 && ax == a)
                    WW[cc] = tb.phic[q * NB + j];
                  const double E0a = W0[a * QC + q];
                  if (c == 0)
                    val = -eps * WW[0] + (kind == 2 ? eps * E0a : 0.0);
                  else if (c == 1)
                    {
                      const double n0 = u0[QC + q];
                      val = -mun * E0a * uu[1] + p.sdrift * n0 * mun * WW[0] + UT * mun * WW[1];
                      if (kind == 2)
                        val += n0 * mun * E0a - UT * mun * W0[(DIM + a) * QC + q];
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
              const int    fs = slot - QC * (1 + DIM), g = fs / QF, s = fs % QF;
              const double nu = (g & 1) ? 1.0 : -1.0;
              double       uu[3] = {0, 0, 0}, WN[3] = {0, 0, 0}, LL[3] = {0, 0, 0};
              if (kind == 0)
                uu[cc] = tb.phif[(g * QF + s) * NB + j];
              else if (kind == 3 && ax == (g >> 1))
                WN[cc] = nu * tb.phif[(g * QF + s) * NB + j];
              else if (kind == 1 && f == g)
                LL[cc] = tb.xi[s * NF + j];
              const int    fi  = g * QF + s;
              const double E0n = W0f[fi];
              const double tc  = tauv[g * 3 + c];
              const bool   up  = ft == 1 || (ft == 2 && E0n <= 0);
              if (c == 0)
                {
                  val = eps * WN[0] + tc * uu[0] - tc * LL[0];
                  if (kind == 2)
                    val += -eps * E0n - tc * u0f[fi] + tc * tr0[fi];
                }
              else if (c == 1)
                {
                  const double n0 = u0f[NFC * QF + fi], l0 = tr0[NFC * QF + fi];
                  const double dr = up ? LL[1] * mun * E0n + l0 * mun * WN[0]
                                       : uu[1] * mun * E0n + n0 * mun * WN[0];
                  val = dr - UT * mun * WN[1] - tc * uu[1] + tc * LL[1];
                  if (kind == 2)
                    val += -((up ? l0 : n0) * mun * E0n - UT * mun * W0f[NFC * QF + fi]) +
                           tc * n0 - tc * l0;
                }
              else
                {
                  const double p0 = u0f[2 * NFC * QF + fi], l0 = tr0[2 * NFC * QF + fi];
                  const double dr = up ? LL[2] * mup * E0n + l0 * mup * WN[0]
                                       : uu[2] * mup * E0n + p0 * mup * WN[0];
                  val = -dr - UT * mup * WN[2] - tc * uu[2] + tc * LL[2];
                  if (kind == 2)
                    val += -(-(up ? l0 : p0) * mup * E0n - UT * mup * W0f[2 * NFC * QF + fi]) +
                           tc * p0 - tc * l0;
                }
            }
          Gc[e] = val;
        }
      __threadfence_block();
      __syncthreads();

      // ---- the rows
      double *A = o.A + (size_t)k * N * N, *CT = o.CT + (size_t)k * N * T;
      double *TC = o.TC + (size_t)k * T * N, *TT = o.TT + (size_t)k * T * T;
      double *cv = o.c + (size_t)k * N, *tv = o.t + (size_t)k * T;
      auto put = [&](int row, bool trace_row, int m, double v) {
        if (!trace_row)
          {
            if (m < N)
              A[(size_t)row * N + m] = v;
            else if (m < N + T)
              CT[(size_t)row * T + (m - N)] = v;
            else
              cv[row] = v;
          }
        else
          {
            // TC is stored negated, as the reference does
            if (m < N)
              TC[(size_t)row * N + m] = -v;
            else if (m < N + T)
              TT[(size_t)row * T + (m - N)] = v;
            else
              tv[row] = v;
          }
      };
      // scalar test rows and trace test rows of each enabled component
      const int nrz = ne * NB, nrt = T;
      for (int e = tid; e < (nrz + nrt) * ncol; e += LT)
        {
          const int rr = e / ncol, m = e % ncol;
          if (rr < nrz)
            {
              const int     ci = rr / NB, i = rr % NB;
              const double *g  = Gc + (size_t)ci * SZ * ncol;
              double        v  = 0;
              for (int sl = 0; sl < SZ; ++sl)
                v += tb.Bz[sl * NB + i] * g[sl * ncol + m];
              put(ci * nb1 + DIM * NB + i, false, m, v);
            }
          else
            {
              const int tr = rr - nrz, lf = tr / tdf, ci = (tr % tdf) / NF, i = tr % NF;
              const int c  = p.encomp[ci];
              double    v  = 0;
              if (!isdir[lf * 3 + c])
                {
                  const double *g = Gc + (size_t)ci * SZ * ncol;
                  for (int s = 0; s < QF; ++s)
                    v += tb.Bt[(lf * QF + s) * NF + i] * g[(QC * (1 + DIM) + lf * QF + s) * ncol + m];
                  if (m == ncol - 1 && isneu[lf * 3 + c])
                    for (int s = 0; s < QF; ++s)
                      v += tb.Bt[(lf * QF + s) * NF + i] * p.neuv[(facev[lf] * 3 + c) * QF + s];
                }
              put(tr, true, m, v);
            }
        }
      // flux test rows: mass, -grad, face coupling; residual rhs_q
      for (int e = tid; e < ne * DIM * NB * ncol; e += LT)
        {
          const int rr = e / ncol, m = e % ncol;
          const int ci = rr / (DIM * NB), a = (rr / NB) % DIM, i = rr % NB;
          const int c  = p.encomp[ci];
          double    v  = 0;
          if (m < N)
            {
              const int r = m % nb1, bj = r / NB, j = r % NB;
              if (m / nb1 == ci)
                {
                  if (bj == a)
                    for (int q = 0; q < QC; ++q)
                      v += tb.phic[q * NB + i] * tb.phic[q * NB + j] * tb.JxWc[q];
                  else if (bj == DIM)
                    for (int q = 0; q < QC; ++q)
                      v -= tb.gradc[(q * NB + i) * DIM + a] * tb.phic[q * NB + j] * tb.JxWc[q];
                }
            }
          else if (m < N + T)
            {
              const int lt = m - N, lf = lt / tdf, cj = (lt % tdf) / NF, j = lt % NF;
              if (cj == ci && (lf >> 1) == a)
                {
                  const double nu = (lf & 1) ? 1.0 : -1.0;
                  for (int s = 0; s < QF; ++s)
                    v += tb.phif[(lf * QF + s) * NB + i] * tb.xi[s * NF + j] * nu *
                         tb.JxWf[lf * QF + s];
                }
            }
          else
            v = rhsq[(c * DIM + a) * NB + i];
          put(ci * nb1 + a * NB + i, false, m, v);
        }
      __threadfence_block();
      __syncthreads();
      // LocalCondenser::condense_ct_matrix (src/local_condenser.cpp:141-160):
      // c -= CT(:, j) g_j and CT(:, j) = 0 for every constrained trace dof j
      for (int i = tid; i < N; i += LT)
        {
          double v = cv[i];
          for (int lf = 0; lf < NFC; ++lf)
            for (int ci = 0; ci < ne; ++ci)
              {
                const int c = p.encomp[ci];
                if (!isdir[lf * 3 + c])
                  continue;
                for (int mm = 0; mm < NF; ++mm)
                  {
                    const int col = lf * tdf + ci * NF + mm;
                    v -= CT[(size_t)i * T + col] * del[(lf * 3 + c) * NF + mm];
                    CT[(size_t)i * T + col] = 0.0;
                  }
              }
          cv[i] = v;
        }
    }

    template <int DIM, int DEG>
    hipError_t launch_local(const KParams &p, const int *cells, int ncell, double *G,
                            const LocalOut &o, hipStream_t s)
    {
      hipLaunchKernelGGL((local_system_kernel<DIM, DEG>), dim3(ncell), dim3(LT), 0, s, p, cells,
                         ncell, G, o);
      return hipGetLastError();
    }
  } // namespace

  size_t local_system_workspace(int dim, int deg, int ne)
  {
    const int nb = dim == 2 ? (deg + 1) * (deg + 1) : (deg + 1) * (deg + 1) * (deg + 1);
    const int nf = dim == 2 ? deg + 1 : (deg + 1) * (deg + 1);
    const int sz = nb * (1 + dim) + 2 * dim * nf;
    const int N = ne * (dim + 1) * nb, T = 2 * dim * ne * nf;
    return (size_t)ne * sz * (N + T + 1) * sizeof(double);
  }

  hipError_t launch_local_system(int dim, int deg, const KParams &p, const int *cells, int ncell,
                                 double *G, double *A, double *CT, double *TC, double *TT,
                                 double *c, double *t, hipStream_t s)
  {
    const LocalOut o{A, CT, TC, TT, c, t};
    if (ncell <= 0)
      return hipSuccess;
    if (dim == 2 && deg == 1)
      return launch_local<2, 1>(p, cells, ncell, G, o, s);
    if (dim == 2 && deg == 2)
      return launch_local<2, 2>(p, cells, ncell, G, o, s);
    if (dim == 2 && deg == 3)
      return launch_local<2, 3>(p, cells, ncell, G, o, s);
    if (dim == 3 && deg == 1)
      return launch_local<3, 1>(p, cells, ncell, G, o, s);
    if (dim == 3 && deg == 2)
      return launch_local<3, 2>(p, cells, ncell, G, o, s);
    return hipErrorInvalidValue;
  }
} // namespace hdg
