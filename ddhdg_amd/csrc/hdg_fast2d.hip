// hdg_fast2d.hip -- register-resident condensation for 2D cells, degree 1-2.
//
// One wave64 per cell, one column of the reduced bordered system per lane:
//   lanes [0, NS)          scalar trial dofs   (NS = ne*(k+1)^2)
//   lanes [NS, NS+T)       trace trial dofs    (T = 4*ne*(k+1))
//   lane  NS+T             right-hand side (residual; Dirichlet update folded in)
// Each lane keeps its whole column (NS+T rows) in VGPRs.  The column is
// assembled from its "bundle" (values, lifted fluxes and traces of the trial
// function at the quadrature points) with the test tables read as uniform
// scalars, exactly the terms of include/assembling_methods.template.h:
// 269-2118 after the flux unknowns are eliminated (hdg_tables.h).  Then NS
// steps of Gaussian elimination with scaled column pivoting: equation k picks
// the scalar unknown (lane) with the largest scaled |entry|, the pivot
// column is broadcast through LDS, every lane updates its own column.  The
// trace rows end as [S | g] and are written straight into the trace CSR
// (two colour launches, deterministic).  Workgroups of 4 waves are
// persistent; the shared tables are staged in LDS once per workgroup.
#include "hdg_device.h"
#include "hdg_kernels.h"

#define WAVE_SYNC()                                          \
  do                                                         \
    {                                                        \
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); \
      __builtin_amdgcn_wave_barrier();                       \
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront"); \
    }                                                        \
  while (0)

namespace hdg
{
  namespace
  {
    constexpr int WPB = 4; // waves per workgroup

    template <int DEG, int NE>
    struct F2
    {
      static constexpr int K1 = DEG + 1, NB = K1 * K1, NF = K1, QC = NB, QF = NF, NFC = 4;
      static constexpr int NCD = 9 * NB;
      static constexpr int NS = NE * NB, TDF = NE * NF, T = NFC * TDF;
      static constexpr int NROW = NS + T, NCOL = NROW + 1;
      static_assert(NCOL <= 64, "one column per lane");
      // workgroup tables (doubles)
      static constexpr int tPHIC  = 0;
      static constexpr int tLDC   = tPHIC + QC * NB;
      static constexpr int tLCC   = tLDC + QC * 2 * NB;
      static constexpr int tPHIF  = tLCC + NFC * QC * NF;
      static constexpr int tLDF   = tPHIF + NFC * QF * NB;
      static constexpr int tLCF   = tLDF + NFC * QF * NB;
      static constexpr int tXI    = tLCF + NFC * NFC * QF * NF;
      static constexpr int tGRADC = tXI + QF * NF;
      static constexpr int tJXWC  = tGRADC + QC * NB * 2;
      static constexpr int tJXWF  = tJXWC + QC;
      static constexpr int tPMC   = tJXWF + NFC * QF;
      static constexpr int tPMF   = tPMC + QC * NB;
      static constexpr int tMFI   = tPMF + NFC * QF * NB;
      static constexpr int TAB    = tMFI + 2 * NF * NF;
      // per-wave scratch (doubles)
      static constexpr int wXS   = 0;
      static constexpr int wXT   = wXS + NCD;
      static constexpr int wU0   = wXT + NFC * 3 * NF;
      static constexpr int wW0   = wU0 + 3 * QC;
      static constexpr int wU0F  = wW0 + 3 * 2 * QC;
      static constexpr int wW0F  = wU0F + 3 * NFC * QF;
      static constexpr int wTR0  = wW0F + 3 * NFC * QF;
      static constexpr int wQD   = wTR0 + 3 * NFC * QF;
      static constexpr int wRHSQ = wQD + 6 * QC;
      static constexpr int wWRC  = wRHSQ + 3 * 2 * NB;
      static constexpr int wWRF  = wWRC + 3 * 2 * QC;
      static constexpr int wLAMF = wWRF + 3 * NFC * QF;
      static constexpr int wLUPD = wLAMF + 3 * NFC * QF;
      static constexpr int wRESC = wLUPD + NFC * 3 * NF;
      static constexpr int wRESF = wRESC + QC * NE * 3;
      static constexpr int wNEU  = wRESF + NFC * QF * NE;
      static constexpr int wBC   = wNEU + NFC * NE * NF;
      static constexpr int wBLK  = wBC + 64;      // 4 long long (blkptr of the faces)
      static constexpr int wFACE = wBLK + NFC;    // 4 long long (face index)
      static constexpr int WDBL  = wFACE + NFC;
      // per-wave ints
      static constexpr int iDIR  = 0;             // NFC*3
      static constexpr int iNEU  = iDIR + NFC * 3;
      static constexpr int iBID  = iNEU + NFC * 3;
      static constexpr int iNBK  = iBID + NFC;
      static constexpr int iSLOT = iNBK + NFC;    // NFC*NFC
      static constexpr int WINT  = iSLOT + NFC * NFC;
      static constexpr size_t bytes =
        (size_t)(TAB + WPB * WDBL) * sizeof(double) + (size_t)WPB * WINT * sizeof(int);
    };

    template <int DEG, int NE, bool THERMO>
    __global__ void __launch_bounds__(64 * WPB)
      fast_condense(KParams p, const int *__restrict__ cells, int ncell, int color, int mode)
    {
      using S              = F2<DEG, NE>;
      constexpr int NB     = S::NB, NF = S::NF, QC = S::QC, QF = S::QF, NFC = S::NFC;
      constexpr int NS     = S::NS, TDF = S::TDF, T = S::T, NROW = S::NROW;
      constexpr int K1     = S::K1;
      extern __shared__ double smem[];
      double       *tab  = smem;
      const int     tid  = threadIdx.x;
      const int     lane = tid & 63, wv = tid >> 6;
      double       *ws   = smem + S::TAB + wv * S::WDBL;
      int *wi = reinterpret_cast<int *>(smem + S::TAB + WPB * S::WDBL) + wv * S::WINT;
      long long *blk   = reinterpret_cast<long long *>(ws + S::wBLK);
      long long *facev = reinterpret_cast<long long *>(ws + S::wFACE);
      const DevTables &tb = p.tb;

      // ---- stage the shared tables once per workgroup
      for (int i = tid; i < QC * NB; i += 64 * WPB)
        {
          tab[S::tPHIC + i] = tb.phic[i];
          tab[S::tPMC + i]  = tb.PhiMinvC[i];
        }
      for (int i = tid; i < QC * 2 * NB; i += 64 * WPB)
        {
          tab[S::tLDC + i]   = tb.LDc[i];
          tab[S::tGRADC + i] = tb.gradc[i];
        }
      for (int i = tid; i < NFC * QC * NF; i += 64 * WPB)
        tab[S::tLCC + i] = tb.LCc[i];
      for (int i = tid; i < NFC * QF * NB; i += 64 * WPB)
        {
          tab[S::tPHIF + i] = tb.phif[i];
          tab[S::tLDF + i]  = tb.LDf[i];
          tab[S::tPMF + i]  = tb.PhiMinvF[i];
        }
      for (int i = tid; i < NFC * NFC * QF * NF; i += 64 * WPB)
        tab[S::tLCF + i] = tb.LCf[i];
      for (int i = tid; i < QF * NF; i += 64 * WPB)
        tab[S::tXI + i] = tb.xi[i];
      for (int i = tid; i < QC; i += 64 * WPB)
        tab[S::tJXWC + i] = tb.JxWc[i];
      for (int i = tid; i < NFC * QF; i += 64 * WPB)
        tab[S::tJXWF + i] = tb.JxWf[i];
      for (int i = tid; i < 2 * NF * NF; i += 64 * WPB)
        tab[S::tMFI + i] = tb.Mfinv[i];
      __syncthreads();

      const double *__restrict__ Bz = tb.Bz;
      const double *__restrict__ Bt = tb.Bt;
      double *u0 = ws + S::wU0, *W0 = ws + S::wW0, *u0f = ws + S::wU0F, *W0f = ws + S::wW0F;
      double *tr0 = ws + S::wTR0, *qd = ws + S::wQD, *rhsq = ws + S::wRHSQ;
      double *Wrc = ws + S::wWRC, *Wrf = ws + S::wWRF, *lamf = ws + S::wLAMF;
      double *lupd = ws + S::wLUPD, *resc = ws + S::wRESC, *resf = ws + S::wRESF;
      double *neu = ws + S::wNEU, *bc = ws + S::wBC, *xs = ws + S::wXS, *xt = ws + S::wXT;
      int    *isdir = wi + S::iDIR, *isneu = wi + S::iNEU, *bidv = wi + S::iBID;
      int    *nbk = wi + S::iNBK, *slot = wi + S::iSLOT;
      const double eps = p.eps, mun = p.mun, mup = p.mup, UT = p.UT;
      const double tV = p.tau[0], tN = p.tau[1], tP = p.tau[2];
      const double sd = p.sdrift;

      // column owned by this lane
      int kind, cc = 0, jj = 0, ff = 0;
      if (lane < NS)
        {
          kind = 0;
          cc   = lane / NB;
          jj   = lane % NB;
        }
      else if (lane < NS + T)
        {
          kind         = 1;
          const int lt = lane - NS;
          ff           = lt / TDF;
          cc           = (lt % TDF) / NF;
          jj           = lt % NF;
        }
      else if (lane == NS + T)
        kind = 2;
      else
        kind = 3;

      for (int idx = blockIdx.x * WPB + wv; idx < ncell; idx += gridDim.x * WPB)
        {
          const long long cell = cells[idx];
          int             ic[2];
          ic[0] = (int)(cell % p.n[0]);
          ic[1] = (int)(cell / p.n[0]);
          // ---- A: topology, flags, state
          if (lane < NFC)
            {
              const int lf = lane, a = lf >> 1, sd = lf & 1;
              const long long F = face_of(p, ic[0], ic[1], 0, lf);
              facev[lf]         = F;
              blk[lf]           = p.blkptr[F];
              nbk[lf]           = p.nbk[F];
              const bool bnd    = sd ? (ic[a] == p.n[a] - 1) : (ic[a] == 0);
              const int  bid    = bnd ? lf : -1;
              bidv[lf]          = bid;
              for (int c = 0; c < 3; ++c)
                {
                  const int bt      = (bid >= 0 && p.en[c]) ? p.bc[bid][c] : 0;
                  isdir[lf * 3 + c] = bt == 1;
                  isneu[lf * 3 + c] = bt == 2;
                }
            }
          if (lane < NFC * NFC)
            slot[lane] = p.cslot[cell * NFC * NFC + lane];
          for (int i = lane; i < S::NCD; i += 64)
            xs[i] = p.xc[cell * S::NCD + i];
          for (int i = lane; i < QC; i += 64)
            qd[5 * QC + i] = p.doping[cell * QC + i];
          WAVE_SYNC();
          for (int i = lane; i < NFC * 3 * NF; i += 64)
            {
              const int lf = i / (3 * NF), r = i % (3 * NF);
              xt[i]        = p.xt[facev[lf] * 3 * NF + r];
            }
          WAVE_SYNC();
          // ---- B: state at quadrature points
          for (int e = lane; e < 3 * 3 * QC; e += 64)
            {
              const int c = e / (3 * QC), r = e % (3 * QC);
              const int blkc = r / QC, q = r % QC;
              const double *x = xs + c * 3 * NB + blkc * NB;
              double        v = 0;
#pragma unroll
              for (int j = 0; j < NB; ++j)
                v += x[j] * tab[S::tPHIC + q * NB + j];
              if (blkc == 2)
                u0[c * QC + q] = v;
              else
                W0[(c * 2 + blkc) * QC + q] = v;
            }
          for (int e = lane; e < 3 * NFC * QF; e += 64)
            {
              const int c = e / (NFC * QF), g = (e / QF) % NFC, s = e % QF;
              const int a = g >> 1;
              const double nu  = (g & 1) ? 1.0 : -1.0;
              const double *xc = xs + c * 3 * NB;
              const double *ph = tab + S::tPHIF + (g * QF + s) * NB;
              double        u = 0, w = 0, t = 0;
#pragma unroll
              for (int j = 0; j < NB; ++j)
                {
                  u += xc[2 * NB + j] * ph[j];
                  w += xc[a * NB + j] * ph[j];
                }
#pragma unroll
              for (int m = 0; m < NF; ++m)
                t += xt[(g * 3 + c) * NF + m] * tab[S::tXI + s * NF + m];
              u0f[e] = u;
              W0f[e] = nu * w;
              tr0[e] = t;
            }
          WAVE_SYNC();
          // ---- C: point data, flux-row residual, Dirichlet update
          for (int q = lane; q < QC; q += 64)
            {
              double r = 0, a = 0, b = 0;
              if (!THERMO)
                recomb(p, u0[QC + q], u0[2 * QC + q], r, a, b);
              qd[q]          = r;
              qd[QC + q]     = a;
              qd[2 * QC + q] = b;
              if (THERMO)
                {
                  const double V0 = u0[q], kT = p.kTq;
                  const double e1 = exp((p.ev_r - V0) / kT);
                  const double e2 = exp((V0 - p.ec_r) / kT);
                  qd[3 * QC + q]  = -1 / p.n_rf * (p.nv / kT * e1 + p.nc / kT * e2);
                  qd[4 * QC + q]  = 1 / p.n_rf * (p.nv * e1 - p.nc * e2);
                }
            }
          for (int e = lane; e < NE * 2 * NB; e += 64)
            {
              const int c = e / (2 * NB), a = (e / NB) % 2, i = e % NB;
              double    v = 0;
#pragma unroll
              for (int q = 0; q < QC; ++q)
                v += (u0[c * QC + q] * tab[S::tGRADC + (q * NB + i) * 2 + a] -
                      W0[(c * 2 + a) * QC + q] * tab[S::tPHIC + q * NB + i]) *
                     tab[S::tJXWC + q];
#pragma unroll
              for (int sd = 0; sd < 2; ++sd)
                {
                  const int    g  = 2 * a + sd;
                  const double nu = sd ? 1.0 : -1.0;
#pragma unroll
                  for (int s = 0; s < QF; ++s)
                    v -= tr0[(c * NFC + g) * QF + s] * tab[S::tPHIF + (g * QF + s) * NB + i] *
                         nu * tab[S::tJXWF + g * QF + s];
                }
              rhsq[e] = v;
            }
          for (int e = lane; e < NFC * 3 * NF; e += 64)
            {
              const int g = e / (3 * NF), c = (e / NF) % 3, m = e % NF;
              double    v = 0;
              if (!THERMO && isdir[g * 3 + c])
                {
                  const int a = g >> 1;
#pragma unroll
                  for (int n2 = 0; n2 < NF; ++n2)
                    {
                      double b = 0;
#pragma unroll
                      for (int s = 0; s < QF; ++s)
                        b += tab[S::tXI + s * NF + n2] * tab[S::tJXWF + g * QF + s] *
                             (p.dirv[(facev[g] * 3 + c) * QF + s] - tr0[(c * NFC + g) * QF + s]);
                      v += tab[S::tMFI + (a * NF + m) * NF + n2] * b;
                    }
                }
              lupd[e] = v;
            }
          WAVE_SYNC();
          // ---- D: lifted residual flux with the Dirichlet update folded in;
          //         residual forms at the quadrature points
          for (int e = lane; e < NE * 2 * QC; e += 64)
            {
              const int c = e / (2 * QC), a = (e / QC) % 2, q = e % QC;
              double    v = 0;
#pragma unroll
              for (int i = 0; i < NB; ++i)
                v += tab[S::tPMC + q * NB + i] * rhsq[(c * 2 + a) * NB + i];
#pragma unroll
              for (int sd = 0; sd < 2; ++sd)
                {
                  const int f = 2 * a + sd;
#pragma unroll
                  for (int m = 0; m < NF; ++m)
                    v += tab[S::tLCC + (f * QC + q) * NF + m] * lupd[(f * 3 + c) * NF + m];
                }
              Wrc[e] = v;
            }
          for (int e = lane; e < NE * NFC * QF; e += 64)
            {
              const int c = e / (NFC * QF), g = (e / QF) % NFC, s = e % QF;
              const int a = g >> 1;
              double    v = 0, l = 0;
#pragma unroll
              for (int i = 0; i < NB; ++i)
                v += tab[S::tPMF + (g * QF + s) * NB + i] * rhsq[(c * 2 + a) * NB + i];
#pragma unroll
              for (int sd = 0; sd < 2; ++sd)
                {
                  const int f = 2 * a + sd;
#pragma unroll
                  for (int m = 0; m < NF; ++m)
                    v += tab[S::tLCF + ((f * NFC + g) * QF + s) * NF + m] *
                         lupd[(f * 3 + c) * NF + m];
                }
#pragma unroll
              for (int m = 0; m < NF; ++m)
                l += lupd[(g * 3 + c) * NF + m] * tab[S::tXI + s * NF + m];
              Wrf[e]  = v;
              lamf[e] = l;
            }
          for (int e = lane; e < QC * NE; e += 64)
            {
              const int q = e / NE, ci = e % NE, c = NE == 1 ? 0 : ci;
              const double E0x = W0[0 * QC + q], E0y = W0[1 * QC + q];
              double       al, bx, by;
              if (c == 0)
                {
                  al = qd[5 * QC + q] + (THERMO ? qd[4 * QC + q] : u0[2 * QC + q] - u0[QC + q]);
                  bx = eps * E0x;
                  by = eps * E0y;
                }
              else if (c == 1)
                {
                  const double n0 = u0[QC + q];
                  al = qd[q];
                  bx = n0 * mun * E0x - UT * mun * W0[2 * QC + q];
                  by = n0 * mun * E0y - UT * mun * W0[3 * QC + q];
                }
              else
                {
                  const double p0 = u0[2 * QC + q];
                  al = qd[q];
                  bx = -p0 * mup * E0x - UT * mup * W0[4 * QC + q];
                  by = -p0 * mup * E0y - UT * mup * W0[5 * QC + q];
                }
              resc[(q * NE + ci) * 3 + 0] = al;
              resc[(q * NE + ci) * 3 + 1] = bx;
              resc[(q * NE + ci) * 3 + 2] = by;
            }
          for (int e = lane; e < NFC * QF * NE; e += 64)
            {
              const int fi = e / NE, ci = e % NE, c = NE == 1 ? 0 : ci;
              const double E0n = W0f[fi];
              double       v;
              if (c == 0)
                v = -eps * E0n - tV * u0f[fi] + tV * tr0[fi];
              else if (c == 1)
                {
                  const double n0 = u0f[NFC * QF + fi];
                  v = -(n0 * mun * E0n - UT * mun * W0f[NFC * QF + fi]) + tN * n0 -
                      tN * tr0[NFC * QF + fi];
                }
              else
                {
                  const double p0 = u0f[2 * NFC * QF + fi];
                  v = -(-p0 * mup * E0n - UT * mup * W0f[2 * NFC * QF + fi]) + tP * p0 -
                      tP * tr0[2 * NFC * QF + fi];
                }
              resf[e] = v;
            }
          for (int e = lane; e < NFC * NE * NF; e += 64)
            {
              const int g = e / (NE * NF), ci = (e / NF) % NE, i = e % NF;
              const int c = NE == 1 ? 0 : ci;
              double    v = 0;
              if (isneu[g * 3 + c] && !isdir[g * 3 + c])
#pragma unroll
                for (int s = 0; s < QF; ++s)
                  v += Bt[(g * QF + s) * NF + i] * p.neuv[(facev[g] * 3 + c) * QF + s];
              neu[e] = v;
            }
          WAVE_SYNC();

          // ---- E: this lane's column of the reduced system, in registers
          double R[NROW];
#pragma unroll
          for (int r = 0; r < NROW; ++r)
            R[r] = 0.0;
          const bool   isU = kind == 0, isL = kind == 1, isR = kind == 2;
          const double isr = isR ? 1.0 : 0.0;
          const int    ju = isU ? jj : 0, jl = isL ? jj : 0, fl = isL ? ff : 0;
#pragma unroll
          for (int q = 0; q < QC; ++q)
            {
              double uu[3] = {0, 0, 0}, WX[3] = {0, 0, 0}, WY[3] = {0, 0, 0};
              const double ph = tab[S::tPHIC + q * NB + ju];
              const double lx = tab[S::tLDC + (q * 2 + 0) * NB + ju];
              const double ly = tab[S::tLDC + (q * 2 + 1) * NB + ju];
              const double lc = tab[S::tLCC + (fl * QC + q) * NF + jl];
#pragma unroll
              for (int ci = 0; ci < NE; ++ci)
                {
                  const bool mine = cc == ci;
                  uu[ci] = (isU && mine) ? ph : 0.0;
                  WX[ci] = (isU && mine) ? lx
                         : (isL && mine && (fl >> 1) == 0) ? lc
                         : isR ? -Wrc[(ci * 2 + 0) * QC + q] : 0.0;
                  WY[ci] = (isU && mine) ? ly
                         : (isL && mine && (fl >> 1) == 1) ? lc
                         : isR ? -Wrc[(ci * 2 + 1) * QC + q] : 0.0;
                }
              const double E0x = W0[q], E0y = W0[QC + q];
#pragma unroll
              for (int ci = 0; ci < NE; ++ci)
                {
                  double al, bx, by;
                  if (ci == 0)
                    {
                      if (THERMO)
                        al = -qd[3 * QC + q] * uu[0];
                      else
                        al = uu[NE > 1 ? 1 : 0] - uu[NE > 2 ? 2 : 0];
                      bx = -eps * WX[0];
                      by = -eps * WY[0];
                    }
                  else if (ci == 1)
                    {
                      const double n0 = u0[QC + q];
                      al = -(qd[QC + q] * uu[1] + qd[2 * QC + q] * uu[NE > 2 ? 2 : 1]);
                      bx = -mun * E0x * uu[1] + sd * n0 * mun * WX[0] + UT * mun * WX[1];
                      by = -mun * E0y * uu[1] + sd * n0 * mun * WY[0] + UT * mun * WY[1];
                    }
                  else
                    {
                      const double p0 = u0[2 * QC + q];
                      al = -(qd[QC + q] * uu[1] + qd[2 * QC + q] * uu[NE > 2 ? 2 : 1]);
                      bx = mup * E0x * uu[NE > 2 ? 2 : 1] - sd * p0 * mup * WX[0] +
                           UT * mup * WX[NE > 2 ? 2 : 1];
                      by = mup * E0y * uu[NE > 2 ? 2 : 1] - sd * p0 * mup * WY[0] +
                           UT * mup * WY[NE > 2 ? 2 : 1];
                    }
                  al += isr * resc[(q * NE + ci) * 3 + 0];
                  bx += isr * resc[(q * NE + ci) * 3 + 1];
                  by += isr * resc[(q * NE + ci) * 3 + 2];
#pragma unroll
                  for (int i = 0; i < NB; ++i)
                    R[ci * NB + i] += Bz[(q * 3 + 0) * NB + i] * al +
                                      Bz[(q * 3 + 1) * NB + i] * bx +
                                      Bz[(q * 3 + 2) * NB + i] * by;
                }
            }
#pragma unroll
          for (int g = 0; g < NFC; ++g)
#pragma unroll
            for (int s = 0; s < QF; ++s)
              {
                double uu[3] = {0, 0, 0}, WN[3] = {0, 0, 0}, LL[3] = {0, 0, 0};
                const int    fi = g * QF + s;
                const double ph = tab[S::tPHIF + fi * NB + ju];
                const double ld = tab[S::tLDF + fi * NB + ju];
                const double lc = tab[S::tLCF + ((fl * NFC + g) * QF + s) * NF + jl];
                const double xv = tab[S::tXI + s * NF + jl];
#pragma unroll
                for (int ci = 0; ci < NE; ++ci)
                  {
                    const bool mine = cc == ci;
                    uu[ci] = (isU && mine) ? ph : 0.0;
                    WN[ci] = (isU && mine) ? ld : (isL && mine) ? lc
                           : isR ? -Wrf[ci * NFC * QF + fi] : 0.0;
                    LL[ci] = (isL && mine && fl == g) ? xv : isR ? -lamf[ci * NFC * QF + fi] : 0.0;
                  }
                const double E0n = W0f[fi];
#pragma unroll
                for (int ci = 0; ci < NE; ++ci)
                  {
                    double Fv;
                    if (ci == 0)
                      Fv = eps * WN[0] + tV * uu[0] - tV * LL[0];
                    else if (ci == 1)
                      {
                        const double n0 = u0f[NFC * QF + fi];
                        Fv = uu[1] * mun * E0n + n0 * mun * WN[0] - UT * mun * WN[1] -
                             tN * uu[1] + tN * LL[1];
                      }
                    else
                      {
                        const int    c2 = NE > 2 ? 2 : 1;
                        const double p0 = u0f[2 * NFC * QF + fi];
                        Fv = -(uu[c2] * mup * E0n + p0 * mup * WN[0]) - UT * mup * WN[c2] -
                             tP * uu[c2] + tP * LL[c2];
                      }
                    Fv += isr * resf[fi * NE + ci];
                    // z rows with support on face g
                    const int a = g >> 1, side = g & 1;
#pragma unroll
                    for (int t = 0; t < K1; ++t)
                      {
                        const int i = a == 0 ? side * DEG + K1 * t : t + K1 * side * DEG;
                        R[ci * NB + i] += Bz[(QC * 3 + fi) * NB + i] * Fv;
                      }
                    const int comp = NE == 1 ? 0 : ci;
                    if (!isdir[g * 3 + comp])
#pragma unroll
                      for (int i = 0; i < NF; ++i)
                        R[NS + g * TDF + ci * NF + i] += Bt[fi * NF + i] * Fv;
                  }
              }
#pragma unroll
          for (int g = 0; g < NFC; ++g)
#pragma unroll
            for (int ci = 0; ci < NE; ++ci)
#pragma unroll
              for (int i = 0; i < NF; ++i)
                R[NS + g * TDF + ci * NF + i] += isr * neu[(g * NE + ci) * NF + i];
          // LocalCondenser::condense_ct_matrix: the Dirichlet update is in the
          // rhs bundle already; the Dirichlet columns are dropped
          if (isL && isdir[ff * 3 + (NE == 1 ? 0 : cc)])
#pragma unroll
            for (int r = 0; r < NROW; ++r)
              R[r] = 0.0;

          // ---- F: elimination, scaled column pivoting
          double cs = 0;
#pragma unroll
          for (int r = 0; r < NS; ++r)
            cs = fmax(cs, fabs(R[r]));
          cs            = cs > 0 ? 1.0 / cs : 0.0;
          bool pivoted = !isU;
#pragma unroll
          for (int k = 0; k < NS; ++k)
            {
              double best = pivoted ? -1.0 : fabs(R[k]) * cs;
              int    bl   = lane;
#pragma unroll
              for (int off = 32; off > 0; off >>= 1)
                {
                  const double ob = __shfl_xor(best, off);
                  const int    ol = __shfl_xor(bl, off);
                  if (ob > best || (ob == best && ol < bl))
                    {
                      best = ob;
                      bl   = ol;
                    }
                }
              const int pl = __builtin_amdgcn_readfirstlane(bl);
              if (lane == pl)
                {
                  pivoted = true;
#pragma unroll
                  for (int r = k; r < NROW; ++r)
                    bc[r] = R[r];
                }
              WAVE_SYNC();
              const double f = R[k] / bc[k];
#pragma unroll
              for (int r = k + 1; r < NROW; ++r)
                R[r] -= bc[r] * f;
              WAVE_SYNC();
            }

          // ---- G: output
          if (mode == 1)
            {
              if (isL)
#pragma unroll
                for (int i = 0; i < T; ++i)
                  p.Sout[(long long)idx * T * T + i * T + (lane - NS)] = R[NS + i];
              if (isR)
#pragma unroll
                for (int i = 0; i < T; ++i)
                  p.gout[(long long)idx * T + i] = R[NS + i];
              continue;
            }
          if (isL)
            {
              const int lt = lane - NS, gj = lt / TDF, rj = lt % TDF;
#pragma unroll
              for (int i = 0; i < T; ++i)
                {
                  const int  gi = i / TDF, ri = i % TDF, ci = ri / NF;
                  const bool dir = isdir[gi * 3 + (NE == 1 ? 0 : ci)];
                  const double v = dir ? (i == lt ? 1.0 : 0.0) : R[NS + i];
                  const long long pos = blk[gi] * TDF * TDF + (long long)ri * nbk[gi] * TDF +
                                        (long long)slot[gi * NFC + gj] * TDF + rj;
                  if (color == 1 && gi == gj && bidv[gi] < 0)
                    p.vals[pos] += v;
                  else
                    p.vals[pos] = v;
                }
            }
          if (isR)
            {
#pragma unroll
              for (int i = 0; i < T; ++i)
                {
                  const int  gi = i / TDF, ri = i % TDF, ci = ri / NF;
                  const int  comp = NE == 1 ? 0 : ci;
                  const bool dir  = isdir[gi * 3 + comp];
                  const long long row = facev[gi] * TDF + ri;
                  const double v = dir ? 0.0 : R[NS + i];
                  if (color == 1 && bidv[gi] < 0)
                    p.rhs[row] += v;
                  else
                    p.rhs[row] = v;
                  if (dir)
                    p.dirout[row] = lupd[(gi * 3 + comp) * NF + (ri % NF)];
                }
            }
          WAVE_SYNC();
        }
    }

    template <int DEG, int NE, bool THERMO>
    hipError_t launch_fast(const KParams &p, const int *cells, int ncell, int color, int mode,
                           hipStream_t s)
    {
      using S = F2<DEG, NE>;
      static bool attr = false;
      if (!attr)
        {
          hipError_t e = hipFuncSetAttribute((const void *)fast_condense<DEG, NE, THERMO>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)S::bytes);
          if (e != hipSuccess)
            return e;
          attr = true;
        }
      if (ncell <= 0)
        return hipSuccess;
      int dev = 0, ncu = 256;
      if (hipGetDevice(&dev) == hipSuccess)
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
      const int want = (ncell + WPB - 1) / WPB;
      const int grid = want < 4 * ncu ? want : 4 * ncu;
      hipLaunchKernelGGL((fast_condense<DEG, NE, THERMO>), dim3(grid), dim3(64 * WPB), S::bytes,
                         s, p, cells, ncell, color, mode);
      return hipGetLastError();
    }
  } // namespace

  bool fast_supported(int dim, int deg, int ne, int thermo, const int en[3])
  {
    if (dim != 2 || deg < 1 || deg > 2)
      return false;
    if (thermo)
      return ne == 1 && en[0];
    return ne == 3;
  }

  hipError_t launch_fast_condense(int deg, const KParams &p, const int *cells, int ncell,
                                  int color, int mode, hipStream_t s)
  {
    if (p.thermo)
      return deg == 1 ? launch_fast<1, 1, true>(p, cells, ncell, color, mode, s)
                      : launch_fast<2, 1, true>(p, cells, ncell, color, mode, s);
    return deg == 1 ? launch_fast<1, 3, false>(p, cells, ncell, color, mode, s)
                    : launch_fast<2, 3, false>(p, cells, ncell, color, mode, s);
  }
} // namespace hdg
