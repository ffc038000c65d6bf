// hdg_fast2dp.hip -- register-resident condensation for 2D cells of degree 3
// (V/n/p), a pair of waves per cell.
//
// A quadrilateral of degree 3 has NS + T + 1 = 48 + 48 + 1 = 97 columns in
// its reduced bordered system, more than one wave's 64 lanes.  Each cell
// gets two waves: wave 0 owns columns 0..63 (the 48 scalar columns and
// trace columns 0..15), wave 1 columns 64..96 (trace columns 16..47 and the
// rhs).  A lane keeps the NS scalar-test rows of its column in VGPRs.
//   A-D. topology, state, point data, lifted residual flux: lane-parallel
//      over the pair (the terms of include/assembling_methods.template.h as
//      in the one-wave kernel)
//   1. scalar-test rows [A' | CT' | c'] of every column, from the column's
//      bundle and the test tables read as scalar loads
//   2. Gauss-Jordan with scaled column pivoting: wave 0 picks the pivot
//      (DPP max over its scalar lanes), the pivot lane writes its column to
//      LDS, and after a workgroup barrier both waves apply the step with the
//      column's entries broadcast by readlane
//   3. face by face: both waves form their trace-test rows (TC' on the
//      scalar lanes, TT'/t' elsewhere); the scalar lanes publish TC' in LDS
//      in pivot order; each wave's TC' X runs on the FP64 matrix cores and
//      [S | g] goes straight into the trace CSR (two colour launches,
//      deterministic)
// A workgroup holds two pairs (two cells) and both run every barrier; in
// the last round a pair without a cell of its own repeats its partner's
// with its stores switched off.  Pass 2 runs the same cell body with the
// trace update folded into the rhs column and writes dx_cell.  Both passes
// in FP64 or with the elimination in FP32 and one FP64 refinement step.
#include <type_traits>
#include <utility>

#include "hdg_device.h"
#include "hdg_kernels.h"
#include "hdg_wave.h"

#define WAVE_SYNCP()                                         \
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
    constexpr int PPB = 2; // cells (wave pairs) per workgroup

    typedef const __attribute__((address_space(4))) double ctabp;

    struct FP
    {
      static constexpr int DEG = 3, K1 = DEG + 1, NB = K1 * K1, NF = K1, QC = NB, QF = NF;
      static constexpr int NFC = 4, NE = 3;
      static constexpr int NCD = 9 * NB;
      static constexpr int NS = NE * NB, TDF = NE * NF, T = NFC * TDF, NCOL = NS + T + 1;
      static_assert(NS <= 64 && NCOL <= 128, "scalar columns in wave 0, two waves");
      static constexpr int GS = 8;                  // elimination steps per unrolled group
      static constexpr int NG = (NS + GS - 1) / GS; // step groups
      // workgroup tables (doubles), the one-wave kernel's layout
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
      // per-cell scratch (doubles), shared by the pair
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
      static constexpr int wBLK  = wNEU + NFC * NE * NF; // NFC long long
      static constexpr int wFACE = wBLK + NFC;           // NFC long long
      static constexpr int wTCF  = wFACE + NFC;          // NS*TDF: TC' rows of one face
      static constexpr int wRING = wTCF + NS * TDF;      // 2*NS: pivot columns of two steps
      static constexpr int wXCH  = wRING + 2 * NS;       // 2 x 4*64: MFMA exchange per wave
      static constexpr int CDBL  = wXCH + 2 * 4 * 64;
      // per-cell ints
      static constexpr int iDIR  = 0;
      static constexpr int iNEU  = iDIR + NFC * 3;
      static constexpr int iBID  = iNEU + NFC * 3;
      static constexpr int iNBK  = iBID + NFC;
      static constexpr int iSLOT = iNBK + NFC;
      static constexpr int iPIV  = iSLOT + NFC * NFC; // NS pivot lanes (FP32 + refinement)
      static constexpr int CINT  = iPIV + NS;
      static constexpr size_t bytes()
      {
        return (size_t)(TAB + PPB * CDBL) * sizeof(double) + (size_t)PPB * CINT * sizeof(int);
      }
    };

    __device__ __forceinline__ void stage_tables_p(const DevTables &tb, double *tab0, int tid)
    {
      using S          = FP;
      constexpr int NB = S::NB, NF = S::NF, QC = S::QC, QF = S::QF, NFC = S::NFC;
      constexpr int NT = 64 * 2 * PPB;
      for (int i = tid; i < QC * NB; i += NT)
        {
          tab0[S::tPHIC + i] = tb.phic[i];
          tab0[S::tPMC + i]  = tb.PhiMinvC[i];
        }
      for (int i = tid; i < QC * 2 * NB; i += NT)
        {
          tab0[S::tLDC + i]   = tb.LDc[i];
          tab0[S::tGRADC + i] = tb.gradc[i];
        }
      for (int i = tid; i < NFC * QC * NF; i += NT)
        tab0[S::tLCC + i] = tb.LCc[i];
      for (int i = tid; i < NFC * QF * NB; i += NT)
        {
          tab0[S::tPHIF + i] = tb.phif[i];
          tab0[S::tLDF + i]  = tb.LDf[i];
          tab0[S::tPMF + i]  = tb.PhiMinvF[i];
        }
      for (int i = tid; i < NFC * NFC * QF * NF; i += NT)
        tab0[S::tLCF + i] = tb.LCf[i];
      for (int i = tid; i < QF * NF; i += NT)
        tab0[S::tXI + i] = tb.xi[i];
      for (int i = tid; i < QC; i += NT)
        tab0[S::tJXWC + i] = tb.JxWc[i];
      for (int i = tid; i < NFC * QF; i += NT)
        tab0[S::tJXWF + i] = tb.JxWf[i];
      for (int i = tid; i < 2 * NF * NF; i += NT)
        tab0[S::tMFI + i] = tb.Mfinv[i];
    }

    // f(integral_constant<int, I>) for I = 0, 1, ..., in order
    template <class F, int... I>
    __device__ __forceinline__ void for_each_group(F &f, std::integer_sequence<int, I...>)
    {
      (f(std::integral_constant<int, I>{}), ...);
    }

    __device__ __forceinline__ unsigned pivot_key_p(double sc, bool cand, int lane)
    {
      return cand ? (leading_bits(sc) & ~63u) + (unsigned)(63 - lane) + 1u : 0u;
    }

    // one cell (one pair of waves), pass 1 (mode 0: CSR scatter; mode 1:
    // S, g of cell idx out) or pass 2 (PASS2: dx_cell of the cell out);
    // live = false: compute, store nothing.  MIXED: the elimination in FP32
    // with one FP64 refinement step
    template <bool PASS2, bool MIXED>
    __device__ __forceinline__ void cell_body_p(const KParams &p, double *smem, long long cell,
                                                long long idx, int color, int mode, bool live)
    {
      using S              = FP;
      constexpr int NB     = S::NB, NF = S::NF, QC = S::QC, QF = S::QF, NFC = S::NFC;
      constexpr int NS     = S::NS, TDF = S::TDF, T = S::T, NE = S::NE;
      constexpr int K1     = S::K1, DEG = S::DEG;
      constexpr bool THERMO = false;
      // the thread index behind an empty asm: everything derived from it is
      // formed again for every cell instead of being hoisted out of the
      // cell loop and kept live (spilled) across it
      int tid_ = threadIdx.x;
      asm volatile("" : "+v"(tid_));
      const int      tid    = tid_;
      // the wave's index is uniform: readfirstlane makes that visible, so
      // branches on half stay scalar
      const int      wv   = __builtin_amdgcn_readfirstlane(tid >> 6);
      const int      lane = tid & 63, pr = wv >> 1, half = wv & 1;
      const int      plane = lane + 64 * half;
      double        *ws  = smem + S::TAB + pr * S::CDBL;
      int *wi = reinterpret_cast<int *>(smem + S::TAB + PPB * S::CDBL) + pr * S::CINT;
      long long *blk   = reinterpret_cast<long long *>(ws + S::wBLK);
      long long *facev = reinterpret_cast<long long *>(ws + S::wFACE);
      const DevTables &tb = p.tb;

      double *u0 = ws + S::wU0, *W0 = ws + S::wW0, *u0f = ws + S::wU0F, *W0f = ws + S::wW0F;
      double *tr0 = ws + S::wTR0, *qd = ws + S::wQD, *rhsq = ws + S::wRHSQ;
      double *Wrc = ws + S::wWRC, *Wrf = ws + S::wWRF, *lamf = ws + S::wLAMF;
      double *lupd = ws + S::wLUPD, *resc = ws + S::wRESC, *resf = ws + S::wRESF;
      double *neu = ws + S::wNEU, *xs = ws + S::wXS, *xt = ws + S::wXT;
      double *tcf = ws + S::wTCF, *ring = ws + S::wRING, *xw = ws + S::wXCH + half * 4 * 64;
      int    *isdir = wi + S::iDIR, *isneu = wi + S::iNEU, *bidv = wi + S::iBID;
      int    *nbk = wi + S::iNBK, *slot = wi + S::iSLOT;
      const double eps = p.eps, mun = p.mun, mup = p.mup, UT = p.UT;
      const double tV = p.tau[0], tN = p.tau[1], tP = p.tau[2];
      const double sd = p.sdrift;

      // column owned by this lane: plane = lane + 64 * half
      int kind, cc = 0, jj = 0, ff = 0;
      if (plane < NS)
        {
          kind = 0;
          cc   = plane / NB;
          jj   = plane % NB;
        }
      else if (plane < NS + T)
        {
          kind         = 1;
          const int lt = plane - NS;
          ff           = lt / TDF;
          cc           = (lt % TDF) / NF;
          jj           = lt % NF;
        }
      else if (plane == NS + T)
        kind = 2;
      else
        kind = 3;

      // the tables are loop invariant; hiding their addresses behind an
      // empty asm keeps the compiler from hoisting every table load out
      // of the cell loop
      double *tab = smem;
      asm volatile("" : "+v"(tab));
      ctabp *Bz = (ctabp *)tb.Bz, *Bt = (ctabp *)tb.Bt;
      int             ic[2];
      ic[0] = (int)(cell % p.n[0]);
      ic[1] = (int)(cell / p.n[0]);
      // ---- A: topology, flags, state
      if (half == 0 && lane < NFC)
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
      if (half == 0 && lane < NFC * NFC)
        slot[lane] = p.cslot[cell * NFC * NFC + lane];
      for (int i = plane; i < S::NCD; i += 128)
        xs[i] = p.xc[cell * S::NCD + i];
      for (int i = plane; i < QC; i += 128)
        qd[5 * QC + i] = p.doping[cell * QC + i];
      __syncthreads();
      for (int i = plane; i < NFC * 3 * NF; i += 128)
        {
          const int lf = i / (3 * NF), r = i % (3 * NF);
          xt[i]        = p.xt[facev[lf] * 3 * NF + r];
        }
      __syncthreads();
      // ---- B: state at quadrature points
      for (int e = plane; e < 3 * 3 * QC; e += 128)
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
      for (int e = plane; e < 3 * NFC * QF; e += 128)
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
      __syncthreads();
      // ---- C: point data, flux-row residual, Dirichlet update
      for (int q = plane; q < QC; q += 128)
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
      for (int e = plane; e < NE * 2 * NB; e += 128)
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
      for (int e = plane; e < NFC * 3 * NF; e += 128)
        {
          const int g = e / (3 * NF), c = (e / NF) % 3, m = e % NF;
          double    v = 0;
          if (PASS2)
            v = p.dxt[facev[g] * 3 * NF + c * NF + m];
          else if (!THERMO && isdir[g * 3 + c])
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
      __syncthreads();
      // ---- D: lifted residual flux with the Dirichlet update folded in;
      //         residual forms at the quadrature points
      for (int e = plane; e < NE * 2 * QC; e += 128)
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
      for (int e = plane; e < NE * NFC * QF; e += 128)
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
      for (int e = plane; e < QC * NE; e += 128)
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
      for (int e = plane; e < NFC * QF * NE; e += 128)
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
      for (int e = plane; e < NFC * NE * NF; e += 128)
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
      __syncthreads();

      // ---- E: bundle of this lane's column (branch-free 0/1 weights)
      const bool   isU = kind == 0, isL = kind == 1, isR = kind == 2;
      const double isr = isR ? 1.0 : 0.0;
      const int    ju = isU ? jj : 0, jl = isL ? jj : 0, fl = isL ? ff : 0;
      // 0/1 lane weights, formed where used (kept live across the
      // elimination they would cost registers it needs)
      auto mU  = [&](int ci) { return (isU && cc == ci) ? 1.0 : 0.0; };
      auto mL  = [&](int ci) { return (isL && cc == ci) ? 1.0 : 0.0; };
      const double mLx = (fl >> 1) == 0 ? 1.0 : 0.0, mLy = 1.0 - mLx;

      // ---- 1a: scalar-test rows, cell quadrature points (q a runtime
      // index, so the uniform Bz loads are not all hoisted)
      double R[NS];
#pragma unroll
      for (int r = 0; r < NS; ++r)
        R[r] = 0.0;
#pragma unroll 1
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
              uu[ci] = mU(ci) * ph;
              WX[ci] = mU(ci) * lx + mL(ci) * mLx * lc - isr * Wrc[(ci * 2 + 0) * QC + q];
              WY[ci] = mU(ci) * ly + mL(ci) * mLy * lc - isr * Wrc[(ci * 2 + 1) * QC + q];
            }
          const double E0x = W0[q], E0y = W0[QC + q];
#pragma unroll
          for (int ci = 0; ci < NE; ++ci)
            {
              double al, bx, by;
              if (ci == 0)
                {
                  al = uu[1] - uu[2];
                  bx = -eps * WX[0];
                  by = -eps * WY[0];
                }
              else if (ci == 1)
                {
                  const double n0 = u0[QC + q];
                  al = -(qd[QC + q] * uu[1] + qd[2 * QC + q] * uu[2]);
                  bx = -mun * E0x * uu[1] + sd * n0 * mun * WX[0] + UT * mun * WX[1];
                  by = -mun * E0y * uu[1] + sd * n0 * mun * WY[0] + UT * mun * WY[1];
                }
              else
                {
                  const double p0 = u0[2 * QC + q];
                  al = -(qd[QC + q] * uu[1] + qd[2 * QC + q] * uu[2]);
                  bx = mup * E0x * uu[2] - sd * p0 * mup * WX[0] + UT * mup * WX[2];
                  by = mup * E0y * uu[2] - sd * p0 * mup * WY[0] + UT * mup * WY[2];
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

      // the numerical flux F_c of this lane's column at face point (g, s)
      // (the HDG flux of component c, use_cell); it feeds the scalar-test
      // rows (phi_i F_c) and the trace-test rows (xi_i F_c)
      auto face_flux = [&](int g, int s, double (&F)[NE]) __attribute__((always_inline)) {
        const int    fi  = g * QF + s;
        const double ph  = tab[S::tPHIF + fi * NB + ju];
        const double ld  = tab[S::tLDF + fi * NB + ju];
        const double lc  = tab[S::tLCF + ((fl * NFC + g) * QF + s) * NF + jl];
        const double xv  = tab[S::tXI + s * NF + jl] * (fl == g ? 1.0 : 0.0);
        const double E0n = W0f[fi];
        const double WN0 = mU(0) * ld + mL(0) * lc - isr * Wrf[fi];
#pragma unroll
        for (int ci = 0; ci < NE; ++ci)
          {
            const double uuc = mU(ci) * ph;
            const double WNc = mU(ci) * ld + mL(ci) * lc - isr * Wrf[ci * NFC * QF + fi];
            const double LLc = mL(ci) * xv - isr * lamf[ci * NFC * QF + fi];
            double       Fv;
            if (ci == 0)
              Fv = eps * WNc + tV * uuc - tV * LLc;
            else if (ci == 1)
              Fv = uuc * mun * E0n + u0f[NFC * QF + fi] * mun * WN0 - UT * mun * WNc -
                   tN * uuc + tN * LLc;
            else
              Fv = -(uuc * mup * E0n + u0f[2 * NFC * QF + fi] * mup * WN0) -
                   UT * mup * WNc - tP * uuc + tP * LLc;
            F[ci] = Fv + isr * resf[fi * NE + ci];
          }
      };
      // trace-test rows of face g: xi_i F_c summed over the face points,
      // Neumann data on the rhs lane, zero rows on Dirichlet faces
      auto trace_rows = [&](int g, double (&acc)[TDF]) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < TDF; ++r)
          acc[r] = 0.0;
#pragma unroll 1
        for (int s = 0; s < QF; ++s)
          {
            double F[NE];
            face_flux(g, s, F);
#pragma unroll
            for (int ci = 0; ci < NE; ++ci)
#pragma unroll
              for (int i = 0; i < NF; ++i)
                acc[ci * NF + i] += Bt[(g * QF + s) * NF + i] * F[ci];
          }
#pragma unroll
        for (int ci = 0; ci < NE; ++ci)
          {
            const double keepr = isdir[g * 3 + ci] ? 0.0 : 1.0;
#pragma unroll
            for (int i = 0; i < NF; ++i)
              acc[ci * NF + i] = keepr * (acc[ci * NF + i] + isr * neu[(g * NE + ci) * NF + i]);
          }
      };

      // ---- 1b: face quadrature points; only the K1 basis functions of the
      // face's support are nonzero there (g unrolled: the row indices must
      // be constants for R to stay in registers)
#pragma unroll
      for (int g = 0; g < NFC; ++g)
        {
          const int a = g >> 1, side = g & 1;
#pragma unroll 1
          for (int s = 0; s < QF; ++s)
            {
              const int fi = g * QF + s;
              double    F[NE];
              face_flux(g, s, F);
#pragma unroll
              for (int ci = 0; ci < NE; ++ci)
#pragma unroll
                for (int t = 0; t < K1; ++t)
                  {
                    const int i = a == 0 ? side * DEG + K1 * t : t + K1 * side * DEG;
                    R[ci * NB + i] += Bz[(QC * 3 + fi) * NB + i] * F[ci];
                  }
            }
        }
      // LocalCondenser::condense_ct_matrix: the Dirichlet update is in the
      // rhs bundle already; the Dirichlet columns are dropped
      const bool   dircol = isL && isdir[ff * 3 + cc];
      const double keep   = dircol ? 0.0 : 1.0;
#pragma unroll
      for (int r = 0; r < NS; ++r)
        R[r] *= keep;

      // ---- 2: Gauss-Jordan with scaled column pivoting.  Step k: wave 0
      // takes its unpivoted scalar lane with the largest |entry| / column
      // max, and that lane writes its column to the ring; after a workgroup
      // barrier both waves read the column back (lane r holds entry r) and
      // apply the same update, each entry broadcast by readlane where it is
      // used.  MIXED: that elimination in FP32 (the pivot lanes recorded),
      // the FP64 residual r = b - A' x0 of every column (the scalar lanes
      // keep their FP64 columns of A' in R and publish them through the ring
      // in pivot order), the same FP32 elimination with the recorded pivots
      // on r, and x = x0 + dx.
      int  myk  = 0; // the equation this scalar lane was the pivot of
      int *pivs = wi + S::iPIV;
      // one elimination pass over Rr (double or float): SEARCH picks the
      // pivots (and records them), otherwise they are read from pivs
      auto elim = [&](auto &Rr, auto search) __attribute__((always_inline)) {
        using Real            = std::remove_reference_t<decltype(Rr[0])>;
        constexpr bool SEARCH = decltype(search)::value;
        Real           cs     = 0;
        bool           pivoted = !isU;
        if constexpr (SEARCH)
          {
#pragma unroll
            for (int r = 0; r < NS; ++r)
              cs = fmax(cs, fabs(Rr[r]));
            cs = cs > Real(0) ? Real(1) / cs : Real(0);
          }
        // the steps in groups of GS, expanded at compile time so that every
        // step index is a constant
        auto group = [&](auto gc) __attribute__((always_inline)) {
          constexpr int G = decltype(gc)::value;
#pragma unroll
          for (int j = 0; j < S::GS; ++j)
            {
              const int k = G * S::GS + j;
              if (k < NS)
                {
                  double *col = ring + (k & 1) * NS;
                  if (half == 0)
                    {
                      int pl;
                      if constexpr (SEARCH)
                        {
                          const unsigned key =
                            wave_max(pivot_key_p(fabs(Rr[k]) * cs, !pivoted, lane));
                          pl = 63 - (int)((key - 1u) & 63u);
                          if (lane == pl)
                            {
                              pivoted = true;
                              myk     = k;
                            }
                          if (lane == 0)
                            pivs[k] = pl;
                        }
                      else
                        pl = __builtin_amdgcn_readfirstlane(pivs[k]);
                      if (lane == pl)
#pragma unroll
                        for (int r = 0; r < NS; ++r)
                          col[r] = (double)Rr[r];
                    }
                  __syncthreads();
                  const Real v = (Real)col[lane < NS ? lane : 0];
                  const Real f = Rr[k] / lane_value(v, k);
#pragma unroll
                  for (int r = 0; r < NS; ++r)
                    if (r != k)
                      Rr[r] -= lane_value(v, r) * f;
                  Rr[k] = f;
                }
            }
        };
        for_each_group(group, std::make_integer_sequence<int, S::NG>{});
      };
      if constexpr (!MIXED)
        elim(R, std::true_type{});
      else
        {
          float R32[NS], X0[NS];
#pragma unroll
          for (int r = 0; r < NS; ++r)
            R32[r] = (float)R[r];
          elim(R32, std::true_type{});
#pragma unroll
          for (int r = 0; r < NS; ++r)
            X0[r] = R32[r];
          // r = b - A' x0: step k takes the FP64 column of the pivot lane
          // of equation k through the ring
          const double mres = isU ? 0.0 : 1.0;
          auto resid = [&](auto gc) __attribute__((always_inline)) {
            constexpr int G = decltype(gc)::value;
#pragma unroll
            for (int j = 0; j < S::GS; ++j)
              {
                const int k = G * S::GS + j;
                if (k < NS)
                  {
                    double *col = ring + (k & 1) * NS;
                    if (half == 0)
                      {
                        const int pl = __builtin_amdgcn_readfirstlane(pivs[k]);
                        if (lane == pl)
#pragma unroll
                          for (int r = 0; r < NS; ++r)
                            col[r] = R[r];
                      }
                    __syncthreads();
                    const double v  = col[lane < NS ? lane : 0];
                    const double xk = mres * (double)X0[k];
#pragma unroll
                    for (int r = 0; r < NS; ++r)
                      R[r] -= lane_value(v, r) * xk;
                  }
              }
          };
          for_each_group(resid, std::make_integer_sequence<int, S::NG>{});
#pragma unroll
          for (int r = 0; r < NS; ++r)
            R32[r] = (float)R[r];
          elim(R32, std::false_type{});
#pragma unroll
          for (int k = 0; k < NS; ++k)
            R[k] = (double)X0[k] + (double)R32[k];
        }

      if constexpr (PASS2)
        {
          // the rhs lane holds u(p_k) in row k: u reaches the scalar lanes
          // through the ring (lane p reads row myk), then W = M^-1 (rhs_q +
          // D_a u - sum_f C_{a,f} lambda_upd,f) and u go out lane-parallel
          // (reference template.h:2688-2697)
          __syncthreads();
          if (isR)
#pragma unroll
            for (int k = 0; k < NS; ++k)
              ring[k] = R[k];
          __syncthreads();
          double *sol = ring + NS;
          if (isU)
            sol[plane] = ring[myk];
          __syncthreads();
          for (int e = plane; e < S::NCD; e += 128)
            {
              const int c = e / (3 * NB), r = e % (3 * NB), blkc = r / NB, i = r % NB;
              double    v = 0;
              const double *u = sol + c * NB;
              if (blkc == 2)
                v = u[i];
              else
                {
                  const int a = blkc;
#pragma unroll
                  for (int l = 0; l < NB; ++l)
                    v += tb.Minv[i * NB + l] * rhsq[(c * 2 + a) * NB + l] +
                         tb.MinvD[(a * NB + i) * NB + l] * u[l];
#pragma unroll
                  for (int sd2 = 0; sd2 < 2; ++sd2)
                    {
                      const int f = 2 * a + sd2;
#pragma unroll
                      for (int m = 0; m < NF; ++m)
                        v -= tb.MinvC[(f * NB + i) * NF + m] * lupd[(f * 3 + c) * NF + m];
                    }
                }
              if (live)
                p.dxc[cell * S::NCD + e] = v;
            }
          // the next cell rewrites the ring and the pair's scratch
          __syncthreads();
          return;
        }

      // ---- 3: [S | g] = [TT' | t'] - TC' X, face by face.  The product
      // runs on the matrix cores (v_mfma_f64_16x16x4f64, K = 4 pivot rows per
      // step) over three 16-lane column tiles per wave from tile t0.  X
      // reaches the B layout through the wave's exchange buffer once, four
      // rows at a time, and stays in registers; per face, TC' (TDF rows
      // padded to 16) is read from tcf in the A layout, and the result
      // returns to column layout through the exchange buffer.
      constexpr int NKK = NS / 4;
      static_assert(NS % 4 == 0, "K steps of four pivot rows");
      const int q16 = lane >> 4, c16 = lane & 15;
      // wave 0's trace and rhs lanes (NS..63) lie in tiles 1..3
      const int t0  = half == 0 ? 1 : 0;
      static_assert(NS >= 16 && NS + T + 1 - 64 <= 48, "tiles 1-3 of wave 0, 0-2 of wave 1");
      double    bop[NKK][3];
#pragma unroll
      for (int kk = 0; kk < NKK; ++kk)
        {
#pragma unroll
          for (int kq = 0; kq < 4; ++kq)
            xw[kq * 64 + lane] = R[4 * kk + kq];
          WAVE_SYNCP();
#pragma unroll
          for (int t = 0; t < 3; ++t)
            bop[kk][t] = xw[q16 * 64 + 16 * (t0 + t) + c16];
          WAVE_SYNCP();
        }
      const bool isC = isL || isR;
      const int  cl  = lane - 16 * t0;
      const int  lt  = plane - NS;
#pragma unroll 1
      for (int g = 0; g < NFC; ++g)
        {
          double acc[TDF];
          trace_rows(g, acc);
          // the scalar columns publish their TC' rows of this face in pivot
          // order (row k of X belongs to the pivot of equation k)
          if (isU)
#pragma unroll
            for (int r = 0; r < TDF; ++r)
              tcf[myk * TDF + r] = acc[r];
          __syncthreads();
          dbl4 dm[3];
#pragma unroll
          for (int t = 0; t < 3; ++t)
            dm[t] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int kk = 0; kk < NKK; ++kk)
            {
              const double v = tcf[(4 * kk + q16) * TDF + (c16 < TDF ? c16 : 0)];
              const double a = c16 < TDF ? v : 0.0;
#pragma unroll
              for (int t = 0; t < 3; ++t)
                dm[t] = mfma_f64_16x16x4(a, bop[kk][t], dm[t]);
            }
          // element e of dm[t] holds row 4e + q16, column 16 (t0 + t) + c16
#pragma unroll
          for (int e = 0; e < (TDF + 3) / 4; ++e)
            {
              WAVE_SYNCP();
#pragma unroll
              for (int t = 0; t < 3; ++t)
                xw[q16 * 48 + 16 * t + c16] = dm[t][e];
              WAVE_SYNCP();
              if (isC)
#pragma unroll
                for (int qq = 0; qq < 4; ++qq)
                  if (4 * e + qq < TDF)
                    acc[4 * e + qq] -= xw[qq * 48 + cl];
            }
          if (live && mode == 1)
            {
              if (isL)
#pragma unroll
                for (int r = 0; r < TDF; ++r)
                  p.Sout[(long long)idx * T * T + (g * TDF + r) * T + lt] = dircol ? 0.0 : acc[r];
              if (isR)
#pragma unroll
                for (int r = 0; r < TDF; ++r)
                  p.gout[(long long)idx * T + g * TDF + r] = acc[r];
            }
          if (live && mode == 0 && isL)
            {
              const int gj = lt / TDF, rj = lt % TDF;
              const long long base = blk[g] * TDF * TDF + (long long)slot[g * NFC + gj] * TDF + rj;
              const long long rs   = (long long)nbk[g] * TDF;
              const bool      addd = color == 1 && g == gj && bidv[g] < 0;
#pragma unroll
              for (int r = 0; r < TDF; ++r)
                {
                  const int  i   = g * TDF + r, ci = r / NF;
                  const bool dir = isdir[g * 3 + ci];
                  const double v = dir ? (i == lt ? 1.0 : 0.0) : (dircol ? 0.0 : acc[r]);
                  const long long pos = base + r * rs;
                  if (addd)
                    p.vals[pos] += v;
                  else
                    p.vals[pos] = v;
                }
            }
          if (live && mode == 0 && isR)
            {
              const bool addr = color == 1 && bidv[g] < 0;
#pragma unroll
              for (int r = 0; r < TDF; ++r)
                {
                  const int  ci  = r / NF;
                  const bool dir = isdir[g * 3 + ci];
                  const long long row = facev[g] * TDF + r;
                  const double v = dir ? 0.0 : acc[r];
                  if (addr)
                    p.rhs[row] += v;
                  else
                    p.rhs[row] = v;
                  if (dir)
                    p.dirout[row] = lupd[(g * 3 + ci) * NF + (r % NF)];
                }
            }
          // the next face rewrites tcf
          __syncthreads();
        }
    }

    template <bool MIXED>
    __global__ void __launch_bounds__(64 * 2 * PPB) __attribute__((amdgpu_waves_per_eu(2)))
      fast_condense_p(KParams p, const int *__restrict__ cells, int ncell, int color, int mode)
    {
      extern __shared__ double smem[];
      stage_tables_p(p.tb, smem, threadIdx.x);
      __syncthreads();
      const int pr = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 7);
      // every pair of the workgroup runs every round (the barriers are
      // workgroup-wide); a pair past the end repeats its partner's cell
      for (int base = blockIdx.x * PPB; base < ncell; base += gridDim.x * PPB)
        {
          const int  idx  = base + pr;
          const bool live = idx < ncell;
          cell_body_p<false, MIXED>(p, smem, (long long)cells[live ? idx : base],
                                    live ? idx : base, color, mode, live);
        }
    }
    // pass 2 over all cells in order (dx_cell out)
    template <bool MIXED>
    __global__ void __launch_bounds__(64 * 2 * PPB) __attribute__((amdgpu_waves_per_eu(2)))
      reconstruct_p(KParams p, int ncell)
    {
      extern __shared__ double smem[];
      stage_tables_p(p.tb, smem, threadIdx.x);
      __syncthreads();
      const int pr = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 7);
      for (int base = blockIdx.x * PPB; base < ncell; base += gridDim.x * PPB)
        {
          const int  idx  = base + pr;
          const bool live = idx < ncell;
          const int  use  = live ? idx : base;
          cell_body_p<true, MIXED>(p, smem, use, use, 0, 0, live);
        }
    }
    // pass 1 over the cells of several degree-3 meshes in one persistent
    // launch per colour: the work list holds chunks of PPB cells of one
    // device (padded with -1 at device boundaries; a chunk's first cell is
    // real); a workgroup restages the tables when its chunk's device changes
    template <bool MIXED>
    __global__ void __launch_bounds__(64 * 2 * PPB) __attribute__((amdgpu_waves_per_eu(2)))
      batch_condense_p(const KParams *__restrict__ plist, const int *__restrict__ chunk_dev,
                       const int *__restrict__ items, int nchunks, int color)
    {
      extern __shared__ double smem[];
      const int pr  = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 7);
      int       cur = -1;
      for (int c = blockIdx.x; c < nchunks; c += gridDim.x)
        {
          const int      dev = chunk_dev[c];
          const KParams &p   = plist[dev];
          if (dev != cur)
            {
              __syncthreads();
              stage_tables_p(p.tb, smem, threadIdx.x);
              __syncthreads();
              cur = dev;
            }
          const int  cell = items[c * PPB + pr];
          const bool live = cell >= 0;
          const int  use  = live ? cell : items[c * PPB];
          cell_body_p<false, MIXED>(p, smem, use, use, color, 0, live);
        }
    }

    // the kernel's dynamic LDS attribute, once per instantiation
    template <class K>
    hipError_t lds_attr(K kernel, bool &done)
    {
      if (done)
        return hipSuccess;
      hipError_t e = hipFuncSetAttribute((const void *)kernel,
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)FP::bytes());
      if (e == hipSuccess)
        done = true;
      return e;
    }

    int persistent_grid(int ncell)
    {
      int dev = 0, ncu = 256;
      if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
      // four rounds of the two workgroups a CU holds (192^2 pass 1: 4.93
      // ms at one round, 4.81 ms at four)
      const int want = (ncell + PPB - 1) / PPB;
      return want < 8 * ncu ? want : 8 * ncu;
    }
  } // namespace

  int batch_pairs_per_group()
  {
    return PPB;
  }

  hipError_t launch_batch_pair(const KParams *plist, const int *chunk_dev, const int *items,
                               int nchunks, int color, int grid, int mixed, hipStream_t s)
  {
    // FP32 + refinement at precision 2 only: three barrier passes of the
    // pair where FP64 has one (slower than FP64)
    static bool a0 = false, a1 = false;
    const bool  mx = mixed == 2;
    hipError_t  e  = mx ? lds_attr(batch_condense_p<true>, a1)
                        : lds_attr(batch_condense_p<false>, a0);
    if (e != hipSuccess)
      return e;
    if (nchunks <= 0)
      return hipSuccess;
    if (mx)
      hipLaunchKernelGGL(batch_condense_p<true>, dim3(grid), dim3(64 * 2 * PPB), FP::bytes(), s,
                         plist, chunk_dev, items, nchunks, color);
    else
      hipLaunchKernelGGL(batch_condense_p<false>, dim3(grid), dim3(64 * 2 * PPB), FP::bytes(), s,
                         plist, chunk_dev, items, nchunks, color);
    return hipGetLastError();
  }

  bool fast2p_supported(int dim, int deg, int ne, int thermo, const int en[3])
  {
    (void)en;
    return dim == 2 && deg == 3 && ne == 3 && !thermo;
  }

  hipError_t launch_fast2p_condense(const KParams &p, const int *cells, int ncell, int color,
                                    int mode, hipStream_t s)
  {
    static bool a0 = false, a1 = false;
    const bool  mx = p.mixed == 2;
    hipError_t  e  = mx ? lds_attr(fast_condense_p<true>, a1)
                        : lds_attr(fast_condense_p<false>, a0);
    if (e != hipSuccess)
      return e;
    if (ncell <= 0)
      return hipSuccess;
    const int grid = persistent_grid(ncell);
    if (mx)
      hipLaunchKernelGGL(fast_condense_p<true>, dim3(grid), dim3(64 * 2 * PPB), FP::bytes(), s,
                         p, cells, ncell, color, mode);
    else
      hipLaunchKernelGGL(fast_condense_p<false>, dim3(grid), dim3(64 * 2 * PPB), FP::bytes(), s,
                         p, cells, ncell, color, mode);
    return hipGetLastError();
  }

  hipError_t launch_fast2p_reconstruct(const KParams &p, int ncells, hipStream_t s)
  {
    static bool a0 = false, a1 = false;
    const bool  mx = p.mixed == 2;
    hipError_t  e  = mx ? lds_attr(reconstruct_p<true>, a1)
                        : lds_attr(reconstruct_p<false>, a0);
    if (e != hipSuccess)
      return e;
    if (ncells <= 0)
      return hipSuccess;
    const int grid = persistent_grid(ncells);
    if (mx)
      hipLaunchKernelGGL(reconstruct_p<true>, dim3(grid), dim3(64 * 2 * PPB), FP::bytes(), s, p,
                         ncells);
    else
      hipLaunchKernelGGL(reconstruct_p<false>, dim3(grid), dim3(64 * 2 * PPB), FP::bytes(), s, p,
                         ncells);
    return hipGetLastError();
  }
} // namespace hdg
