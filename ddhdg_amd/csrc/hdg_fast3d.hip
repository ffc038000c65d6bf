// hdg_fast3d.hip -- register-resident condensation for 3D cells of degree 1.
//
// One wave64 per cell.  The reduced bordered system of a hexahedron has
// NS + T + 1 columns (24 scalar, 72 trace and the rhs with V/n/p: 97), so a
// lane owns up to two columns ("slots": column lane and column lane + 64).
// Each slot keeps the NS scalar-test rows of its column in VGPRs; the
// scalar columns are all in slot 0 (lanes 0..NS-1).
//   1. scalar-test rows [A' | CT' | c'] of both slots in registers, built
//      from the slot's bundle with the test tables read as scalar loads
//      (the terms of include/assembling_methods.template.h:269-2118 after
//      the flux unknowns are eliminated, hdg_tables.h)
//   2. Gauss-Jordan on the scalar rows with scaled column pivoting: the
//      pivot lane by a DPP max over the slot-0 scalar columns, its column to
//      every lane by readlane; both slots are updated.  Afterwards row k of a
//      trace/rhs slot holds X(p_k, m), X = A'^-1 [CT' | c']
//   3. face by face: every slot forms its trace-test rows (TC' for the
//      scalar columns, TT'/t' for the others); the scalar lanes publish
//      their TC' rows of the face in LDS, the product TC' X runs on the
//      FP64 matrix cores, the trace/rhs slots subtract it and write [S | g]
//      straight into the trace CSR (two colour launches, deterministic)
// TC' is recomputed face by face instead of being kept for the whole cell:
// that keeps the per-wave LDS at 11 KB and two workgroups of four waves per
// CU.  Optional FP32 elimination with one FP64 refinement step (MIXED).
#include "hdg_device.h"
#include "hdg_kernels.h"
#include "hdg_wave.h"

#define WAVE_SYNC3()                                         \
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
    // waves per workgroup: four, three with the larger FP32+refine scratch
    // (two workgroups per CU either way)
    __host__ __device__ constexpr int wpb3(bool mixed) { return mixed ? 3 : 4; }

    typedef const __attribute__((address_space(4))) double ctab3;

    template <int NE>
    struct F3
    {
      static constexpr int DIM = 3, K1 = 2, NB = 8, NF = 4, QC = 8, QF = 4, NFC = 6;
      static constexpr int NCD = 3 * (DIM + 1) * NB;
      static constexpr int NS = NE * NB, TDF = NE * NF, T = NFC * TDF;
      static constexpr int NCOL = NS + T + 1;
      static constexpr int NH   = (NCOL + 63) / 64; // column slots per lane
      static_assert(NH <= 2, "two columns per lane at most");
      static_assert(NS <= 64, "the scalar columns live in slot 0");
      // workgroup tables (doubles), read at lane-dependent indices
      static constexpr int tPHIC  = 0;
      static constexpr int tLDC   = tPHIC + QC * NB;
      static constexpr int tLCC   = tLDC + QC * DIM * NB;
      static constexpr int tPHIF  = tLCC + NFC * QC * NF;
      static constexpr int tLDF   = tPHIF + NFC * QF * NB;
      static constexpr int tLCF   = tLDF + NFC * QF * NB;
      static constexpr int tXI    = tLCF + NFC * NFC * QF * NF;
      static constexpr int tGRADC = tXI + QF * NF;
      static constexpr int tJXWC  = tGRADC + QC * NB * DIM;
      static constexpr int tJXWF  = tJXWC + QC;
      static constexpr int tPMC   = tJXWF + NFC * QF;
      static constexpr int tPMF   = tPMC + QC * NB;
      static constexpr int tMFI   = tPMF + NFC * QF * NB;
      static constexpr int TAB    = tMFI + DIM * NF * NF;
      // per-wave scratch (doubles)
      static constexpr int wXS   = 0;
      static constexpr int wXT   = wXS + NCD;
      static constexpr int wU0   = wXT + NFC * 3 * NF;
      static constexpr int wW0   = wU0 + 3 * QC;
      static constexpr int wU0F  = wW0 + 3 * DIM * QC;
      static constexpr int wW0F  = wU0F + 3 * NFC * QF;
      static constexpr int wTR0  = wW0F + 3 * NFC * QF;
      static constexpr int wQD   = wTR0 + 3 * NFC * QF;
      static constexpr int wRHSQ = wQD + 6 * QC;
      static constexpr int wWRC  = wRHSQ + 3 * DIM * NB;
      static constexpr int wWRF  = wWRC + 3 * DIM * QC;
      static constexpr int wLAMF = wWRF + 3 * NFC * QF;
      static constexpr int wLUPD = wLAMF + 3 * NFC * QF;
      static constexpr int wRESC = wLUPD + NFC * 3 * NF;
      static constexpr int wRESF = wRESC + QC * NE * (DIM + 1);
      static constexpr int wNEU  = wRESF + NFC * QF * NE;
      static constexpr int wBLK  = wNEU + NFC * NE * NF; // NFC long long
      static constexpr int wFACE = wBLK + NFC;           // NFC long long
      static constexpr int wTCF  = wFACE + NFC;          // TC' rows of one face; the MFMA exchange
      static constexpr int TCFN  = NS * TDF > 4 * 64 ? NS * TDF : 4 * 64;
      static constexpr int wGJR  = wTCF + TCFN;          // NS*NS: pivot columns (FP64 mode)
      static constexpr int wAPR  = wGJR;                 // NS*NS: A' columns (MIXED)
      static constexpr int wGJ32 = wAPR + NS * NS;       // NS*NS floats: pivot columns (MIXED)
      static constexpr int WDBL_FP64  = wGJR + NS * NS;
      static constexpr int WDBL_MIXED = wGJ32 + (NS * NS + 1) / 2;
      // per-wave ints
      static constexpr int iDIR  = 0;
      static constexpr int iNEU  = iDIR + NFC * 3;
      static constexpr int iBID  = iNEU + NFC * 3;
      static constexpr int iNBK  = iBID + NFC;
      static constexpr int iSLOT = iNBK + NFC;
      static constexpr int iPIV  = iSLOT + NFC * NFC;
      static constexpr int WINT  = iPIV + NS;
      static constexpr int wdbl(bool mixed) { return mixed ? WDBL_MIXED : WDBL_FP64; }
      static constexpr size_t bytes(bool mixed)
      {
        return (size_t)(TAB + wpb3(mixed) * wdbl(mixed)) * sizeof(double) +
               (size_t)wpb3(mixed) * WINT * sizeof(int);
      }
    };

    template <int NE, int NT>
    __device__ __forceinline__ void stage_tables3(const DevTables &tb, double *tab0, int tid)
    {
      using S          = F3<NE>;
      constexpr int NB = S::NB, NF = S::NF, QC = S::QC, QF = S::QF, NFC = S::NFC, D = S::DIM;
      for (int i = tid; i < QC * NB; i += NT)
        {
          tab0[S::tPHIC + i] = tb.phic[i];
          tab0[S::tPMC + i]  = tb.PhiMinvC[i];
        }
      for (int i = tid; i < QC * D * NB; i += NT)
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
      for (int i = tid; i < D * NF * NF; i += NT)
        tab0[S::tMFI + i] = tb.Mfinv[i];
    }

    // Gauss-Jordan over the NS scalar rows of the lanes' slot-0 columns
    // with scaled column pivoting: equation k takes the unpivoted scalar
    // lane (isU) with the largest |entry| / column max (DPP max), its column
    // reaches every lane by readlane and is recorded in rec[k*NS..] by the
    // pivot lane (for the slot-1 replay).  pivs (LDS) by lane 0; myk: the
    // equation a scalar lane was the pivot of.
    template <int NS, typename Real>
    __device__ __forceinline__ void gauss_jordan_rec(Real (&R)[NS], bool isU, int lane,
                                                     int *pivs, int &myk, Real *rec)
    {
      Real cs = 0;
#pragma unroll
      for (int r = 0; r < NS; ++r)
        cs = fmax(cs, fabs(R[r]));
      cs           = cs > Real(0) ? Real(1) / cs : Real(0);
      bool pivoted = !isU;
#pragma unroll
      for (int k = 0; k < NS; ++k)
        {
          const Real     sc  = fabs(R[k]) * cs;
          const unsigned key =
            wave_max(pivoted ? 0u
                             : ((sizeof(Real) == 8 ? leading_bits((double)sc)
                                                   : __float_as_uint((float)sc)) &
                                ~63u) +
                                 (unsigned)(63 - lane) + 1u);
          const int pl = 63 - (int)((key - 1u) & 63u);
          if (lane == pl)
            {
              pivoted = true;
              myk     = k;
#pragma unroll
              for (int r = 0; r < NS; ++r)
                rec[k * NS + r] = R[r];
            }
          if (lane == 0)
            pivs[k] = pl;
          Real bq[NS];
#pragma unroll
          for (int r = 0; r < NS; ++r)
            bq[r] = lane_value(R[r], pl);
          const Real f = pivot_div(R[k], bq[k]);
#pragma unroll
          for (int r = 0; r < NS; ++r)
            if (r != k)
              R[r] -= bq[r] * f;
          R[k] = f;
        }
    }

    // the recorded elimination applied to another column
    template <int NS, typename Real>
    __device__ __forceinline__ void gauss_jordan_replay(Real (&R)[NS], const Real *rec)
    {
#pragma unroll
      for (int k = 0; k < NS; ++k)
        {
          const Real *bq = rec + k * NS;
          const Real  f  = pivot_div(R[k], bq[k]);
#pragma unroll
          for (int r = 0; r < NS; ++r)
            if (r != k)
              R[r] -= bq[r] * f;
          R[k] = f;
        }
    }

    // one cell (one wave), pass 1 (mode 0: CSR scatter; mode 1: S, g out),
    // or PASS2: the reconstruction, columns = the scalar dofs and one rhs
    // column in slot 0 (the rhs bundle carries the whole trace update)
    template <int NE, bool THERMO, bool MIXED, bool PASS2 = false>
    __device__ __forceinline__ void cell_body3(const KParams &p, double *smem, long long cell,
                                               long long idx, int color, int mode)
    {
      using S          = F3<NE>;
      constexpr int D  = S::DIM;
      constexpr int NB = S::NB, NF = S::NF, QC = S::QC, QF = S::QF, NFC = S::NFC;
      constexpr int NS = S::NS, TDF = S::TDF, T = S::T, NH = S::NH, NCOL = S::NCOL;
      // the tables are cell-loop invariant; hiding their address behind an
      // empty asm keeps the compiler from hoisting every table load out of
      // the persistent cell loop (hundreds of live values)
      double *tab = smem;
      asm volatile("" : "+v"(tab));
      // the thread index behind an empty asm as well: everything derived
      // from it is formed again for every cell instead of being hoisted
      // out of the cell loop and kept live (spilled) across it
      int tid_ = threadIdx.x;
      asm volatile("" : "+v"(tid_));
      const int     tid = tid_;
      const int     lane = tid & 63, wv = tid >> 6;
      constexpr int WST = S::wdbl(MIXED);
      double       *ws  = smem + S::TAB + wv * WST;
      int *wi = reinterpret_cast<int *>(smem + S::TAB + wpb3(MIXED) * WST) + wv * S::WINT;
      long long *blk   = reinterpret_cast<long long *>(ws + S::wBLK);
      long long *facev = reinterpret_cast<long long *>(ws + S::wFACE);
      const DevTables &tb = p.tb;
      double *u0 = ws + S::wU0, *W0 = ws + S::wW0, *u0f = ws + S::wU0F, *W0f = ws + S::wW0F;
      double *tr0 = ws + S::wTR0, *qd = ws + S::wQD, *rhsq = ws + S::wRHSQ;
      double *Wrc = ws + S::wWRC, *Wrf = ws + S::wWRF, *lamf = ws + S::wLAMF;
      double *lupd = ws + S::wLUPD, *resc = ws + S::wRESC, *resf = ws + S::wRESF;
      double *neu = ws + S::wNEU, *xs = ws + S::wXS, *xt = ws + S::wXT, *tcf = ws + S::wTCF;
      double *gjr = ws + S::wGJR, *apr = ws + S::wAPR;
      float  *gjr32 = reinterpret_cast<float *>(ws + S::wGJ32);
      int    *isdir = wi + S::iDIR, *isneu = wi + S::iNEU, *bidv = wi + S::iBID;
      int    *nbk = wi + S::iNBK, *slot = wi + S::iSLOT, *pivs = wi + S::iPIV;
      const double eps = p.eps, mun = p.mun, mup = p.mup, UT = p.UT;
      const double tV = p.tau[0], tN = p.tau[1], tP = p.tau[2];
      const double sd = p.sdrift;
      ctab3 *Bz = (ctab3 *)tb.Bz, *Bt = (ctab3 *)tb.Bt;

      int ic[3];
      ic[0] = (int)(cell % p.n[0]);
      ic[1] = (int)((cell / p.n[0]) % p.n[1]);
      ic[2] = (int)(cell / ((long long)p.n[0] * p.n[1]));
      // ---- A: topology, flags, state
      if (lane < NFC)
        {
          const int lf = lane, a = lf >> 1, s1 = lf & 1;
          const long long F = face_of(p, ic[0], ic[1], ic[2], lf);
          facev[lf]         = F;
          blk[lf]           = p.blkptr[F];
          nbk[lf]           = p.nbk[F];
          const bool bnd    = s1 ? (ic[a] == p.n[a] - 1) : (ic[a] == 0);
          const int  bid    = bnd ? lf : -1;
          bidv[lf]          = bid;
          for (int c = 0; c < 3; ++c)
            {
              const int bt      = (bid >= 0 && p.en[c]) ? p.bc[bid][c] : 0;
              isdir[lf * 3 + c] = bt == 1;
              isneu[lf * 3 + c] = bt == 2 && !THERMO;  // zero Neumann data at equilibrium
            }
        }
      if (lane < NFC * NFC)
        slot[lane] = p.cslot[cell * NFC * NFC + lane];
      for (int i = lane; i < S::NCD; i += 64)
        xs[i] = p.xc[cell * S::NCD + i];
      for (int i = lane; i < QC; i += 64)
        qd[5 * QC + i] = p.doping[cell * QC + i];
      WAVE_SYNC3();
      for (int i = lane; i < NFC * 3 * NF; i += 64)
        {
          const int lf = i / (3 * NF), r = i % (3 * NF);
          xt[i]        = p.xt[facev[lf] * 3 * NF + r];
        }
      WAVE_SYNC3();
      // ---- B: state at quadrature points
      for (int e = lane; e < 3 * (D + 1) * QC; e += 64)
        {
          const int c = e / ((D + 1) * QC), r = e % ((D + 1) * QC);
          const int blkc = r / QC, q = r % QC;
          const double *x = xs + c * (D + 1) * NB + blkc * NB;
          double        v = 0;
#pragma unroll
          for (int j = 0; j < NB; ++j)
            v += x[j] * tab[S::tPHIC + q * NB + j];
          if (blkc == D)
            u0[c * QC + q] = v;
          else
            W0[(c * D + blkc) * QC + q] = v;
        }
      for (int e = lane; e < 3 * NFC * QF; e += 64)
        {
          const int c = e / (NFC * QF), g = (e / QF) % NFC, s = e % QF;
          const int a = g >> 1;
          const double nu  = (g & 1) ? 1.0 : -1.0;
          const double *xc = xs + c * (D + 1) * NB;
          const double *ph = tab + S::tPHIF + (g * QF + s) * NB;
          double        u = 0, w = 0, t = 0;
#pragma unroll
          for (int j = 0; j < NB; ++j)
            {
              u += xc[D * NB + j] * ph[j];
              w += xc[a * NB + j] * ph[j];
            }
#pragma unroll
          for (int m = 0; m < NF; ++m)
            t += xt[(g * 3 + c) * NF + m] * tab[S::tXI + s * NF + m];
          u0f[e] = u;
          W0f[e] = nu * w;
          tr0[e] = t;
        }
      WAVE_SYNC3();
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
      for (int e = lane; e < NE * D * NB; e += 64)
        {
          const int c = e / (D * NB), a = (e / NB) % D, i = e % NB;
          double    v = 0;
#pragma unroll
          for (int q = 0; q < QC; ++q)
            v += (u0[c * QC + q] * tab[S::tGRADC + (q * NB + i) * D + a] -
                  W0[(c * D + a) * QC + q] * tab[S::tPHIC + q * NB + i]) *
                 tab[S::tJXWC + q];
#pragma unroll
          for (int s1 = 0; s1 < 2; ++s1)
            {
              const int    g  = 2 * a + s1;
              const double nu = s1 ? 1.0 : -1.0;
#pragma unroll
              for (int s = 0; s < QF; ++s)
                v -= tr0[(c * NFC + g) * QF + s] * tab[S::tPHIF + (g * QF + s) * NB + i] * nu *
                     tab[S::tJXWF + g * QF + s];
            }
          rhsq[e] = v;
        }
      for (int e = lane; e < NFC * 3 * NF; e += 64)
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
      WAVE_SYNC3();
      // ---- D: lifted residual flux with the Dirichlet update folded in;
      //         residual forms at the quadrature points
      for (int e = lane; e < NE * D * QC; e += 64)
        {
          const int c = e / (D * QC), a = (e / QC) % D, q = e % QC;
          double    v = 0;
#pragma unroll
          for (int i = 0; i < NB; ++i)
            v += tab[S::tPMC + q * NB + i] * rhsq[(c * D + a) * NB + i];
#pragma unroll
          for (int s1 = 0; s1 < 2; ++s1)
            {
              const int f = 2 * a + s1;
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
            v += tab[S::tPMF + (g * QF + s) * NB + i] * rhsq[(c * D + a) * NB + i];
#pragma unroll
          for (int s1 = 0; s1 < 2; ++s1)
            {
              const int f = 2 * a + s1;
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
          double    al, bv[D];
          if (c == 0)
            {
              al = qd[5 * QC + q] + (THERMO ? qd[4 * QC + q] : u0[2 * QC + q] - u0[QC + q]);
#pragma unroll
              for (int a = 0; a < D; ++a)
                bv[a] = eps * W0[a * QC + q];
            }
          else if (c == 1)
            {
              const double n0 = u0[QC + q];
              al              = qd[q];
#pragma unroll
              for (int a = 0; a < D; ++a)
                bv[a] = n0 * mun * W0[a * QC + q] - UT * mun * W0[(D + a) * QC + q];
            }
          else
            {
              const double p0 = u0[2 * QC + q];
              al              = qd[q];
#pragma unroll
              for (int a = 0; a < D; ++a)
                bv[a] = -p0 * mup * W0[a * QC + q] - UT * mup * W0[(2 * D + a) * QC + q];
            }
          resc[(q * NE + ci) * (D + 1)] = al;
#pragma unroll
          for (int a = 0; a < D; ++a)
            resc[(q * NE + ci) * (D + 1) + 1 + a] = bv[a];
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
      WAVE_SYNC3();

      // ---- E: the columns of the two slots.  Column col = lane + 64 h:
      // scalar dof (kind 0), trace dof (1), rhs (2) or none (3).  The bundle
      // is branch-free: table values are loaded at valid indices for every
      // lane and blended with 0/1 weights.
      int ju[NH], jl[NH], fl[NH], kind[NH], cc[NH], jj[NH], ff[NH];
#pragma unroll
      for (int h = 0; h < NH; ++h)
        {
          const int col = lane + 64 * h;
          cc[h] = jj[h] = ff[h] = 0;
          if (PASS2)
            kind[h] = col < NS ? 0 : (col == NS ? 2 : 3);
          if (PASS2 && col < NS)
            {
              cc[h] = col / NB;
              jj[h] = col % NB;
            }
          else if (PASS2)
            ;
          else if (col < NS)
            {
              kind[h] = 0;
              cc[h]   = col / NB;
              jj[h]   = col % NB;
            }
          else if (col < NS + T)
            {
              kind[h]      = 1;
              const int lt = col - NS;
              ff[h]        = lt / TDF;
              cc[h]        = (lt % TDF) / NF;
              jj[h]        = lt % NF;
            }
          else
            kind[h] = col == NCOL - 1 ? 2 : 3;
          ju[h] = kind[h] == 0 ? jj[h] : 0;
          jl[h] = kind[h] == 1 ? jj[h] : 0;
          fl[h] = kind[h] == 1 ? ff[h] : 0;
        }
      // 0/1 lane weights of the bundle, formed where used
      auto mU  = [&](int h, int ci) { return (kind[h] == 0 && cc[h] == ci) ? 1.0 : 0.0; };
      auto mL  = [&](int h, int ci) { return (kind[h] == 1 && cc[h] == ci) ? 1.0 : 0.0; };
      auto isr = [&](int h) { return kind[h] == 2 ? 1.0 : 0.0; };
      auto mLa = [&](int h, int a) { return (fl[h] >> 1) == a ? 1.0 : 0.0; };
      const bool isU = kind[0] == 0;

      // the numerical flux F_c of slot h's column at face point (g, s)
      // (the HDG flux of component c, use_cell)
      auto face_flux = [&](int h, int g, int s, double (&F)[NE]) __attribute__((always_inline)) {
        const int    fi  = g * QF + s;
        const double ph  = tab[S::tPHIF + fi * NB + ju[h]];
        const double ld  = tab[S::tLDF + fi * NB + ju[h]];
        const double lc  = tab[S::tLCF + ((fl[h] * NFC + g) * QF + s) * NF + jl[h]];
        const double xv  = tab[S::tXI + s * NF + jl[h]] * (fl[h] == g ? 1.0 : 0.0);
        const double E0n = W0f[fi];
        const double WN0 = mU(h, 0) * ld + mL(h, 0) * lc - isr(h) * Wrf[fi];
#pragma unroll
        for (int ci = 0; ci < NE; ++ci)
          {
            const double uuc = mU(h, ci) * ph;
            const double WNc = mU(h, ci) * ld + mL(h, ci) * lc - isr(h) * Wrf[ci * NFC * QF + fi];
            const double LLc = mL(h, ci) * xv - isr(h) * lamf[ci * NFC * QF + fi];
            double       Fv;
            if (ci == 0)
              Fv = eps * WNc + tV * uuc - tV * LLc;
            else if (ci == 1)
              Fv = uuc * mun * E0n + u0f[NFC * QF + fi] * mun * WN0 - UT * mun * WNc -
                   tN * uuc + tN * LLc;
            else
              Fv = -(uuc * mup * E0n + u0f[2 * NFC * QF + fi] * mup * WN0) - UT * mup * WNc -
                   tP * uuc + tP * LLc;
            F[ci] = Fv + isr(h) * resf[fi * NE + ci];
          }
      };

      // trace-test rows of face g for slot h's column (TC' for a scalar
      // column, TT'/t' otherwise), Neumann data on the rhs, zero rows on
      // Dirichlet faces
      auto trace_rows = [&](int h, int g, double (&acc)[TDF]) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < TDF; ++r)
          acc[r] = 0.0;
#pragma unroll 1
        for (int s = 0; s < QF; ++s)
          {
            double F[NE];
            face_flux(h, g, s, F);
#pragma unroll
            for (int ci = 0; ci < NE; ++ci)
#pragma unroll
              for (int i = 0; i < NF; ++i)
                acc[ci * NF + i] += Bt[(g * QF + s) * NF + i] * F[ci];
          }
#pragma unroll
        for (int ci = 0; ci < NE; ++ci)
          {
            const double keepr = isdir[g * 3 + (NE == 1 ? 0 : ci)] ? 0.0 : 1.0;
#pragma unroll
            for (int i = 0; i < NF; ++i)
              acc[ci * NF + i] = keepr * (acc[ci * NF + i] + isr(h) * neu[(g * NE + ci) * NF + i]);
          }
      };

      constexpr int NKK = NS / 4;
      static_assert(NS % 4 == 0, "K steps of four pivot rows");
      static_assert(NS / 16 * 16 + 48 >= (NCOL < 64 ? NCOL : 64) && NCOL - 64 <= 48,
                    "three 16-lane column tiles cover the trace and rhs columns of a slot");
      const int q16 = lane >> 4, c16 = lane & 15;
      int myk = 0; // the equation this scalar lane was the pivot of
      // the slots one after the other: only one slot's rows are live (pass
      // 2 has slot 0 only)
      constexpr int NHX = PASS2 ? 1 : NH;
#pragma unroll
      for (int h = 0; h < NHX; ++h)
        {
          // ---- 1a: scalar-test rows, cell quadrature points (q stays a
          // runtime index so the uniform Bz loads are not all hoisted)
          double R[NS];
#pragma unroll
          for (int r = 0; r < NS; ++r)
            R[r] = 0.0;
#pragma unroll 1
          for (int q = 0; q < QC; ++q)
            {
              double       E0[D], uu[3] = {0, 0, 0}, W[D][3];
              const double ph = tab[S::tPHIC + q * NB + ju[h]];
              const double lc = tab[S::tLCC + (fl[h] * QC + q) * NF + jl[h]];
#pragma unroll
              for (int a = 0; a < D; ++a)
                {
                  E0[a]            = W0[a * QC + q];
                  const double ldv = tab[S::tLDC + (q * D + a) * NB + ju[h]];
#pragma unroll
                  for (int ci = 0; ci < NE; ++ci)
                    W[a][ci] = mU(h, ci) * ldv + mL(h, ci) * mLa(h, a) * lc -
                               isr(h) * Wrc[(ci * D + a) * QC + q];
                }
#pragma unroll
              for (int ci = 0; ci < NE; ++ci)
                uu[ci] = mU(h, ci) * ph;
#pragma unroll
              for (int ci = 0; ci < NE; ++ci)
                {
                  double al, bv[D];
                  if (ci == 0)
                    {
                      if (THERMO)
                        al = -qd[3 * QC + q] * uu[0];
                      else
                        al = uu[NE > 1 ? 1 : 0] - uu[NE > 2 ? 2 : 0];
#pragma unroll
                      for (int a = 0; a < D; ++a)
                        bv[a] = -eps * W[a][0];
                    }
                  else if (ci == 1)
                    {
                      const double n0 = u0[QC + q];
                      al = -(qd[QC + q] * uu[1] + qd[2 * QC + q] * uu[NE > 2 ? 2 : 1]);
#pragma unroll
                      for (int a = 0; a < D; ++a)
                        bv[a] = -mun * E0[a] * uu[1] + sd * n0 * mun * W[a][0] +
                                UT * mun * W[a][1];
                    }
                  else
                    {
                      const double p0 = u0[2 * QC + q];
                      al = -(qd[QC + q] * uu[1] + qd[2 * QC + q] * uu[NE > 2 ? 2 : 1]);
#pragma unroll
                      for (int a = 0; a < D; ++a)
                        bv[a] = mup * E0[a] * uu[NE > 2 ? 2 : 1] - sd * p0 * mup * W[a][0] +
                                UT * mup * W[a][NE > 2 ? 2 : 1];
                    }
                  al += isr(h) * resc[(q * NE + ci) * (D + 1)];
#pragma unroll
                  for (int a = 0; a < D; ++a)
                    bv[a] += isr(h) * resc[(q * NE + ci) * (D + 1) + 1 + a];
#pragma unroll
                  for (int i = 0; i < NB; ++i)
                    {
                      double v = Bz[(q * (D + 1)) * NB + i] * al;
#pragma unroll
                      for (int a = 0; a < D; ++a)
                        v += Bz[(q * (D + 1) + 1 + a) * NB + i] * bv[a];
                      R[ci * NB + i] += v;
                    }
                }
            }
          // ---- 1b: face quadrature points, all NB rows per point (the
          // table is zero off the face's support)
#pragma unroll 1
          for (int fi = 0; fi < NFC * QF; ++fi)
            {
              const int g = fi / QF, s = fi % QF;
              double    F[NE];
              face_flux(h, g, s, F);
#pragma unroll
              for (int ci = 0; ci < NE; ++ci)
#pragma unroll
                for (int i = 0; i < NB; ++i)
                  R[ci * NB + i] += Bz[(QC * (D + 1) + fi) * NB + i] * F[ci];
            }
          // LocalCondenser::condense_ct_matrix: the Dirichlet update is in
          // the rhs bundle already; the Dirichlet columns are dropped
          const bool dircol = kind[h] == 1 && isdir[ff[h] * 3 + (NE == 1 ? 0 : cc[h])];
          if (dircol)
#pragma unroll
            for (int r = 0; r < NS; ++r)
              R[r] = 0.0;

          // ---- 2: Gauss-Jordan with scaled column pivoting on slot 0 (it
          // holds the scalar columns), recording each step's pivot column;
          // slot 1 replays the recorded steps
          if (h == 0)
            {
              if constexpr (PASS2)
                {
                  if constexpr (!MIXED)
                    gauss_jordan_rec<NS>(R, isU, lane, pivs, myk, gjr);
                  else
                    {
                      // the FP32 elimination, the FP64 residual of the rhs
                      // column (the scalar lanes keep their FP64 columns of
                      // A' in R), the FP32 elimination again on it
                      float R32[NS], X0[NS];
#pragma unroll
                      for (int r = 0; r < NS; ++r)
                        R32[r] = (float)R[r];
                      gauss_jordan_rec<NS>(R32, isU, lane, pivs, myk, gjr32);
#pragma unroll
                      for (int r = 0; r < NS; ++r)
                        X0[r] = R32[r];
                      WAVE_SYNC3();
                      const double mres = isU ? 0.0 : 1.0;
#pragma unroll
                      for (int k = 0; k < NS; ++k)
                        {
                          const int    pl = __builtin_amdgcn_readfirstlane(pivs[k]);
                          const double xk = mres * (double)X0[k];
#pragma unroll
                          for (int r = 0; r < NS; ++r)
                            R[r] -= lane_value(R[r], pl) * xk;
                        }
#pragma unroll
                      for (int r = 0; r < NS; ++r)
                        R32[r] = (float)R[r];
                      gauss_jordan_replay<NS>(R32, gjr32);
#pragma unroll
                      for (int k = 0; k < NS; ++k)
                        R[k] = (double)X0[k] + (double)R32[k];
                    }
                  WAVE_SYNC3();
                  // the rhs lane holds u(p_k) in row k: publish u, then
                  // W = M^-1 (rhs_q + D_a u - sum_f C_{a,f} lambda_upd,f)
                  // and u go out lane-parallel (reference template.h:
                  // 2688-2697)
                  double *sol = tcf;
                  if (kind[0] == 2)
#pragma unroll
                    for (int k = 0; k < NS; ++k)
                      sol[pivs[k]] = R[k];
                  WAVE_SYNC3();
                  for (int e = lane; e < S::NCD; e += 64)
                    {
                      const int c = e / ((D + 1) * NB), r = e % ((D + 1) * NB);
                      const int blkc = r / NB, i = r % NB;
                      double    v = 0;
                      if (c < NE)
                        {
                          const double *u = sol + c * NB;
                          if (blkc == D)
                            v = u[i];
                          else
                            {
                              const int a = blkc;
#pragma unroll
                              for (int l = 0; l < NB; ++l)
                                v += tb.Minv[i * NB + l] * rhsq[(c * D + a) * NB + l] +
                                     tb.MinvD[(a * NB + i) * NB + l] * u[l];
#pragma unroll
                              for (int s1 = 0; s1 < 2; ++s1)
                                {
                                  const int f = 2 * a + s1;
#pragma unroll
                                  for (int m = 0; m < NF; ++m)
                                    v -= tb.MinvC[(f * NB + i) * NF + m] *
                                         lupd[(f * 3 + c) * NF + m];
                                }
                            }
                        }
                      p.dxc[cell * S::NCD + e] = v;
                    }
                  WAVE_SYNC3();
                  return;
                }
              else if constexpr (!MIXED)
                gauss_jordan_rec<NS>(R, isU, lane, pivs, myk, gjr);
              else
                {
                  // FP32 elimination, the FP64 residual r = b - A' x0 (the
                  // scalar lanes keep their FP64 columns of A' in R), a
                  // second FP32 elimination with the same pivots
                  float R32[NS], X0[NS];
#pragma unroll
                  for (int r = 0; r < NS; ++r)
                    R32[r] = (float)R[r];
                  gauss_jordan_rec<NS>(R32, isU, lane, pivs, myk, gjr32);
                  // the scalar lanes' FP64 columns of A', in pivot order,
                  // for the slot-1 residual
                  if (isU)
#pragma unroll
                    for (int r = 0; r < NS; ++r)
                      apr[myk * NS + r] = R[r];
#pragma unroll
                  for (int r = 0; r < NS; ++r)
                    X0[r] = R32[r];
                  WAVE_SYNC3();
                  const double mres = isU ? 0.0 : 1.0;
#pragma unroll
                  for (int k = 0; k < NS; ++k)
                    {
                      const int    pl = __builtin_amdgcn_readfirstlane(pivs[k]);
                      const double xk = mres * (double)X0[k];
#pragma unroll
                      for (int r = 0; r < NS; ++r)
                        R[r] -= lane_value(R[r], pl) * xk;
                    }
#pragma unroll
                  for (int r = 0; r < NS; ++r)
                    R32[r] = (float)R[r];
                  gauss_jordan_replay<NS>(R32, gjr32);
#pragma unroll
                  for (int k = 0; k < NS; ++k)
                    R[k] = (double)X0[k] + (double)R32[k];
                }
            }
          else
            {
              WAVE_SYNC3();
              if constexpr (!MIXED)
                gauss_jordan_replay<NS>(R, gjr);
              else
                {
                  float R32[NS], X0[NS];
#pragma unroll
                  for (int r = 0; r < NS; ++r)
                    R32[r] = (float)R[r];
                  gauss_jordan_replay<NS>(R32, gjr32);
#pragma unroll
                  for (int r = 0; r < NS; ++r)
                    X0[r] = R32[r];
#pragma unroll
                  for (int k = 0; k < NS; ++k)
                    {
                      const double *ak = apr + k * NS;
                      const double  xk = (double)X0[k];
#pragma unroll
                      for (int r = 0; r < NS; ++r)
                        R[r] -= ak[r] * xk;
                      // pace the NS*NS LDS loads of A'
                      asm volatile("" ::: "memory");
                    }
#pragma unroll
                  for (int r = 0; r < NS; ++r)
                    R32[r] = (float)R[r];
                  gauss_jordan_replay<NS>(R32, gjr32);
#pragma unroll
                  for (int k = 0; k < NS; ++k)
                    R[k] = (double)X0[k] + (double)R32[k];
                }
            }
          WAVE_SYNC3();

          // ---- 3: [S | g] = [TT' | t'] - TC' X of this slot, face by face.
          // The product runs on the matrix cores (v_mfma_f64_16x16x4f64, K =
          // 4 pivot rows per step) over the slot's three 16-lane column tiles
          // from tile tb.  X reaches the B layout through tcf once per slot,
          // four rows at a time, and stays in registers; per face, TC' (TDF
          // rows padded to 16) is gathered from tcf in the A layout, and the
          // result returns to column layout through tcf, four rows at a time.
          const int tb = h == 0 ? NS / 16 : 0;
          double    bop[NKK][3];
#pragma unroll
          for (int kk = 0; kk < NKK; ++kk)
            {
#pragma unroll
              for (int kq = 0; kq < 4; ++kq)
                tcf[kq * 64 + lane] = R[4 * kk + kq];
              WAVE_SYNC3();
#pragma unroll
              for (int t = 0; t < 3; ++t)
                bop[kk][t] = tcf[q16 * 64 + 16 * (tb + t) + c16];
              WAVE_SYNC3();
            }
          const bool isC = kind[h] == 1 || kind[h] == 2;
          const int  cl  = lane - 16 * tb;
#pragma unroll 1
          for (int g = 0; g < NFC; ++g)
            {
              double acc[TDF];
              // the scalar columns (slot 0) publish their TC' rows of this
              // face in pivot order (row k of X belongs to the pivot of
              // equation k)
              if (h == 0)
                {
                  trace_rows(0, g, acc);
                  if (isU)
#pragma unroll
                    for (int r = 0; r < TDF; ++r)
                      tcf[myk * TDF + r] = acc[r];
                }
              else
                {
                  double tcu[TDF];
                  trace_rows(0, g, tcu);
                  if (isU)
#pragma unroll
                    for (int r = 0; r < TDF; ++r)
                      tcf[myk * TDF + r] = tcu[r];
                  trace_rows(h, g, acc);
                }
              WAVE_SYNC3();
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
              // element e of dm[t] holds row 4e + q16, column 16 (tb + t) + c16
#pragma unroll
              for (int e = 0; e < (TDF + 3) / 4; ++e)
                {
                  WAVE_SYNC3();
#pragma unroll
                  for (int t = 0; t < 3; ++t)
                    tcf[q16 * 48 + 16 * t + c16] = dm[t][e];
                  WAVE_SYNC3();
                  if (isC)
#pragma unroll
                    for (int qq = 0; qq < 4; ++qq)
                      if (4 * e + qq < TDF)
                        acc[4 * e + qq] -= tcf[qq * 48 + cl];
                }
              // the next face overwrites tcf
              WAVE_SYNC3();
              const bool isL = kind[h] == 1, isR = kind[h] == 2;
              const int  lt  = lane + 64 * h - NS;
              if (mode == 1)
                {
                  if (isL)
#pragma unroll
                    for (int r = 0; r < TDF; ++r)
                      p.Sout[(long long)idx * T * T + (g * TDF + r) * T + lt] =
                        dircol ? 0.0 : acc[r];
                  if (isR)
#pragma unroll
                    for (int r = 0; r < TDF; ++r)
                      p.gout[(long long)idx * T + g * TDF + r] = acc[r];
                  continue;
                }
              if (isL)
                {
                  const int gj = lt / TDF, rj = lt % TDF;
                  const long long base = blk[g] * TDF * TDF + (long long)slot[g * NFC + gj] * TDF + rj;
                  const long long rs   = (long long)nbk[g] * TDF;
                  const bool      addd = color == 1 && g == gj && bidv[g] < 0;
#pragma unroll
                  for (int r = 0; r < TDF; ++r)
                    {
                      const int  i   = g * TDF + r, ci = r / NF;
                      const bool dir = isdir[g * 3 + (NE == 1 ? 0 : ci)];
                      const double v = dir ? (i == lt ? 1.0 : 0.0) : (dircol ? 0.0 : acc[r]);
                      const long long pos = base + r * rs;
                      if (addd)
                        p.vals[pos] += v;
                      else
                        p.vals[pos] = v;
                    }
                }
              if (isR)
                {
                  const bool addr = color == 1 && bidv[g] < 0;
#pragma unroll
                  for (int r = 0; r < TDF; ++r)
                    {
                      const int  ci   = r / NF;
                      const int  comp = NE == 1 ? 0 : ci;
                      const bool dir  = isdir[g * 3 + comp];
                      const long long row = facev[g] * TDF + r;
                      const double v = dir ? 0.0 : acc[r];
                      if (addr)
                        p.rhs[row] += v;
                      else
                        p.rhs[row] = v;
                      if (dir)
                        p.dirout[row] = lupd[(g * 3 + comp) * NF + (r % NF)];
                    }
                }
            }
        }
    }

    // PASS2: cells = all cells in order (cells unused)
    template <int NE, bool THERMO, bool MIXED, bool PASS2 = false>
    __global__ void __launch_bounds__(64 * wpb3(MIXED)) __attribute__((amdgpu_waves_per_eu(2)))
      fast_condense3(KParams p, const int *__restrict__ cells, int ncell, int color, int mode)
    {
      constexpr int WPB = wpb3(MIXED);
      extern __shared__ double smem[];
      stage_tables3<NE, 64 * WPB>(p.tb, smem, threadIdx.x);
      __syncthreads();
      const int wv = threadIdx.x >> 6;
      for (int idx = blockIdx.x * WPB + wv; idx < ncell; idx += gridDim.x * WPB)
        cell_body3<NE, THERMO, MIXED, PASS2>(p, smem, PASS2 ? (long long)idx : (long long)cells[idx],
                                             idx, color, mode);
    }

    template <int NE, bool THERMO, bool MIXED, bool PASS2 = false>
    hipError_t launch_fast3(const KParams &p, const int *cells, int ncell, int color, int mode,
                            hipStream_t s)
    {
      constexpr size_t bytes = F3<NE>::bytes(MIXED);
      static bool      attr  = false;
      if (!attr)
        {
          hipError_t e = hipFuncSetAttribute((const void *)fast_condense3<NE, THERMO, MIXED, PASS2>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)bytes);
          if (e != hipSuccess)
            return e;
          attr = true;
        }
      if (ncell <= 0)
        return hipSuccess;
      int dev = 0, ncu = 256;
      if (hipGetDevice(&dev) == hipSuccess)
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
      constexpr int WPB  = wpb3(MIXED);
      const int     want = (ncell + WPB - 1) / WPB;
      // eight rounds of the two workgroups a CU holds (48^3 pass 1: 8.53
      // ms at two rounds, 8.30 ms at eight)
      const int     grid = want < 16 * ncu ? want : 16 * ncu;
      hipLaunchKernelGGL((fast_condense3<NE, THERMO, MIXED, PASS2>), dim3(grid), dim3(64 * WPB),
                         bytes, s, p, cells, ncell, color, mode);
      return hipGetLastError();
    }
  } // namespace

  bool fast3_supported(int dim, int deg, int ne, int thermo, const int en[3])
  {
    if (dim != 3 || deg != 1)
      return false;
    if (thermo)
      return ne == 1 && en[0];
    return ne == 3;
  }

  hipError_t launch_fast3_condense(const KParams &p, const int *cells, int ncell, int color,
                                   int mode, hipStream_t s)
  {
    if (p.thermo)
      return p.mixed ? launch_fast3<1, true, true>(p, cells, ncell, color, mode, s)
                     : launch_fast3<1, true, false>(p, cells, ncell, color, mode, s);
    return p.mixed ? launch_fast3<3, false, true>(p, cells, ncell, color, mode, s)
                   : launch_fast3<3, false, false>(p, cells, ncell, color, mode, s);
  }

  hipError_t launch_fast3_reconstruct(const KParams &p, int ncells, hipStream_t s)
  {
    if (p.mixed)
      return p.thermo ? launch_fast3<1, true, true, true>(p, nullptr, ncells, 0, 0, s)
                      : launch_fast3<3, false, true, true>(p, nullptr, ncells, 0, 0, s);
    if (p.thermo)
      return launch_fast3<1, true, false, true>(p, nullptr, ncells, 0, 0, s);
    return launch_fast3<3, false, false, true>(p, nullptr, ncells, 0, 0, s);
  }
} // namespace hdg
