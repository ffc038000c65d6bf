// hdg_fast2d.hip -- register-resident condensation for 2D cells, degree 1-2.
//
// One wave64 per cell, one column of the reduced bordered system per lane:
//   lanes [0, NS)          scalar trial dofs   (NS = ne*(k+1)^2)
//   lanes [NS, NS+T)       trace trial dofs    (T = 4*ne*(k+1))
//   lane  NS+T             right-hand side (residual; Dirichlet update folded in)
// A lane keeps only the NS scalar-test rows of its column in VGPRs.  Its
// column is assembled from its "bundle" (values, lifted fluxes and traces of
// the trial function at the quadrature points) with the test tables read as
// uniform scalars: the terms of include/assembling_methods.template.h:
// 269-2118 after the flux unknowns are eliminated (hdg_tables.h).
//   1. scalar-test rows [A' | CT' | c'] in registers
//   2. (no storage of the trace-test rows: step 4 forms them face by face)
//   3. Gauss-Jordan on the scalar rows with scaled column pivoting: equation k
//      picks the unknown (lane) with the largest scaled |entry| (a DPP
//      max-reduction), the pivot column reaches the other lanes by readlane;
//      afterwards row k of a trace/rhs lane holds X(p_k, m),
//      X = A'^-1 [CT' | c']
//   4. every lane forms its trace-test rows of one face at a time (TT', t'
//      on the trace/rhs lanes, TC' on the scalar lanes, published in LDS
//      for the matrix-core product) and the trace/rhs lanes subtract TC' X: [S | g] = [TT' - TC' A'^-1 CT' | t' - TC' A'^-1 c'],
//      written straight into the trace CSR (two colour launches,
//      deterministic)
// Workgroups of 4 waves are persistent.  Tables read at lane-dependent
// indices are staged in LDS once per workgroup; the test tables read at
// uniform indices (Bz, Bt) come through scalar loads.
#include "hdg_device.h"
#include "hdg_kernels.h"
#include "hdg_wave.h"

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
    // cells per chunk of a batched work list; a batched workgroup of 4 or
    // 8 waves (the tile the autotuner picks) takes one chunk at a time
    constexpr int BCH = 8;
    // waves per SIMD the register budget is set for, and persistent
    // workgroups per CU (four rounds of the resident ones), by degree and
    // pass (256^2 diode: pass 1 at k = 2 1.88 -> 1.55 ms, k = 1 0.83 ->
    // 0.62 ms, pass 2 at k = 2 1.20 -> 1.00 ms against two waves and four
    // workgroups per CU).  The batched kernel (degree 1 and 2 in one body)
    // keeps two waves: built for three, its FP32 + refinement form gave
    // blocks different from the per-mesh kernel's in the batch test (cause
    // not found)
    constexpr int BATCH_WPE = 2;
    constexpr int waves_per_simd(int deg, bool pass2)
    {
      return pass2 ? 3 : deg == 1 ? 4 : 3;
    }
    constexpr int groups_per_cu(int deg, bool pass2)
    {
      return pass2 ? 12 : deg == 1 ? 16 : 12;
    }

    template <int DEG, int NE>
    struct F2
    {
      static constexpr int K1 = DEG + 1, NB = K1 * K1, NF = K1, QC = NB, QF = NF, NFC = 4;
      static constexpr int NCD = 9 * NB;
      static constexpr int NS = NE * NB, TDF = NE * NF, T = NFC * TDF;
      static constexpr int NROW = NS + T, NCOL = NROW + 1;
      static_assert(NCOL <= 64, "one column per lane");
      static constexpr int NROW2 = NS; // rows held in registers
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
      // per-wave scratch (doubles).  The first part lives through the
      // whole cell; the second (cell state, point data, the residual forms
      // of the scalar rows, pass 2's solution) is dead once the scalar-test
      // rows are formed, and step 4's TC' rows and X / Schur exchange reuse
      // it, so that three workgroups fit one CU's LDS
      static constexpr int wU0F  = 0;
      static constexpr int wW0F  = wU0F + 3 * NFC * QF;
      static constexpr int wWRF  = wW0F + 3 * NFC * QF;
      static constexpr int wLAMF = wWRF + 3 * NFC * QF;
      static constexpr int wLUPD = wLAMF + 3 * NFC * QF;
      static constexpr int wRESF = wLUPD + NFC * 3 * NF;
      static constexpr int wNEU  = wRESF + NFC * QF * NE;
      static constexpr int wBLK  = wNEU + NFC * NE * NF; // 4 long long (blkptr of the faces)
      static constexpr int wFACE = wBLK + NFC;           // 4 long long (face index)
      static constexpr int wEARLY = wFACE + NFC;
      static constexpr int wXS   = wEARLY;
      static constexpr int wXT   = wXS + NCD;
      static constexpr int wU0   = wXT + NFC * 3 * NF;
      static constexpr int wW0   = wU0 + 3 * QC;
      static constexpr int wTR0  = wW0 + 3 * 2 * QC;
      static constexpr int wQD   = wTR0 + 3 * NFC * QF;
      static constexpr int wRHSQ = wQD + 6 * QC;
      static constexpr int wWRC  = wRHSQ + 3 * 2 * NB;
      static constexpr int wRESC = wWRC + 3 * 2 * QC;
      static constexpr int wBC   = wRESC + QC * NE * 3;
      static constexpr int wEND  = wBC + 64;
      static constexpr int wTCL  = wEARLY;           // NS*TDF: TC' rows of one face (step 4)
      static constexpr int wXCH  = wTCL + NS * TDF;  // X rows / Schur rows to and from the MFMA layouts
      static constexpr int XCH   = (TDF > 4 ? TDF : 4) * (T + 1);
      static constexpr int WDBL  = wXCH + XCH > wEND ? wXCH + XCH : wEND; // per-wave doubles, pass 1
      static constexpr int WDBL2 = wEND;             // pass 2 needs no TC'
      // per-wave ints
      static constexpr int iDIR  = 0;             // NFC*3
      static constexpr int iNEU  = iDIR + NFC * 3;
      static constexpr int iBID  = iNEU + NFC * 3;
      static constexpr int iNBK  = iBID + NFC;
      static constexpr int iSLOT = iNBK + NFC;    // NFC*NFC
      static constexpr int iPIV  = iSLOT + NFC * NFC; // NS pivot lanes
      static constexpr int WINT  = iPIV + NS;
      static constexpr size_t bytes_for(bool pass2, int nw = WPB)
      {
        return (size_t)(TAB + nw * (pass2 ? WDBL2 : WDBL)) * sizeof(double) +
               (size_t)nw * WINT * sizeof(int);
      }
    };

    // read-only tables addressed uniformly by the whole wave: loads through
    // the constant address space become scalar loads (no LDS traffic, no
    // VGPRs for the values)
    typedef const __attribute__((address_space(4))) double ctab;

    // pivot-search key: the leading bits of the non-negative scaled
    // candidate (they order as integers) with 63 - lane in the low 6 bits
    // (ties go to the lowest lane); 0 for a lane that is not a candidate
    __device__ __forceinline__ unsigned pivot_key(double sc, bool cand, int lane)
    {
      return cand ? (leading_bits(sc) & ~63u) + (unsigned)(63 - lane) + 1u : 0u;
    }
    __device__ __forceinline__ unsigned pivot_key(float sc, bool cand, int lane)
    {
      return cand ? (__float_as_uint(sc) & ~63u) + (unsigned)(63 - lane) + 1u : 0u;
    }

    // Gauss-Jordan over the NS scalar rows of the lanes' columns with
    // scaled column pivoting: equation k takes the unpivoted scalar lane
    // with the largest |entry| / column max; that column reaches every lane
    // by readlane.  Lane 0 records the pivot lanes in pivs (LDS).
    template <int NS, typename Real>
    __device__ __forceinline__ void gauss_jordan(Real (&R)[NS], bool isU, int lane, int *pivs)
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
          const unsigned key = wave_max(pivot_key(fabs(R[k]) * cs, !pivoted, lane));
          const int      pl  = 63 - (int)((key - 1u) & 63u);
          if (lane == pl)
            pivoted = true;
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

    // the same elimination with the pivot lanes of an earlier one
    template <int NS, typename Real>
    __device__ __forceinline__ void gauss_jordan_fixed(Real (&R)[NS], const int *pivs)
    {
#pragma unroll
      for (int k = 0; k < NS; ++k)
        {
          const int pl = __builtin_amdgcn_readfirstlane(pivs[k]);
          Real      bq[NS];
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

    // the reference-cell tables of one mesh into LDS (one workgroup)
    template <int DEG, int NE>
    __device__ __forceinline__ void stage_tables(const DevTables &tb, double *tab0, int tid,
                                                 int nthreads = 64 * WPB)
    {
      using S          = F2<DEG, NE>;
      constexpr int NB = S::NB, NF = S::NF, QC = S::QC, QF = S::QF, NFC = S::NFC;
      for (int i = tid; i < QC * NB; i += nthreads)
        {
          tab0[S::tPHIC + i] = tb.phic[i];
          tab0[S::tPMC + i]  = tb.PhiMinvC[i];
        }
      for (int i = tid; i < QC * 2 * NB; i += nthreads)
        {
          tab0[S::tLDC + i]   = tb.LDc[i];
          tab0[S::tGRADC + i] = tb.gradc[i];
        }
      for (int i = tid; i < NFC * QC * NF; i += nthreads)
        tab0[S::tLCC + i] = tb.LCc[i];
      for (int i = tid; i < NFC * QF * NB; i += nthreads)
        {
          tab0[S::tPHIF + i] = tb.phif[i];
          tab0[S::tLDF + i]  = tb.LDf[i];
          tab0[S::tPMF + i]  = tb.PhiMinvF[i];
        }
      for (int i = tid; i < NFC * NFC * QF * NF; i += nthreads)
        tab0[S::tLCF + i] = tb.LCf[i];
      for (int i = tid; i < QF * NF; i += nthreads)
        tab0[S::tXI + i] = tb.xi[i];
      for (int i = tid; i < QC; i += nthreads)
        tab0[S::tJXWC + i] = tb.JxWc[i];
      for (int i = tid; i < NFC * QF; i += nthreads)
        tab0[S::tJXWF + i] = tb.JxWf[i];
      for (int i = tid; i < 2 * NF * NF; i += nthreads)
        tab0[S::tMFI + i] = tb.Mfinv[i];
    }

    // one cell (one wave): assembly, condensation and output of pass 1, or
    // the reconstruction of pass 2; tables staged in smem, per-wave
    // scratch of wave wv at stride wstr
    template <int DEG, int NE, bool THERMO, bool PASS2, bool MIXED>
    __device__ __forceinline__ void cell_body(const KParams &p, double *smem, int wstr,
                                              long long cell, long long idx, int color,
                                              int mode, int nw = WPB)
    {
      using S              = F2<DEG, NE>;
      constexpr int NB     = S::NB, NF = S::NF, QC = S::QC, QF = S::QF, NFC = S::NFC;
      constexpr int NS     = S::NS, TDF = S::TDF, T = S::T, NROW = S::NROW;
      constexpr int K1     = S::K1;
      double       *tab0 = smem;
      // the thread index behind an empty asm: everything derived from it is
      // formed again for every cell instead of being hoisted out of the
      // cell loop and kept live across it (spilled, at three waves per SIMD)
      int tid_ = threadIdx.x;
      asm volatile("" : "+v"(tid_));
      const int     tid  = tid_;
      const int     lane = tid & 63, wv = tid >> 6;
      double       *ws   = smem + S::TAB + wv * wstr;
      int *wi = reinterpret_cast<int *>(smem + S::TAB + nw * wstr) + wv * S::WINT;
      long long *blk   = reinterpret_cast<long long *>(ws + S::wBLK);
      long long *facev = reinterpret_cast<long long *>(ws + S::wFACE);
      const DevTables &tb = p.tb;

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

      // the tables are loop invariant; hiding their addresses behind an
      // empty asm keeps the compiler from hoisting every table load out
      // of the cell loop (hundreds of live values, all spilled)
      double *tab = tab0;
      asm volatile("" : "+v"(tab));
      ctab *Bz = (ctab *)tb.Bz, *Bt = (ctab *)tb.Bt;
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
              isneu[lf * 3 + c] = bt == 2 && !THERMO;  // zero Neumann data at equilibrium
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

      // ---- E: bundle of this lane's column and the forms it feeds.
      // Branch-free: every table value is loaded (indices valid for all
      // lanes) and blended with the lane's 0/1 weights, so no per-lane
      // masks have to be kept alive.
      const bool   isU = kind == 0, isL = kind == 1, isR = kind == 2;
      const double isr = isR ? 1.0 : 0.0;
      const int    ju = isU ? jj : 0, jl = isL ? jj : 0, fl = isL ? ff : 0;
      double       mU[NE], mL[NE];
#pragma unroll
      for (int ci = 0; ci < NE; ++ci)
        {
          mU[ci] = (isU && cc == ci) ? 1.0 : 0.0;
          mL[ci] = (isL && cc == ci) ? 1.0 : 0.0;
        }
      const double mLx = (fl >> 1) == 0 ? 1.0 : 0.0, mLy = 1.0 - mLx;
      double      *tcl  = ws + S::wTCL;
      int         *pivs = wi + S::iPIV;

      // ---- 1a: scalar-test rows, cell quadrature points.  q stays a
      // runtime index: unrolled, the scheduler hoists every (uniform) Bz
      // load of all quadrature points at once
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
              uu[ci] = mU[ci] * ph;
              WX[ci] = mU[ci] * lx + mL[ci] * mLx * lc - isr * Wrc[(ci * 2 + 0) * QC + q];
              WY[ci] = mU[ci] * ly + mL[ci] * mLy * lc - isr * Wrc[(ci * 2 + 1) * QC + q];
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

      // ---- 1b: face quadrature points.  The numerical flux F_c of the
      // column (the HDG flux of component c, use_cell) at face point
      // (g, s); it feeds the scalar-test rows (phi_i F_c) and the
      // trace-test rows (xi_i F_c).
      auto face_flux = [&](int g, int s, double (&F)[NE]) {
        const int    fi  = g * QF + s;
        const double ph  = tab[S::tPHIF + fi * NB + ju];
        const double ld  = tab[S::tLDF + fi * NB + ju];
        const double lc  = tab[S::tLCF + ((fl * NFC + g) * QF + s) * NF + jl];
        const double xv  = tab[S::tXI + s * NF + jl] * (fl == g ? 1.0 : 0.0);
        const double E0n = W0f[fi];
        const double WN0 = mU[0] * ld + mL[0] * lc - isr * Wrf[fi];
#pragma unroll
        for (int ci = 0; ci < NE; ++ci)
          {
            const double uuc = mU[ci] * ph;
            const double WNc = mU[ci] * ld + mL[ci] * lc - isr * Wrf[ci * NFC * QF + fi];
            const double LLc = mL[ci] * xv - isr * lamf[ci * NFC * QF + fi];
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
      auto trace_rows = [&](int g, double (&acc)[TDF]) {
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
            const double keepr = isdir[g * 3 + (NE == 1 ? 0 : ci)] ? 0.0 : 1.0;
#pragma unroll
            for (int i = 0; i < NF; ++i)
              acc[ci * NF + i] = keepr * (acc[ci * NF + i] + isr * neu[(g * NE + ci) * NF + i]);
          }
      };
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
      const bool   dircol = isL && isdir[ff * 3 + (NE == 1 ? 0 : cc)];
      const double keep   = dircol ? 0.0 : 1.0;
#pragma unroll
      for (int r = 0; r < NS; ++r)
        R[r] *= keep;

      // The second colour adds its part of the diagonal block of an interior
      // face to the first colour's.  A trace lane's column of that block
      // lies in the rows of its own face only, so its TDF earlier values
      // are loaded here, their latency hidden behind the elimination.
      double     dold[TDF];
      const bool dadd = !PASS2 && mode == 0 && isL && color == 1 && bidv[fl] < 0;
      {
        const int       rj0 = isL ? (lane - NS) % TDF : 0;
        const long long b0  = blk[fl] * TDF * TDF + (long long)slot[fl * NFC + fl] * TDF + rj0;
        const long long rs0 = (long long)nbk[fl] * TDF;
#pragma unroll
        for (int r = 0; r < TDF; ++r)
          dold[r] = dadd ? p.vals[b0 + r * rs0] : 0.0;
      }

      // ---- 3: Gauss-Jordan, scaled column pivoting
      if constexpr (!MIXED)
        gauss_jordan<NS>(R, isU, lane, pivs);
      else
        {
          // FP32 elimination with one FP64 refinement step: x0 from the
          // FP32 elimination, r = b - A' x0 in FP64 (the scalar lanes keep
          // their FP64 columns of A' untouched in R), the correction from
          // a second FP32 elimination with the same pivots, x = x0 + dx
          float R32[NS], X0[NS];
#pragma unroll
          for (int r = 0; r < NS; ++r)
            R32[r] = (float)R[r];
          gauss_jordan<NS>(R32, isU, lane, pivs);
#pragma unroll
          for (int k = 0; k < NS; ++k)
            X0[k] = R32[k];
          WAVE_SYNC();
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
          gauss_jordan_fixed<NS>(R32, pivs);
#pragma unroll
          for (int k = 0; k < NS; ++k)
            R[k] = (double)X0[k] + (double)R32[k];
        }
      // the pivot lanes (LDS, lane 0) are read below
      WAVE_SYNC();

      if (PASS2)
        {
          // the rhs lane holds u(p_k) in row k: publish u, then
          // W = M^-1 (rhs_q + D_a u - sum_f C_{a,f} lambda_upd,f) and u
          // go out lane-parallel (reference template.h:2688-2697)
          double *sol = bc;
          if (isR)
#pragma unroll
            for (int k = 0; k < NS; ++k)
              sol[pivs[k]] = R[k];
          WAVE_SYNC();
          for (int e = lane; e < S::NCD; e += 64)
            {
              const int c = e / (3 * NB), r = e % (3 * NB), blkc = r / NB, i = r % NB;
              double    v = 0;
              if (c < NE)
                {
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
                }
              p.dxc[cell * S::NCD + e] = v;
            }
          WAVE_SYNC();
          return;
        }
      // ---- 4: [S | g] = [TT' | t'] - TC' X, face by face.  The product
      // runs on the matrix cores (v_mfma_f64_16x16x4f64, K = 4 pivot rows
      // per step).  X, NS rows (pivot order) by NC trace and rhs columns,
      // reaches the B layout through LDS once, four rows at a time, and
      // stays in registers; per face, TC' (TDF rows padded to 16, columns
      // in pivot order) is gathered from LDS in the A layout, and the TDF x
      // NC result returns to column layout through the same buffer.
      constexpr int NC = T + 1, NKK = (NS + 3) / 4, NCT = (NC + 15) / 16;
      double       *xch = ws + S::wXCH;
      const int     q16 = lane >> 4, c16 = lane & 15;
      const bool    isC = lane >= NS && lane < NS + NC;
      double        bop[NKK][NCT];
      int           pka[NKK];
#pragma unroll
      for (int kk = 0; kk < NKK; ++kk)
        {
          const int k = 4 * kk + q16;
          pka[kk]     = pivs[k < NS ? k : 0];
          if (isC)
#pragma unroll
            for (int kq = 0; kq < 4; ++kq)
              if (4 * kk + kq < NS)
                xch[kq * NC + lane - NS] = R[4 * kk + kq];
          WAVE_SYNC();
#pragma unroll
          for (int tj = 0; tj < NCT; ++tj)
            {
              const int    c  = 16 * tj + c16;
              const bool   ok = k < NS && c < NC;
              const double v  = xch[q16 * NC + (ok ? c : 0)];
              bop[kk][tj]     = ok ? v : 0.0;
            }
          WAVE_SYNC();
        }
#pragma unroll 1
      for (int g = 0; g < NFC; ++g)
        {
          // the trace-test rows of face g: TT', t' on the trace and rhs
          // lanes, TC' on the scalar lanes (published for the A operand)
          double acc[TDF];
          trace_rows(g, acc);
          if (isU)
#pragma unroll
            for (int r = 0; r < TDF; ++r)
              tcl[lane * TDF + r] = acc[r];
          WAVE_SYNC();
          dbl4 dm[NCT];
#pragma unroll
          for (int tj = 0; tj < NCT; ++tj)
            dm[tj] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int kk = 0; kk < NKK; ++kk)
            {
              const bool   ok = 4 * kk + q16 < NS && c16 < TDF;
              const double v  = tcl[pka[kk] * TDF + (ok ? c16 : 0)];
              const double a  = ok ? v : 0.0;
#pragma unroll
              for (int tj = 0; tj < NCT; ++tj)
                dm[tj] = mfma_f64_16x16x4(a, bop[kk][tj], dm[tj]);
            }
          // element e of dm[tj] holds row 4e + q16, column 16 tj + c16
#pragma unroll
          for (int tj = 0; tj < NCT; ++tj)
#pragma unroll
            for (int e = 0; e < 4; ++e)
              {
                const int r = 4 * e + q16, c = 16 * tj + c16;
                if (4 * e < TDF && r < TDF && c < NC)
                  xch[r * NC + c] = dm[tj][e];
              }
          WAVE_SYNC();
          if (isC)
#pragma unroll
            for (int r = 0; r < TDF; ++r)
              acc[r] -= xch[r * NC + lane - NS];
          WAVE_SYNC();
          if (mode == 1)
            {
              if (isL)
#pragma unroll
                for (int r = 0; r < TDF; ++r)
                  p.Sout[(long long)idx * T * T + (g * TDF + r) * T + (lane - NS)] =
                    dircol ? 0.0 : acc[r];
              if (isR)
#pragma unroll
                for (int r = 0; r < TDF; ++r)
                  p.gout[(long long)idx * T + g * TDF + r] = acc[r];
              continue;
            }
          if (isL)
            {
              const int lt = lane - NS, gj = lt / TDF, rj = lt % TDF;
              const long long base = blk[g] * TDF * TDF + (long long)slot[g * NFC + gj] * TDF + rj;
              const long long rs   = (long long)nbk[g] * TDF;
              const bool      addd = dadd && g == gj;
#pragma unroll
              for (int r = 0; r < TDF; ++r)
                {
                  const int  i   = g * TDF + r, ci = r / NF;
                  const bool dir = isdir[g * 3 + (NE == 1 ? 0 : ci)];
                  const double v = dir ? (i == lt ? 1.0 : 0.0) : (dircol ? 0.0 : acc[r]);
                  const long long pos = base + r * rs;
                  p.vals[pos] = addd ? dold[r] + v : v;
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
      WAVE_SYNC();
    }

    // PASS2: the reconstruction pass (cells = all cells in order, columns =
    // scalar dofs + rhs with the trace update folded in, output dx_cell)
    // MIXED: the elimination of the scalar block in FP32 with one FP64
    // refinement step; assembly, residual, refinement residual and Schur
    // product in FP64
    template <int DEG, int NE, bool THERMO, bool PASS2, bool MIXED = false>
    __global__ void __launch_bounds__(64 * WPB)
      __attribute__((amdgpu_waves_per_eu(waves_per_simd(DEG, PASS2))))
      fast_condense(KParams p, const int *__restrict__ cells, int ncell, int color, int mode)
    {
      using S            = F2<DEG, NE>;
      constexpr int WSTR = PASS2 ? S::WDBL2 : S::WDBL;
      extern __shared__ double smem[];
      stage_tables<DEG, NE>(p.tb, smem, threadIdx.x);
      __syncthreads();
      const int wv = threadIdx.x >> 6;
      for (int idx = blockIdx.x * WPB + wv; idx < ncell; idx += gridDim.x * WPB)
        {
          const long long cell = PASS2 ? (long long)idx : (long long)cells[idx];
          cell_body<DEG, NE, THERMO, PASS2, MIXED>(p, smem, WSTR, cell, idx, color, mode);
        }
    }


    // Pass 1 over the cells of several meshes (devices) of degree 1 or 2 in
    // one persistent launch per colour: the work list holds chunks of BCH
    // cells of one device (padded with -1 at device boundaries); a workgroup
    // of NW waves (the tile: 4 or 8) takes one chunk at a time, wave w its
    // cells w, w + NW, ..., and restages the reference-cell tables when its
    // chunk's device changes.  LDS is sized for degree 2; a degree-1 device
    // uses the first part.
    template <int NE, bool THERMO, bool MIXED, int NW>
    __global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(BATCH_WPE)))
      batch_condense(const KParams *__restrict__ plist, const int *__restrict__ chunk_dev,
                     const int *__restrict__ items, int nchunks, int color)
    {
      static_assert(BCH % NW == 0, "a chunk splits evenly over the waves");
      extern __shared__ double smem[];
      const int wv  = threadIdx.x >> 6;
      int       cur = -1, deg = 0;
      for (int c = blockIdx.x; c < nchunks; c += gridDim.x)
        {
          const int      dev = chunk_dev[c];
          const KParams &p   = plist[dev];
          if (dev != cur)
            {
              __syncthreads();
              deg = p.deg;
              if (deg == 1)
                stage_tables<1, NE>(p.tb, smem, threadIdx.x, 64 * NW);
              else
                stage_tables<2, NE>(p.tb, smem, threadIdx.x, 64 * NW);
              __syncthreads();
              cur = dev;
            }
#pragma unroll 1
          for (int w = wv; w < BCH; w += NW)
            {
              const int cell = items[c * BCH + w];
              if (cell < 0)
                continue;
              if (deg == 1)
                cell_body<1, NE, THERMO, false, MIXED>(p, smem, F2<1, NE>::WDBL, cell, cell,
                                                       color, 0, NW);
              else
                cell_body<2, NE, THERMO, false, MIXED>(p, smem, F2<2, NE>::WDBL, cell, cell,
                                                       color, 0, NW);
            }
        }
    }

    template <int NE, bool THERMO, bool MIXED, int NW>
    hipError_t launch_batch_nw(const KParams *plist, const int *chunk_dev, const int *items,
                               int nchunks, int color, int grid, hipStream_t s)
    {
      constexpr size_t bytes = F2<2, NE>::bytes_for(false, NW);
      static_assert(bytes <= 160 * 1024, "LDS of one workgroup");
      static bool attr = false;
      if (!attr)
        {
          hipError_t e =
            hipFuncSetAttribute((const void *)batch_condense<NE, THERMO, MIXED, NW>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
          if (e != hipSuccess)
            return e;
          attr = true;
        }
      if (nchunks <= 0)
        return hipSuccess;
      hipLaunchKernelGGL((batch_condense<NE, THERMO, MIXED, NW>), dim3(grid), dim3(64 * NW),
                         bytes, s, plist, chunk_dev, items, nchunks, color);
      return hipGetLastError();
    }

    template <int NE, bool THERMO, bool MIXED>
    hipError_t launch_batch_t(const KParams *plist, const int *chunk_dev, const int *items,
                              int nchunks, int color, int grid, int tile, hipStream_t s)
    {
      if (tile == 8)
        return launch_batch_nw<NE, THERMO, MIXED, 8>(plist, chunk_dev, items, nchunks, color,
                                                     grid, s);
      if (tile == 4)
        return launch_batch_nw<NE, THERMO, MIXED, 4>(plist, chunk_dev, items, nchunks, color,
                                                     grid, s);
      return hipErrorInvalidValue;
    }

    template <int DEG, int NE, bool THERMO, bool PASS2 = false, bool MIXED = false>
    hipError_t launch_fast(const KParams &p, const int *cells, int ncell, int color, int mode,
                           hipStream_t s)
    {
      using S                = F2<DEG, NE>;
      constexpr size_t bytes = S::bytes_for(PASS2);
      static bool attr = false;
      if (!attr)
        {
          hipError_t e = hipFuncSetAttribute((const void *)fast_condense<DEG, NE, THERMO, PASS2, MIXED>,
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
      const int want = (ncell + WPB - 1) / WPB;
      const int gmax = groups_per_cu(DEG, PASS2) * ncu;
      const int grid = want < gmax ? want : gmax;
      hipLaunchKernelGGL((fast_condense<DEG, NE, THERMO, PASS2, MIXED>), dim3(grid), dim3(64 * WPB),
                         bytes, s, p, cells, ncell, color, mode);
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
    if (p.mixed)
      {
        if (p.thermo)
          return deg == 1 ? launch_fast<1, 1, true, false, true>(p, cells, ncell, color, mode, s)
                          : launch_fast<2, 1, true, false, true>(p, cells, ncell, color, mode, s);
        return deg == 1 ? launch_fast<1, 3, false, false, true>(p, cells, ncell, color, mode, s)
                        : launch_fast<2, 3, false, false, true>(p, cells, ncell, color, mode, s);
      }
    if (p.thermo)
      return deg == 1 ? launch_fast<1, 1, true>(p, cells, ncell, color, mode, s)
                      : launch_fast<2, 1, true>(p, cells, ncell, color, mode, s);
    return deg == 1 ? launch_fast<1, 3, false>(p, cells, ncell, color, mode, s)
                    : launch_fast<2, 3, false>(p, cells, ncell, color, mode, s);
  }

  int batch_waves_per_group()
  {
    return BCH;
  }

  hipError_t launch_batch_condense(int ne, int thermo, int mixed, const KParams *plist,
                                   const int *chunk_dev, const int *items, int nchunks,
                                   int color, int grid, int tile, hipStream_t s)
  {
    if (thermo)
      return mixed ? launch_batch_t<1, true, true>(plist, chunk_dev, items, nchunks, color, grid,
                                                   tile, s)
                   : launch_batch_t<1, true, false>(plist, chunk_dev, items, nchunks, color,
                                                    grid, tile, s);
    if (ne != 3)
      return hipErrorInvalidValue;
    return mixed ? launch_batch_t<3, false, true>(plist, chunk_dev, items, nchunks, color, grid,
                                                  tile, s)
                 : launch_batch_t<3, false, false>(plist, chunk_dev, items, nchunks, color, grid,
                                                   tile, s);
  }

  hipError_t launch_fast_reconstruct(int deg, const KParams &p, int ncells, hipStream_t s)
  {
    if (p.mixed)
      {
        if (p.thermo)
          return deg == 1 ? launch_fast<1, 1, true, true, true>(p, nullptr, ncells, 0, 2, s)
                          : launch_fast<2, 1, true, true, true>(p, nullptr, ncells, 0, 2, s);
        return deg == 1 ? launch_fast<1, 3, false, true, true>(p, nullptr, ncells, 0, 2, s)
                        : launch_fast<2, 3, false, true, true>(p, nullptr, ncells, 0, 2, s);
      }
    if (p.thermo)
      return deg == 1 ? launch_fast<1, 1, true, true>(p, nullptr, ncells, 0, 2, s)
                      : launch_fast<2, 1, true, true>(p, nullptr, ncells, 0, 2, s);
    return deg == 1 ? launch_fast<1, 3, false, true>(p, nullptr, ncells, 0, 2, s)
                    : launch_fast<2, 3, false, true>(p, nullptr, ncells, 0, 2, s);
  }
} // namespace hdg
