// hdg_trace_solver.hip -- the global trace solve on the GPU: a nested-
// dissection multifrontal LU of the condensed trace system.
//
// The reference factors the restricted trace matrix with SparseDirectUMFPACK
// and solves once per Newton step (src/np_solver.cpp:1326-1354).  Here the
// plan of hdg_nd_plan.h (tree, fronts, assembly and extend-add maps) is built
// once per mesh and component set; each factorisation then runs level by
// level.  A front's eliminated block is taken in panels of at most 72 dofs
// (whole faces): per panel, the LU of its diagonal block with partial
// pivoting inside the panel (one workgroup per front, in LDS), the row
// interchanges applied to the panel's rows right of it, the two
// triangular solves (U12 = L11^-1 P F12, L21 = F21 U11^-1; blocked on the
// matrix cores in groups of few fronts) and one rocBLAS
// strided-batched gemm (F22 -= L21 U12) on the rest of the front.  That is
// the elimination of a chain of nodes of at most 72 dofs each, with
// pivoting inside each node as in any multifrontal code.  Before a level is
// factored its fronts are assembled in one pass (pull form of the
// extend-add): every entry is written once with the sum of its children's
// update-block entries (slot 0, then slot 1; deterministic, no atomics, no
// zero fill), then the level's matrix entries and padded diagonals are
// added.  The fronts stay
// in HBM as the factors for the forward and backward sweeps, which run the
// same panels.
#include <algorithm>
#include <cstdint>
#include <limits>
#include <string>
#include <unordered_map>
#include <vector>

// the solution-index GEMM interface (rocblas_gemm_strided_batched_ex_get_solutions)
#define ROCBLAS_BETA_FEATURES_API
#include <rocblas/rocblas.h>
#include <chrono>
#include <map>
#include <mutex>
#include <tuple>

#include "hdg_ctx.h"
#include "hdg_nd_plan.h"
#include "hdg_wave.h"

namespace
{
  using hdg::lane_value;
  using hdg::leading_bits;
  using hdg::wave_max;
  using hdg::dbl4;
  using hdg::mfma_f64_16x16x4;

  struct DevGroup
  {
    long long *d_pf = nullptr, *d_pw = nullptr, *d_edof = nullptr, *d_idof = nullptr;
    int       *d_pn = nullptr, *d_slot = nullptr, *d_pos = nullptr;
    // assembly of the group's fronts from its nodes' children (null for a
    // group of leaves): per node and child slot, the child's Schur block
    // (front offset, leading dimension) and, per front index of the node,
    // the child's interface index there (-1: none)
    long long *d_cbase = nullptr;
    int       *d_cld = nullptr, *d_inv = nullptr;
    // the group's matrix entries and padded unit diagonals: ranges of the
    // solver's grouped lists
    long long csr0 = 0, ncsr = 0, pad0 = 0, npad = 0;
  };

  constexpr int TPB = 256;

  __global__ void k_fill(double *a, long long n, double v)
  {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
      a[i] = v;
  }

  __global__ void k_fill_int(int *a, long long n, int v)
  {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
      a[i] = v;
  }

  // the matrix entries of one group added into its fronts (entry src[k]
  // of the CSR values at front offset dst[k]; no two share an offset)
  // (ADD false: the fronts hold zeros there, a group without children)
  template <bool ADD>
  __global__ void k_scatter_add(const double *__restrict__ vals, const int *__restrict__ src,
                                const long long *__restrict__ dst, long long n,
                                double *__restrict__ F)
  {
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x)
      {
        if (ADD)
          F[dst[k]] += vals[src[k]];
        else
          F[dst[k]] = vals[src[k]];
      }
  }

  // every entry of the fronts of one group, written once: the sum of its
  // children's Schur-complement entries that land there (child slot 0,
  // then slot 1), zero where none does.  One column per wave, the rows over
  // the lanes.  cbase null: a group of leaves (zeros).
  __global__ void k_assemble(double *F, long long fbase, int n, int batch,
                             const long long *__restrict__ cbase, const int *__restrict__ cld,
                             const int *__restrict__ inv, int yoff)
  {
    const int b = blockIdx.y + yoff;
    if (b >= batch)
      return;
    double         *Fb  = F + fbase + (long long)b * n * n;
    const int       wpb = blockDim.x >> 6, l = threadIdx.x & 63;
    const long long c0  = cbase ? cbase[2 * b] : -1, c1 = cbase ? cbase[2 * b + 1] : -1;
    const int       ld0 = cbase ? cld[2 * b] : 0, ld1 = cbase ? cld[2 * b + 1] : 0;
    const int      *i0  = cbase ? inv + (long long)(2 * b) * n : nullptr;
    const int      *i1  = cbase ? i0 + n : nullptr;
    for (int c = blockIdx.x * wpb + (threadIdx.x >> 6); c < n; c += gridDim.x * wpb)
      {
        const int k0 = c0 >= 0 ? __builtin_amdgcn_readfirstlane(i0[c]) : -1;
        const int k1 = c1 >= 0 ? __builtin_amdgcn_readfirstlane(i1[c]) : -1;
        double   *Fc = Fb + (long long)c * n;
        if (k0 < 0 && k1 < 0)
          {
            for (int r = l; r < n; r += 64)
              Fc[r] = 0.0;
            continue;
          }
        const double *U0 = F + (k0 >= 0 ? c0 + (long long)k0 * ld0 : 0);
        const double *U1 = F + (k1 >= 0 ? c1 + (long long)k1 * ld1 : 0);
        // four rows per lane at a time: their index and value loads are all
        // in flight together
        constexpr int RU = 4;
        for (int rb = 0; rb < n; rb += 64 * RU)
          {
            int r0[RU], r1[RU];
#pragma unroll
            for (int u = 0; u < RU; ++u)
              {
                const int  r  = rb + l + 64 * u;
                const bool in = r < n;
                r0[u]         = (in && k0 >= 0) ? i0[r] : -1;
                r1[u]         = (in && k1 >= 0) ? i1[r] : -1;
              }
            double v0[RU], v1[RU];
#pragma unroll
            for (int u = 0; u < RU; ++u)
              {
                v0[u] = r0[u] >= 0 ? U0[r0[u]] : 0.0;
                v1[u] = r1[u] >= 0 ? U1[r1[u]] : 0.0;
              }
#pragma unroll
            for (int u = 0; u < RU; ++u)
              {
                const int r = rb + l + 64 * u;
                if (r < n)
                  Fc[r] = r1[u] >= 0 ? v0[u] + v1[u] : v0[u];
              }
          }
      }
  }

  __global__ void k_set_one(double *F, const long long *__restrict__ idx, long long n)
  {
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x)
      F[idx[k]] = 1.0;
  }

  // ---------------------------------------------------------------------
  // One panel of every front of a group: its S <= PW_MAX eliminated dofs
  // are the leading block of the trailing front at the panel's corner
  // (address fbase + b*n*n, leading dimension n).
  constexpr int PW_MAX = 64;

  // the leading S x S block at Fb (leading dimension n, S <= 64) into LDS,
  // one wave: lane t loads row t of G columns at a time (G loads in flight
  // from HBM, not one), dst(ii, jj) = out[ii * rs + jj * cs]
  template <int G>
  __device__ __forceinline__ void stage_block(const double *Fb, long long n, int S, double *out,
                                              int rs, int cs, int t)
  {
    const int ti = t < S ? t : 0;
#pragma unroll 1
    for (int j0 = 0; j0 < S; j0 += G)
      {
        double v[G];
#pragma unroll
        for (int u = 0; u < G; ++u)
          v[u] = Fb[ti + (long long)(j0 + u < S ? j0 + u : 0) * n];
        if (t < S)
#pragma unroll
          for (int u = 0; u < G; ++u)
            if (j0 + u < S)
              out[t * rs + (j0 + u) * cs] = v[u];
      }
  }

  // pivot-search key of a row lane: the leading bits of |a| (they order as
  // the values, to about 5 significant digits) with 63 - lane in the low 6
  // bits (ties go to the lowest row); 0 for a row that is not a candidate
  __device__ __forceinline__ unsigned row_key(double a, bool cand, int lane)
  {
    return cand ? (leading_bits(fabs(a)) & ~63u) + (unsigned)(63 - lane) + 1u : 0u;
  }

  // F11 = P L11 U11 of the panel with partial pivoting inside it, one wave
  // per front: lane i holds row i of the block in registers; the pivot of
  // column k is the unpivoted row with the largest |a_ik| (DPP max), its
  // row reaches every lane by readlane and is not moved.  Row k of the
  // result is [L(k, 0..k-1) | U(k, k..)] of the k-th pivot row, written in
  // pivot order; perm[k] = that row's index in the panel (front b's at
  // pbase + b*ps).  An exactly zero pivot is reported as column o + k + 1
  // (the first one of the front is kept) and its multipliers are 0.  SM >=
  // S bounds the unrolled loops: columns S..SM-1 are zero and are updated
  // without a guard (a branch per entry cost more than the extra FMAs).
  template <int SM>
  __global__ void __launch_bounds__(64)
    k_panel_lu(double *F, long long fbase, int n, int S, int *perm, long long pbase, int ps,
               int o, int *info, int batch, int yoff)
  {
    const int b = blockIdx.x + yoff;
    if (b >= batch)
      return;
    extern __shared__ double blk[]; // S x S, column-major
    const int i  = threadIdx.x;
    double   *Fb = F + fbase + (long long)b * n * n;
    // the block goes through LDS: loads from HBM in flight eight at a time
    // (row by row into registers they were issued one after the other)
    stage_block<16>(Fb, n, S, blk, 1, S, i);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double A[SM];
#pragma unroll
    for (int j = 0; j < SM; ++j)
      A[j] = (j < S && i < S) ? blk[i + j * S] : 0.0;
    bool pivoted = i >= S;
    int  myk = 0, bad = 0;
#pragma unroll
    for (int k = 0; k < SM; ++k)
      if (k < S)
      {
        const unsigned key = wave_max(row_key(A[k], !pivoted, i));
        const int      p   = 63 - (int)((key - 1u) & 63u);
        const double   d   = lane_value(A[k], p);
        if (d == 0.0 && bad == 0)
          bad = k + 1;
        const bool   upd = !pivoted && i != p;
        const double l   = (upd && d != 0.0) ? A[k] / d : 0.0;
        if (upd)
          A[k] = l;
        if (i == p)
          {
            pivoted = true;
            myk     = k;
          }
#pragma unroll
        for (int j = k + 1; j < SM; ++j)
          A[j] -= l * lane_value(A[j], p);
      }
    if (i < S)
      {
#pragma unroll
        for (int j = 0; j < SM; ++j)
          if (j < S)
            Fb[myk + (long long)j * n] = A[j];
        perm[pbase + (long long)b * ps + myk] = i;
      }
    if (i == 0 && bad && info[b] == 0)
      info[b] = o + bad;
  }

  // The same factorisation with four waves per front, for groups of few
  // fronts (where one wave per front leaves the GPU idle).  Lane i of every
  // wave is row i; wave w holds columns w, w+4, w+8, ... in SM/4 registers,
  // loaded straight from the front (coalesced down each column).  Step k:
  // the wave holding column k picks the pivot (DPP max over the unpivoted
  // rows), stores its multipliers l_i = a_ik / a_pk in place and publishes
  // them and p in LDS (two buffers, one barrier a step); then every wave
  // updates its columns right of k, a_ij -= l_i a_pj (a_pj by readlane).
  // The steps run four at a time in a loop that is not unrolled (the
  // unrolled 64 steps outgrow the instruction cache): after each four, the
  // finished column of every wave goes to LDS and its registers rotate by
  // one, so that a wave's next column is always register 0.  Same pivots,
  // multipliers, output rows and zero-pivot report as k_panel_lu.
  template <int SM>
  __global__ void __launch_bounds__(256)
    k_panel_lu4(double *F, long long fbase, int n, int S, int *perm, long long pbase, int ps,
                int o, int *info, int batch)
  {
    constexpr int NM = SM / 4, LC = 65;
    __shared__ double lb[2][64];
    __shared__ double Cs[SM * LC]; // finished columns, Cs[j * LC + i]
    __shared__ int    pb[2], badk;
    const int b = blockIdx.x;
    if (b >= batch)
      return;
    const int tid = threadIdx.x, i = tid & 63, w = tid >> 6;
    double   *Fb = F + fbase + (long long)b * n * n;
    double    A[NM];
    {
      const int ic = i < S ? i : 0;
#pragma unroll
      for (int m = 0; m < NM; ++m)
        {
          const int    j = 4 * m + w;
          const double v = Fb[ic + (long long)(j < S ? j : 0) * n];
          A[m]           = (j < S && i < S) ? v : 0.0;
        }
    }
    if (tid == 0)
      badk = 0;
    bool pivoted = i >= S;
    int  myk     = 0;
    __syncthreads();
    const int nk4 = (S + 3) / 4;
#pragma unroll 1
    for (int k4 = 0; k4 < nk4; ++k4)
      {
        // this wave's register m holds column 4 (k4 + m) + w
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
          {
            const int k = 4 * k4 + kk;
            if (k >= S)
              break;
            if (w == kk)
              {
                const double   v   = A[0];
                const unsigned key = wave_max(row_key(v, !pivoted, i));
                const int      p   = 63 - (int)((key - 1u) & 63u);
                const double   d   = lane_value(v, p);
                const bool     upd = !pivoted && i != p;
                const double   l   = (upd && d != 0.0) ? v / d : 0.0;
                if (upd)
                  A[0] = l;
                lb[k & 1][i] = l;
                if (i == 0)
                  {
                    pb[k & 1] = p;
                    if (d == 0.0 && badk == 0)
                      badk = k + 1;
                  }
              }
            __syncthreads();
            const int    p = __builtin_amdgcn_readfirstlane(pb[k & 1]);
            const double l = lb[k & 1][i];
            if (i == p)
              {
                pivoted = true;
                myk     = k;
              }
#pragma unroll
            for (int m = 0; m < NM; ++m)
              if (4 * m + w > kk && m < NM - k4)
                A[m] -= l * lane_value(A[m], p);
          }
        // column 4 k4 + w is final: out to LDS, then the rotation
        Cs[(4 * k4 + w) * LC + i] = A[0];
#pragma unroll
        for (int m = 0; m + 1 < NM; ++m)
          A[m] = A[m + 1];
      }
    __syncthreads();
    if (i < S)
      {
        for (int j = w; j < S; j += 4)
          Fb[myk + (long long)j * n] = Cs[j * LC + i];
        if (w == 0)
          perm[pbase + (long long)b * ps + myk] = i;
      }
    if (tid == 0 && badk && info[b] == 0)
      info[b] = o + badk;
  }

  // The off-diagonal blocks of one panel, one launch: blocks x < nbl (one
  // lane per column, the Ic columns from column c0 of the panel's trailing
  // front) form U12 = L11^-1 P F12 (the column's rows gathered in pivot
  // order), blocks x >= nbl (one lane per row, the Ir rows below the panel)
  // form L21 = F21 U11^-1.  The factor is staged in LDS and read at
  // wave-uniform addresses; the lane's column or row stays in registers.
  __global__ void __launch_bounds__(64)
    k_panel_trsm(double *F, long long fbase, int n, int S, int Ic, int c0, int Ir,
                 const int *__restrict__ perm, long long pbase, int ps, int nbl, int batch,
                 int yoff)
  {
    // the factor, S x S (left: L row-major; right: U column-major, leading
    // dimension S): dynamic, so a small panel leaves room for more waves
    extern __shared__ double Ts[];
    __shared__ int pm[PW_MAX];
    const int b = blockIdx.y + yoff;
    if (b >= batch)
      return;
    const bool left = (int)blockIdx.x < nbl;
    const int  t    = threadIdx.x;
    const int  r    = (left ? (int)blockIdx.x : (int)blockIdx.x - nbl) * 64 + t;
    double    *Fb   = F + fbase + (long long)b * n * n;
    if (left)
      stage_block<16>(Fb, n, S, Ts, S, 1, t);
    else
      stage_block<16>(Fb, n, S, Ts, 1, S, t);
    if (left && t < S)
      pm[t] = perm[pbase + (long long)b * ps + t];
    __syncthreads();
    const bool act = r < (left ? Ic : Ir);
    double     y[PW_MAX];
    if (left)
      {
        double *col = Fb + (long long)(c0 + (act ? r : 0)) * n;
#pragma unroll
        for (int j = 0; j < PW_MAX; ++j)
          {
            const double v = col[pm[j < S ? j : 0]];
            y[j]           = v * ((j < S && act) ? 1.0 : 0.0);
          }
#pragma unroll
        for (int j = 1; j < PW_MAX; ++j)
          if (j < S)
          {
            double a0 = y[j], a1 = 0.0;
#pragma unroll
            for (int l = 0; l < j; ++l)
              if (l & 1)
                a1 -= Ts[j * S + l] * y[l];
              else
                a0 -= Ts[j * S + l] * y[l];
            y[j] = a0 + a1;
          }
        if (act)
#pragma unroll
          for (int j = 0; j < PW_MAX; ++j)
            if (j < S)
              col[j] = y[j];
      }
    else
      {
        double *row = Fb + S + (act ? r : 0);
#pragma unroll
        for (int j = 0; j < PW_MAX; ++j)
          {
            const double v = row[(long long)(j < S ? j : 0) * n];
            y[j]           = v * ((j < S && act) ? 1.0 : 0.0);
          }
#pragma unroll
        for (int j = 0; j < PW_MAX; ++j)
          if (j < S)
          {
            double a0 = y[j], a1 = 0.0;
#pragma unroll
            for (int l = 0; l < j; ++l)
              if (l & 1)
                a1 -= y[l] * Ts[l + j * S];
              else
                a0 -= y[l] * Ts[l + j * S];
            y[j] = (a0 + a1) / Ts[j + j * S];
          }
        if (act)
#pragma unroll
          for (int j = 0; j < PW_MAX; ++j)
            if (j < S)
              row[(long long)j * n] = y[j];
      }
  }

  // The same blocks on the matrix cores, for groups of at most
  // TRSM_MFMA_MAX_BATCH fronts (where one lane per column leaves the GPU
  // idle), one launch:
  // one workgroup (four waves) per tile of 64 columns of F12 (blocks x <
  // nbl) or 64 rows of F21 of one front.  Both are a lower-triangular solve M Y = B with the panel's
  // factor staged in LDS:
  //   U12: M = L11 (unit lower), B = P F12, the rows gathered in pivot order;
  //   L21: M = U11^T, B = F21^T (L21^T = U11^-T F21^T).
  // Blocked by 16: wave 0 inverts the diagonal 16 x 16 blocks M_ii (lane =
  // 16 i + column, a 16-step substitution); then wave w solves columns
  // 16w..16w+15 of the tile block row by block row,
  //   Y_i = M_ii^-1 (B_i - sum_{k<i} M_ik Y_k),
  // every product on v_mfma_f64_16x16x4f64 (A[i][k] in lane i + 16k, B[k][j]
  // in lane j + 16k, element e of the result = D[4e + lane/16][lane%16]).
  // Element kk of a result is its B operand for the K step kk (rows
  // 4kk..4kk+3), so the Y_k stay in registers.  The tile is loaded and stored
  // coalesced through LDS.  SM >= S bounds the blocks; the factor is padded
  // with the identity and the tile with zeros.
  template <int SM>
  __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
    k_trsm_mfma(double *F, long long fbase, int n, int S, int Ic, int c0, int Ir,
                const int *__restrict__ perm, long long pbase, int ps, int nbl, int batch,
                int yoff)
  {
    constexpr int LF = SM + 1, LB = 65, NB = SM / 16, LD = 17;
    extern __shared__ double sm[];
    double       *Fs = sm;                 // SM x LF: the factored block, row-major
    double       *Ds = Fs + SM * LF;       // NB x 16 x LD: inverses of the diagonal blocks
    double       *Bs = Ds + NB * 16 * LD;  // SM x LB: B (row j, column c of the tile)
    const int b = blockIdx.y + yoff;
    if (b >= batch)
      return;
    const bool left = (int)blockIdx.x < nbl;
    const int  tile = (left ? (int)blockIdx.x : (int)blockIdx.x - nbl) * 64;
    const int  nvr  = (left ? Ic : Ir) - tile;
    const int  nv   = nvr < 64 ? nvr : 64;
    const int  tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    double    *Fb  = F + fbase + (long long)b * n * n;
    // every load of a wave is issued before its LDS stores (one HBM round
    // trip, not SM/4)
    constexpr int NJ = SM / 4;
    {
      double v[NJ];
      const int li = lane < S ? lane : 0;
#pragma unroll
      for (int jj = 0; jj < NJ; ++jj)
        {
          const int j = 4 * jj + w;
          v[jj]       = Fb[li + (long long)(j < S ? j : 0) * n];
        }
      if (lane < SM)
#pragma unroll
        for (int jj = 0; jj < NJ; ++jj)
          {
            const int j        = 4 * jj + w;
            Fs[lane * LF + j] = (lane < S && j < S) ? v[jj] : (lane == j ? 1.0 : 0.0);
          }
    }
    if (left)
      {
        // lane j: row perm[j] of column c (a permutation inside the column)
        const int pj = lane < S ? perm[pbase + (long long)b * ps + lane] : 0;
        double    v[16];
#pragma unroll
        for (int cc = 0; cc < 16; ++cc)
          {
            const int c = 4 * cc + w;
            v[cc]       = Fb[pj + (long long)(c0 + tile + (c < nv ? c : 0)) * n];
          }
        if (lane < SM)
#pragma unroll
          for (int cc = 0; cc < 16; ++cc)
            {
              const int c       = 4 * cc + w;
              Bs[lane * LB + c] = (lane < S && c < nv) ? v[cc] : 0.0;
            }
      }
    else
      {
        // lane r: row S + tile + r of F21, column j (transposed into Bs)
        const int ri = lane < nv ? lane : 0;
        double    v[NJ];
#pragma unroll
        for (int jj = 0; jj < NJ; ++jj)
          {
            const int j = 4 * jj + w;
            v[jj]       = Fb[S + tile + ri + (long long)(j < S ? j : 0) * n];
          }
#pragma unroll
        for (int jj = 0; jj < NJ; ++jj)
          {
            const int j      = 4 * jj + w;
            Bs[j * LB + lane] = (lane < nv && j < S) ? v[jj] : 0.0;
          }
      }
    __syncthreads();
    // M[r][c] (r >= c) from the factored block
    auto mval = [&](int r, int c) -> double {
      return left ? (r == c ? 1.0 : Fs[r * LF + c]) : Fs[c * LF + r];
    };
    if (w == 0)
      {
        const int bi = lane >> 4, j = lane & 15, o = 16 * bi;
        if (bi < NB)
          {
            double x[16];
#pragma unroll
            for (int r = 0; r < 16; ++r)
              {
                double a = r == j ? 1.0 : 0.0;
#pragma unroll
                for (int m = 0; m < r; ++m)
                  a = fma(-mval(o + r, o + m), x[m], a);
                x[r] = a / mval(o + r, o + r);
              }
#pragma unroll
            for (int r = 0; r < 16; ++r)
              Ds[(o + r) * LD + j] = x[r];
          }
      }
    __syncthreads();
    const int q = lane >> 4, c16 = lane & 15;
    dbl4      Y[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i)
      {
        dbl4 acc;
#pragma unroll
        for (int e = 0; e < 4; ++e)
          acc[e] = Bs[(16 * i + 4 * e + q) * LB + 16 * w + c16];
#pragma unroll
        for (int k = 0; k < i; ++k)
#pragma unroll
          for (int kk = 0; kk < 4; ++kk)
            acc = mfma_f64_16x16x4(-mval(16 * i + c16, 16 * k + 4 * kk + q), Y[k][kk], acc);
        Y[i] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
          Y[i] = mfma_f64_16x16x4(Ds[(16 * i + c16) * LD + 4 * kk + q], acc[kk], Y[i]);
      }
    // each wave wrote and reads only its own 16 columns of Bs
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        Bs[(16 * i + 4 * e + q) * LB + 16 * w + c16] = Y[i][e];
    __syncthreads();
    if (left)
      {
        for (int c = w; c < nv; c += 4)
          if (lane < S)
            Fb[lane + (long long)(c0 + tile + c) * n] = Bs[lane * LB + c];
      }
    else
      {
        for (int j = w; j < S; j += 4)
          if (lane < nv)
            Fb[S + tile + lane + (long long)j * n] = Bs[j * LB + lane];
      }
  }

  // forward sweep of one panel of a front, one wave: z = L11^-1 P w_s with
  // L11 staged in LDS and z in registers (rows lane and lane + 64; each
  // solved entry reaches the other lanes by readlane), then w_i -= L21 z.
  // perm: row k of the factored block is row perm[k] of the panel; front
  // b's at pbase + b*ps.
  // the S x S block at Fb (leading dimension n) into LDS column-major by
  // the whole workgroup (NT threads), eight loads in flight per thread
  template <int NT>
  __device__ __forceinline__ void stage_square(const double *Fb, long long n, int S, double *out,
                                               int t)
  {
    const int SS = S * S;
#pragma unroll 1
    for (int e0 = 0; e0 < SS; e0 += 8 * NT)
      {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
          {
            const int e = e0 + u * NT + t;
            v[u]        = e < SS ? Fb[(e % S) + (long long)(e / S) * n] : 0.0;
          }
#pragma unroll
        for (int u = 0; u < 8; ++u)
          {
            const int e = e0 + u * NT + t;
            if (e < SS)
              out[e] = v[u];
          }
      }
  }

  // forward sweep of one panel of a front, four waves: the factor staged by
  // all of them; wave 0 solves z = L11^-1 P w_s (z in registers, entries by
  // readlane); then every thread updates rows of the short rest,
  // w_rest -= L21 z (I rows, one per thread; 0 when a separate kernel does
  // it)
  constexpr int SWEEP_T = 256;
  __global__ void __launch_bounds__(SWEEP_T)
    k_fwd_small(const double *F, double *w, long long fbase, long long wbase, int n, int S,
                int I, const int *perm, long long pbase, int ps, int batch, int yoff)
  {
    extern __shared__ double L[]; // S x S, column-major
    const int b = blockIdx.x + yoff;
    if (b >= batch)
      return;
    const int     t  = threadIdx.x;
    const double *Fb = F + fbase + (long long)b * n * n;
    double       *wb = w + wbase + (long long)b * n;
    const int    *pv = perm + pbase + (long long)b * ps;
    stage_square<SWEEP_T>(Fb, n, S, L, t);
    __syncthreads();
    if (t < 64)
      {
        const int i0 = t, i1 = t + 64;
        double    z0 = i0 < S ? wb[pv[i0]] : 0.0;
        double    z1 = i1 < S ? wb[pv[i1]] : 0.0;
        for (int j = 0; j + 1 < S; ++j)
          {
            const double y = j < 64 ? lane_value(z0, j) : lane_value(z1, j - 64);
            if (i0 > j && i0 < S)
              z0 -= L[i0 + j * S] * y;
            if (i1 > j && i1 < S)
              z1 -= L[i1 + j * S] * y;
          }
        // every lane has read its pivoted entries of w_s: overwrite them
        __builtin_amdgcn_wave_barrier();
        if (i0 < S)
          wb[i0] = z0;
        if (i1 < S)
          wb[i1] = z1;
      }
    __syncthreads();
    for (int r = t; r < I; r += SWEEP_T)
      {
        double a0 = wb[S + r], a1 = 0, a2 = 0, a3 = 0;
        int    j  = 0;
        for (; j + 3 < S; j += 4)
          {
            a0 -= Fb[S + r + (long long)j * n] * wb[j];
            a1 -= Fb[S + r + (long long)(j + 1) * n] * wb[j + 1];
            a2 -= Fb[S + r + (long long)(j + 2) * n] * wb[j + 2];
            a3 -= Fb[S + r + (long long)(j + 3) * n] * wb[j + 3];
          }
        for (; j < S; ++j)
          a0 -= Fb[S + r + (long long)j * n] * wb[j];
        wb[S + r] = (a0 + a1) + (a2 + a3);
      }
  }

  // backward sweep of one panel of a front, four waves: z = w_s - U12 x_i
  // (wave q sums the columns c = q, q + 4, ... of each row; the four
  // partial sums are added in order), then wave 0 applies U11^-1 with U11
  // staged in LDS and z in registers
  __global__ void __launch_bounds__(SWEEP_T)
    k_bwd_small(const double *F, double *w, long long fbase, long long wbase, int n, int S,
                int I, int batch, int yoff)
  {
    extern __shared__ double U[]; // S x S column-major, then 4 x S partial sums
    const int b = blockIdx.x + yoff;
    if (b >= batch)
      return;
    const int     t  = threadIdx.x, l = t & 63, q = t >> 6;
    const double *Fb = F + fbase + (long long)b * n * n;
    double       *wb = w + wbase + (long long)b * n;
    double       *part = U + S * S;
    stage_square<SWEEP_T>(Fb, n, S, U, t);
    for (int h = 0; h < 2; ++h)
      {
        const int i = h ? l + 64 : l;
        if (i >= S)
          continue;
        double a0 = q == 0 ? wb[i] : 0.0, a1 = 0;
        int    c  = q;
        for (; c + 4 < I; c += 8)
          {
            a0 -= Fb[i + (long long)(S + c) * n] * wb[S + c];
            a1 -= Fb[i + (long long)(S + c + 4) * n] * wb[S + c + 4];
          }
        for (; c < I; c += 4)
          a0 -= Fb[i + (long long)(S + c) * n] * wb[S + c];
        part[q * S + i] = a0 + a1;
      }
    __syncthreads();
    if (t < 64)
      {
        const int i0 = t, i1 = t + 64;
        double    z0 = 0.0, z1 = 0.0;
        if (i0 < S)
          z0 = ((part[i0] + part[S + i0]) + part[2 * S + i0]) + part[3 * S + i0];
        if (i1 < S)
          z1 = ((part[i1] + part[S + i1]) + part[2 * S + i1]) + part[3 * S + i1];
        for (int j = S - 1; j >= 0; --j)
          {
            const double y = (j < 64 ? lane_value(z0, j) : lane_value(z1, j - 64)) / U[j + j * S];
            if (i0 == j)
              z0 = y;
            else if (i0 < j)
              z0 -= U[i0 + j * S] * y;
            if (i1 == j)
              z1 = y;
            else if (i1 < j)
              z1 -= U[i1 + j * S] * y;
          }
        if (i0 < S)
          wb[i0] = z0;
        if (i1 < S)
          wb[i1] = z1;
      }
  }


  // ---------------------------------------------------------------------
  // The sweeps' off-diagonal updates where the rest of a panel's trailing
  // front is long.  Forward: w_rest -= L21 z, one thread per row, z (the
  // panel's solved entries) in LDS, each L21 column read coalesced across
  // the rows.  SM >= S bounds the unrolled loop; z is zero past S.
  template <int SM>
  __global__ void __launch_bounds__(256)
    k_sweep_rows(const double *F, double *w, long long fbase, long long wbase, int n, int S,
                 int I, int batch, int yoff)
  {
    __shared__ double z[SM];
    const int b = blockIdx.y + yoff;
    if (b >= batch)
      return;
    const double *Fb = F + fbase + (long long)b * n * n;
    double       *wb = w + wbase + (long long)b * n;
    const int     t  = threadIdx.x;
    if (t < SM)
      z[t] = t < S ? wb[t] : 0.0;
    __syncthreads();
    const int r = blockIdx.x * 256 + t;
    if (r >= I)
      return;
    const double *col = Fb + S + r;
    double        a0 = 0.0, a1 = 0.0;
#pragma unroll
    for (int j = 0; j < SM; ++j)
      {
        const double v = col[(long long)(j < S ? j : 0) * n] * z[j];
        if (j & 1)
          a1 += v;
        else
          a0 += v;
      }
    wb[S + r] -= a0 + a1;
  }

  // Backward, first half: partial sums of U12 x_rest over chunks of 256
  // columns, one workgroup per (chunk, front): lane j of each wave takes
  // row j (coalesced down a column), the four waves every fourth column;
  // the chunk's sums go to part[(b * nch + chunk) * 64 + j]
  __global__ void __launch_bounds__(256)
    k_sweep_cols_part(const double *F, const double *w, long long fbase, long long wbase, int n,
                      int S, int I, int nch, double *part, int batch, int yoff)
  {
    __shared__ double red[4][64];
    const int b = blockIdx.y + yoff;
    if (b >= batch)
      return;
    const double *Fb = F + fbase + (long long)b * n * n;
    const double *wb = w + wbase + (long long)b * n;
    const int     t = threadIdx.x, j = t & 63, cg = t >> 6;
    const int     jc = j < S ? j : 0;
    const double  mj = j < S ? 1.0 : 0.0;
    const int     c0 = blockIdx.x * 256;
    double        a0 = 0.0, a1 = 0.0;
#pragma unroll 8
    for (int i = 0; i < 64; ++i)
      {
        const int c  = c0 + cg + 4 * i;
        const int cc = c < I ? c : 0;
        const double v = Fb[jc + (long long)(S + cc) * n] * (c < I ? wb[S + cc] : 0.0);
        if (i & 1)
          a1 += v;
        else
          a0 += v;
      }
    red[cg][j] = (a0 + a1) * mj;
    __syncthreads();
    if (t < 64)
      part[((long long)b * nch + blockIdx.x) * 64 + t] =
        (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
  }

  // Backward, second half: w_j -= the chunks' partial sums, in chunk order
  // (deterministic), one wave per front
  __global__ void __launch_bounds__(64)
    k_sweep_cols_sum(double *w, long long wbase, int n, int S, int nch, const double *part,
                     int batch, int yoff)
  {
    const int b = blockIdx.x + yoff;
    if (b >= batch)
      return;
    const int t = threadIdx.x;
    if (t >= S)
      return;
    double a = 0.0;
    for (int k = 0; k < nch; ++k)
      a += part[((long long)b * nch + k) * 64 + t];
    w[wbase + (long long)b * n + t] -= a;
  }

  __global__ void k_gather_rhs(const double *__restrict__ b, const long long *__restrict__ wmap,
                               long long ndof, double *__restrict__ w)
  {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < ndof;
         i += (long long)gridDim.x * blockDim.x)
      w[wmap[i]] = b[i];
  }

  __global__ void k_vec_extend_add(double *w, long long wbase, int n, int S, int I, int batch,
                                   const long long *__restrict__ pw,
                                   const long long *__restrict__ pf, const int *__restrict__ slot,
                                   const int *__restrict__ pos, int want, int yoff)
  {
    const int b = blockIdx.y + yoff;
    if (b >= batch || slot[b] != want || pf[b] < 0)
      return;
    const double *u  = w + wbase + (long long)b * n + S;
    const int    *ps = pos + (long long)b * I;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < I; r += gridDim.x * blockDim.x)
      if (ps[r] >= 0)
        w[pw[b] + ps[r]] += u[r];
  }

  __global__ void k_gather_x(const double *__restrict__ x, double *w, long long wbase, int n,
                             int S, int I, int batch, const long long *__restrict__ idof, int yoff)
  {
    const int b = blockIdx.y + yoff;
    if (b >= batch)
      return;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < I; r += gridDim.x * blockDim.x)
      {
        const long long dof = idof[(long long)b * I + r];
        w[wbase + (long long)b * n + S + r] = dof >= 0 ? x[dof] : 0.0;
      }
  }

  __global__ void k_scatter_x(double *x, const double *__restrict__ w, long long wbase, int n,
                              int S, int batch, const long long *__restrict__ edof, int yoff)
  {
    const int b = blockIdx.y + yoff;
    if (b >= batch)
      return;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < S; j += gridDim.x * blockDim.x)
      {
        const long long dof = edof[(long long)b * S + j];
        if (dof >= 0)
          x[dof] = w[wbase + (long long)b * n + j];
      }
  }

  // dynamic LDS of k_trsm_mfma<SM>: the factor SM x (SM+1), the diagonal
  // inverses SM x 17 and the tile SM x 65 doubles
  constexpr size_t trsm_mfma_lds(int SM)
  {
    return (size_t)(SM * (SM + 1) + SM * 17 + SM * 65) * sizeof(double);
  }

  // groups of at most this many fronts take the triangular solves and the
  // panel LU in four-wave workgroups (k_trsm_mfma, k_panel_lu4); larger
  // groups keep one wave per front and panel (enough waves to fill the GPU)
  constexpr int TRSM_MFMA_MAX_BATCH = 1024;

  constexpr int YCHUNK = 32768; // nodes per launch along grid y
  // a sweep's off-diagonal update stays in the one-wave panel kernel up to
  // this many trailing rows, and goes to a batched gemv above
  constexpr int FUSE_I = 256;
  // panels per super-panel of the factorisation (rank of its trailing GEMM:
  // SP panel widths)
  constexpr int SP = 4;
  // super-panels where the trailing fronts of a group hold at least this
  // many entries (64 MB)
  constexpr long long SP_MIN_ENTRIES = 8ll << 20;

  // C -= A B over the batch (leading dimension ld, batch stride st): with
  // the rocBLAS solution sol from the tuning below, or rocBLAS's own choice
  // (sol 0, or a solution that refuses the shape)
  rocblas_status strided_gemm(rocblas_handle h, int sol, int M, int N, int K, const double *A,
                              const double *B, double *C, int ld, long long st, int batch)
  {
    const double one = 1.0, mone = -1.0;
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wdeprecated-declarations"
    if (sol != 0)
      {
        const rocblas_status r = rocblas_gemm_strided_batched_ex(
          h, rocblas_operation_none, rocblas_operation_none, M, N, K, &mone, A,
          rocblas_datatype_f64_r, ld, st, B, rocblas_datatype_f64_r, ld, st, &one, C,
          rocblas_datatype_f64_r, ld, st, C, rocblas_datatype_f64_r, ld, st, batch,
          rocblas_datatype_f64_r, rocblas_gemm_algo_solution_index, sol, rocblas_gemm_flags_none);
        if (r == rocblas_status_success)
          return r;
      }
#pragma clang diagnostic pop
    return rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, M, N,
                                         K, &mone, A, ld, st, B, ld, st, &one, C, ld, st, batch);
  }

  // groups whose factorisation holds at least this many FLOPs have their
  // main trailing GEMM tuned
  constexpr double GEMM_TUNE_MIN_FLOP = 5e9;
  // wall time one solver may spend tuning (it is created once per mesh and
  // component set; a served request creates one per stage)
  constexpr double GEMM_TUNE_BUDGET_S = 0.3;

  // The fastest rocBLAS solution for one GEMM shape, timed on scratch
  // operands (A, B, C inside a front buffer; values do not matter): every
  // solution rocBLAS lists for the shape against rocBLAS's default, one
  // warm-up and one timed call each, on at most 2^26 / (M N) of the batch.
  // 0 (the default) unless a solution is at least 5 % faster.  Results are
  // kept per shape for the process.
  int tune_gemm(rocblas_handle h, int M, int N, int K, const double *A, const double *B,
                double *C, int ld, long long st, int batch_all)
  {
    const int batch = (int)std::max<long long>(
      1, std::min<long long>(batch_all, (1ll << 26) / std::max<long long>(1, (long long)M * N)));
    typedef std::tuple<int, int, int, int, long long, int> Key;
    static std::map<Key, int> cache;
    static std::mutex         mu;
    const Key                 key{M, N, K, ld, st, batch};
    {
      std::lock_guard<std::mutex> lk(mu);
      auto                        it = cache.find(key);
      if (it != cache.end())
        return it->second;
    }
    const double one = 1.0, mone = -1.0;
    rocblas_int  nsol = 0;
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wdeprecated-declarations"
    auto list = [&](rocblas_int *arr) {
      return rocblas_gemm_strided_batched_ex_get_solutions(
        h, rocblas_operation_none, rocblas_operation_none, M, N, K, &mone, A,
        rocblas_datatype_f64_r, ld, st, B, rocblas_datatype_f64_r, ld, st, &one, C,
        rocblas_datatype_f64_r, ld, st, C, rocblas_datatype_f64_r, ld, st, batch,
        rocblas_datatype_f64_r, rocblas_gemm_algo_solution_index, rocblas_gemm_flags_none, arr,
        &nsol);
    };
#pragma clang diagnostic pop
    int best = 0;
    if (list(nullptr) == rocblas_status_success && nsol > 0)
      {
        std::vector<rocblas_int> sols((size_t)nsol);
        if (list(sols.data()) == rocblas_status_success)
          {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            rocblas_set_stream(h, nullptr);
            if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess)
              {
                auto ms = [&](int sol) -> float {
                  if (strided_gemm(h, sol, M, N, K, A, B, C, ld, st, batch) !=
                      rocblas_status_success)
                    return -1.0f;
                  (void)hipEventRecord(e0, nullptr);
                  (void)strided_gemm(h, sol, M, N, K, A, B, C, ld, st, batch);
                  (void)hipEventRecord(e1, nullptr);
                  float t = -1.0f;
                  if (hipEventSynchronize(e1) != hipSuccess ||
                      hipEventElapsedTime(&t, e0, e1) != hipSuccess)
                    return -1.0f;
                  return t;
                };
                const float tdef = ms(0);
                float       tb   = tdef > 0 ? 0.95f * tdef : 0.0f;
                for (rocblas_int sol : sols)
                  {
                    const float t = ms(sol);
                    if (t > 0 && t < tb)
                      {
                        tb   = t;
                        best = sol;
                      }
                  }
              }
            if (e0)
              (void)hipEventDestroy(e0);
            if (e1)
              (void)hipEventDestroy(e1);
          }
      }
    (void)hipGetLastError();
    std::lock_guard<std::mutex> lk(mu);
    cache[key] = best;
    return best;
  }

  int grid_for(long long work)
  {
    const long long g = (work + TPB - 1) / TPB;
    return (int)std::max<long long>(1, std::min<long long>(g, 8192));
  }

  template <class T>
  int upload(T **dst, const std::vector<T> &src, const char *what)
  {
    *dst = nullptr;
    if (src.empty())
      return 0;
    hipError_t e = hipMalloc((void **)dst, src.size() * sizeof(T));
    if (e == hipSuccess)
      e = hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess)
      return hdg::set_error(std::string(what) + ": " + hipGetErrorString(e), -3);
    return 0;
  }
} // namespace

struct hdg_trace_solver
{
  hdg::NDPlan           P;
  std::vector<DevGroup> dg;
  rocblas_handle        h      = nullptr;
  double               *d_F    = nullptr, *d_w = nullptr;
  int                  *d_info = nullptr;
  int                   pw     = PW_MAX; // panel width: whole faces, at most PW_MAX dofs
  double               *d_part = nullptr; // backward sweep: chunk partial sums
  int                  *d_perm = nullptr; // row order of the small fronts' factored blocks
  long long            *d_map = nullptr, *d_pad = nullptr, *d_wmap = nullptr;
  int                  *d_csrc = nullptr; // CSR entry of each d_map offset (both grouped)
  std::vector<int>      gemm_sol; // per group: the tuned rocBLAS solution (0: rocBLAS's)
};

namespace
{
  // The tables that assemble each group's fronts just before the group is
  // factored (hdg_trace_factor): per node, its children's Schur blocks and
  // index maps (pull form of the extend-add: every front entry is written
  // once, no zero fill, no read-modify-write of the parent), and the
  // matrix entries and padded diagonals of each group as contiguous ranges.
  int build_assembly(hdg_trace_solver *T)
  {
    const hdg::NDPlan &P  = T->P;
    const int          ng = (int)P.groups.size();
    std::vector<std::pair<long long, int>> by_base(ng);
    for (int gi = 0; gi < ng; ++gi)
      by_base[gi] = {P.groups[gi].fbase, gi};
    std::sort(by_base.begin(), by_base.end());
    // group and node of a front entry's offset
    auto locate = [&](long long off, int &gi, long long &b, long long &e) {
      auto it = std::upper_bound(by_base.begin(), by_base.end(),
                                 std::make_pair(off, std::numeric_limits<int>::max()));
      if (it == by_base.begin())
        return false;
      gi = std::prev(it)->second;
      const hdg::NDGroup &G  = P.groups[gi];
      const long long     nn = (long long)G.n * G.n;
      b = (off - G.fbase) / nn;
      e = (off - G.fbase) % nn;
      return b >= 0 && b < G.batch;
    };
    std::vector<std::vector<long long>> cbase(ng);
    std::vector<std::vector<int>>       cld(ng), inv(ng);
    for (int gc = 0; gc < ng; ++gc)
      {
        const hdg::NDGroup &C = P.groups[gc];
        for (int b = 0; b < C.batch; ++b)
          {
            if (C.pf[b] < 0 || C.I == 0)
              continue;
            int       gp;
            long long bp, e;
            if (!locate(C.pf[b], gp, bp, e) || e != 0 || C.slot[b] < 0 || C.slot[b] > 1)
              return hdg::set_error("hdg_trace_solver_create: a parent front outside the plan", -4);
            const hdg::NDGroup &G = P.groups[gp];
            if (cbase[gp].empty())
              {
                cbase[gp].assign(2 * (size_t)G.batch, -1);
                cld[gp].assign(2 * (size_t)G.batch, 0);
                inv[gp].assign(2 * (size_t)G.batch * G.n, -1);
              }
            const long long k = 2 * bp + C.slot[b];
            if (cbase[gp][k] >= 0)
              return hdg::set_error("hdg_trace_solver_create: two children in one slot", -4);
            cbase[gp][k] = C.fbase + (long long)b * C.n * C.n + C.S + (long long)C.S * C.n;
            cld[gp][k]   = C.n;
            for (int r = 0; r < C.I; ++r)
              {
                const int q = C.pos[(long long)b * C.I + r];
                if (q < -1 || q >= G.n)
                  return hdg::set_error("hdg_trace_solver_create: interface index out of range", -4);
                if (q >= 0)
                  inv[gp][k * G.n + q] = r;
              }
          }
      }
    // the matrix entries and the padded diagonals, grouped (stable)
    auto group_lists = [&](const std::vector<long long> &offs, std::vector<long long> &out,
                           std::vector<int> *src, bool pads) {
      std::vector<long long> count(ng + 1, 0);
      std::vector<int>       gof(offs.size());
      for (size_t k = 0; k < offs.size(); ++k)
        {
          int       gi;
          long long b, e;
          if (!locate(offs[k], gi, b, e))
            return false;
          gof[k] = gi;
          ++count[gi + 1];
          if (pads)
            {
              // a padded diagonal receives nothing from the children
              const hdg::NDGroup &G = P.groups[gi];
              const int           j = (int)(e % G.n);
              if (!inv[gi].empty())
                for (int sl = 0; sl < 2; ++sl)
                  if (inv[gi][(2 * b + sl) * G.n + j] >= 0)
                    return false;
            }
        }
      for (int gi = 0; gi < ng; ++gi)
        count[gi + 1] += count[gi];
      out.assign(offs.size(), 0);
      if (src)
        src->assign(offs.size(), 0);
      std::vector<long long> at(count.begin(), count.end() - 1);
      for (size_t k = 0; k < offs.size(); ++k)
        {
          const long long q = at[gof[k]]++;
          out[q]            = offs[k];
          if (src)
            (*src)[q] = (int)k;
        }
      for (int gi = 0; gi < ng; ++gi)
        {
          DevGroup &D = T->dg[gi];
          (pads ? D.pad0 : D.csr0) = count[gi];
          (pads ? D.npad : D.ncsr) = count[gi + 1] - count[gi];
        }
      return true;
    };
    if (P.map.size() >= (size_t)std::numeric_limits<int>::max())
      return hdg::set_error("hdg_trace_solver_create: more than 2^31 matrix entries", -4);
    std::vector<long long> dst, pad;
    std::vector<int>       src;
    if (!group_lists(P.map, dst, &src, false))
      return hdg::set_error("hdg_trace_solver_create: a matrix entry outside the fronts", -4);
    if (!group_lists(P.pad, pad, nullptr, true))
      return hdg::set_error("hdg_trace_solver_create: a padded diagonal outside the fronts or under a child", -4);
    int rc = upload(&T->d_map, dst, "trace solver map");
    if (rc == 0)
      rc = upload(&T->d_csrc, src, "trace solver map");
    if (rc == 0)
      rc = upload(&T->d_pad, pad, "trace solver padding");
    for (int gi = 0; gi < ng && rc == 0; ++gi)
      {
        DevGroup &D = T->dg[gi];
        rc          = upload(&D.d_cbase, cbase[gi], "trace solver assembly");
        if (rc == 0)
          rc = upload(&D.d_cld, cld[gi], "trace solver assembly");
        if (rc == 0)
          rc = upload(&D.d_inv, inv[gi], "trace solver assembly");
      }
    return rc;
  }
} // namespace

extern "C" {

int hdg_trace_solver_create(const hdg_ctx *C, hdg_trace_solver **out)
{
  if (!C || !out)
    return hdg::set_error("hdg_trace_solver_create: null argument", -1);
  *out                = nullptr;
  hdg_trace_solver *T = new hdg_trace_solver();
  std::string       err;
  if (!hdg::build_nd_plan(C->topo, C->ne * C->nf, T->P, err))
    {
      delete T;
      return hdg::set_error("hdg_trace_solver_create: " + err, -4);
    }
  const hdg::NDPlan &P = T->P;
  int                rc = 0;
  T->dg.resize(P.groups.size());
  for (size_t gi = 0; gi < P.groups.size() && rc == 0; ++gi)
    {
      const hdg::NDGroup &G = P.groups[gi];
      DevGroup           &D = T->dg[gi];
      if (rc == 0)
        rc = upload(&D.d_pf, G.pf, "trace solver tables");
      if (rc == 0)
        rc = upload(&D.d_pw, G.pw, "trace solver tables");
      if (rc == 0)
        rc = upload(&D.d_pn, G.pn, "trace solver tables");
      if (rc == 0)
        rc = upload(&D.d_slot, G.slot, "trace solver tables");
      if (rc == 0)
        rc = upload(&D.d_pos, G.pos, "trace solver tables");
      if (rc == 0)
        rc = upload(&D.d_edof, G.edof, "trace solver tables");
      if (rc == 0)
        rc = upload(&D.d_idof, G.idof, "trace solver tables");
    }
  if (rc == 0)
    rc = build_assembly(T);
  if (rc == 0)
    rc = upload(&T->d_wmap, P.wmap, "trace solver rhs map");
  if (rc == 0)
    {
      hipError_t e = hipMalloc((void **)&T->d_F, (size_t)P.Ftot * sizeof(double));
      if (e == hipSuccess)
        e = hipMalloc((void **)&T->d_w, (size_t)std::max<long long>(1, P.wtot) * sizeof(double));
      long long npart = 1;
      for (const hdg::NDGroup &G : P.groups)
        npart = std::max<long long>(npart, (long long)G.batch * ((G.n + 255) / 256) * 64);
      if (e == hipSuccess)
        e = hipMalloc((void **)&T->d_part, (size_t)npart * sizeof(double));
      if (e == hipSuccess)
        e = hipMalloc((void **)&T->d_perm, (size_t)std::max<long long>(1, P.ptot) * sizeof(int));
      if (e == hipSuccess)
        e = hipMalloc((void **)&T->d_info, (size_t)P.nnodes * sizeof(int));
      if (e != hipSuccess)
        rc = hdg::set_error(std::string("hdg_trace_solver_create: fronts: ") +
                              hipGetErrorString(e),
                            -3);
    }
  if (rc == 0)
    T->pw = std::max(1, PW_MAX / std::max(1, P.tdf)) * std::max(1, P.tdf);
  if (rc == 0 && T->pw > PW_MAX)
    rc = hdg::set_error("hdg_trace_solver_create: more than 64 dofs per face", -4);
  if (rc == 0)
    {
      // k_trsm_mfma<64> takes more than the default 64 KiB of dynamic LDS
      static bool attr = false;
      if (!attr)
        {
          if (hipFuncSetAttribute((const void *)k_trsm_mfma<64>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)trsm_mfma_lds(64)) != hipSuccess)
            rc = hdg::set_error("hdg_trace_solver_create: LDS attribute", -3);
          else
            attr = true;
        }
    }
  if (rc == 0 && rocblas_create_handle(&T->h) != rocblas_status_success)
    rc = hdg::set_error("hdg_trace_solver_create: rocblas_create_handle failed", -3);
  if (rc == 0)
    {
      // the main trailing GEMM of the groups with enough work, tuned on the
      // front buffer (its contents are written by each factorisation): the
      // largest groups first, within GEMM_TUNE_BUDGET_S of wall time
      T->gemm_sol.assign(P.groups.size(), 0);
      std::vector<std::pair<double, size_t>> order;
      for (size_t gi = 0; gi < P.groups.size(); ++gi)
        {
          const hdg::NDGroup &G = P.groups[gi];
          const double        S = G.S, I = G.I;
          const double flop = G.batch * (2.0 / 3.0 * S * S * S + 2.0 * S * S * I + 2.0 * S * I * I);
          if (flop >= GEMM_TUNE_MIN_FLOP)
            order.emplace_back(-flop, gi);
        }
      std::sort(order.begin(), order.end());
      const auto t0 = std::chrono::steady_clock::now();
      for (const auto &og : order)
        {
          if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() >
              GEMM_TUNE_BUDGET_S)
            break;
          const size_t        gi = og.second;
          const hdg::NDGroup &G  = P.groups[gi];
          const long long rest = (long long)G.n * G.n * G.batch;
          const int       SB   = std::min((rest >= SP_MIN_ENTRIES ? SP : 1) * T->pw, G.S);
          const int       M    = G.n - SB;
          if (M <= 0)
            continue;
          double *Fg = T->d_F + G.fbase;
          const long long nn = (long long)G.n * G.n;
          T->gemm_sol[gi] = tune_gemm(T->h, M, M, SB, Fg + SB, Fg + (long long)SB * G.n,
                                      Fg + SB + (long long)SB * G.n, G.n, nn, G.batch);
        }
    }
  if (rc != 0)
    {
      hdg_trace_solver_destroy(T);
      return rc;
    }
  *out = T;
  return 0;
}

void hdg_trace_solver_destroy(hdg_trace_solver *T)
{
  if (!T)
    return;
  for (DevGroup &D : T->dg)
    for (void *p : {(void *)D.d_pf, (void *)D.d_pw, (void *)D.d_pn, (void *)D.d_slot,
                    (void *)D.d_pos, (void *)D.d_edof, (void *)D.d_idof, (void *)D.d_cbase,
                    (void *)D.d_cld, (void *)D.d_inv})
      if (p)
        (void)hipFree(p);
  for (void *p : {(void *)T->d_F, (void *)T->d_w, (void *)T->d_part, (void *)T->d_perm,
                  (void *)T->d_info,
                  (void *)T->d_map, (void *)T->d_pad, (void *)T->d_wmap, (void *)T->d_csrc})
    if (p)
      (void)hipFree(p);
  if (T->h)
    rocblas_destroy_handle(T->h);
  delete T;
}

int hdg_trace_solver_info(const hdg_trace_solver *T, double out[HDG_TS_NINFO])
{
  if (!T || !out)
    return hdg::set_error("hdg_trace_solver_info: null argument", -1);
  out[HDG_TS_FLOPS]     = T->P.flops;
  out[HDG_TS_BYTES]     = (double)T->P.Ftot * sizeof(double);
  out[HDG_TS_LEVELS]    = (double)T->P.groups.size();
  out[HDG_TS_MAX_FRONT] = (double)T->P.max_front;
  out[HDG_TS_NODES]     = (double)T->P.nnodes;
  return 0;
}

int hdg_trace_factor(hdg_trace_solver *T, const double *csr_vals, void *stream)
{
  if (!T || !csr_vals)
    return hdg::set_error("hdg_trace_factor: null argument", -1);
  hipStream_t s = (hipStream_t)stream;
  rocblas_set_stream(T->h, s);
  rocblas_set_pointer_mode(T->h, rocblas_pointer_mode_host);
  hipLaunchKernelGGL(k_fill_int, dim3(grid_for(T->P.nnodes)), dim3(TPB), 0, s,
                     T->d_info, T->P.nnodes, 0);
  long long info_off = 0;
  for (size_t gi = 0; gi < T->P.groups.size(); ++gi)
    {
      const hdg::NDGroup &G = T->P.groups[gi];
      const DevGroup     &D = T->dg[gi];
      // the group's fronts: its children's Schur blocks (every child group
      // is factored by now), then its matrix entries and padded diagonals
      for (int y0 = 0; y0 < G.batch; y0 += YCHUNK)
        hipLaunchKernelGGL(k_assemble,
                           dim3(std::max(1, std::min(1024, (G.n + TPB / 64 - 1) / (TPB / 64))),
                                std::min(YCHUNK, G.batch - y0)),
                           dim3(TPB), 0, s, T->d_F, G.fbase, G.n, G.batch, D.d_cbase, D.d_cld,
                           D.d_inv, y0);
      if (D.ncsr > 0 && D.d_cbase)
        hipLaunchKernelGGL(k_scatter_add<true>, dim3(grid_for(D.ncsr)), dim3(TPB), 0, s,
                           csr_vals, T->d_csrc + D.csr0, T->d_map + D.csr0, D.ncsr, T->d_F);
      else if (D.ncsr > 0)
        hipLaunchKernelGGL(k_scatter_add<false>, dim3(grid_for(D.ncsr)), dim3(TPB), 0, s,
                           csr_vals, T->d_csrc + D.csr0, T->d_map + D.csr0, D.ncsr, T->d_F);
      if (D.npad > 0)
        hipLaunchKernelGGL(k_set_one, dim3(grid_for(D.npad)), dim3(TPB), 0, s, T->d_F,
                           T->d_pad + D.pad0, D.npad);
      // the eliminated block in panels of at most pw dofs (whole faces),
      // SP panels to a super-panel: within a super-panel each panel's LU,
      // its triangular solves and a GEMM update of the super-panel's own
      // columns; then the super-panel's U12 for the columns beyond it, and
      // one GEMM of rank up to SP*pw for the trailing front (the rank-pw
      // updates of single panels each re-read the whole trailing front)
      const long long nn = (long long)G.n * G.n;
      auto lu = [&](int q0, int Sp) {
        const long long fb = G.fbase + (long long)q0 * G.n + q0;
        if (G.batch <= TRSM_MFMA_MAX_BATCH)
          {
            if (Sp <= 32)
              hipLaunchKernelGGL(k_panel_lu4<32>, dim3(G.batch), dim3(256), 0, s, T->d_F, fb,
                                 G.n, Sp, T->d_perm, G.pbase + q0, G.S, q0, T->d_info + info_off,
                                 G.batch);
            else
              hipLaunchKernelGGL(k_panel_lu4<PW_MAX>, dim3(G.batch), dim3(256), 0, s, T->d_F, fb,
                                 G.n, Sp, T->d_perm, G.pbase + q0, G.S, q0, T->d_info + info_off,
                                 G.batch);
            return;
          }
        for (int y0 = 0; y0 < G.batch; y0 += YCHUNK)
          {
            const dim3   grid(std::min(YCHUNK, G.batch - y0));
            const size_t lds = (size_t)Sp * Sp * sizeof(double);
            if (Sp <= 16)
              hipLaunchKernelGGL(k_panel_lu<16>, grid, dim3(64), lds, s, T->d_F, fb, G.n, Sp,
                                 T->d_perm, G.pbase + q0, G.S, q0, T->d_info + info_off,
                                 G.batch, y0);
            else if (Sp <= 32)
              hipLaunchKernelGGL(k_panel_lu<32>, grid, dim3(64), lds, s, T->d_F, fb, G.n, Sp,
                                 T->d_perm, G.pbase + q0, G.S, q0, T->d_info + info_off,
                                 G.batch, y0);
            else
              hipLaunchKernelGGL(k_panel_lu<PW_MAX>, grid, dim3(64), lds, s, T->d_F, fb, G.n,
                                 Sp, T->d_perm, G.pbase + q0, G.S, q0, T->d_info + info_off,
                                 G.batch, y0);
          }
      };
      // U12 of panel q0 on Ic columns from column c0 of its corner, L21 on
      // the Ir rows below it
      auto trsm = [&](int q0, int Sp, int Ic, int c0, int Ir) {
        const long long fb  = G.fbase + (long long)q0 * G.n + q0;
        const int       nbc = (Ic + 63) / 64, nbr = (Ir + 63) / 64;
        if (nbc + nbr == 0)
          return;
        if (G.batch <= TRSM_MFMA_MAX_BATCH)
          {
            const dim3 gt(nbc + nbr, G.batch);
            if (Sp <= 16)
              hipLaunchKernelGGL(k_trsm_mfma<16>, gt, dim3(256), trsm_mfma_lds(16), s, T->d_F,
                                 fb, G.n, Sp, Ic, c0, Ir, T->d_perm, G.pbase + q0, G.S, nbc,
                                 G.batch, 0);
            else if (Sp <= 32)
              hipLaunchKernelGGL(k_trsm_mfma<32>, gt, dim3(256), trsm_mfma_lds(32), s, T->d_F,
                                 fb, G.n, Sp, Ic, c0, Ir, T->d_perm, G.pbase + q0, G.S, nbc,
                                 G.batch, 0);
            else if (Sp <= 48)
              hipLaunchKernelGGL(k_trsm_mfma<48>, gt, dim3(256), trsm_mfma_lds(48), s, T->d_F,
                                 fb, G.n, Sp, Ic, c0, Ir, T->d_perm, G.pbase + q0, G.S, nbc,
                                 G.batch, 0);
            else
              hipLaunchKernelGGL(k_trsm_mfma<64>, gt, dim3(256), trsm_mfma_lds(64), s, T->d_F,
                                 fb, G.n, Sp, Ic, c0, Ir, T->d_perm, G.pbase + q0, G.S, nbc,
                                 G.batch, 0);
            return;
          }
        for (int y0 = 0; y0 < G.batch; y0 += YCHUNK)
          hipLaunchKernelGGL(k_panel_trsm, dim3(nbc + nbr, std::min(YCHUNK, G.batch - y0)),
                             dim3(64), (size_t)Sp * Sp * sizeof(double), s, T->d_F, fb, G.n, Sp, Ic, c0, Ir, T->d_perm,
                             G.pbase + q0, G.S, nbc, G.batch, y0);
      };
      // F(r0.., c0..) -= F(r0.., k0..k0+K) F(k0..k0+K, c0..), M x N, every front
      auto gemm = [&](int r0, int c0, int k0, int M, int N, int K) -> bool {
        if (M <= 0 || N <= 0 || K <= 0)
          return true;
        double *Fg = T->d_F + G.fbase;
        return strided_gemm(T->h, T->gemm_sol[gi], M, N, K, Fg + r0 + (long long)k0 * G.n,
                            Fg + k0 + (long long)c0 * G.n, Fg + r0 + (long long)c0 * G.n, G.n,
                            nn, G.batch) == rocblas_status_success;
      };
      bool ok = true;
      for (int o = 0, SBo = 0; o < G.S && ok; o += SBo)
        {
          // a super-panel pays when the trailing fronts are large (each
          // rank-pw pass re-reads them); otherwise its extra launches cost more
          const long long rest = (long long)(G.n - o) * (G.n - o) * G.batch;
          const int SB  = std::min((rest >= SP_MIN_ENTRIES ? SP : 1) * T->pw, G.S - o);
          const int end = o + SB;
          SBo           = SB;
          if (SB <= T->pw)
            {
              // a single panel: both solves over the whole trailing front
              const int Ip = G.n - end;
              lu(o, SB);
              trsm(o, SB, Ip, SB, Ip);
              ok = gemm(end, end, o, Ip, Ip, SB);
              continue;
            }
          for (int q0 = o; q0 < end && ok; q0 += T->pw)
            {
              const int Sp = std::min(T->pw, end - q0);
              lu(q0, Sp);
              trsm(q0, Sp, end - q0 - Sp, Sp, G.n - q0 - Sp);
              ok = gemm(q0 + Sp, q0 + Sp, q0, G.n - q0 - Sp, end - q0 - Sp, Sp);
            }
          const int Icols = G.n - end;
          for (int q0 = o; q0 < end && ok && Icols > 0; q0 += T->pw)
            {
              const int Sp = std::min(T->pw, end - q0);
              trsm(q0, Sp, Icols, end - q0, 0);
              ok = gemm(q0 + Sp, end, q0, end - q0 - Sp, Icols, Sp);
            }
          if (ok)
            ok = gemm(end, end, o, G.n - end, Icols, SB);
        }
      if (!ok)
        return hdg::set_error("hdg_trace_factor: rocblas gemm failed", -3);
      info_off += G.batch;
    }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : hdg::set_error(std::string("hdg_trace_factor: ") + hipGetErrorString(e), -3);
}

int hdg_trace_factor_status(hdg_trace_solver *T, void *stream)
{
  if (!T)
    return hdg::set_error("hdg_trace_factor_status: null argument", -1);
  std::vector<int> info((size_t)T->P.nnodes);
  hipError_t e = hipMemcpyAsync(info.data(), T->d_info, info.size() * sizeof(int),
                                hipMemcpyDeviceToHost, (hipStream_t)stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess)
    return hdg::set_error(std::string("hdg_trace_factor_status: ") + hipGetErrorString(e), -3);
  size_t k = 0;
  for (size_t gi = 0; gi < T->P.groups.size(); ++gi)
    for (int b = 0; b < T->P.groups[gi].batch; ++b, ++k)
      if (info[k] != 0)
        return hdg::set_error("hdg_trace_factor: exactly zero pivot at level " +
                                std::to_string(gi) + ", front " + std::to_string(b) +
                                ", column " + std::to_string(info[k]),
                              -5);
  return 0;
}

int hdg_trace_solve(hdg_trace_solver *T, const double *rhs, double *x, void *stream)
{
  if (!T || !rhs || !x)
    return hdg::set_error("hdg_trace_solve: null argument", -1);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(T->P.wtot)), dim3(TPB), 0, s, T->d_w, T->P.wtot, 0.0);
  hipLaunchKernelGGL(k_gather_rhs, dim3(grid_for(T->P.ndof)), dim3(TPB), 0, s, rhs, T->d_wmap,
                     T->P.ndof, T->d_w);
  // forward: z = L^-1 P b, update vectors to the parents
  for (size_t gi = 0; gi < T->P.groups.size(); ++gi)
    {
      const hdg::NDGroup &G = T->P.groups[gi];
      const DevGroup     &D = T->dg[gi];
      // the panels in order: z_p = L_p^-1 P_p w_p, then the rest of the
      // trailing front's vector -= L21_p z_p (in the same one-wave kernel
      // while that rest is short, else by a batched rocBLAS gemv)
      for (int o = 0; o < G.S; o += T->pw)
        {
          const int       Sp   = std::min(T->pw, G.S - o), Ip = G.n - o - Sp;
          const bool      fuse = Ip <= FUSE_I;
          const long long fb   = G.fbase + (long long)o * G.n + o;
          for (int y0 = 0; y0 < G.batch; y0 += YCHUNK)
            hipLaunchKernelGGL(k_fwd_small, dim3(std::min(YCHUNK, G.batch - y0)), dim3(SWEEP_T),
                               (size_t)Sp * Sp * sizeof(double), s, T->d_F, T->d_w, fb,
                               G.wbase + o, G.n, Sp, fuse ? Ip : 0, T->d_perm, G.pbase + o, G.S,
                               G.batch, y0);
          if (!fuse)
            for (int y0 = 0; y0 < G.batch; y0 += YCHUNK)
              {
                const dim3 grid((Ip + 255) / 256, std::min(YCHUNK, G.batch - y0));
                if (Sp <= 16)
                  hipLaunchKernelGGL(k_sweep_rows<16>, grid, dim3(256), 0, s, T->d_F, T->d_w, fb,
                                     G.wbase + o, G.n, Sp, Ip, G.batch, y0);
                else if (Sp <= 32)
                  hipLaunchKernelGGL(k_sweep_rows<32>, grid, dim3(256), 0, s, T->d_F, T->d_w, fb,
                                     G.wbase + o, G.n, Sp, Ip, G.batch, y0);
                else
                  hipLaunchKernelGGL(k_sweep_rows<PW_MAX>, grid, dim3(256), 0, s, T->d_F, T->d_w,
                                     fb, G.wbase + o, G.n, Sp, Ip, G.batch, y0);
              }
        }
      if (G.I > 0)
        for (int want = 0; want < 2; ++want)
          for (int y0 = 0; y0 < G.batch; y0 += YCHUNK)
            hipLaunchKernelGGL(k_vec_extend_add,
                               dim3((G.I + TPB - 1) / TPB, std::min(YCHUNK, G.batch - y0)),
                               dim3(TPB), 0, s, T->d_w, G.wbase, G.n, G.S, G.I, G.batch, D.d_pw,
                               D.d_pf, D.d_slot, D.d_pos, want, y0);
    }
  // backward: x_s = U11^-1 (z - U12 x_i), root first
  for (int gi = (int)T->P.groups.size() - 1; gi >= 0; --gi)
    {
      const hdg::NDGroup &G = T->P.groups[gi];
      const DevGroup     &D = T->dg[gi];
      if (G.I > 0)
        for (int y0 = 0; y0 < G.batch; y0 += YCHUNK)
          hipLaunchKernelGGL(k_gather_x,
                             dim3((G.I + TPB - 1) / TPB, std::min(YCHUNK, G.batch - y0)),
                             dim3(TPB), 0, s, x, T->d_w, G.wbase, G.n, G.S, G.I, G.batch,
                             D.d_idof, y0);
      if (G.S > 0)
        {
          // the panels in reverse order: x_p = U_p^-1 (z_p - U12_p x_rest)
          // (U12_p x_rest in the same kernel while the rest is short)
          const int np = (G.S + T->pw - 1) / T->pw;
          for (int pi = np - 1; pi >= 0; --pi)
            {
              const int       o    = pi * T->pw;
              const int       Sp   = std::min(T->pw, G.S - o), Ip = G.n - o - Sp;
              const bool      fuse = Ip <= FUSE_I;
              const long long fb   = G.fbase + (long long)o * G.n + o;
              if (!fuse)
                for (int y0 = 0; y0 < G.batch; y0 += YCHUNK)
                  {
                    const int nb  = std::min(YCHUNK, G.batch - y0);
                    const int nch = (Ip + 255) / 256;
                    hipLaunchKernelGGL(k_sweep_cols_part, dim3(nch, nb), dim3(256), 0, s, T->d_F,
                                       T->d_w, fb, G.wbase + o, G.n, Sp, Ip, nch, T->d_part,
                                       G.batch, y0);
                    hipLaunchKernelGGL(k_sweep_cols_sum, dim3(nb), dim3(64), 0, s, T->d_w,
                                       G.wbase + o, G.n, Sp, nch, T->d_part, G.batch, y0);
                  }
              for (int y0 = 0; y0 < G.batch; y0 += YCHUNK)
                hipLaunchKernelGGL(k_bwd_small, dim3(std::min(YCHUNK, G.batch - y0)),
                                   dim3(SWEEP_T), (size_t)(Sp * Sp + 4 * Sp) * sizeof(double), s, T->d_F, T->d_w, fb,
                                   G.wbase + o, G.n, Sp, fuse ? Ip : 0, G.batch, y0);
            }
          for (int y0 = 0; y0 < G.batch; y0 += YCHUNK)
            hipLaunchKernelGGL(k_scatter_x,
                               dim3((G.S + TPB - 1) / TPB, std::min(YCHUNK, G.batch - y0)),
                               dim3(TPB), 0, s, x, T->d_w, G.wbase, G.n, G.S, G.batch, D.d_edof,
                               y0);
        }
    }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : hdg::set_error(std::string("hdg_trace_solve: ") + hipGetErrorString(e), -3);
}

} // extern "C"
