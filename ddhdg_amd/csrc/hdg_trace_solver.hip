// hdg_trace_solver.hip -- the global trace solve on the GPU: a nested-
// dissection multifrontal LU of the condensed trace system.
//
// The reference factors the restricted trace matrix with SparseDirectUMFPACK
// and solves once per Newton step (src/np_solver.cpp:1326-1354).  Here the
// plan of hdg_nd_plan.h (tree, fronts, assembly and extend-add maps) is built
// once per mesh and component set; each factorisation then runs level by
// level: one strided-batched rocSOLVER getrf of every front's F11 (partial
// pivoting inside the eliminated block), its row interchanges applied to F12
// in HIP, two rocBLAS trsm (U12 = L11^-1 P F12, L21 = F21 U11^-1), one gemm
// (F22 -= L21 U12), and an extend-add of the
// update blocks into the parents written in HIP, one launch per child slot so
// that no two writes meet (deterministic sums, no atomics).  The fronts stay
// in HBM as the factors for the forward and backward sweeps.
#include <algorithm>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "hdg_ctx.h"
#include "hdg_nd_plan.h"
#include "hdg_wave.h"

namespace
{
  using hdg::lane_value;
  using hdg::leading_bits;
  using hdg::wave_max;

  struct DevGroup
  {
    long long *d_pf = nullptr, *d_pw = nullptr, *d_edof = nullptr, *d_idof = nullptr;
    int       *d_pn = nullptr, *d_slot = nullptr, *d_pos = nullptr;
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

  __global__ void k_scatter(const double *__restrict__ vals, const long long *__restrict__ map,
                            long long nnz, double *__restrict__ F)
  {
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < nnz;
         k += (long long)gridDim.x * blockDim.x)
      F[map[k]] = vals[k];
  }

  __global__ void k_set_one(double *F, const long long *__restrict__ idx, long long n)
  {
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x)
      F[idx[k]] = 1.0;
  }

  // F22 of each node of one group (child slot want) added into its parent
  __global__ void k_extend_add(double *F, long long fbase, int n, int S, int I, int batch,
                               const long long *__restrict__ pf, const int *__restrict__ pn,
                               const int *__restrict__ slot, const int *__restrict__ pos,
                               int want, int yoff)
  {
    const int b = blockIdx.y + yoff;
    if (b >= batch || slot[b] != want || pf[b] < 0)
      return;
    const double *U  = F + fbase + (long long)b * n * n + S + (long long)S * n;
    double       *P  = F + pf[b];
    const long long np = pn[b];
    const int    *ps = pos + (long long)b * I;
    const long long II = (long long)I * I;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < II;
         e += (long long)gridDim.x * blockDim.x)
      {
        const int r = (int)(e % I), c = (int)(e / I);
        const int pr = ps[r], pc = ps[c];
        if (pr >= 0 && pc >= 0)
          P[pr + pc * np] += U[r + (long long)c * n];
      }
  }

  // getrf's row interchanges applied to F12 (columns S..n-1) of every front
  // of a group; one thread per column, the swaps in order
  __global__ void k_laswp_cols(double *F, long long fbase, int n, int S, int I,
                               const int *__restrict__ ipiv, long long pbase, int batch, int yoff)
  {
    const int b = blockIdx.y + yoff;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch || c >= I)
      return;
    double    *col = F + fbase + (long long)b * n * n + (long long)(S + c) * n;
    const int *pv  = ipiv + pbase + (long long)b * S;
    for (int j = 0; j < S; ++j)
      {
        const int p = pv[j] - 1;
        if (p != j)
          {
            const double t = col[j];
            col[j]         = col[p];
            col[p]         = t;
          }
      }
  }

  // Batched triangular solves for fronts with a small eliminated block
  // (S <= TRSM_SMAX), one launch per level.  A lane owns one row (right solve)
  // or one column (left solve) of the off-diagonal block; the values it has
  // solved stay in LDS.  The triangular factor is read at wave-uniform
  // addresses through the constant address space (scalar loads into
  // SGPRs): this kernel writes only the off-diagonal block, never the
  // factor it reads.
  constexpr int TRSM_SMAX = 160;
  typedef const __attribute__((address_space(4))) double cdouble;

  // B := B U^-1, B = F21 (I x S, rows S.. of columns 0..S-1), U = upper(F11)
  __global__ void __launch_bounds__(64)
    k_trsm_right_upper(double *F, long long fbase, int n, int S, int I, int batch, int yoff)
  {
    extern __shared__ double xs[]; // S x 64
    const int b = blockIdx.y + yoff;
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (b >= batch)
      return;
    cdouble   *U   = (cdouble *)(F + fbase + (long long)b * n * n);
    double    *B   = F + fbase + (long long)b * n * n + S;
    const bool act = i < I;
    for (int j = 0; j < S; ++j)
      {
        cdouble *Uj = U + (long long)j * n;
        // four partial sums: the FMA chain is the latency bound
        double a0 = act ? B[i + (long long)j * n] : 0.0, a1 = 0, a2 = 0, a3 = 0;
        int    l  = 0;
        for (; l + 3 < j; l += 4)
          {
            a0 -= xs[l * 64 + threadIdx.x] * Uj[l];
            a1 -= xs[(l + 1) * 64 + threadIdx.x] * Uj[l + 1];
            a2 -= xs[(l + 2) * 64 + threadIdx.x] * Uj[l + 2];
            a3 -= xs[(l + 3) * 64 + threadIdx.x] * Uj[l + 3];
          }
        for (; l < j; ++l)
          a0 -= xs[l * 64 + threadIdx.x] * Uj[l];
        const double acc            = (a0 + a1) + (a2 + a3);
        const double x              = acc / Uj[j];
        xs[j * 64 + threadIdx.x]    = x;
        if (act)
          B[i + (long long)j * n] = x;
      }
  }

  // B := L^-1 B, B = F12 (S x I, rows 0..S-1 of columns S..), L = unit lower(F11)
  __global__ void __launch_bounds__(64)
    k_trsm_left_lower_unit(double *F, long long fbase, int n, int S, int I, int batch, int yoff)
  {
    extern __shared__ double ys[]; // S x 64
    const int b = blockIdx.y + yoff;
    const int c = blockIdx.x * 64 + threadIdx.x;
    if (b >= batch)
      return;
    cdouble      *L   = (cdouble *)(F + fbase + (long long)b * n * n);
    double       *B   = F + fbase + (long long)b * n * n + (long long)S * n;
    const bool    act = c < I;
    double       *Bc  = B + (long long)(act ? c : 0) * n;
    for (int j = 0; j < S; ++j)
      {
        double a0 = act ? Bc[j] : 0.0, a1 = 0, a2 = 0, a3 = 0;
        int    l  = 0;
        for (; l + 3 < j; l += 4)
          {
            a0 -= L[j + (long long)l * n] * ys[l * 64 + threadIdx.x];
            a1 -= L[j + (long long)(l + 1) * n] * ys[(l + 1) * 64 + threadIdx.x];
            a2 -= L[j + (long long)(l + 2) * n] * ys[(l + 2) * 64 + threadIdx.x];
            a3 -= L[j + (long long)(l + 3) * n] * ys[(l + 3) * 64 + threadIdx.x];
          }
        for (; l < j; ++l)
          a0 -= L[j + (long long)l * n] * ys[l * 64 + threadIdx.x];
        const double acc         = (a0 + a1) + (a2 + a3);
        ys[j * 64 + threadIdx.x] = acc;
        if (act)
          Bc[j] = acc;
      }
  }

  // ---------------------------------------------------------------------
  // Fronts with a small eliminated block (S <= SMALL_S): F11 = P L11 U11
  // (partial pivoting) in one workgroup per front, F11 in LDS, ipiv in
  // LAPACK's 1-based form.  The off-diagonal blocks and the Schur update
  // follow as for the other fronts.
  constexpr int SMALL_S = 72;
  constexpr int FT      = 256; // threads of the small-front kernels

  __global__ void __launch_bounds__(FT)
    k_front_lu_small(double *F, long long fbase, int n, int S, int I, int *ipiv, int *perm,
                     long long pbase, int *info, int batch, int yoff)
  {
    __shared__ double A[SMALL_S * SMALL_S]; // F11, column-major, ld S
    __shared__ int    piv[SMALL_S], pm[SMALL_S];
    __shared__ int    pk;
    const int b = blockIdx.x + yoff;
    if (b >= batch)
      return;
    const int t  = threadIdx.x;
    double   *Fb = F + fbase + (long long)b * n * n;
    for (int e = t; e < S * S; e += FT)
      A[e] = Fb[(e % S) + (long long)(e / S) * n];
    for (int k = t; k < S; k += FT)
      pm[k] = k;
    __syncthreads();
    int bad = 0;
    for (int k = 0; k < S; ++k)
      {
        // pivot: largest |A(i,k)|, i >= k, to the leading bits of the
        // value (lowest row on ties), found by wave 0 with a DPP max
        if (t < 64)
          {
            unsigned key = 0;
            for (int i = k + t; i < S; i += 64)
              {
                const unsigned c = (leading_bits(fabs(A[i + k * S])) & ~127u) | (unsigned)(127 - i);
                key              = c > key ? c : key;
              }
            key = wave_max(key);
            if (t == 0)
              pk = 127 - (int)(key & 127u);
          }
        __syncthreads();
        const int p = pk;
        if (t == 0)
          {
            piv[k] = p;
            if (p != k)
              {
                const int tmp = pm[k];
                pm[k]         = pm[p];
                pm[p]         = tmp;
              }
          }
        if (p != k)
          for (int j = t; j < S; j += FT)
            {
              const double tmp = A[k + j * S];
              A[k + j * S]     = A[p + j * S];
              A[p + j * S]     = tmp;
            }
        __syncthreads();
        const double d = A[k + k * S];
        if (t == 0 && d == 0.0 && !bad)
          bad = k + 1;
        for (int i = k + 1 + t; i < S; i += FT)
          A[i + k * S] = d != 0.0 ? A[i + k * S] / d : 0.0;
        __syncthreads();
        const int m = S - k - 1;
        for (int e = t; e < m * m; e += FT)
          {
            const int i = k + 1 + e % m, j = k + 1 + e / m;
            A[i + j * S] -= A[i + k * S] * A[k + j * S];
          }
        __syncthreads();
      }
    if (t == 0)
      info[b] = bad;
    for (int e = t; e < S * S; e += FT)
      Fb[(e % S) + (long long)(e / S) * n] = A[e];
    for (int k = t; k < S; k += FT)
      {
        ipiv[pbase + (long long)b * S + k] = piv[k] + 1;
        perm[pbase + (long long)b * S + k] = pm[k];
      }
  }

  // forward sweep of one small front, one wave: z = L11^-1 P w_s with L11
  // staged in LDS and z in registers (rows lane and lane + 64; each
  // solved entry reaches the other lanes by readlane), then w_i -= L21 z.
  // perm: row k of the factored block is row perm[k] of the front.
  __global__ void __launch_bounds__(64)
    k_fwd_small(const double *F, double *w, long long fbase, long long wbase, int n, int S,
                int I, const int *perm, long long pbase, int batch, int yoff)
  {
    extern __shared__ double L[]; // S x S, column-major
    const int b = blockIdx.x + yoff;
    if (b >= batch)
      return;
    const int     t  = threadIdx.x;
    const double *Fb = F + fbase + (long long)b * n * n;
    double       *wb = w + wbase + (long long)b * n;
    const int    *pv = perm + pbase + (long long)b * S;
    for (int e = t; e < S * S; e += 64)
      L[e] = Fb[(e % S) + (long long)(e / S) * n];
    const int i0 = t, i1 = t + 64;
    double    z0 = i0 < S ? wb[pv[i0]] : 0.0;
    double    z1 = i1 < S ? wb[pv[i1]] : 0.0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int r = t; r < I; r += 64)
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

  // backward sweep of one small front, one wave: z = w_s - U12 x_i, then
  // U11^-1 z with U11 staged in LDS and z in registers
  __global__ void __launch_bounds__(64)
    k_bwd_small(const double *F, double *w, long long fbase, long long wbase, int n, int S,
                int I, int batch, int yoff)
  {
    extern __shared__ double U[]; // S x S, column-major
    const int b = blockIdx.x + yoff;
    if (b >= batch)
      return;
    const int     t  = threadIdx.x;
    const double *Fb = F + fbase + (long long)b * n * n;
    double       *wb = w + wbase + (long long)b * n;
    for (int e = t; e < S * S; e += 64)
      U[e] = Fb[(e % S) + (long long)(e / S) * n];
    const int i0 = t, i1 = t + 64;
    double    z0 = 0.0, z1 = 0.0;
    for (int h = 0; h < 2; ++h)
      {
        const int i = h ? i1 : i0;
        if (i >= S)
          continue;
        double a0 = wb[i], a1 = 0, a2 = 0, a3 = 0;
        int    c  = 0;
        for (; c + 3 < I; c += 4)
          {
            a0 -= Fb[i + (long long)(S + c) * n] * wb[S + c];
            a1 -= Fb[i + (long long)(S + c + 1) * n] * wb[S + c + 1];
            a2 -= Fb[i + (long long)(S + c + 2) * n] * wb[S + c + 2];
            a3 -= Fb[i + (long long)(S + c + 3) * n] * wb[S + c + 3];
          }
        for (; c < I; ++c)
          a0 -= Fb[i + (long long)(S + c) * n] * wb[S + c];
        (h ? z1 : z0) = (a0 + a1) + (a2 + a3);
      }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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

  __global__ void k_gather_rhs(const double *__restrict__ b, const long long *__restrict__ wmap,
                               long long ndof, double *__restrict__ w)
  {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < ndof;
         i += (long long)gridDim.x * blockDim.x)
      w[wmap[i]] = b[i];
  }

  // row interchanges of getrf applied to the eliminated part of each work vector
  __global__ void k_laswp(double *w, long long wbase, int n, int S, const int *__restrict__ ipiv,
                          long long pbase, int batch)
  {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch)
      return;
    double    *x  = w + wbase + (long long)b * n;
    const int *pv = ipiv + pbase + (long long)b * S;
    for (int j = 0; j < S; ++j)
      {
        const int p = pv[j] - 1;
        if (p != j)
          {
            const double t = x[j];
            x[j]           = x[p];
            x[p]           = t;
          }
      }
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

  constexpr int YCHUNK = 32768; // nodes per launch along grid y

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
  int                  *d_ipiv = nullptr, *d_info = nullptr;
  int                  *d_perm = nullptr; // row order of the small fronts' factored blocks
  long long            *d_map = nullptr, *d_pad = nullptr, *d_wmap = nullptr;
};

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
    rc = upload(&T->d_map, P.map, "trace solver map");
  if (rc == 0)
    rc = upload(&T->d_pad, P.pad, "trace solver padding");
  if (rc == 0)
    rc = upload(&T->d_wmap, P.wmap, "trace solver rhs map");
  if (rc == 0)
    {
      hipError_t e = hipMalloc((void **)&T->d_F, (size_t)P.Ftot * sizeof(double));
      if (e == hipSuccess)
        e = hipMalloc((void **)&T->d_w, (size_t)std::max<long long>(1, P.wtot) * sizeof(double));
      if (e == hipSuccess)
        e = hipMalloc((void **)&T->d_ipiv, (size_t)std::max<long long>(1, P.ptot) * sizeof(int));
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
    {
      const int  lds = TRSM_SMAX * 64 * (int)sizeof(double);
      hipError_t e   = hipFuncSetAttribute((const void *)k_trsm_right_upper,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);
      if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_trsm_left_lower_unit,
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds);
      if (e != hipSuccess)
        rc = hdg::set_error(std::string("hdg_trace_solver_create: ") + hipGetErrorString(e), -3);
    }
  if (rc == 0 && rocblas_create_handle(&T->h) != rocblas_status_success)
    rc = hdg::set_error("hdg_trace_solver_create: rocblas_create_handle failed", -3);
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
                    (void *)D.d_pos, (void *)D.d_edof, (void *)D.d_idof})
      if (p)
        (void)hipFree(p);
  for (void *p : {(void *)T->d_F, (void *)T->d_w, (void *)T->d_ipiv, (void *)T->d_perm,
                  (void *)T->d_info,
                  (void *)T->d_map, (void *)T->d_pad, (void *)T->d_wmap})
    if (p)
      hipFree(p);
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
  hipLaunchKernelGGL(k_fill, dim3(grid_for(T->P.Ftot)), dim3(TPB), 0, s, T->d_F, T->P.Ftot, 0.0);
  hipLaunchKernelGGL(k_fill_int, dim3(grid_for(T->P.nnodes)), dim3(TPB), 0, s,
                     T->d_info, T->P.nnodes, 0);
  hipLaunchKernelGGL(k_scatter, dim3(grid_for(T->P.nnz)), dim3(TPB), 0, s, csr_vals, T->d_map,
                     T->P.nnz, T->d_F);
  if ((long long)T->P.pad.size() > 0)
    hipLaunchKernelGGL(k_set_one, dim3(grid_for((long long)T->P.pad.size())), dim3(TPB), 0, s, T->d_F, T->d_pad,
                       (long long)T->P.pad.size());
  const double one = 1.0, mone = -1.0;
  long long    info_off = 0;
  for (size_t gi = 0; gi < T->P.groups.size(); ++gi)
    {
      const hdg::NDGroup &G = T->P.groups[gi];
      const DevGroup     &D = T->dg[gi];
      double             *F = T->d_F + G.fbase;
      if (G.S > 0)
        {
          // F11 = P L11 U11 (partial pivoting inside the eliminated block)
          rocblas_status st = rocblas_status_success;
          if (G.S <= SMALL_S)
            for (int y0 = 0; y0 < G.batch; y0 += YCHUNK)
              hipLaunchKernelGGL(k_front_lu_small, dim3(std::min(YCHUNK, G.batch - y0)),
                                 dim3(FT), 0, s, T->d_F, G.fbase, G.n, G.S, G.I, T->d_ipiv,
                                 T->d_perm, G.pbase, T->d_info + info_off, G.batch, y0);
          else
            st = rocsolver_dgetrf_strided_batched(
              T->h, G.S, G.S, F, G.n, (rocblas_stride)G.n * G.n, T->d_ipiv + G.pbase,
              (rocblas_stride)G.S, T->d_info + info_off, G.batch);
          if (st != rocblas_status_success)
            return hdg::set_error("hdg_trace_factor: rocsolver getrf failed", -3);
          if (G.I > 0)
            {
              // U12 = L11^-1 P F12, L21 = F21 U11^-1, F22 -= L21 U12
              for (int y0 = 0; y0 < G.batch; y0 += YCHUNK)
                hipLaunchKernelGGL(k_laswp_cols,
                                   dim3((G.I + 63) / 64, std::min(YCHUNK, G.batch - y0)),
                                   dim3(64), 0, s, T->d_F, G.fbase, G.n, G.S, G.I, T->d_ipiv,
                                   G.pbase, G.batch, y0);
              if (G.S <= TRSM_SMAX)
                {
                  // rocBLAS's strided-batched trsm above 64 loops over the
                  // batch; these are one launch per level
                  const size_t lds = (size_t)G.S * 64 * sizeof(double);
                  for (int y0 = 0; y0 < G.batch; y0 += YCHUNK)
                    {
                      const dim3 grid((G.I + 63) / 64, std::min(YCHUNK, G.batch - y0));
                      hipLaunchKernelGGL(k_trsm_left_lower_unit, grid, dim3(64), lds, s, T->d_F,
                                         G.fbase, G.n, G.S, G.I, G.batch, y0);
                      hipLaunchKernelGGL(k_trsm_right_upper, grid, dim3(64), lds, s, T->d_F,
                                         G.fbase, G.n, G.S, G.I, G.batch, y0);
                    }
                }
              else
                {
                  st = rocblas_dtrsm_strided_batched(
                    T->h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                    rocblas_diagonal_unit, G.S, G.I, &one, F, G.n, (rocblas_stride)G.n * G.n,
                    F + (long long)G.S * G.n, G.n, (rocblas_stride)G.n * G.n, G.batch);
                  if (st == rocblas_status_success)
                    st = rocblas_dtrsm_strided_batched(
                      T->h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                      rocblas_diagonal_non_unit, G.I, G.S, &one, F, G.n,
                      (rocblas_stride)G.n * G.n, F + G.S, G.n, (rocblas_stride)G.n * G.n,
                      G.batch);
                }
              if (st == rocblas_status_success)
                st = rocblas_dgemm_strided_batched(
                  T->h, rocblas_operation_none, rocblas_operation_none, G.I, G.I, G.S, &mone,
                  F + G.S, G.n, (rocblas_stride)G.n * G.n, F + (long long)G.S * G.n, G.n,
                  (rocblas_stride)G.n * G.n, &one, F + G.S + (long long)G.S * G.n, G.n,
                  (rocblas_stride)G.n * G.n, G.batch);
              if (st != rocblas_status_success)
                return hdg::set_error("hdg_trace_factor: rocblas trsm/gemm failed", -3);
            }
        }
      info_off += G.batch;
      if (G.I > 0)
        for (int want = 0; want < 2; ++want)
          for (int y0 = 0; y0 < G.batch; y0 += YCHUNK)
            hipLaunchKernelGGL(
              k_extend_add,
              dim3(std::max(1, std::min(4096, (int)(((long long)G.I * G.I + TPB - 1) / TPB))),
                   std::min(YCHUNK, G.batch - y0)),
              dim3(TPB), 0, s, T->d_F, G.fbase, G.n, G.S, G.I, G.batch, D.d_pf, D.d_pn,
              D.d_slot, D.d_pos, want, y0);
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
  rocblas_set_stream(T->h, s);
  rocblas_set_pointer_mode(T->h, rocblas_pointer_mode_host);
  const double one = 1.0, mone = -1.0;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(T->P.wtot)), dim3(TPB), 0, s, T->d_w, T->P.wtot, 0.0);
  hipLaunchKernelGGL(k_gather_rhs, dim3(grid_for(T->P.ndof)), dim3(TPB), 0, s, rhs, T->d_wmap,
                     T->P.ndof, T->d_w);
  // forward: z = L^-1 P b, update vectors to the parents
  for (size_t gi = 0; gi < T->P.groups.size(); ++gi)
    {
      const hdg::NDGroup &G = T->P.groups[gi];
      const DevGroup     &D = T->dg[gi];
      const double       *F = T->d_F + G.fbase;
      double       *w = T->d_w + G.wbase;
      if (G.S > 0 && G.S <= SMALL_S)
        {
          for (int y0 = 0; y0 < G.batch; y0 += YCHUNK)
            hipLaunchKernelGGL(k_fwd_small, dim3(std::min(YCHUNK, G.batch - y0)), dim3(64),
                               (size_t)G.S * G.S * sizeof(double), s, T->d_F, T->d_w, G.fbase,
                               G.wbase, G.n, G.S, G.I, T->d_perm, G.pbase, G.batch, y0);
        }
      else if (G.S > 0)
        {
          hipLaunchKernelGGL(k_laswp, dim3((G.batch + 63) / 64), dim3(64), 0, s, T->d_w,
                             G.wbase, G.n, G.S, T->d_ipiv, G.pbase, G.batch);
          rocblas_status st = rocblas_dtrsv_strided_batched(
            T->h, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_unit, G.S, F, G.n,
            (rocblas_stride)G.n * G.n, w, 1, (rocblas_stride)G.n, G.batch);
          if (st == rocblas_status_success && G.I > 0)
            st = rocblas_dgemv_strided_batched(T->h, rocblas_operation_none, G.I, G.S, &mone,
                                               F + G.S, G.n, (rocblas_stride)G.n * G.n, w, 1,
                                               (rocblas_stride)G.n, &one, w + G.S, 1,
                                               (rocblas_stride)G.n, G.batch);
          if (st != rocblas_status_success)
            return hdg::set_error("hdg_trace_solve: rocblas forward sweep failed", -3);
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
      const double *F = T->d_F + G.fbase;
      double       *w = T->d_w + G.wbase;
      if (G.I > 0)
        for (int y0 = 0; y0 < G.batch; y0 += YCHUNK)
          hipLaunchKernelGGL(k_gather_x,
                             dim3((G.I + TPB - 1) / TPB, std::min(YCHUNK, G.batch - y0)),
                             dim3(TPB), 0, s, x, T->d_w, G.wbase, G.n, G.S, G.I, G.batch,
                             D.d_idof, y0);
      if (G.S > 0 && G.S <= SMALL_S)
        {
          for (int y0 = 0; y0 < G.batch; y0 += YCHUNK)
            hipLaunchKernelGGL(k_bwd_small, dim3(std::min(YCHUNK, G.batch - y0)), dim3(64),
                               (size_t)G.S * G.S * sizeof(double), s, T->d_F, T->d_w, G.fbase,
                               G.wbase, G.n, G.S, G.I, G.batch, y0);
          for (int y0 = 0; y0 < G.batch; y0 += YCHUNK)
            hipLaunchKernelGGL(k_scatter_x,
                               dim3((G.S + TPB - 1) / TPB, std::min(YCHUNK, G.batch - y0)),
                               dim3(TPB), 0, s, x, T->d_w, G.wbase, G.n, G.S, G.batch, D.d_edof,
                               y0);
        }
      else if (G.S > 0)
        {
          rocblas_status st = rocblas_status_success;
          if (G.I > 0)
            st = rocblas_dgemv_strided_batched(T->h, rocblas_operation_none, G.S, G.I, &mone,
                                               F + (long long)G.S * G.n, G.n,
                                               (rocblas_stride)G.n * G.n, w + G.S, 1,
                                               (rocblas_stride)G.n, &one, w, 1,
                                               (rocblas_stride)G.n, G.batch);
          if (st == rocblas_status_success)
            st = rocblas_dtrsv_strided_batched(T->h, rocblas_fill_upper, rocblas_operation_none,
                                               rocblas_diagonal_non_unit, G.S, F, G.n,
                                               (rocblas_stride)G.n * G.n, w, 1,
                                               (rocblas_stride)G.n, G.batch);
          if (st != rocblas_status_success)
            return hdg::set_error("hdg_trace_solve: rocblas backward sweep failed", -3);
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
