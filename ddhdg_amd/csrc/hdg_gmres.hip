// hdg_gmres.hip -- restarted GMRES without preconditioner on a device CSR
// matrix: the reference's iterative_linear_solver branch of
// solve_linear_problem (src/np_solver.cpp:1326-1354: SolverGMRES with
// PreconditionIdentity, SolverControl(10 m, 1e-10 |rhs|), zero initial
// guess).  deal.II's default GMRES keeps 30 temporary vectors, i.e. restarts
// every 28 iterations, left preconditioned (here: none), and stops on the
// residual estimate of the Arnoldi process.
//
// Device kernels: the sparse product (one wave per row), batched dot
// products against the Krylov basis with a fixed-order two-stage reduction
// (deterministic), basis updates.  The Arnoldi vectors are orthogonalised
// by classical Gram-Schmidt applied twice; the small Hessenberg least-squares
// problem (Givens rotations) lives on the host.
#include <cmath>
#include <string>
#include <vector>

#include "hdg_kernels.h"
#include "hdg_wave.h"

namespace hdg
{
  namespace
  {
    constexpr int GT = 256, GB = 512; // threads per block, blocks of the dot stage 1

    __global__ void __launch_bounds__(GT)
      k_spmv(const long long *__restrict__ rp, const int *__restrict__ ci,
             const double *__restrict__ v, long long n, const double *__restrict__ x,
             double *__restrict__ y)
    {
      const long long row = (long long)blockIdx.x * (GT / 64) + (threadIdx.x >> 6);
      const int       lane = threadIdx.x & 63;
      if (row >= n)
        return;
      double s = 0;
      for (long long k = rp[row] + lane; k < rp[row + 1]; k += 64)
        s += v[k] * x[ci[k]];
      // fixed-order wave sum
      for (int off = 32; off > 0; off >>= 1)
        s += __shfl_down(s, off, 64);
      if (lane == 0)
        y[row] = s;
    }

    // y = b - y
    __global__ void __launch_bounds__(GT) k_bminus(const double *b, double *y, long long n)
    {
      const long long i = (long long)blockIdx.x * GT + threadIdx.x;
      if (i < n)
        y[i] = b[i] - y[i];
    }

    // part[j * GB + block] = sum over the block's chunk of V_j . w, j < m
    __global__ void __launch_bounds__(GT)
      k_dots1(const double *V, long long ld, int m, const double *w, long long n, double *part)
    {
      __shared__ double red[GT];
      for (int j = 0; j < m; ++j)
        {
          const double *vj = V + (size_t)j * ld;
          double        s  = 0;
          for (long long i = (long long)blockIdx.x * GT + threadIdx.x; i < n;
               i += (long long)gridDim.x * GT)
            s += vj[i] * w[i];
          red[threadIdx.x] = s;
          __syncthreads();
          for (int off = GT / 2; off > 0; off >>= 1)
            {
              if ((int)threadIdx.x < off)
                red[threadIdx.x] += red[threadIdx.x + off];
              __syncthreads();
            }
          if (threadIdx.x == 0)
            part[j * GB + blockIdx.x] = red[0];
          __syncthreads();
        }
    }

    // out[j] = sum of the GB partials of dot j (one block per dot)
    __global__ void __launch_bounds__(GT) k_dots2(const double *part, double *out)
    {
      __shared__ double red[GT];
      double            s = 0;
      for (int b = threadIdx.x; b < GB; b += GT)
        s += part[blockIdx.x * GB + b];
      red[threadIdx.x] = s;
      __syncthreads();
      for (int off = GT / 2; off > 0; off >>= 1)
        {
          if ((int)threadIdx.x < off)
            red[threadIdx.x] += red[threadIdx.x + off];
          __syncthreads();
        }
      if (threadIdx.x == 0)
        out[blockIdx.x] = red[0];
    }

    // w -= sum_j h[j] V_j   (h on the device)
    __global__ void __launch_bounds__(GT)
      k_subtract(const double *V, long long ld, int m, const double *h, double *w, long long n)
    {
      const long long i = (long long)blockIdx.x * GT + threadIdx.x;
      if (i >= n)
        return;
      double s = w[i];
      for (int j = 0; j < m; ++j)
        s -= h[j] * V[(size_t)j * ld + i];
      w[i] = s;
    }

    // y = a * x
    __global__ void __launch_bounds__(GT) k_scale(const double *x, double a, double *y, long long n)
    {
      const long long i = (long long)blockIdx.x * GT + threadIdx.x;
      if (i < n)
        y[i] = a * x[i];
    }

    __global__ void __launch_bounds__(GT) k_zero(double *y, long long n)
    {
      const long long i = (long long)blockIdx.x * GT + threadIdx.x;
      if (i < n)
        y[i] = 0;
    }

    // x += sum_j c[j] V_j   (c on the device)
    __global__ void __launch_bounds__(GT)
      k_update(const double *V, long long ld, int m, const double *c, double *x, long long n)
    {
      const long long i = (long long)blockIdx.x * GT + threadIdx.x;
      if (i >= n)
        return;
      double s = x[i];
      for (int j = 0; j < m; ++j)
        s += c[j] * V[(size_t)j * ld + i];
      x[i] = s;
    }

    unsigned blocks(long long n) { return (unsigned)((n + GT - 1) / GT); }
  } // namespace

  int gmres_solve(const long long *rp, const int *ci, const double *vals, long long n,
                  const double *b, double *x, int restart, double tol, long long max_it,
                  long long *iterations, double *residual, hipStream_t s, std::string &err)
  {
    *iterations = 0;
    *residual   = 0;
    if (n <= 0)
      return 0;
    const int m = restart;
    double *V = nullptr, *w = nullptr, *part = nullptr, *dh = nullptr;
    hipError_t e = hipMalloc((void **)&V, sizeof(double) * (size_t)n * (m + 1));
    if (e == hipSuccess)
      e = hipMalloc((void **)&w, sizeof(double) * (size_t)n);
    if (e == hipSuccess)
      e = hipMalloc((void **)&part, sizeof(double) * (size_t)GB * (m + 1));
    if (e == hipSuccess)
      e = hipMalloc((void **)&dh, sizeof(double) * (m + 1));
    auto cleanup = [&]() {
      for (double *p : {V, w, part, dh})
        if (p)
          (void)hipFree(p);
    };
    if (e != hipSuccess)
      {
        cleanup();
        err = std::string("gmres workspace: ") + hipGetErrorString(e);
        return -3;
      }
    const unsigned nb  = blocks(n);
    const unsigned nbs = (unsigned)((n + GT / 64 - 1) / (GT / 64));
    std::vector<double> h(m + 1), H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
    // dots of w with V_0..V_{k-1} into h (host), through dh
    auto dots = [&](int k, const double *vec) -> bool {
      hipLaunchKernelGGL(k_dots1, dim3(GB), dim3(GT), 0, s, V, n, k, vec, n, part);
      hipLaunchKernelGGL(k_dots2, dim3(k), dim3(GT), 0, s, part, dh);
      if (hipMemcpyAsync(h.data(), dh, sizeof(double) * k, hipMemcpyDeviceToHost, s) !=
          hipSuccess)
        return false;
      return hipStreamSynchronize(s) == hipSuccess;
    };
    auto norm_of = [&](const double *vec, double &out) -> bool {
      hipLaunchKernelGGL(k_dots1, dim3(GB), dim3(GT), 0, s, vec, n, 1, vec, n, part);
      hipLaunchKernelGGL(k_dots2, dim3(1), dim3(GT), 0, s, part, dh);
      double v2 = 0;
      if (hipMemcpyAsync(&v2, dh, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess)
        return false;
      out = std::sqrt(v2);
      return true;
    };
    int  rc   = 0;
    bool done = false;
    // x = 0 (the reference's zeroed system_solution)
    hipLaunchKernelGGL(k_zero, dim3(nb), dim3(GT), 0, s, x, n);
    while (!done && rc == 0)
      {
        // r = b - A x into V_0
        hipLaunchKernelGGL(k_spmv, dim3(nbs), dim3(GT), 0, s, rp, ci, vals, n, x, V);
        hipLaunchKernelGGL(k_bminus, dim3(nb), dim3(GT), 0, s, b, V, n);
        double beta = 0;
        if (!norm_of(V, beta))
          {
            rc = -3;
            break;
          }
        *residual = beta;
        if (std::isnan(beta))
          {
            err = "gmres: the residual is not a number";
            rc  = -5;
            break;
          }
        if (beta <= tol)
          break;
        hipLaunchKernelGGL(k_scale, dim3(nb), dim3(GT), 0, s, V, 1.0 / beta, V, n);
        std::fill(g.begin(), g.end(), 0.0);
        g[0]  = beta;
        int k = 0;
        for (; k < m; ++k)
          {
            if (*iterations >= max_it)
              break;
            ++*iterations;
            double *vk = V + (size_t)(k + 1) * n;
            hipLaunchKernelGGL(k_spmv, dim3(nbs), dim3(GT), 0, s, rp, ci, vals, n,
                               V + (size_t)k * n, vk);
            // classical Gram-Schmidt, twice
            for (int j = 0; j <= k; ++j)
              H[(size_t)j * m + k] = 0;
            for (int pass = 0; pass < 2; ++pass)
              {
                if (!dots(k + 1, vk))
                  {
                    rc = -3;
                    break;
                  }
                for (int j = 0; j <= k; ++j)
                  H[(size_t)j * m + k] += h[j];
                hipLaunchKernelGGL(k_subtract, dim3(nb), dim3(GT), 0, s, V, n, k + 1, dh, vk, n);
              }
            if (rc)
              break;
            double hn = 0;
            if (!norm_of(vk, hn))
              {
                rc = -3;
                break;
              }
            H[(size_t)(k + 1) * m + k] = hn;
            if (hn > 0)
              hipLaunchKernelGGL(k_scale, dim3(nb), dim3(GT), 0, s, vk, 1.0 / hn, vk, n);
            // apply the previous rotations, then a new one
            for (int j = 0; j < k; ++j)
              {
                const double a = H[(size_t)j * m + k], c2 = H[(size_t)(j + 1) * m + k];
                H[(size_t)j * m + k]       = cs[j] * a + sn[j] * c2;
                H[(size_t)(j + 1) * m + k] = -sn[j] * a + cs[j] * c2;
              }
            const double a = H[(size_t)k * m + k], c2 = H[(size_t)(k + 1) * m + k];
            const double r = std::hypot(a, c2);
            cs[k]          = r > 0 ? a / r : 1.0;
            sn[k]          = r > 0 ? c2 / r : 0.0;
            H[(size_t)k * m + k]       = r;
            H[(size_t)(k + 1) * m + k] = 0;
            g[k + 1] = -sn[k] * g[k];
            g[k]     = cs[k] * g[k];
            *residual = std::fabs(g[k + 1]);
            if (std::isnan(*residual))
              {
                err = "gmres: the residual is not a number";
                rc  = -5;
                break;
              }
            if (*residual <= tol || hn == 0)
              {
                ++k;
                done = true;
                break;
              }
          }
        if (rc)
          break;
        // y = H^-1 g on the k columns, x += V y
        for (int i = k - 1; i >= 0; --i)
          {
            double t = g[i];
            for (int j = i + 1; j < k; ++j)
              t -= H[(size_t)i * m + j] * y[j];
            y[i] = t / H[(size_t)i * m + i];
          }
        if (k > 0)
          {
            if (hipMemcpyAsync(dh, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, s) !=
                hipSuccess)
              {
                rc = -3;
                break;
              }
            hipLaunchKernelGGL(k_update, dim3(nb), dim3(GT), 0, s, V, n, k, dh, x, n);
            if (hipStreamSynchronize(s) != hipSuccess)
              {
                rc = -3;
                break;
              }
          }
        if (!done && *iterations >= max_it)
          {
            err = "gmres: no convergence in " + std::to_string(max_it) +
                  " iterations (residual " + std::to_string(*residual) + ", tolerance " +
                  std::to_string(tol) + ")";
            rc = -6;
          }
      }
    if (rc == -3 && err.empty())
      err = "gmres: a device call failed";
    cleanup();
    return rc;
  }
} // namespace hdg
