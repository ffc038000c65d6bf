// hdg_wave.h -- wave64 helpers for gfx950: lane broadcasts by readlane,
// wave-wide max reductions by DPP (no LDS round trips), the FP64 MFMA.
#pragma once
#include <hip/hip_runtime.h>

namespace hdg
{
  // the value v of lane l (l uniform), in every lane
  __device__ __forceinline__ double lane_value(double v, int l)
  {
    const long long b  = __double_as_longlong(v);
    const int       lo = __builtin_amdgcn_readlane((int)b, l);
    const int       hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
  }
  __device__ __forceinline__ float lane_value(float v, int l)
  {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
  }

  // wave-wide max of an unsigned key with DPP row shifts and row
  // broadcasts: lane 15 of every row of 16 ends with its row's max, rows 1
  // and 3 then take rows 0 and 2, rows 2 and 3 take lane 31; lane 63 holds
  // the max of the wave, returned to every lane
  template <int CTRL, int ROWS>
  __device__ __forceinline__ unsigned dpp_max_step(unsigned v)
  {
    const unsigned o = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xf, false);
    return o > v ? o : v;
  }
  __device__ __forceinline__ unsigned wave_max(unsigned v)
  {
    v = dpp_max_step<0x111, 0xf>(v); // row_shr:1
    v = dpp_max_step<0x112, 0xf>(v); // row_shr:2
    v = dpp_max_step<0x114, 0xf>(v); // row_shr:4
    v = dpp_max_step<0x118, 0xf>(v); // row_shr:8
    v = dpp_max_step<0x142, 0xa>(v); // row_bcast:15
    v = dpp_max_step<0x143, 0xc>(v); // row_bcast:31
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
  }

  typedef double dbl4 __attribute__((ext_vector_type(4)));

  // v_mfma_f64_16x16x4f64, D = A B + C (64 lanes active): A[i][k] in lane
  // i + 16k, B[k][j] in lane j + 16k; element e of the result holds
  // D[4e + lane/16][lane%16]
  __device__ __forceinline__ dbl4 mfma_f64_16x16x4(double a, double b, dbl4 c)
  {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }

  // leading 32 bits of a non-negative double: they order as unsigned
  // integers, as the values do (to about 5 significant digits)
  __device__ __forceinline__ unsigned leading_bits(double v)
  {
    return (unsigned)(__double_as_longlong(v) >> 32);
  }

  // a / b for the eliminations' multipliers: the hardware reciprocal, one
  // Newton step on it, the quotient and one residual correction (within an
  // ulp or two of the IEEE quotient, about half the instructions of the
  // full division sequence); b = 0 gives inf or nan as the division does
  __device__ __forceinline__ double pivot_div(double a, double b)
  {
    double r = __builtin_amdgcn_rcp(b);
    r        = fma(fma(-b, r, 1.0), r, r);
    double q = a * r;
    return fma(fma(-b, q, a), r, q);
  }
  __device__ __forceinline__ float pivot_div(float a, float b)
  {
    return a / b;
  }
} // namespace hdg
