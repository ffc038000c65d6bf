// hdg_device.h -- device helpers shared by the generic and the
// register-resident kernels.
#pragma once
#include "hdg_kernels.h"

namespace hdg
{
  namespace
  {
    __device__ inline long long face_of(const KParams &p, int ix, int iy, int iz,
                                        int lf)
    {
      int       pos[3] = {ix, iy, iz};
      const int a      = lf >> 1;
      pos[a] += lf & 1;
      const long long *m = p.face_dims[a];
      return p.face_off[a] + pos[0] + m[0] * (pos[1] + m[1] * (long long)pos[2]);
    }

    // SRH / linear recombination with the reference's scaling
    // (src/recombination_term.cpp:433-480, src/adimensionalizer.cpp:476-497)
    __device__ inline void recomb(const KParams &p, double n, double pp, double &r,
                                  double &drn, double &drp)
    {
      if (p.rec == 1)
        {
          const double ni  = p.rec0 / p.n_rf;
          const double num = n * pp - ni * ni;
          const double den = p.rec2 * (n + ni) + p.rec1 * (pp + ni);
          r                = num * p.n_rf / den;
          drn              = pp / den - num * p.rec2 / (den * den);
          drp              = n / den - num * p.rec1 / (den * den);
        }
      else if (p.rec == 2)
        {
          r   = p.rec0 + (p.rec1 * n + p.rec2 * pp) * p.n_rf;
          drn = p.rec1;
          drp = p.rec2;
        }
      else
        {
          r = drn = drp = 0;
        }
      r /= p.rec_rf;
      drn /= p.rec_rf;
      drp /= p.rec_rf;
    }

  } // namespace
} // namespace hdg
