// hdg_tables.cpp -- reference-cell tables and flux-lifting operators.
//
// FE spaces re-created for axis-aligned box cells, following the
// reference's choice of elements (src/np_solver.cpp:292-325):
//   FE_DGQ(k)   Lagrange on equidistant nodes, lexicographic
//   FE_FaceQ(k) Lagrange on Gauss-Lobatto nodes of the face
//   QGauss(k+1) on cells and faces (src/np_solver.cpp:848-859)
// Face 2a+s has normal axis a on side s; its local coordinates are the
// tangent axes (y) / (x) in 2D and (y,z) / (z,x) / (x,y) in 3D, first
// fastest, as deal.II's reference hexahedron faces.
#include "hdg_tables.h"

#include <cmath>

namespace hdg
{
  namespace
  {
    double legendre(int n, double x, double *dp)
    {
      double p0 = 1, p1 = x;
      if (n == 0)
        {
          *dp = 0;
          return 1;
        }
      for (int m = 2; m <= n; ++m)
        {
          const double p2 = ((2 * m - 1) * x * p1 - (m - 1) * p0) / m;
          p0              = p1;
          p1              = p2;
        }
      *dp = n * (x * p1 - p0) / (x * x - 1);
      return p1;
    }

    void gauss01(int n, std::vector<double> &x, std::vector<double> &w)
    {
      x.assign(n, 0);
      w.assign(n, 0);
      for (int i = 0; i < n; ++i)
        {
          double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 0;
          for (int it = 0; it < 100; ++it)
            {
              const double dz = legendre(n, z, &dp) / dp;
              z -= dz;
              if (std::fabs(dz) < 1e-16)
                break;
            }
          legendre(n, z, &dp);
          x[n - 1 - i] = 0.5 * (z + 1);
          w[n - 1 - i] = 1.0 / ((1 - z * z) * dp * dp);
        }
    }

    std::vector<double> lobatto01(int n)
    {
      std::vector<double> x(n);
      x[0]        = 0;
      x[n - 1]    = 1;
      const int m = n - 1;
      for (int i = 1; i < n - 1; ++i)
        {
          double z = -std::cos(M_PI * i / m), dp = 0;
          for (int it = 0; it < 100; ++it)
            {
              const double p   = legendre(m, z, &dp);
              const double d2p = (2 * z * dp - m * (m + 1) * p) / (1 - z * z);
              const double dz  = dp / d2p;
              z -= dz;
              if (std::fabs(dz) < 1e-16)
                break;
            }
          x[i] = 0.5 * (z + 1);
        }
      return x;
    }

    double lag(const std::vector<double> &nd, int i, double x)
    {
      double v = 1;
      for (int m = 0; m < (int)nd.size(); ++m)
        if (m != i)
          v *= (x - nd[m]) / (nd[i] - nd[m]);
      return v;
    }

    double lag_der(const std::vector<double> &nd, int i, double x)
    {
      double s = 0;
      for (int l = 0; l < (int)nd.size(); ++l)
        {
          if (l == i)
            continue;
          double v = 1 / (nd[i] - nd[l]);
          for (int m = 0; m < (int)nd.size(); ++m)
            if (m != i && m != l)
              v *= (x - nd[m]) / (nd[i] - nd[m]);
          s += v;
        }
      return s;
    }

    // dense inverse (partial pivoting); small SPD mass matrices only
    bool invert(std::vector<double> A, int n, std::vector<double> &X)
    {
      X.assign(n * n, 0);
      for (int i = 0; i < n; ++i)
        X[i * n + i] = 1;
      for (int c = 0; c < n; ++c)
        {
          int p = c;
          for (int r = c + 1; r < n; ++r)
            if (std::fabs(A[r * n + c]) > std::fabs(A[p * n + c]))
              p = r;
          if (A[p * n + c] == 0)
            return false;
          for (int j = 0; j < n; ++j)
            {
              std::swap(A[c * n + j], A[p * n + j]);
              std::swap(X[c * n + j], X[p * n + j]);
            }
          const double inv = 1 / A[c * n + c];
          for (int j = 0; j < n; ++j)
            {
              A[c * n + j] *= inv;
              X[c * n + j] *= inv;
            }
          for (int r = 0; r < n; ++r)
            {
              if (r == c)
                continue;
              const double f = A[r * n + c];
              if (f == 0)
                continue;
              for (int j = 0; j < n; ++j)
                {
                  A[r * n + j] -= f * A[c * n + j];
                  X[r * n + j] -= f * X[c * n + j];
                }
            }
        }
      return true;
    }
  } // namespace

  bool build_tables(int dim, int deg, const double h[3], HostTables &T)
  {
    const int d = dim, k = deg;
    T.dim = d;
    T.deg = k;
    T.nb  = 1;
    T.nf  = 1;
    for (int a = 0; a < d; ++a)
      T.nb *= k + 1;
    for (int a = 0; a < d - 1; ++a)
      T.nf *= k + 1;
    T.qc  = T.nb;
    T.qf  = T.nf;
    T.nfc = 2 * d;
    T.sz  = T.qc * (1 + d) + T.nfc * T.qf;
    const int nb = T.nb, nf = T.nf, qc = T.qc, qf = T.qf, nfc = T.nfc;
    if (d == 2)
      {
        T.tang[0][0] = 1;
        T.tang[1][0] = 0;
      }
    else
      {
        T.tang[0][0] = 1;
        T.tang[0][1] = 2;
        T.tang[1][0] = 2;
        T.tang[1][1] = 0;
        T.tang[2][0] = 0;
        T.tang[2][1] = 1;
      }
    std::vector<double> qw1;
    gauss01(k + 1, T.qx1, qw1);
    std::vector<double> cn(k + 1);
    for (int i = 0; i <= k; ++i)
      cn[i] = (double)i / k;
    const std::vector<double> tn = lobatto01(k + 1);

    auto digits = [&](int idx, int nd, int *out) {
      for (int a = 0; a < nd; ++a)
        {
          out[a] = idx % (k + 1);
          idx /= (k + 1);
        }
    };
    double vol = 1;
    for (int a = 0; a < d; ++a)
      vol *= h[a];
    double area[3];
    for (int a = 0; a < d; ++a)
      {
        area[a] = 1;
        for (int b = 0; b < d; ++b)
          if (b != a)
            area[a] *= h[b];
      }

    // cell quadrature
    T.JxWc.assign(qc, 0);
    T.phic.assign(qc * nb, 0);
    T.gradc.assign(qc * nb * d, 0);
    for (int q = 0; q < qc; ++q)
      {
        int qi[3] = {0, 0, 0};
        digits(q, d, qi);
        double w = vol;
        for (int a = 0; a < d; ++a)
          w *= qw1[qi[a]];
        T.JxWc[q] = w;
        for (int j = 0; j < nb; ++j)
          {
            int ji[3] = {0, 0, 0};
            digits(j, d, ji);
            double v = 1;
            for (int a = 0; a < d; ++a)
              v *= lag(cn, ji[a], T.qx1[qi[a]]);
            T.phic[q * nb + j] = v;
            for (int a = 0; a < d; ++a)
              {
                double g = 1;
                for (int b = 0; b < d; ++b)
                  g *= (a == b) ? lag_der(cn, ji[b], T.qx1[qi[b]])
                                : lag(cn, ji[b], T.qx1[qi[b]]);
                T.gradc[(q * nb + j) * d + a] = g / h[a];
              }
          }
      }
    // face quadrature and bases
    T.phif.assign(nfc * qf * nb, 0);
    T.JxWf.assign(nfc * qf, 0);
    T.xi.assign(qf * nf, 0);
    for (int s = 0; s < qf; ++s)
      {
        int si[2] = {0, 0};
        digits(s, d - 1, si);
        for (int j = 0; j < nf; ++j)
          {
            int ji[2] = {0, 0};
            digits(j, d - 1, ji);
            double v = 1;
            for (int m = 0; m < d - 1; ++m)
              v *= lag(tn, ji[m], T.qx1[si[m]]);
            T.xi[s * nf + j] = v;
          }
      }
    for (int g = 0; g < nfc; ++g)
      {
        const int a = g / 2, side = g % 2;
        for (int s = 0; s < qf; ++s)
          {
            int si[2] = {0, 0};
            digits(s, d - 1, si);
            double w = area[a];
            for (int m = 0; m < d - 1; ++m)
              w *= qw1[si[m]];
            T.JxWf[g * qf + s] = w;
            double pt[3] = {0, 0, 0};
            pt[a]        = side;
            for (int m = 0; m < d - 1; ++m)
              pt[T.tang[a][m]] = T.qx1[si[m]];
            for (int j = 0; j < nb; ++j)
              {
                int ji[3] = {0, 0, 0};
                digits(j, d, ji);
                double v = 1;
                for (int b = 0; b < d; ++b)
                  v *= lag(cn, ji[b], pt[b]);
                T.phif[(g * qf + s) * nb + j] = v;
              }
          }
      }
    // mass, divergence and face coupling
    std::vector<double> M(nb * nb, 0);
    for (int q = 0; q < qc; ++q)
      for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j)
          M[i * nb + j] += T.phic[q * nb + i] * T.phic[q * nb + j] * T.JxWc[q];
    if (!invert(M, nb, T.Minv))
      return false;
    T.MinvD.assign(d * nb * nb, 0);
    for (int a = 0; a < d; ++a)
      {
        std::vector<double> D(nb * nb, 0);
        for (int q = 0; q < qc; ++q)
          for (int i = 0; i < nb; ++i)
            for (int j = 0; j < nb; ++j)
              D[i * nb + j] += T.phic[q * nb + j] *
                               T.gradc[(q * nb + i) * d + a] * T.JxWc[q];
        for (int i = 0; i < nb; ++i)
          for (int j = 0; j < nb; ++j)
            {
              double s = 0;
              for (int l = 0; l < nb; ++l)
                s += T.Minv[i * nb + l] * D[l * nb + j];
              T.MinvD[(a * nb + i) * nb + j] = s;
            }
      }
    T.MinvC.assign(nfc * nb * nf, 0);
    for (int f = 0; f < nfc; ++f)
      {
        const double nu = (f % 2) ? 1.0 : -1.0;
        std::vector<double> Cm(nb * nf, 0);
        for (int s = 0; s < qf; ++s)
          for (int i = 0; i < nb; ++i)
            for (int m = 0; m < nf; ++m)
              Cm[i * nf + m] += T.xi[s * nf + m] * T.phif[(f * qf + s) * nb + i] *
                                nu * T.JxWf[f * qf + s];
        for (int i = 0; i < nb; ++i)
          for (int m = 0; m < nf; ++m)
            {
              double s = 0;
              for (int l = 0; l < nb; ++l)
                s += T.Minv[i * nb + l] * Cm[l * nf + m];
              T.MinvC[(f * nb + i) * nf + m] = s;
            }
      }
    // lifted values at quadrature points
    T.LDc.assign(qc * d * nb, 0);
    for (int q = 0; q < qc; ++q)
      for (int a = 0; a < d; ++a)
        for (int j = 0; j < nb; ++j)
          {
            double s = 0;
            for (int i = 0; i < nb; ++i)
              s += T.phic[q * nb + i] * T.MinvD[(a * nb + i) * nb + j];
            T.LDc[(q * d + a) * nb + j] = s;
          }
    T.LDf.assign(nfc * qf * nb, 0);
    for (int g = 0; g < nfc; ++g)
      {
        const int    a  = g / 2;
        const double nu = (g % 2) ? 1.0 : -1.0;
        for (int s = 0; s < qf; ++s)
          for (int j = 0; j < nb; ++j)
            {
              double v = 0;
              for (int i = 0; i < nb; ++i)
                v += T.phif[(g * qf + s) * nb + i] * T.MinvD[(a * nb + i) * nb + j];
              T.LDf[(g * qf + s) * nb + j] = nu * v;
            }
      }
    T.LCc.assign(nfc * qc * nf, 0);
    for (int f = 0; f < nfc; ++f)
      for (int q = 0; q < qc; ++q)
        for (int m = 0; m < nf; ++m)
          {
            double v = 0;
            for (int i = 0; i < nb; ++i)
              v += T.phic[q * nb + i] * T.MinvC[(f * nb + i) * nf + m];
            T.LCc[(f * qc + q) * nf + m] = -v;
          }
    T.LCf.assign(nfc * nfc * qf * nf, 0);
    for (int f = 0; f < nfc; ++f)
      for (int g = 0; g < nfc; ++g)
        {
          if (f / 2 != g / 2)
            continue;
          const double nu = (g % 2) ? 1.0 : -1.0;
          for (int s = 0; s < qf; ++s)
            for (int m = 0; m < nf; ++m)
              {
                double v = 0;
                for (int i = 0; i < nb; ++i)
                  v += T.phif[(g * qf + s) * nb + i] * T.MinvC[(f * nb + i) * nf + m];
                T.LCf[((f * nfc + g) * qf + s) * nf + m] = -nu * v;
              }
        }
    T.PhiMinvC.assign(qc * nb, 0);
    for (int q = 0; q < qc; ++q)
      for (int i = 0; i < nb; ++i)
        {
          double v = 0;
          for (int l = 0; l < nb; ++l)
            v += T.phic[q * nb + l] * T.Minv[l * nb + i];
          T.PhiMinvC[q * nb + i] = v;
        }
    T.PhiMinvF.assign(nfc * qf * nb, 0);
    for (int g = 0; g < nfc; ++g)
      {
        const double nu = (g % 2) ? 1.0 : -1.0;
        for (int s = 0; s < qf; ++s)
          for (int i = 0; i < nb; ++i)
            {
              double v = 0;
              for (int l = 0; l < nb; ++l)
                v += T.phif[(g * qf + s) * nb + l] * T.Minv[l * nb + i];
              T.PhiMinvF[(g * qf + s) * nb + i] = nu * v;
            }
      }
    // test tables
    T.Bz.assign(T.sz * nb, 0);
    for (int q = 0; q < qc; ++q)
      for (int i = 0; i < nb; ++i)
        {
          T.Bz[(q * (1 + d)) * nb + i] = T.phic[q * nb + i] * T.JxWc[q];
          for (int a = 0; a < d; ++a)
            T.Bz[(q * (1 + d) + 1 + a) * nb + i] =
              T.gradc[(q * nb + i) * d + a] * T.JxWc[q];
        }
    for (int g = 0; g < nfc; ++g)
      for (int s = 0; s < qf; ++s)
        for (int i = 0; i < nb; ++i)
          T.Bz[(qc * (1 + d) + g * qf + s) * nb + i] =
            T.phif[(g * qf + s) * nb + i] * T.JxWf[g * qf + s];
    T.Bt.assign(nfc * qf * nf, 0);
    for (int g = 0; g < nfc; ++g)
      for (int s = 0; s < qf; ++s)
        for (int i = 0; i < nf; ++i)
          T.Bt[(g * qf + s) * nf + i] = T.xi[s * nf + i] * T.JxWf[g * qf + s];
    // face mass inverse per axis (Dirichlet L2 projection)
    T.Mfinv.assign(d * nf * nf, 0);
    for (int a = 0; a < d; ++a)
      {
        std::vector<double> Mf(nf * nf, 0), X;
        for (int s = 0; s < qf; ++s)
          for (int i = 0; i < nf; ++i)
            for (int j = 0; j < nf; ++j)
              Mf[i * nf + j] +=
                T.xi[s * nf + i] * T.xi[s * nf + j] * T.JxWf[(2 * a) * qf + s];
        if (!invert(Mf, nf, X))
          return false;
        for (int i = 0; i < nf * nf; ++i)
          T.Mfinv[a * nf * nf + i] = X[i];
      }
    return true;
  }
} // namespace hdg
