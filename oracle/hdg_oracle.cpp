/*
 * hdg_oracle.cpp -- CPU restatement of ddhdg's per-cell HDG hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the timed CPU
 * baseline.  The product never links it.
 *
 * The per-cell operator follows include/assembling_methods.template.h of
 * the reference term by term (citations per function).  FE semantics of
 * deal.II are re-created here for axis-aligned box cells:
 *   cell space   FESystem(FESystem(FE_DGQ(k),dim),FE_DGQ(k)) per component
 *                (src/np_solver.cpp:292-325): Lagrange on equidistant nodes,
 *                lexicographic, flux dofs (axis-major) before scalar dofs
 *   trace space  FE_FaceQ(k): Lagrange on Gauss-Lobatto nodes of each face
 *   quadrature   QGauss(k+1) cell and face (src/np_solver.cpp:848-859)
 *   face order   deal.II: face 2a+s is the face with normal axis a on side s;
 *                face-local tangent axes (y), (x) in 2D and (y,z), (z,x),
 *                (x,y) in 3D, first one fastest.
 */
#include <cstdio>
#include "hdg_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace
{
  // reference include/constants.h:12-15
  constexpr double kQ  = 1.602176634E-19;
  constexpr double kKB = 1.380649E-23;
  constexpr double kEV = 1.602176565E-19;

  thread_local std::string g_err;

  // ---------------------------------------------------------------- 1D rules
  double legendre(int n, double x, double *dp)
  {
    double p0 = 1, p1 = x;
    if (n == 0)
      {
        if (dp)
          *dp = 0;
        return 1;
      }
    for (int m = 2; m <= n; ++m)
      {
        const double p2 = ((2 * m - 1) * x * p1 - (m - 1) * p0) / m;
        p0              = p1;
        p1              = p2;
      }
    if (dp)
      *dp = n * (x * p1 - p0) / (x * x - 1);
    return p1;
  }

  // Gauss-Legendre on [0,1], ascending
  void gauss01(int n, std::vector<double> &x, std::vector<double> &w)
  {
    x.resize(n);
    w.resize(n);
    for (int i = 0; i < n; ++i)
      {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it)
          {
            double dp;
            const double p  = legendre(n, z, &dp);
            const double dz = p / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16)
              break;
          }
        double dp;
        legendre(n, z, &dp);
        x[n - 1 - i] = 0.5 * (z + 1);
        w[n - 1 - i] = 1.0 / ((1 - z * z) * dp * dp);
      }
  }

  // Gauss-Lobatto nodes on [0,1] (n >= 2): endpoints + roots of P'_{n-1}
  std::vector<double> lobatto01(int n)
  {
    std::vector<double> x(n);
    x[0]     = 0;
    x[n - 1] = 1;
    const int m = n - 1;
    for (int i = 1; i < n - 1; ++i)
      {
        double z = -std::cos(M_PI * i / m);
        for (int it = 0; it < 100; ++it)
          {
            // f = P'_m(z); f' = P''_m(z) from the Legendre ODE
            double dp;
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

  struct Lagrange1D
  {
    std::vector<double> nd;
    double val(int i, double x) const
    {
      double v = 1;
      for (int m = 0; m < (int)nd.size(); ++m)
        if (m != i)
          v *= (x - nd[m]) / (nd[i] - nd[m]);
      return v;
    }
    double der(int i, double x) const
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
  };

  // -------------------------------------------------- dense linear algebra
  // In-place inverse by Gauss-Jordan elimination with partial (row)
  // pivoting; stands for FullMatrix::gauss_jordan (reference
  // include/assembling_methods.template.h:2654-2662).
  bool gj_invert(std::vector<double> &A, int n, std::vector<int> &piv,
                 std::vector<double> &tmp)
  {
    piv.resize(n);
    tmp.resize(n);
    for (int i = 0; i < n; ++i)
      piv[i] = i;
    for (int col = 0; col < n; ++col)
      {
        int    best = col;
        double bmax = std::fabs(A[col * n + col]);
        for (int r = col + 1; r < n; ++r)
          if (std::fabs(A[r * n + col]) > bmax)
            {
              bmax = std::fabs(A[r * n + col]);
              best = r;
            }
        if (!(bmax > 0))
          return false;
        if (best != col)
          {
            for (int c = 0; c < n; ++c)
              std::swap(A[col * n + c], A[best * n + c]);
            std::swap(piv[col], piv[best]);
          }
        const double inv = 1.0 / A[col * n + col];
        // eliminate every other row with the pivot row
        for (int r = 0; r < n; ++r)
          {
            if (r == col)
              continue;
            const double f = A[r * n + col] * inv;
            if (f == 0.0)
              continue;
            for (int c = 0; c < n; ++c)
              if (c != col)
                A[r * n + c] -= f * A[col * n + c];
            A[r * n + col] = -f;
          }
        for (int c = 0; c < n; ++c)
          if (c != col)
            A[col * n + c] *= inv;
        A[col * n + col] = inv;
      }
    // undo the row interchanges as column interchanges of the inverse
    for (int r = 0; r < n; ++r)
      {
        for (int c = 0; c < n; ++c)
          tmp[piv[c]] = A[r * n + c];
        for (int c = 0; c < n; ++c)
          A[r * n + c] = tmp[c];
      }
    return true;
  }

  // small SPD solve (Cholesky); LAPACKFullMatrix::compute_cholesky_
  // factorization + solve in apply_dbc_on_face (template.h:1676-1678)
  void chol_solve(std::vector<double> &M, int n, std::vector<double> &b)
  {
    for (int j = 0; j < n; ++j)
      {
        double s = M[j * n + j];
        for (int k = 0; k < j; ++k)
          s -= M[j * n + k] * M[j * n + k];
        const double l = std::sqrt(s);
        M[j * n + j]   = l;
        for (int i = j + 1; i < n; ++i)
          {
            double t = M[i * n + j];
            for (int k = 0; k < j; ++k)
              t -= M[i * n + k] * M[j * n + k];
            M[i * n + j] = t / l;
          }
      }
    for (int i = 0; i < n; ++i)
      {
        double s = b[i];
        for (int k = 0; k < i; ++k)
          s -= M[i * n + k] * b[k];
        b[i] = s / M[i * n + i];
      }
    for (int i = n - 1; i >= 0; --i)
      {
        double s = b[i];
        for (int k = i + 1; k < n; ++k)
          s -= M[k * n + i] * b[k];
        b[i] = s / M[i * n + i];
      }
  }

  struct Vec3
  {
    double v[3] = {0, 0, 0};
  };
  inline double dot(const Vec3 &a, const Vec3 &b, int d)
  {
    double s = 0;
    for (int i = 0; i < d; ++i)
      s += a.v[i] * b.v[i];
    return s;
  }
  inline Vec3 scale(const Vec3 &a, double s)
  {
    Vec3 r;
    for (int i = 0; i < 3; ++i)
      r.v[i] = a.v[i] * s;
    return r;
  }

  enum
  {
    CV = 0,
    CN = 1,
    CP = 2
  };
} // namespace

struct orc_ctx
{
  orc_problem P;
  int         d, k, nb, nf, Qc, Qf, nfaces_cell, NC;
  long long   ncells, nfaces;
  int         n[3];
  double      h[3], lo[3];
  // face grid per axis
  long long face_off[4];
  long long face_dims[3][3];

  // adimensionalizer-derived constants (include/adimensionalizer.h)
  double V_rf, n_rf, perm_rf, UT_rf, mu_rf, rec_rf;
  double tau_rf[3], neu_rf[3];
  double eps_r, mun_r, mup_r, tau_r[3];
  // CellFaceTauComputer values, rescaled, ncells*2d*3 (cell, face, comp);
  // empty: FixedTauComputer (tau_r)
  std::vector<double> tau_cf;

  // reference shape tables on [0,1]^d
  std::vector<double>         qx1, qw1;  // 1D gauss
  std::vector<double>         qw_cell;   // Qc weights on unit cell
  std::vector<double>         qw_face;   // Qf weights on unit face
  std::vector<double>         phi;       // Qc*nb
  std::vector<double>         dphi;      // Qc*nb*d (reference derivs)
  std::vector<double>         phif;      // nfaces_cell*Qf*nb
  std::vector<double>         xi;        // Qf*nf (trace basis)
  std::vector<std::vector<int>> cell_support_on_face; // per face: enabled-idx
  int tang[3][2];

  // enabled components
  int              en[3], ne, N, T;
  std::vector<int> enabled_idx; // enabled cell dofs (indices 0..NC-1)

  // data
  std::vector<double> doping;    // rescaled, ncells*Qc
  std::vector<double> dir_vals;  // nfaces*3*Qf physical
  std::vector<double> neu_vals;  // nfaces*3*Qf physical
  std::vector<double> x_cell, x_trace;

  // global restricted system
  bool                   pattern_ready = false;
  std::vector<long long> rowptr;
  std::vector<int>       colind;
  std::vector<double>    vals, rhs;
  std::vector<long long> dir_idx;
  std::vector<double>    dir_val;
};

namespace
{
  long long face_index(const orc_ctx *C, int a, const int pos[3])
  {
    const long long *m = C->face_dims[a];
    return C->face_off[a] + pos[0] + m[0] * (pos[1] + m[1] * (long long)pos[2]);
  }

  void cell_coords(const orc_ctx *C, long long c, int ic[3])
  {
    ic[0] = (int)(c % C->n[0]);
    long long r = c / C->n[0];
    ic[1]       = (int)(r % C->n[1]);
    ic[2]       = (int)(r / C->n[1]);
  }

  long long cell_face(const orc_ctx *C, long long c, int lf)
  {
    int ic[3];
    cell_coords(C, c, ic);
    const int a = lf / 2, s = lf % 2;
    ic[a] += s;
    return face_index(C, a, ic);
  }

  // boundary id of a cell face (colorized), -1 if interior
  int boundary_id(const orc_ctx *C, long long c, int lf)
  {
    int ic[3];
    cell_coords(C, c, ic);
    const int a = lf / 2, s = lf % 2;
    if (s == 0 && ic[a] == 0)
      return lf;
    if (s == 1 && ic[a] == C->n[a] - 1)
      return lf;
    return -1;
  }

  void setup_enabled(orc_ctx *C)
  {
    const int d = C->d, nb = C->nb;
    C->ne        = 0;
    for (int c = 0; c < 3; ++c)
      if (C->en[c])
        C->ne++;
    C->N = C->ne * (d + 1) * nb;
    C->T = C->nfaces_cell * C->ne * C->nf;
    C->enabled_idx.clear();
    for (int kk = 0; kk < C->NC; ++kk)
      if (C->en[kk / ((d + 1) * nb)])
        C->enabled_idx.push_back(kk);
    // ScratchData::check_dofs_on_faces_for_cells (src/np_solver.cpp:185-205)
    C->cell_support_on_face.assign(C->nfaces_cell, {});
    for (int lf = 0; lf < C->nfaces_cell; ++lf)
      {
        const int a = lf / 2, s = lf % 2;
        for (int i = 0; i < C->N; ++i)
          {
            const int kk = C->enabled_idx[i];
            int       j  = kk % ((d + 1) * nb) % nb;
            int       ja = j;
            for (int b = 0; b < a; ++b)
              ja /= (C->k + 1);
            ja %= (C->k + 1);
            if (ja == (s ? C->k : 0))
              C->cell_support_on_face[lf].push_back(i);
          }
      }
    // the trace CSR pattern is built when first needed (pass 1, the system,
    // the pattern itself), so per-cell calls on large meshes stay cheap
    C->pattern_ready = false;
    std::vector<long long>().swap(C->rowptr);
    std::vector<int>().swap(C->colind);
    std::vector<double>().swap(C->vals);
    std::vector<double>().swap(C->rhs);
  }

  // restricted CSR entries: a boundary face couples to the faces of its one
  // cell, an interior face to those of its two cells (one shared)
  long long csr_entries(const orc_ctx *C)
  {
    long long nbnd = 0;
    for (int a = 0; a < C->d; ++a)
      {
        long long m = 2;
        for (int b = 0; b < C->d; ++b)
          if (b != a)
            m *= C->n[b];
        nbnd += m;
      }
    const long long tdf = (long long)C->ne * C->nf;
    return tdf * tdf * (nbnd * C->nfaces_cell + (C->nfaces - nbnd) * (2 * C->nfaces_cell - 1));
  }

  void ensure_pattern(orc_ctx *C)
  {
    if (C->pattern_ready)
      return;
    // restricted trace CSR pattern: a face's rows couple to the trace dofs
    // of the cells adjacent to it (DoFTools::make_sparsity_pattern,
    // src/np_solver.cpp:644-652)
    const int        tdf = C->ne * C->nf;
    const long long  nr  = C->nfaces * tdf;
    C->rowptr.assign(nr + 1, 0);
    std::vector<std::vector<long long>> face_cells(C->nfaces);
    for (long long c = 0; c < C->ncells; ++c)
      for (int lf = 0; lf < C->nfaces_cell; ++lf)
        face_cells[cell_face(C, c, lf)].push_back(c);
    std::vector<std::vector<int>> cols_of_face(C->nfaces);
    for (long long f = 0; f < C->nfaces; ++f)
      {
        std::vector<long long> fs;
        for (long long c : face_cells[f])
          for (int lf = 0; lf < C->nfaces_cell; ++lf)
            fs.push_back(cell_face(C, c, lf));
        std::sort(fs.begin(), fs.end());
        fs.erase(std::unique(fs.begin(), fs.end()), fs.end());
        for (long long g : fs)
          for (int t = 0; t < tdf; ++t)
            cols_of_face[f].push_back((int)(g * tdf + t));
      }
    for (long long f = 0; f < C->nfaces; ++f)
      for (int t = 0; t < tdf; ++t)
        C->rowptr[f * tdf + t + 1] = (long long)cols_of_face[f].size();
    for (long long r = 0; r < nr; ++r)
      C->rowptr[r + 1] += C->rowptr[r];
    C->colind.resize(C->rowptr[nr]);
    for (long long f = 0; f < C->nfaces; ++f)
      for (int t = 0; t < tdf; ++t)
        std::copy(cols_of_face[f].begin(), cols_of_face[f].end(),
                  C->colind.begin() + C->rowptr[f * tdf + t]);
    C->vals.assign(C->colind.size(), 0.0);
    C->rhs.assign(nr, 0.0);
    if ((long long)C->colind.size() != csr_entries(C))
      {
        std::fprintf(stderr, "oracle: CSR pattern has %lld entries, %lld expected\n",
                     (long long)C->colind.size(), csr_entries(C));
        std::abort();
      }
    C->pattern_ready = true;
  }

  // SRH / linear models (src/recombination_term.cpp:433-480, 123-253)
  void recombination(const orc_ctx *C, double n, double p, double rf,
                     double &r, double &drn, double &drp)
  {
    const double *rp = C->P.rec_params;
    if (C->P.recombination == 1)
      {
        const double ni  = rp[0] / rf;
        const double tn  = rp[1];
        const double tp  = rp[2];
        const double num = n * p - ni * ni;
        const double den = tp * (n + ni) + tn * (p + ni);
        r                = num * rf / den;
        drn              = p / den - num * tp / (den * den);
        drp              = n / den - num * tn / (den * den);
      }
    else if (C->P.recombination == 2)
      {
        r   = rp[0] + (rp[1] * n + rp[2] * p) * rf;
        drn = rp[1];
        drp = rp[2];
      }
    else
      {
        r = drn = drp = 0;
      }
  }

  // ------------------------------------------------------- per-cell work
  struct Work
  {
    const orc_ctx *C;
    int            d, N, T, NC;
    // matrices (row-major)
    std::vector<double> A, CT, TC, TT, c, t, tmp;
    std::vector<double> Sg;
    std::vector<int>    piv;
    std::vector<double> gjtmp;
    // per-q cell data
    std::vector<double> JxW, V0, n0, p0, UT, dop, r, drn, drp, phin, phip;
    std::vector<Vec3>   E0, Wn0, Wp0;
    // per-k shape buffers for one quadrature point
    std::vector<Vec3>   q1, q2, q3, g1, g2, g3;
    std::vector<double> q1d, q2d, q3d, z1, z2, z3;
    // face data
    std::vector<double> fJxW, fV0, fn0, fp0, trV0, trn0, trp0, UTf;
    std::vector<Vec3>   fE0, fWn0, fWp0;
    std::vector<double> upd_tr[3];
    Vec3                normal;
    long long           cur = 0; // cell being assembled
    // dirichlet condenser
    int                 n_constr;
    std::vector<int>    constr_dofs;
    std::vector<double> constr_vals;

    explicit Work(const orc_ctx *C_) : C(C_)
    {
      d  = C->d;
      N  = C->N;
      T  = C->T;
      NC = C->NC;
      A.resize(N * N);
      CT.resize(N * T);
      TC.resize(T * N);
      TT.resize(T * T);
      c.resize(N);
      t.resize(T);
      tmp.resize(T * N);
      const int Qc = C->Qc, Qf = C->Qf;
      for (auto *v : {&JxW, &V0, &n0, &p0, &UT, &dop, &r, &drn, &drp, &phin,
                      &phip})
        v->resize(Qc);
      E0.resize(Qc);
      Wn0.resize(Qc);
      Wp0.resize(Qc);
      for (auto *v : {&q1, &q2, &q3, &g1, &g2, &g3})
        v->resize(NC);
      for (auto *v : {&q1d, &q2d, &q3d, &z1, &z2, &z3})
        v->resize(NC);
      for (auto *v : {&fJxW, &fV0, &fn0, &fp0, &trV0, &trn0, &trp0, &UTf})
        v->resize(Qf);
      fE0.resize(Qf);
      fWn0.resize(Qf);
      fWp0.resize(Qf);
      for (auto &u : upd_tr)
        u.resize(Qf);
      constr_dofs.resize(T);
      constr_vals.resize(T);
    }

    // tau of component comp on local face lf of the current cell:
    // FixedTauComputer::compute_tau or CellFaceTauComputer::compute_tau
    // (include/np_solver_parameters.h:92-115, 157-236), rescaled
    inline double tau(int lf, int comp) const
    {
      if (C->tau_cf.empty())
        return C->tau_r[comp];
      return C->tau_cf[(cur * C->nfaces_cell + lf) * 3 + comp];
    }

    // dof decode: component, flux axis (-1 scalar), basis index
    inline void decode(int kk, int &comp, int &axis, int &j) const
    {
      const int bs = (d + 1) * C->nb;
      comp         = kk / bs;
      int rr       = kk % bs;
      if (rr < d * C->nb)
        {
          axis = rr / C->nb;
          j    = rr % C->nb;
        }
      else
        {
          axis = -1;
          j    = rr - d * C->nb;
        }
    }

    // fills q*/z* shape buffers (all NC dofs) for cell qp q
    void cell_shapes(int q)
    {
      const int nb = C->nb;
      for (int kk = 0; kk < NC; ++kk)
        {
          int comp, axis, j;
          decode(kk, comp, axis, j);
          Vec3   qv, gv;
          double qd = 0, zv = 0;
          if (axis >= 0)
            {
              qv.v[axis] = C->phi[q * nb + j];
              qd         = C->dphi[(q * nb + j) * d + axis] / C->h[axis];
            }
          else
            {
              zv = C->phi[q * nb + j];
              for (int a = 0; a < d; ++a)
                gv.v[a] = C->dphi[(q * nb + j) * d + a] / C->h[a];
            }
          Vec3 zero;
          q1[kk]  = comp == CV ? qv : zero;
          q2[kk]  = comp == CN ? qv : zero;
          q3[kk]  = comp == CP ? qv : zero;
          q1d[kk] = comp == CV ? qd : 0;
          q2d[kk] = comp == CN ? qd : 0;
          q3d[kk] = comp == CP ? qd : 0;
          z1[kk]  = comp == CV ? zv : 0;
          z2[kk]  = comp == CN ? zv : 0;
          z3[kk]  = comp == CP ? zv : 0;
          g1[kk]  = comp == CV ? gv : zero;
          g2[kk]  = comp == CN ? gv : zero;
          g3[kk]  = comp == CP ? gv : zero;
        }
    }

    // copy_fe_values_on_scratch (template.h:822-849): face values of cell
    // shape functions for the enabled dofs with support on the face, packed
    // by position in cell_support_on_face
    void face_shapes(int lf, int s)
    {
      const int nb = C->nb;
      const auto &sup = C->cell_support_on_face[lf];
      for (size_t i = 0; i < sup.size(); ++i)
        {
          const int kk = C->enabled_idx[sup[i]];
          int       comp, axis, j;
          decode(kk, comp, axis, j);
          const double v = C->phif[(lf * C->Qf + s) * nb + j];
          Vec3         qv;
          double       zv = 0;
          if (axis >= 0)
            qv.v[axis] = v;
          else
            zv = v;
          Vec3 zero;
          q1[i] = comp == CV ? qv : zero;
          q2[i] = comp == CN ? qv : zero;
          q3[i] = comp == CP ? qv : zero;
          z1[i] = comp == CV ? zv : 0;
          z2[i] = comp == CN ? zv : 0;
          z3[i] = comp == CP ? zv : 0;
        }
    }

    // value of restricted trace dof lt at face lf qp s for component c
    inline double tr_value(int lt, int lf, int s, int c) const
    {
      const int tdf = C->ne * C->nf;
      if (lt / tdf != lf)
        return 0;
      const int ci = (lt % tdf) / C->nf;
      int       cc = -1, m = 0;
      for (int q = 0; q < 3; ++q)
        if (C->en[q])
          {
            if (m == ci)
              cc = q;
            ++m;
          }
      if (cc != c)
        return 0;
      return C->xi[s * C->nf + (lt % tdf) % C->nf];
    }

    // prepare_data_on_cell_quadrature_points (template.h:134-265)
    void prepare_cell(long long cell, bool thermo)
    {
      const int     nb = C->nb, Qc = C->Qc;
      const double *x  = &C->x_cell[cell * NC];
      double        vol = 1;
      for (int a = 0; a < d; ++a)
        vol *= C->h[a];
      for (int q = 0; q < Qc; ++q)
        {
          JxW[q] = C->qw_cell[q] * vol;
          double vals[3];
          Vec3   fl[3];
          for (int comp = 0; comp < 3; ++comp)
            {
              const double *xc = x + comp * (d + 1) * nb;
              double        s  = 0;
              for (int j = 0; j < nb; ++j)
                s += xc[d * nb + j] * C->phi[q * nb + j];
              vals[comp] = s;
              for (int a = 0; a < d; ++a)
                {
                  double w = 0;
                  for (int j = 0; j < nb; ++j)
                    w += xc[a * nb + j] * C->phi[q * nb + j];
                  fl[comp].v[a] = w;
                }
            }
          V0[q]  = vals[0];
          n0[q]  = vals[1];
          p0[q]  = vals[2];
          E0[q]  = fl[0];
          Wn0[q] = fl[1];
          Wp0[q] = fl[2];
          UT[q]  = (kKB * C->P.temperature / kQ) / C->UT_rf;
          dop[q] = C->en[CV] ? C->doping[cell * Qc + q] : 0;
          if (C->en[CN] || C->en[CP])
            {
              double rr, a, b;
              recombination(C, n0[q], p0[q], C->n_rf, rr, a, b);
              // Adimensionalizer::adimensionalize_recombination_term
              // (src/adimensionalizer.cpp:476-497); derivative left
              // without the x doping_magnitude factor (reference quirk)
              r[q]   = rr / C->rec_rf;
              drn[q] = a / C->rec_rf;
              drp[q] = b / C->rec_rf;
            }
          else
            r[q] = drn[q] = drp[q] = 0;
          if (C->P.phi_linearize && !thermo)
            {
              // compute_quasi_fermi_potential (include/solver.h:343-375) of
              // the enabled carriers (template.h:229-264); a disabled
              // carrier keeps the scratch's zero (src/np_solver.cpp:138)
              const double UTp = C->P.temperature * (kKB / kQ);
              phin[q] = C->en[CN] ? V0[q] * C->V_rf - C->P.ec * kEV / kQ -
                                      UTp * std::log(n0[q] * C->n_rf / C->P.nc)
                                  : 0.0;
              phip[q] = C->en[CP] ? V0[q] * C->V_rf - C->P.ev * kEV / kQ +
                                      UTp * std::log(p0[q] * C->n_rf / C->P.nv)
                                  : 0.0;
            }
        }
    }

    double phi_term(int q, bool thermo, bool der) const
    {
      // template.h:442-459 (matrix) and 677-693 (rhs)
      const double KbT_over_q = kKB * C->P.temperature / (kQ * C->V_rf);
      const double ev_r       = C->P.ev * kEV / (kQ * C->V_rf);
      const double ec_r       = C->P.ec * kEV / (kQ * C->V_rf);
      const double pn         = thermo ? 0 : phin[q] / C->V_rf;
      const double pp         = thermo ? 0 : phip[q] / C->V_rf;
      const double nc = C->P.nc, nv = C->P.nv;
      if (der)
        return -1 / C->n_rf *
               (nv / KbT_over_q * std::exp((ev_r - V0[q] + pp) / KbT_over_q) +
                nc / KbT_over_q * std::exp((V0[q] - pn - ec_r) / KbT_over_q));
      return 1 / C->n_rf *
             (nv * std::exp((ev_r - V0[q] + pp) / KbT_over_q) -
              nc * std::exp((V0[q] - pn - ec_r) / KbT_over_q));
    }

    // add_cell_products_to_cc_matrix (template.h:269-532)
    void cell_matrix(bool thermo)
    {
      const bool eV = C->en[CV], eN = C->en[CN], eP = C->en[CP];
      const bool phimode = thermo || C->P.phi_linearize;
      const double eps = C->eps_r, mun = C->mun_r, mup = C->mup_r;
      // volume drift coupling to the E trial function: the reference adds
      // +n0 (mu_n E . grad z2) and -p0 (mu_p E . grad z3) (template.h:500,
      // 521-522); the derivative of its residual J . grad z (:656, :673) has
      // the opposite signs
      const double sdrift = C->P.drift_jacobian ? -1.0 : 1.0;
      for (int q = 0; q < C->Qc; ++q)
        {
          const double JW = JxW[q];
          cell_shapes(q);
          const Vec3   munE0 = scale(E0[q], mun);
          const Vec3   mupE0 = scale(E0[q], mup);
          const double phiV  = phimode ? phi_term(q, thermo, true) : 0;
          for (int i = 0; i < N; ++i)
            {
              const int ii = C->enabled_idx[i];
              for (int j = 0; j < N; ++j)
                {
                  const int jj = C->enabled_idx[j];
                  double   &a  = A[i * N + j];
                  if (eV)
                    {
                      const Vec3 epsE = scale(q1[jj], eps);
                      a += (-z1[jj] * q1d[ii] + dot(q1[jj], q1[ii], d) -
                            dot(epsE, g1[ii], d)) *
                           JW;
                      if (phimode)
                        a += -phiV * z1[jj] * z1[ii] * JW;
                      else
                        a += (z2[jj] - z3[jj]) * z1[ii] * JW;
                    }
                  if (eN)
                    {
                      const Vec3 munE = scale(q1[jj], mun);
                      const Vec3 Dn   = scale(scale(q2[jj], mun), UT[q]);
                      a += (-z2[jj] * q2d[ii] + dot(q2[jj], q2[ii], d) -
                            z2[jj] * dot(munE0, g2[ii], d) +
                            sdrift * n0[q] * dot(munE, g2[ii], d) + dot(Dn, g2[ii], d) -
                            (drn[q] * z2[jj] + drp[q] * z3[jj]) * z2[ii]) *
                           JW;
                    }
                  if (eP)
                    {
                      const Vec3 mupE = scale(q1[jj], mup);
                      const Vec3 Dp   = scale(scale(q3[jj], mup), UT[q]);
                      a += (-z3[jj] * q3d[ii] + dot(q3[jj], q3[ii], d) +
                            z3[jj] * dot(mupE0, g3[ii], d) -
                            sdrift * p0[q] * dot(mupE, g3[ii], d) + dot(Dp, g3[ii], d) -
                            (drn[q] * z2[jj] + drp[q] * z3[jj]) * z3[ii]) *
                           JW;
                    }
                }
            }
        }
    }

    // add_cell_products_to_cc_rhs (template.h:536-760)
    void cell_rhs(bool thermo)
    {
      const bool eV = C->en[CV], eN = C->en[CN], eP = C->en[CP];
      const bool phimode = thermo || C->P.phi_linearize;
      for (int q = 0; q < C->Qc; ++q)
        {
          const double JW = JxW[q];
          cell_shapes(q);
          const Vec3 epsE = scale(E0[q], C->eps_r);
          Vec3       Jn, Jp;
          if (eN)
            {
              const Vec3 mE = scale(E0[q], C->mun_r);
              const Vec3 Dn = scale(scale(Wn0[q], C->mun_r), UT[q]);
              for (int a = 0; a < 3; ++a)
                Jn.v[a] = n0[q] * mE.v[a] - Dn.v[a];
            }
          if (eP)
            {
              const Vec3 mE = scale(E0[q], C->mup_r);
              const Vec3 Dp = scale(scale(Wp0[q], C->mup_r), UT[q]);
              for (int a = 0; a < 3; ++a)
                Jp.v[a] = -p0[q] * mE.v[a] - Dp.v[a];
            }
          const double prhs = phimode ? phi_term(q, thermo, false) : 0;
          for (int i = 0; i < N; ++i)
            {
              const int ii = C->enabled_idx[i];
              double   &ci = c[i];
              if (eV)
                {
                  ci += (V0[q] * q1d[ii] - dot(E0[q], q1[ii], d) +
                         dot(epsE, g1[ii], d) + dop[q] * z1[ii]) *
                        JW;
                  if (phimode)
                    ci += prhs * z1[ii] * JW;
                  else
                    ci += (-n0[q] + p0[q]) * z1[ii] * JW;
                }
              if (eN)
                ci += (n0[q] * q2d[ii] - dot(Wn0[q], q2[ii], d) +
                       r[q] * z2[ii] + dot(Jn, g2[ii], d)) *
                      JW;
              if (eP)
                ci += (p0[q] * q3d[ii] - dot(Wp0[q], q3[ii], d) +
                       r[q] * z3[ii] + dot(Jp, g3[ii], d)) *
                      JW;
            }
        }
    }

    // prepare_data_on_face_quadrature_points (template.h:764-820)
    void prepare_face(long long cell, int lf, bool reconstruct,
                      const double *upd_trace)
    {
      const int nb = C->nb, Qf = C->Qf, nf = C->nf, a = lf / 2;
      double    area = 1;
      for (int b = 0; b < d; ++b)
        if (b != a)
          area *= C->h[b];
      normal = Vec3();
      normal.v[a] = (lf % 2) ? 1.0 : -1.0;
      const double   *x  = &C->x_cell[cell * NC];
      const long long f  = cell_face(C, cell, lf);
      const double   *tr = &C->x_trace[f * 3 * nf];
      for (int s = 0; s < Qf; ++s)
        {
          fJxW[s] = C->qw_face[s] * area;
          double vals[3], trv[3], up[3];
          Vec3   fl[3];
          for (int comp = 0; comp < 3; ++comp)
            {
              const double *xc = x + comp * (d + 1) * nb;
              double        sv = 0;
              for (int j = 0; j < nb; ++j)
                sv += xc[d * nb + j] * C->phif[(lf * Qf + s) * nb + j];
              vals[comp] = sv;
              for (int b = 0; b < d; ++b)
                {
                  double w = 0;
                  for (int j = 0; j < nb; ++j)
                    w += xc[b * nb + j] * C->phif[(lf * Qf + s) * nb + j];
                  fl[comp].v[b] = w;
                }
              double tv = 0, uv = 0;
              for (int j = 0; j < nf; ++j)
                {
                  tv += tr[comp * nf + j] * C->xi[s * nf + j];
                  if (reconstruct)
                    uv += upd_trace[f * 3 * nf + comp * nf + j] *
                          C->xi[s * nf + j];
                }
              trv[comp] = tv;
              up[comp]  = uv;
            }
          fV0[s]  = vals[0];
          fn0[s]  = vals[1];
          fp0[s]  = vals[2];
          fE0[s]  = fl[0];
          fWn0[s] = fl[1];
          fWp0[s] = fl[2];
          trV0[s] = trv[0];
          trn0[s] = trv[1];
          trp0[s] = trv[2];
          for (int comp = 0; comp < 3; ++comp)
            upd_tr[comp][s] = up[comp];
          UTf[s] = (kKB * C->P.temperature / kQ) / C->UT_rf;
        }
    }

    // add_ct_matrix_terms_to_cc_rhs (template.h:1023-1141)
    void ct_rhs(int lf)
    {
      const auto &sup = C->cell_support_on_face[lf];
      for (int s = 0; s < C->Qf; ++s)
        {
          const double JW = fJxW[s];
          face_shapes(lf, s);
          for (size_t i = 0; i < sup.size(); ++i)
            {
              const int ii = sup[i];
              if (C->en[CV])
                c[ii] += (-trV0[s] * dot(q1[i], normal, d) +
                          tau(lf, CV) * trV0[s] * z1[i]) *
                         JW;
              if (C->en[CN])
                c[ii] += (-trn0[s] * dot(q2[i], normal, d) -
                          tau(lf, CN) * trn0[s] * z2[i]) *
                         JW;
              if (C->en[CP])
                c[ii] += (-trp0[s] * dot(q3[i], normal, d) -
                          tau(lf, CP) * trp0[s] * z3[i]) *
                         JW;
            }
        }
    }

    // assemble_ct_matrix (template.h:885-1019)
    void ct_matrix(int lf)
    {
      const auto &sup  = C->cell_support_on_face[lf];
      const int   tdf  = C->ne * C->nf;
      const int   ft   = C->P.flux_type;
      for (int s = 0; s < C->Qf; ++s)
        {
          const double JW = fJxW[s];
          face_shapes(lf, s);
          const double E0n  = dot(fE0[s], normal, d);
          const double munE = C->mun_r * E0n, mupE = C->mup_r * E0n;
          for (size_t i = 0; i < sup.size(); ++i)
            {
              const int ii = sup[i];
              for (int j = 0; j < tdf; ++j)
                {
                  const int    jj  = lf * tdf + j;
                  const double trV = tr_value(jj, lf, s, CV);
                  const double trN = tr_value(jj, lf, s, CN);
                  const double trP = tr_value(jj, lf, s, CP);
                  double      &e   = CT[ii * T + jj];
                  if (C->en[CV])
                    e += (trV * dot(q1[i], normal, d) -
                          tau(lf, CV) * trV * z1[i]) *
                         JW;
                  if (C->en[CN])
                    {
                      if (ft == 1 || (ft == 2 && E0n <= 0))
                        e += munE * trN * z2[i] * JW;
                      e += (trN * dot(q2[i], normal, d) +
                            tau(lf, CN) * (trN * z2[i])) *
                           JW;
                    }
                  if (C->en[CP])
                    {
                      if (ft == 1 || (ft == 2 && E0n <= 0))
                        e += -mupE * trP * z3[i] * JW;
                      e += (trP * dot(q3[i], normal, d) +
                            tau(lf, CP) * (trP * z3[i])) *
                           JW;
                    }
                }
            }
        }
    }

    // assemble_tc_matrix (template.h:1145-1307); tc stored negated
    void tc_matrix(int lf, int comp, int ft)
    {
      const auto &sup = C->cell_support_on_face[lf];
      const int   tdf = C->ne * C->nf;
      const double tf = tau(lf, comp);
      for (int s = 0; s < C->Qf; ++s)
        {
          const double JW = fJxW[s];
          face_shapes(lf, s);
          const double E0n = dot(fE0[s], normal, d);
          for (int i = 0; i < tdf; ++i)
            {
              const int    ii  = lf * tdf + i;
              const double xiv = tr_value(ii, lf, s, comp);
              for (size_t j = 0; j < sup.size(); ++j)
                {
                  const int jj = sup[j];
                  double   &e  = TC[ii * N + jj];
                  if (comp == CV)
                    {
                      const double epsEn =
                        C->eps_r * dot(q1[j], normal, d);
                      e -= (epsEn + tf * z1[j]) * xiv * JW;
                    }
                  else if (comp == CN)
                    {
                      const double mE = C->mun_r * dot(q1[j], normal, d);
                      const double Df =
                        C->mun_r * UTf[s] * dot(q2[j], normal, d);
                      double cE;
                      if (ft == 0 || (ft == 2 && E0n > 0))
                        cE = z2[j] * C->mun_r * E0n + fn0[s] * mE;
                      else
                        cE = trn0[s] * mE;
                      e -= (cE - Df - tf * z2[j]) * xiv * JW;
                    }
                  else
                    {
                      const double mE = C->mup_r * dot(q1[j], normal, d);
                      const double Df =
                        C->mup_r * UTf[s] * dot(q3[j], normal, d);
                      double cE;
                      if (ft == 0 || (ft == 2 && E0n > 0))
                        cE = -z3[j] * C->mup_r * E0n - fp0[s] * mE;
                      else
                        cE = -trp0[s] * mE;
                      e -= (cE - Df - tf * z3[j]) * xiv * JW;
                    }
                }
            }
        }
    }

    // add_tc_matrix_terms_to_tt_rhs (template.h:1311-1447)
    void tc_rhs(int lf, int comp, int ft)
    {
      const int    tdf = C->ne * C->nf;
      const double tf = tau(lf, comp);
      for (int s = 0; s < C->Qf; ++s)
        {
          const double JW  = fJxW[s];
          const double E0n = dot(fE0[s], normal, d);
          for (int i = 0; i < tdf; ++i)
            {
              const int    ii  = lf * tdf + i;
              const double xiv = tr_value(ii, lf, s, comp);
              if (comp == CV)
                t[ii] += (-C->eps_r * E0n - tf * fV0[s]) * xiv * JW;
              else if (comp == CN)
                {
                  const double nn =
                    (ft == 0 || (ft == 2 && E0n > 0)) ? fn0[s] : trn0[s];
                  const double Dn =
                    C->mun_r * UTf[s] * dot(fWn0[s], normal, d);
                  t[ii] += (-nn * C->mun_r * E0n + Dn + tf * fn0[s]) *
                           xiv * JW;
                }
              else
                {
                  const double pp =
                    (ft == 0 || (ft == 2 && E0n > 0)) ? fp0[s] : trp0[s];
                  const double Dp =
                    C->mup_r * UTf[s] * dot(fWp0[s], normal, d);
                  t[ii] += (pp * C->mup_r * E0n + Dp + tf * fp0[s]) *
                           xiv * JW;
                }
            }
        }
    }

    // assemble_tt_matrix (template.h:1451-1547)
    void tt_matrix(int lf, int comp, int ft)
    {
      const int    tdf  = C->ne * C->nf;
      const double tf   = tau(lf, comp);
      const double sign = comp == CV ? -1 : 1;
      for (int s = 0; s < C->Qf; ++s)
        {
          const double JW  = fJxW[s];
          const double E0n = dot(fE0[s], normal, d);
          for (int i = 0; i < tdf; ++i)
            {
              const int    ii  = lf * tdf + i;
              const double xii = tr_value(ii, lf, s, comp);
              for (int j = 0; j < tdf; ++j)
                {
                  const int    jj  = lf * tdf + j;
                  const double trj = tr_value(jj, lf, s, comp);
                  double      &e   = TT[ii * T + jj];
                  if (comp == CN && (ft == 1 || (ft == 2 && E0n <= 0)))
                    e += C->mun_r * E0n * trj * xii * JW;
                  if (comp == CP && (ft == 1 || (ft == 2 && E0n <= 0)))
                    e += -C->mup_r * E0n * trj * xii * JW;
                  e += sign * tf * trj * xii * JW;
                }
            }
        }
    }

    // add_tt_matrix_terms_to_tt_rhs (template.h:1551-1588)
    void tt_rhs(int lf, int comp)
    {
      const int    tdf  = C->ne * C->nf;
      const double tf   = tau(lf, comp);
      const double sign = comp == CV ? -1 : 1;
      const std::vector<double> &tr0 =
        comp == CV ? trV0 : (comp == CN ? trn0 : trp0);
      for (int s = 0; s < C->Qf; ++s)
        for (int i = 0; i < tdf; ++i)
          {
            const int ii = lf * tdf + i;
            t[ii] += -sign * tf * tr0[s] * tr_value(ii, lf, s, comp) *
                     fJxW[s];
          }
    }

    // apply_dbc_on_face (template.h:1592-1691)
    void dbc(long long cell, int lf, int comp, bool thermo)
    {
      const int tdf = C->ne * C->nf, nf = C->nf, Qf = C->Qf;
      int       ci  = 0;
      for (int q = 0; q < comp; ++q)
        ci += C->en[q];
      const long long f  = cell_face(C, cell, lf);
      const double    rf = comp == CV ? C->V_rf : C->n_rf;
      const std::vector<double> &tr0 =
        comp == CV ? trV0 : (comp == CN ? trn0 : trp0);
      std::vector<double> M(nf * nf, 0.0), b(nf, 0.0);
      for (int s = 0; s < Qf; ++s)
        {
          const double gval =
            thermo ? 0 : C->dir_vals[(f * 3 + comp) * Qf + s] / rf - tr0[s];
          for (int i = 0; i < nf; ++i)
            {
              const double bi = C->xi[s * nf + i];
              for (int j = 0; j < nf; ++j)
                M[i * nf + j] += bi * C->xi[s * nf + j] * fJxW[s];
              b[i] += bi * gval * fJxW[s];
            }
        }
      chol_solve(M, nf, b);
      for (int i = 0; i < nf; ++i)
        {
          constr_dofs[n_constr] = lf * tdf + ci * nf + i;
          constr_vals[n_constr] = b[i];
          ++n_constr;
        }
    }

    // apply_nbc_on_face (template.h:1695-1734)
    void nbc(long long cell, int lf, int comp)
    {
      const int       tdf = C->ne * C->nf, Qf = C->Qf;
      const long long f   = cell_face(C, cell, lf);
      for (int s = 0; s < Qf; ++s)
        {
          const double hv = C->neu_vals[(f * 3 + comp) * Qf + s] / C->neu_rf[comp];
          for (int i = 0; i < tdf; ++i)
            {
              const int ii = lf * tdf + i;
              t[ii] += tr_value(ii, lf, s, comp) * hv * fJxW[s];
            }
        }
    }

    // add_border_products_to_cc_matrix (template.h:1737-1923)
    void border_matrix(int lf)
    {
      const auto &sup = C->cell_support_on_face[lf];
      const int   ft  = C->P.flux_type;
      for (int s = 0; s < C->Qf; ++s)
        {
          const double JW = fJxW[s];
          face_shapes(lf, s);
          const double E0n = dot(fE0[s], normal, d);
          for (size_t i = 0; i < sup.size(); ++i)
            {
              const int ii = sup[i];
              for (size_t j = 0; j < sup.size(); ++j)
                {
                  const int jj = sup[j];
                  double   &a  = A[ii * N + jj];
                  if (C->en[CV])
                    a += (C->eps_r * dot(q1[j], normal, d) * z1[i] +
                          tau(lf, CV) * z1[j] * z1[i]) *
                         JW;
                  if (C->en[CN])
                    {
                      const double mEn = C->mun_r * dot(q1[j], normal, d);
                      const double Dn =
                        C->mun_r * UTf[s] * dot(q2[j], normal, d);
                      double cE;
                      if (ft == 0 || (ft == 2 && E0n > 0))
                        cE = z2[j] * C->mun_r * E0n + mEn * fn0[s];
                      else
                        cE = mEn * trn0[s];
                      a += (cE * z2[i] - Dn * z2[i] -
                            tau(lf, CN) * z2[j] * z2[i]) *
                           JW;
                    }
                  if (C->en[CP])
                    {
                      const double mEp = C->mup_r * dot(q1[j], normal, d);
                      const double Dp =
                        C->mup_r * UTf[s] * dot(q3[j], normal, d);
                      double cE;
                      if (ft == 0 || (ft == 2 && E0n > 0))
                        cE = -(z3[j] * C->mup_r * E0n + mEp * fp0[s]);
                      else
                        cE = -mEp * trp0[s];
                      a += (cE * z3[i] - Dp * z3[i] -
                            tau(lf, CP) * z3[j] * z3[i]) *
                           JW;
                    }
                }
            }
        }
    }

    // add_border_products_to_cc_rhs (template.h:1927-2118)
    void border_rhs(int lf)
    {
      const auto &sup = C->cell_support_on_face[lf];
      const int   ft  = C->P.flux_type;
      for (int s = 0; s < C->Qf; ++s)
        {
          const double JW  = fJxW[s];
          face_shapes(lf, s);
          const double E0n = dot(fE0[s], normal, d);
          double       Jn = 0, Jp = 0;
          if (C->en[CN])
            {
              const double nn =
                (ft == 0 || (ft == 2 && E0n > 0)) ? fn0[s] : trn0[s];
              Jn = nn * C->mun_r * E0n -
                   C->mun_r * UTf[s] * dot(fWn0[s], normal, d);
            }
          if (C->en[CP])
            {
              const double pp =
                (ft == 0 || (ft == 2 && E0n > 0)) ? fp0[s] : trp0[s];
              Jp = -pp * C->mup_r * E0n -
                   C->mup_r * UTf[s] * dot(fWp0[s], normal, d);
            }
          for (size_t i = 0; i < sup.size(); ++i)
            {
              const int ii = sup[i];
              if (C->en[CV])
                c[ii] += (-C->eps_r * E0n * z1[i] -
                          tau(lf, CV) * fV0[s] * z1[i]) *
                         JW;
              if (C->en[CN])
                c[ii] += (-Jn * z2[i] + tau(lf, CN) * fn0[s] * z2[i]) * JW;
              if (C->en[CP])
                c[ii] += (-Jp * z3[i] + tau(lf, CP) * fp0[s] * z3[i]) * JW;
            }
        }
    }

    // add_trace_terms_to_cc_rhs (template.h:2122-2160)
    void trace_terms_rhs(int lf)
    {
      const auto &sup = C->cell_support_on_face[lf];
      for (int s = 0; s < C->Qf; ++s)
        {
          const double JW = fJxW[s];
          face_shapes(lf, s);
          for (int comp = 0; comp < 3; ++comp)
            {
              if (!C->en[comp])
                continue;
              const std::vector<Vec3>   &fq = comp == CV ? q1 : (comp == CN ? q2 : q3);
              const std::vector<double> &fz = comp == CV ? z1 : (comp == CN ? z2 : z3);
              const double sign = comp == CV ? 1 : -1;
              for (size_t i = 0; i < sup.size(); ++i)
                c[sup[i]] += (-dot(fq[i], normal, d) +
                              sign * tau(lf, comp) * fz[i]) *
                             upd_tr[comp][s] * JW;
            }
        }
    }

    // assemble_system_one_cell + _internal (template.h:2290-2748) up to the
    // LocalCondenser (pass 1) or the cell rhs of the reconstruct pass
    void assemble(long long cell, bool thermo, bool reconstruct,
                  const double *upd_trace)
    {
      cur = cell;
      std::fill(A.begin(), A.end(), 0.0);
      std::fill(c.begin(), c.end(), 0.0);
      if (!reconstruct)
        {
          std::fill(CT.begin(), CT.end(), 0.0);
          std::fill(TC.begin(), TC.end(), 0.0);
          std::fill(TT.begin(), TT.end(), 0.0);
          std::fill(t.begin(), t.end(), 0.0);
        }
      n_constr = 0;
      bool has_bc = false;
      for (int lf = 0; lf < C->nfaces_cell && !has_bc; ++lf)
        {
          const int id = boundary_id(C, cell, lf);
          if (id < 0)
            continue;
          for (int comp = 0; comp < 3; ++comp)
            if (C->en[comp] && C->P.bc_type[id][comp] != 0)
              has_bc = true;
        }
      prepare_cell(cell, thermo);
      cell_matrix(thermo);
      cell_rhs(thermo);
      const int ft = C->P.flux_type;
      for (int lf = 0; lf < C->nfaces_cell; ++lf)
        {
          prepare_face(cell, lf, reconstruct, upd_trace);
          ct_rhs(lf);
          if (!reconstruct)
            {
              ct_matrix(lf);
              const int id = has_bc ? boundary_id(C, cell, lf) : -1;
              for (int comp = 0; comp < 3; ++comp)
                {
                  if (!C->en[comp])
                    continue;
                  const int bt = id >= 0 ? C->P.bc_type[id][comp] : 0;
                  if (bt == 1)
                    {
                      dbc(cell, lf, comp, thermo);
                      continue;
                    }
                  tc_matrix(lf, comp, ft);
                  tc_rhs(lf, comp, ft);
                  // Q6: V keeps the use_cell tt inside the use_trace branch
                  tt_matrix(lf, comp, ft);
                  tt_rhs(lf, comp);
                  // in thermodynamic equilibrium the Neumann data is the zero
                  // function (template.h:2229-2236)
                  if (bt == 2 && !thermo)
                    nbc(cell, lf, comp);
                }
            }
          border_matrix(lf);
          border_rhs(lf);
          if (reconstruct)
            trace_terms_rhs(lf);
        }
      // LocalCondenser::condense_ct_matrix (src/local_condenser.cpp:141-160)
      if (has_bc && !reconstruct)
        for (int i = 0; i < N; ++i)
          for (int j = 0; j < n_constr; ++j)
            {
              const int    jj = constr_dofs[j];
              const double mv = CT[i * T + jj];
              CT[i * T + jj]  = 0;
              c[i] -= mv * constr_vals[j];
            }
    }

    // Schur complement (template.h:2654-2668): A^-1 by Gauss-Jordan, then
    // tmp = tc*A^-1, t += tmp*c, TT += tmp*CT
    bool condense()
    {
      if (!gj_invert(A, N, piv, gjtmp))
        return false;
      for (int i = 0; i < T; ++i)
        for (int j = 0; j < N; ++j)
          {
            double s = 0;
            for (int m = 0; m < N; ++m)
              s += TC[i * N + m] * A[m * N + j];
            tmp[i * N + j] = s;
          }
      for (int i = 0; i < T; ++i)
        {
          double s = 0;
          for (int m = 0; m < N; ++m)
            s += tmp[i * N + m] * c[m];
          t[i] += s;
        }
      for (int i = 0; i < T; ++i)
        for (int j = 0; j < T; ++j)
          {
            double s = 0;
            for (int m = 0; m < N; ++m)
              s += tmp[i * N + m] * CT[m * T + j];
            TT[i * T + j] += s;
          }
      return true;
    }
  };

  struct CopyData
  {
    std::vector<double>    S, g;
    std::vector<long long> dofs;
    int                    ndir = 0;
    std::vector<long long> didx;
    std::vector<double>    dval;
    bool                   ok = true;
  };

  void restricted_dofs(const orc_ctx *C, long long cell, long long *dofs)
  {
    const int tdf = C->ne * C->nf;
    for (int lf = 0; lf < C->nfaces_cell; ++lf)
      {
        const long long f = cell_face(C, cell, lf);
        for (int i = 0; i < tdf; ++i)
          dofs[lf * tdf + i] = f * tdf + i;
      }
  }

  bool worker_pass1(Work &W, long long cell, bool thermo, CopyData &cd)
  {
    const int T = W.T;
    W.assemble(cell, thermo, false, nullptr);
    cd.S.resize(T * T);
    cd.g.resize(T);
    cd.dofs.resize(T);
    cd.didx.resize(T);
    cd.dval.resize(T);
    cd.ok = W.condense();
    std::copy(W.TT.begin(), W.TT.end(), cd.S.begin());
    std::copy(W.t.begin(), W.t.end(), cd.g.begin());
    restricted_dofs(W.C, cell, cd.dofs.data());
    cd.ndir = W.n_constr;
    for (int i = 0; i < W.n_constr; ++i)
      {
        cd.didx[i] = cd.dofs[W.constr_dofs[i]];
        cd.dval[i] = W.constr_vals[i];
      }
    return cd.ok;
  }

  // copy_local_to_global (src/np_solver.cpp:1293-1322) with
  // AffineConstraints::distribute_local_to_global semantics for the
  // homogeneous Dirichlet lines: constrained rows and columns are dropped,
  // the constrained diagonal receives the mean |diagonal| of the cell block
  void copier(orc_ctx *C, const CopyData &cd, std::vector<char> &is_dir)
  {
    const int T = (int)cd.dofs.size();
    for (int i = 0; i < cd.ndir; ++i)
      is_dir[cd.didx[i]] = 1;
    double diag = 0;
    for (int i = 0; i < T; ++i)
      diag += std::fabs(cd.S[i * T + i]);
    diag = diag > 0 ? diag / T : 1.0;
    for (int i = 0; i < T; ++i)
      {
        const long long gi = cd.dofs[i];
        const long long rb = C->rowptr[gi], re = C->rowptr[gi + 1];
        if (is_dir[gi])
          {
            const int *pos = std::lower_bound(&C->colind[rb], &C->colind[re],
                                              (int)gi);
            C->vals[pos - C->colind.data()] += diag;
            continue;
          }
        C->rhs[gi] += cd.g[i];
        for (int j = 0; j < T; ++j)
          {
            const long long gj = cd.dofs[j];
            if (is_dir[gj])
              continue;
            const int *pos = std::lower_bound(&C->colind[rb], &C->colind[re],
                                              (int)gj);
            C->vals[pos - C->colind.data()] += cd.S[i * T + j];
          }
      }
    for (int i = 0; i < cd.ndir; ++i)
      {
        C->dir_idx.push_back(cd.didx[i]);
        C->dir_val.push_back(cd.dval[i]);
      }
  }

  std::vector<char> dirichlet_mask(const orc_ctx *C)
  {
    // get_dofs_constrained_by_dirichlet_conditions (src/np_solver.cpp:500-548)
    const int         tdf = C->ne * C->nf;
    std::vector<char> m(C->nfaces * tdf, 0);
    for (long long cell = 0; cell < C->ncells; ++cell)
      for (int lf = 0; lf < C->nfaces_cell; ++lf)
        {
          const int id = boundary_id(C, cell, lf);
          if (id < 0)
            continue;
          int ci = 0;
          for (int comp = 0; comp < 3; ++comp)
            {
              if (!C->en[comp])
                continue;
              if (C->P.bc_type[id][comp] == 1)
                {
                  const long long f = cell_face(C, cell, lf);
                  for (int i = 0; i < C->nf; ++i)
                    m[f * tdf + ci * C->nf + i] = 1;
                }
              ++ci;
            }
        }
    return m;
  }

  // WorkStream::run stand-in (template.h:67-77): nthreads workers over
  // chunks of cells, results copied serially in cell order
  template <class WorkerFn, class CopierFn>
  void work_stream(const orc_ctx *C, long long ncells, int nthreads,
                   WorkerFn worker, CopierFn copy)
  {
    const long long chunk   = 32;
    const long long nchunks = (ncells + chunk - 1) / chunk;
    if (nthreads <= 1)
      {
        Work     W(C);
        CopyData cd;
        for (long long cell = 0; cell < ncells; ++cell)
          {
            worker(W, cell, cd);
            copy(cd);
          }
        return;
      }
    std::vector<std::vector<CopyData>> buf(nchunks);
    std::vector<std::atomic<int>>      done(nchunks);
    for (auto &x : done)
      x.store(0);
    std::atomic<long long>  next(0);
    std::mutex              mu;
    std::condition_variable cv;
    // bound the number of chunks in flight (WorkStream's queue length)
    const long long         window = 4 * nthreads;
    std::atomic<long long>  copied(0);
    auto run = [&]() {
      Work W(C);
      for (;;)
        {
          const long long ch = next.fetch_add(1);
          if (ch >= nchunks)
            break;
          while (ch - copied.load() >= window)
            std::this_thread::yield();
          const long long b = ch * chunk, e = std::min(ncells, b + chunk);
          buf[ch].resize(e - b);
          for (long long cell = b; cell < e; ++cell)
            worker(W, cell, buf[ch][cell - b]);
          {
            std::lock_guard<std::mutex> lk(mu);
            done[ch].store(1);
          }
          cv.notify_all();
        }
    };
    std::vector<std::thread> th;
    for (int i = 0; i < nthreads; ++i)
      th.emplace_back(run);
    for (long long ch = 0; ch < nchunks; ++ch)
      {
        {
          std::unique_lock<std::mutex> lk(mu);
          cv.wait(lk, [&] { return done[ch].load() == 1; });
        }
        for (auto &cd : buf[ch])
          copy(cd);
        std::vector<CopyData>().swap(buf[ch]);
        copied.store(ch + 1);
      }
    for (auto &x : th)
      x.join();
  }
} // namespace

extern "C" {

const char *orc_last_error(void)
{
  return g_err.c_str();
}

int orc_create(const orc_problem *prob, orc_ctx **out)
{
  if (!prob || !out)
    return -1;
  const orc_problem &P = *prob;
  if ((P.dim != 2 && P.dim != 3) || P.degree < 1 || P.degree > 4)
    {
      g_err = "oracle: dim must be 2 or 3 and degree in 1..4";
      return -2;
    }
  orc_ctx *C = new orc_ctx();
  C->P       = P;
  C->d       = P.dim;
  C->k       = P.degree;
  const int d = C->d, k = C->k;
  C->nb          = 1;
  C->nf          = 1;
  for (int a = 0; a < d; ++a)
    C->nb *= (k + 1);
  for (int a = 0; a < d - 1; ++a)
    C->nf *= (k + 1);
  C->Qc          = C->nb;
  C->Qf          = C->nf;
  C->nfaces_cell = 2 * d;
  C->NC          = 3 * (d + 1) * C->nb;
  C->ncells      = 1;
  for (int a = 0; a < 3; ++a)
    {
      C->n[a]  = a < d ? P.n_cells[a] : 1;
      C->lo[a] = a < d ? P.lower[a] : 0;
      C->h[a]  = a < d ? (P.upper[a] - P.lower[a]) / P.n_cells[a] : 1;
      C->ncells *= C->n[a];
    }
  C->face_off[0] = 0;
  for (int a = 0; a < 3; ++a)
    {
      for (int b = 0; b < 3; ++b)
        C->face_dims[a][b] = C->n[b] + (a == b ? 1 : 0);
      long long cnt = 0;
      if (a < d)
        cnt = C->face_dims[a][0] * C->face_dims[a][1] * C->face_dims[a][2];
      C->face_off[a + 1] = C->face_off[a] + cnt;
    }
  C->nfaces = C->face_off[d];
  if (d == 2)
    {
      C->tang[0][0] = 1;
      C->tang[1][0] = 0;
    }
  else
    {
      C->tang[0][0] = 1;
      C->tang[0][1] = 2;
      C->tang[1][0] = 2;
      C->tang[1][1] = 0;
      C->tang[2][0] = 0;
      C->tang[2][1] = 1;
    }

  // Adimensionalizer (include/adimensionalizer.h:28-262)
  const double l = P.adim[0], Tm = P.adim[1], Nm = P.adim[2], mm = P.adim[3];
  C->V_rf     = Tm * kKB / kQ;
  C->n_rf     = Nm;
  C->perm_rf  = l * l * kQ * kQ * Nm / (Tm * kKB);
  C->UT_rf    = Tm * kKB / kQ;
  C->mu_rf    = mm;
  C->rec_rf   = kKB * Nm * Tm * mm / (l * l * kQ);
  const double Tv = Tm * kKB / kQ;
  C->tau_rf[0] = kQ * Nm / Tv;
  C->tau_rf[1] = C->tau_rf[2] = Tv * mm / (l * l);
  C->neu_rf[0]                = kQ;
  C->neu_rf[1] = C->neu_rf[2] = kKB * Nm * Tm * mm / (l * l);
  C->eps_r                    = P.epsilon / C->perm_rf;
  C->mun_r                    = P.mu_n / C->mu_rf;
  C->mup_r                    = P.mu_p / C->mu_rf;
  for (int c = 0; c < 3; ++c)
    C->tau_r[c] = P.tau[c] / C->tau_rf[c];

  // shape tables
  gauss01(k + 1, C->qx1, C->qw1);
  Lagrange1D cellb, traceb;
  for (int i = 0; i <= k; ++i)
    cellb.nd.push_back((double)i / k);
  traceb.nd  = lobatto01(k + 1);
  const int nb = C->nb, nf = C->nf, Qc = C->Qc, Qf = C->Qf;
  C->qw_cell.resize(Qc);
  C->phi.resize(Qc * nb);
  C->dphi.resize(Qc * nb * d);
  for (int q = 0; q < Qc; ++q)
    {
      int qi[3] = {0, 0, 0}, r = q;
      for (int a = 0; a < d; ++a)
        {
          qi[a] = r % (k + 1);
          r /= (k + 1);
        }
      double w = 1;
      for (int a = 0; a < d; ++a)
        w *= C->qw1[qi[a]];
      C->qw_cell[q] = w;
      for (int j = 0; j < nb; ++j)
        {
          int ji[3] = {0, 0, 0}, rr = j;
          for (int a = 0; a < d; ++a)
            {
              ji[a] = rr % (k + 1);
              rr /= (k + 1);
            }
          double v = 1;
          for (int a = 0; a < d; ++a)
            v *= cellb.val(ji[a], C->qx1[qi[a]]);
          C->phi[q * nb + j] = v;
          for (int a = 0; a < d; ++a)
            {
              double g = 1;
              for (int b = 0; b < d; ++b)
                g *= (a == b) ? cellb.der(ji[b], C->qx1[qi[b]])
                              : cellb.val(ji[b], C->qx1[qi[b]]);
              C->dphi[(q * nb + j) * d + a] = g;
            }
        }
    }
  C->qw_face.resize(Qf);
  C->xi.resize(Qf * nf);
  for (int s = 0; s < Qf; ++s)
    {
      int si[2] = {0, 0}, r = s;
      for (int m = 0; m < d - 1; ++m)
        {
          si[m] = r % (k + 1);
          r /= (k + 1);
        }
      double w = 1;
      for (int m = 0; m < d - 1; ++m)
        w *= C->qw1[si[m]];
      C->qw_face[s] = w;
      for (int j = 0; j < nf; ++j)
        {
          int ji[2] = {0, 0}, rr = j;
          for (int m = 0; m < d - 1; ++m)
            {
              ji[m] = rr % (k + 1);
              rr /= (k + 1);
            }
          double v = 1;
          for (int m = 0; m < d - 1; ++m)
            v *= traceb.val(ji[m], C->qx1[si[m]]);
          C->xi[s * nf + j] = v;
        }
    }
  C->phif.resize(2 * d * Qf * nb);
  for (int lf = 0; lf < 2 * d; ++lf)
    {
      const int a = lf / 2, side = lf % 2;
      for (int s = 0; s < Qf; ++s)
        {
          double pt[3] = {0, 0, 0};
          pt[a]        = side;
          int r        = s;
          for (int m = 0; m < d - 1; ++m)
            {
              pt[C->tang[a][m]] = C->qx1[r % (k + 1)];
              r /= (k + 1);
            }
          for (int j = 0; j < nb; ++j)
            {
              int ji[3] = {0, 0, 0}, rr = j;
              for (int b = 0; b < d; ++b)
                {
                  ji[b] = rr % (k + 1);
                  rr /= (k + 1);
                }
              double v = 1;
              for (int b = 0; b < d; ++b)
                v *= cellb.val(ji[b], pt[b]);
              C->phif[(lf * Qf + s) * nb + j] = v;
            }
        }
    }
  C->en[0] = C->en[1] = C->en[2] = 1;
  C->doping.assign(C->ncells * Qc, 0.0);
  C->dir_vals.assign(C->nfaces * 3 * Qf, 0.0);
  C->neu_vals.assign(C->nfaces * 3 * Qf, 0.0);
  C->x_cell.assign(C->ncells * C->NC, 0.0);
  C->x_trace.assign(C->nfaces * 3 * nf, 0.0);
  setup_enabled(C);
  *out = C;
  return 0;
}

void orc_destroy(orc_ctx *ctx)
{
  delete ctx;
}

void orc_sizes(const orc_ctx *C, long long out[12])
{
  out[0]  = C->ncells;
  out[1]  = C->nfaces;
  out[2]  = C->NC;
  out[3]  = C->Qc;
  out[4]  = C->Qf;
  out[5]  = C->nb;
  out[6]  = C->nf;
  out[7]  = C->ne;
  out[8]  = C->N;
  out[9]  = C->T;
  out[10] = C->nfaces * C->ne * C->nf;
  out[11] = csr_entries(C);
}

int orc_set_enabled(orc_ctx *C, int V, int n, int p)
{
  if (!(V || n || p))
    {
      g_err = "oracle: no component enabled";
      return -1;
    }
  C->en[0] = V != 0;
  C->en[1] = n != 0;
  C->en[2] = p != 0;
  setup_enabled(C);
  return 0;
}

void orc_cell_qpoints(const orc_ctx *C, double *xyz)
{
  const int d = C->d, k = C->k;
  for (long long cell = 0; cell < C->ncells; ++cell)
    {
      int ic[3];
      cell_coords(C, cell, ic);
      for (int q = 0; q < C->Qc; ++q)
        {
          int r = q;
          for (int a = 0; a < d; ++a)
            {
              const int qa = r % (k + 1);
              r /= (k + 1);
              xyz[(cell * C->Qc + q) * d + a] =
                C->lo[a] + (ic[a] + C->qx1[qa]) * C->h[a];
            }
        }
    }
}

void orc_face_qpoints(const orc_ctx *C, double *xyz)
{
  const int d = C->d, k = C->k;
  for (int a = 0; a < d; ++a)
    {
      const long long *m = C->face_dims[a];
      for (long long fi = 0; fi < C->face_off[a + 1] - C->face_off[a]; ++fi)
        {
          int pos[3];
          pos[0]            = (int)(fi % m[0]);
          pos[1]            = (int)((fi / m[0]) % m[1]);
          pos[2]            = (int)(fi / (m[0] * m[1]));
          const long long f = C->face_off[a] + fi;
          for (int s = 0; s < C->Qf; ++s)
            {
              double pt[3];
              for (int b = 0; b < d; ++b)
                pt[b] = C->lo[b] + pos[b] * C->h[b];
              int r = s;
              for (int mm = 0; mm < d - 1; ++mm)
                {
                  const int b = C->tang[a][mm];
                  pt[b] += C->qx1[r % (k + 1)] * C->h[b];
                  r /= (k + 1);
                }
              for (int b = 0; b < d; ++b)
                xyz[(f * C->Qf + s) * d + b] = pt[b];
            }
        }
    }
}

void orc_set_doping(orc_ctx *C, const double *doping_qp)
{
  // Adimensionalizer::adimensionalize_doping_function (cpp:450-456)
  for (long long i = 0; i < C->ncells * C->Qc; ++i)
    C->doping[i] = doping_qp[i] / C->n_rf;
}

void orc_set_bc_values(orc_ctx *C, const double *dir, const double *neu)
{
  const long long n = C->nfaces * 3 * C->Qf;
  if (dir)
    std::copy(dir, dir + n, C->dir_vals.begin());
  if (neu)
    std::copy(neu, neu + n, C->neu_vals.begin());
}

// CellFaceTauNPSolverParameters::set_face values for every (cell, face):
// tau[(cell*2d + face)*3 + comp], physical; rescaled as CellFaceTauComputer
// does (src/np_solver_parameters.cpp:79-95, include/np_solver_parameters.h:
// 157-236).  NULL returns to the fixed tau of the problem.
void orc_set_cell_face_tau(orc_ctx *C, const double *tau)
{
  if (!tau)
    {
      C->tau_cf.clear();
      return;
    }
  const long long n = C->ncells * C->nfaces_cell * 3;
  C->tau_cf.resize(n);
  for (long long i = 0; i < n; ++i)
    C->tau_cf[i] = tau[i] / C->tau_rf[i % 3];
}

void orc_set_state(orc_ctx *C, const double *cell, const double *trace)
{
  if (cell)
    std::copy(cell, cell + C->x_cell.size(), C->x_cell.begin());
  if (trace)
    std::copy(trace, trace + C->x_trace.size(), C->x_trace.begin());
}

void orc_get_state(const orc_ctx *C, double *cell, double *trace)
{
  if (cell)
    std::copy(C->x_cell.begin(), C->x_cell.end(), cell);
  if (trace)
    std::copy(C->x_trace.begin(), C->x_trace.end(), trace);
}

int orc_local_system(orc_ctx *C, int cell, int thermo, double *A, double *CT,
                     double *TC, double *TT, double *c, double *t)
{
  if (cell < 0 || cell >= C->ncells)
    return -1;
  Work W(C);
  W.assemble(cell, thermo != 0, false, nullptr);
  std::copy(W.A.begin(), W.A.end(), A);
  std::copy(W.CT.begin(), W.CT.end(), CT);
  std::copy(W.TC.begin(), W.TC.end(), TC);
  std::copy(W.TT.begin(), W.TT.end(), TT);
  std::copy(W.c.begin(), W.c.end(), c);
  std::copy(W.t.begin(), W.t.end(), t);
  return 0;
}

int orc_condensed_cell(orc_ctx *C, int cell, int thermo, double *S, double *g,
                       long long *dofs, int *n_dir, long long *dir_idx,
                       double *dir_val)
{
  if (cell < 0 || cell >= C->ncells)
    return -1;
  Work     W(C);
  CopyData cd;
  if (!worker_pass1(W, cell, thermo != 0, cd))
    {
      g_err = "oracle: singular cell matrix";
      return -2;
    }
  std::copy(cd.S.begin(), cd.S.end(), S);
  std::copy(cd.g.begin(), cd.g.end(), g);
  std::copy(cd.dofs.begin(), cd.dofs.end(), dofs);
  *n_dir = cd.ndir;
  for (int i = 0; i < cd.ndir; ++i)
    {
      dir_idx[i] = cd.didx[i];
      dir_val[i] = cd.dval[i];
    }
  return 0;
}

void orc_csr_pattern(const orc_ctx *C, long long *rowptr, int *colind)
{
  ensure_pattern(const_cast<orc_ctx *>(C));
  std::copy(C->rowptr.begin(), C->rowptr.end(), rowptr);
  std::copy(C->colind.begin(), C->colind.end(), colind);
}

double orc_pass1(orc_ctx *C, int thermo, int nthreads, long long limit)
{
  ensure_pattern(C);
  std::fill(C->vals.begin(), C->vals.end(), 0.0);
  std::fill(C->rhs.begin(), C->rhs.end(), 0.0);
  C->dir_idx.clear();
  C->dir_val.clear();
  const long long ncells = (limit < 0 || limit > C->ncells) ? C->ncells : limit;
  std::vector<char> is_dir = dirichlet_mask(C);
  bool              ok     = true;
  const auto        t0     = std::chrono::steady_clock::now();
  work_stream(
    C, ncells, nthreads,
    [&](Work &W, long long cell, CopyData &cd) {
      worker_pass1(W, cell, thermo != 0, cd);
    },
    [&](const CopyData &cd) {
      if (!cd.ok)
        ok = false;
      copier(C, cd, is_dir);
    });
  const auto t1 = std::chrono::steady_clock::now();
  if (!ok)
    {
      g_err = "oracle: singular cell matrix";
      return -1;
    }
  return std::chrono::duration<double>(t1 - t0).count();
}

void orc_get_system(const orc_ctx *C, double *vals, double *rhs)
{
  ensure_pattern(const_cast<orc_ctx *>(C));
  if (vals)
    std::copy(C->vals.begin(), C->vals.end(), vals);
  if (rhs)
    std::copy(C->rhs.begin(), C->rhs.end(), rhs);
}

int orc_n_dirichlet(const orc_ctx *C)
{
  return (int)C->dir_idx.size();
}

void orc_get_dirichlet(const orc_ctx *C, long long *idx, double *val)
{
  std::copy(C->dir_idx.begin(), C->dir_idx.end(), idx);
  std::copy(C->dir_val.begin(), C->dir_val.end(), val);
}

double orc_pass2(orc_ctx *C, int thermo, int nthreads,
                 const double *upd_trace, double *upd_cell)
{
  const auto t0 = std::chrono::steady_clock::now();
  bool       ok = true;
  std::fill(upd_cell, upd_cell + C->ncells * C->NC, 0.0);
  // cell writes are disjoint, so the copier has nothing to do
  work_stream(
    C, C->ncells, nthreads,
    [&](Work &W, long long cell, CopyData &cd) {
      W.assemble(cell, thermo != 0, true, upd_trace);
      if (!gj_invert(W.A, W.N, W.piv, W.gjtmp))
        {
          cd.ok = false;
          return;
        }
      double *out = upd_cell + cell * C->NC;
      for (int i = 0; i < W.N; ++i)
        {
          double s = 0;
          for (int j = 0; j < W.N; ++j)
            s += W.A[i * W.N + j] * W.c[j];
          out[C->enabled_idx[i]] = s;
        }
      cd.ok = true;
    },
    [&](const CopyData &cd) {
      if (!cd.ok)
        ok = false;
    });
  const auto t1 = std::chrono::steady_clock::now();
  if (!ok)
    {
      g_err = "oracle: singular cell matrix";
      return -1;
    }
  return std::chrono::duration<double>(t1 - t0).count();
}

// Terminal current (a quantity this build defines; the reference has no
// contact-current integral): the numerical fluxes of its trace equations
// (add_tc_matrix_terms_to_tt_rhs, template.h:1311-1447)
//   F_n =  n^ mu_n E.nu - U_T mu_n W_n.nu - tau_n (n - lambda_n)
//   F_p = -p^ mu_p E.nu - U_T mu_p W_p.nu - tau_p (p - lambda_p)
// with n^ = n (use_cell), lambda_n (use_trace) or, for qiu_shi, n where
// E.nu > 0 and lambda_n elsewhere (template.h:1417, 1435); integrated over
// the faces of one boundary id, outward normal.
int orc_terminal_current(orc_ctx *C, int bid, double out[2])
{
  if (bid < 0 || bid >= 2 * C->P.dim)
    {
      g_err = "oracle: boundary id out of range";
      return -1;
    }
  Work   W(C);
  double sn = 0, sp = 0;
  for (long long cell = 0; cell < C->ncells; ++cell)
    {
      if (boundary_id(C, cell, bid) != bid)
        continue;
      W.cur = cell;
      W.prepare_face(cell, bid, false, nullptr);
      const int ft = C->P.flux_type;
      for (int s = 0; s < C->Qf; ++s)
        {
          const double En  = dot(W.fE0[s], W.normal, C->P.dim);
          const double Wnn = dot(W.fWn0[s], W.normal, C->P.dim);
          const double Wpn = dot(W.fWp0[s], W.normal, C->P.dim);
          const bool   cl  = ft == 0 || (ft == 2 && En > 0);
          const double nn = cl ? W.fn0[s] : W.trn0[s];
          const double pp = cl ? W.fp0[s] : W.trp0[s];
          sn += W.fJxW[s] * (nn * C->mun_r * En - W.UTf[s] * C->mun_r * Wnn -
                             W.tau(bid, CN) * (W.fn0[s] - W.trn0[s]));
          sp += W.fJxW[s] * (-pp * C->mup_r * En - W.UTf[s] * C->mup_r * Wpn -
                             W.tau(bid, CP) * (W.fp0[s] - W.trp0[s]));
        }
    }
  out[0] = sn;
  out[1] = sp;
  return 0;
}

void orc_recombination(const orc_problem *prob, double n, double p, double rf,
                       double out[3])
{
  orc_ctx C;
  C.P = *prob;
  recombination(&C, n, p, rf, out[0], out[1], out[2]);
}

} // extern "C"
