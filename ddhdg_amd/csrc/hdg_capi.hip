// hdg_capi.hip -- C ABI of the gfx950 per-cell HDG path (include/ddhdg_hip.h).
//
// Host side of a context: structured-mesh topology, the restricted trace
// matrix pattern and its per-cell scatter map, the reference-cell tables and
// the launches of the kernels in hdg_kernels.hip.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "hdg_ctx.h"
#include "hdg_nd_plan.h"

namespace
{
  thread_local std::string g_err;

  int fail(const std::string &m, int code = -1)
  {
    g_err = m;
    return code;
  }
} // namespace

int hdg::set_error(const std::string &msg, int code)
{
  g_err = msg;
  return code;
}

namespace
{

  int hip_fail(hipError_t e, const char *what)
  {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return -3;
  }
} // namespace

namespace
{
  using hdg::boundary_id;
  using hdg::cell_coords;
  using hdg::cell_face;

  void set_enabled(hdg_ctx *C, int V, int n, int p)
  {
    C->en[0] = V != 0;
    C->en[1] = n != 0;
    C->en[2] = p != 0;
    C->ne    = 0;
    for (int c = 0; c < 3; ++c)
      if (C->en[c])
        C->encomp[C->ne++] = c;
    for (int c = C->ne; c < 3; ++c)
      C->encomp[c] = 0;
  }

  template <class T>
  int upload(T **dst, const std::vector<T> &src)
  {
    if (*dst)
      {
        hipFree(*dst);
        *dst = nullptr;
      }
    if (src.empty())
      return 0;
    hipError_t e = hipMalloc((void **)dst, src.size() * sizeof(T));
    if (e != hipSuccess)
      return hip_fail(e, "hipMalloc");
    e = hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess)
      return hip_fail(e, "hipMemcpy");
    return 0;
  }

  // the default operator: use_cell flux, fixed tau, no phi_linearize (the
  // equilibrium operator ignores the flux and phi_linearize)
  bool default_variant(const hdg_ctx *C, int thermodyn)
  {
    if (C->d_taucf)
      return false;
    return thermodyn || (C->phys.flux_type == 0 && !C->phys.phi_linearize);
  }

  // pass 1 on a register-resident kernel (2D degree 1-3, 3D degree 1)
  bool use_fast(const hdg_ctx *C, int thermodyn)
  {
    if (C->variant == 1 || !default_variant(C, thermodyn))
      return false;
    return hdg::fast_supported(C->d, C->k, C->ne, thermodyn, C->en) ||
           hdg::fast2p_supported(C->d, C->k, C->ne, thermodyn, C->en) ||
           hdg::fast3_supported(C->d, C->k, C->ne, thermodyn, C->en);
  }

  // pass 2 on the register-resident kernels (2D degree 1-3, 3D degree 1)
  bool use_fast_pass2(const hdg_ctx *C, int thermodyn)
  {
    return use_fast(C, thermodyn) && ((C->d == 2 && C->k <= 3) || (C->d == 3 && C->k == 1));
  }

  hipError_t launch_pass1(const hdg_ctx *C, bool fast, const hdg::KParams &p, const int *cells,
                          int ncell, int color, int mode, hipStream_t s)
  {
    if (!fast)
      return hdg::launch_condense(C->d, C->k, p, cells, ncell, color, mode, s);
    if (C->d == 3)
      return hdg::launch_fast3_condense(p, cells, ncell, color, mode, s);
    if (C->k == 3)
      return hdg::launch_fast2p_condense(p, cells, ncell, color, mode, s);
    return hdg::launch_fast_condense(C->k, p, cells, ncell, color, mode, s);
  }

  hdg::KParams params(const hdg_ctx *C, int thermodyn)
  {
    hdg::KParams p{};
    for (int a = 0; a < 3; ++a)
      p.n[a] = C->n[a];
    for (int a = 0; a < 4; ++a)
      p.face_off[a] = C->face_off[a];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        p.face_dims[a][b] = C->face_dims[a][b];
    for (int c = 0; c < 3; ++c)
      {
        p.en[c]     = C->en[c];
        p.encomp[c] = C->encomp[c];
        p.tau[c]    = C->phys.tau[c];
      }
    p.ne     = C->ne;
    p.thermo = thermodyn ? 1 : 0;
    p.eps    = C->phys.eps;
    p.mun    = C->phys.mu_n;
    p.mup    = C->phys.mu_p;
    p.UT     = C->phys.U_T;
    p.rec    = C->phys.recombination;
    p.rec0   = C->phys.rec[0];
    p.rec1   = C->phys.rec[1];
    p.rec2   = C->phys.rec[2];
    p.n_rf   = C->phys.n_rescale;
    p.rec_rf = C->phys.rec_rescale;
    p.kTq    = C->phys.kT_over_q;
    p.ec_r   = C->phys.ec_r;
    p.ev_r   = C->phys.ev_r;
    p.nc     = C->phys.nc;
    p.nv     = C->phys.nv;
    for (int b = 0; b < 6; ++b)
      for (int c = 0; c < 3; ++c)
        p.bc[b][c] = C->phys.bc_type[b][c];
    p.sdrift = C->phys.drift_jacobian ? -1.0 : 1.0;
    p.mixed  = C->precision;
    p.deg    = C->k;
    p.flux_type = C->phys.flux_type;
    p.phi_lin   = C->phys.phi_linearize ? 1 : 0;
    p.tau_cf    = C->d_taucf;
    p.rgscr     = C->d_rgscr;
    p.tb     = C->tb;
    p.doping = C->d_doping;
    p.dirv   = C->d_dirv;
    p.neuv   = C->d_neuv;
    p.blkptr = C->d_blkptr;
    p.nbk    = C->d_nbk;
    p.cslot  = C->d_cslot;
    return p;
  }
} // namespace

extern "C" {

const char *hdg_last_error(void)
{
  return g_err.c_str();
}

int hdg_ctx_create(const hdg_mesh_desc *mesh, const hdg_phys *phys, hdg_ctx **out)
{
  if (!mesh || !phys || !out)
    return fail("hdg_ctx_create: null argument");
  if (!hdg::supported(mesh->dim, mesh->degree))
    return fail("hdg_ctx_create: (dim, degree) not supported: 2D k=1..3, 3D k=1..2", -2);
  if (phys->flux_type < 0 || phys->flux_type > 2)
    return fail("hdg_ctx_create: flux_type must be 0 (use_cell), 1 (use_trace) or 2 "
                "(qiu_shi_stabilization)", -2);
  for (int a = 0; a < mesh->dim; ++a)
    if (mesh->n_cells[a] < 1 || !(mesh->upper[a] > mesh->lower[a]))
      return fail("hdg_ctx_create: invalid mesh extent", -2);
  hdg_ctx *C = new hdg_ctx();
  C->mesh    = *mesh;
  C->phys    = *phys;
  C->d       = mesh->dim;
  C->k       = mesh->degree;
  C->ncells  = 1;
  for (int a = 0; a < 3; ++a)
    {
      C->n[a]  = a < C->d ? mesh->n_cells[a] : 1;
      C->lo[a] = a < C->d ? mesh->lower[a] : 0;
      C->h[a]  = a < C->d ? (mesh->upper[a] - mesh->lower[a]) / mesh->n_cells[a] : 1;
      C->ncells *= C->n[a];
    }
  hdg::build_topology(C->d, mesh->n_cells, C->topo);
  C->nfaces = C->topo.nfaces;
  for (int a = 0; a < 4; ++a)
    C->face_off[a] = C->topo.face_off[a];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      C->face_dims[a][b] = C->topo.face_dims[a][b];
  if (!hdg::build_tables(C->d, C->k, C->h, C->ht))
    {
      delete C;
      return fail("hdg_ctx_create: singular reference-cell mass matrix");
    }
  const hdg::HostTables &T = C->ht;
  C->nb  = T.nb;
  C->nf  = T.nf;
  C->qc  = T.qc;
  C->qf  = T.qf;
  C->nfc = T.nfc;
  C->NC  = 3 * (C->d + 1) * C->nb;
  set_enabled(C, 1, 1, 1);
  for (long long c = 0; c < C->ncells; ++c)
    {
      int ic[3];
      cell_coords(C, c, ic);
      C->color_cells[(ic[0] + ic[1] + ic[2]) & 1].push_back((int)c);
    }

  // device tables, one allocation
  std::vector<const std::vector<double> *> parts = {
    &T.JxWc, &T.phic, &T.gradc, &T.phif, &T.JxWf, &T.xi, &T.Bz, &T.Bt, &T.LDc,
    &T.LDf,  &T.LCc,  &T.LCf,   &T.PhiMinvC, &T.PhiMinvF, &T.Minv, &T.MinvD, &T.MinvC,
    &T.Mfinv};
  std::vector<double> all;
  std::vector<size_t> off;
  for (auto *v : parts)
    {
      off.push_back(all.size());
      all.insert(all.end(), v->begin(), v->end());
      all.resize((all.size() + 31) / 32 * 32, 0.0); // 256-byte alignment
    }
  int rc = upload(&C->d_tab, all);
  if (rc == 0)
    {
      const double *b    = C->d_tab;
      hdg::DevTables &tb = C->tb;
      tb.JxWc            = b + off[0];
      tb.phic            = b + off[1];
      tb.gradc           = b + off[2];
      tb.phif            = b + off[3];
      tb.JxWf            = b + off[4];
      tb.xi              = b + off[5];
      tb.Bz              = b + off[6];
      tb.Bt              = b + off[7];
      tb.LDc             = b + off[8];
      tb.LDf             = b + off[9];
      tb.LCc             = b + off[10];
      tb.LCf             = b + off[11];
      tb.PhiMinvC        = b + off[12];
      tb.PhiMinvF        = b + off[13];
      tb.Minv            = b + off[14];
      tb.MinvD           = b + off[15];
      tb.MinvC           = b + off[16];
      tb.Mfinv           = b + off[17];
    }
  std::vector<long long> bp(C->topo.blkptr.begin(), C->topo.blkptr.end() - 1);
  if (rc == 0)
    rc = upload(&C->d_blkptr, bp);
  if (rc == 0)
    rc = upload(&C->d_nbk, C->topo.nbk);
  if (rc == 0)
    rc = upload(&C->d_cslot, C->topo.cslot);
  for (int col = 0; col < 2 && rc == 0; ++col)
    rc = upload(&C->d_cells[col], C->color_cells[col]);
  if (rc == 0)
    rc = upload(&C->d_doping, std::vector<double>(C->ncells * C->qc, 0.0));
  if (rc == 0)
    rc = upload(&C->d_dirv, std::vector<double>(C->nfaces * 3 * C->qf, 0.0));
  if (rc == 0)
    rc = upload(&C->d_neuv, std::vector<double>(C->nfaces * 3 * C->qf, 0.0));
  if (rc == 0 && hdg::generic_scratch_bytes(C->d, C->k) > 0)
    {
      // the reduced systems of the generic kernel do not fit LDS at this
      // shape: one slice per resident workgroup, allocated with the context
      hipError_t e =
        hipMalloc((void **)&C->d_rgscr, hdg::generic_scratch_bytes(C->d, C->k));
      if (e != hipSuccess)
        rc = hip_fail(e, "hdg_ctx_create: generic-kernel scratch");
    }
  if (rc != 0)
    {
      std::string m = g_err;
      hdg_ctx_destroy(C);
      return fail(m, rc);
    }
  *out = C;
  return 0;
}

void hdg_ctx_destroy(hdg_ctx *C)
{
  if (!C)
    return;
  for (void *p : {(void *)C->d_tab, (void *)C->d_doping, (void *)C->d_dirv,
                  (void *)C->d_neuv, (void *)C->d_taucf, (void *)C->d_rgscr, (void *)C->d_blkptr,
                  (void *)C->d_nbk,
                  (void *)C->d_cslot, (void *)C->d_cells[0], (void *)C->d_cells[1]})
    if (p)
      hipFree(p);
  delete C;
}

int hdg_sizes(const hdg_ctx *C, long long out[HDG_NSIZES])
{
  if (!C || !out)
    return fail("hdg_sizes: null argument");
  const long long tdf = (long long)C->ne * C->nf;
  out[HDG_NCELLS] = C->ncells;
  out[HDG_NFACES] = C->nfaces;
  out[HDG_NC]     = C->NC;
  out[HDG_QC]     = C->qc;
  out[HDG_QF]     = C->qf;
  out[HDG_NB]     = C->nb;
  out[HDG_NF]     = C->nf;
  out[HDG_NE]     = C->ne;
  out[HDG_NS]     = (long long)C->ne * C->nb;
  out[HDG_T]      = C->nfc * tdf;
  out[HDG_NTRACE] = C->nfaces * tdf;
  out[HDG_NNZ]    = C->topo.blkptr[C->nfaces] * tdf * tdf;
  return 0;
}

int hdg_set_enabled(hdg_ctx *C, int V, int n, int p)
{
  if (!C)
    return fail("hdg_set_enabled: null context");
  if (!(V || n || p))
    return fail("hdg_set_enabled: no component enabled", -2);
  set_enabled(C, V, n, p);
  return 0;
}

int hdg_set_kernel_variant(hdg_ctx *C, int variant)
{
  if (!C)
    return fail("hdg_set_kernel_variant: null context");
  if (variant < 0 || variant > 2)
    return fail("hdg_set_kernel_variant: variant must be 0, 1 or 2", -2);
  if (variant == 2 && !hdg::fast_supported(C->d, C->k, C->ne, 0, C->en) &&
      !hdg::fast_supported(C->d, C->k, C->ne, 1, C->en) &&
      !hdg::fast2p_supported(C->d, C->k, C->ne, 0, C->en) &&
      !hdg::fast3_supported(C->d, C->k, C->ne, 0, C->en) &&
      !hdg::fast3_supported(C->d, C->k, C->ne, 1, C->en))
    return fail("hdg_set_kernel_variant: no register-resident kernel for this shape", -2);
  if (variant == 2 && !default_variant(C, 1))
    return fail("hdg_set_kernel_variant: the register-resident kernel runs the default "
                "operator only (use_cell flux, fixed tau, no phi_linearize)", -2);
  C->variant = variant;
  return 0;
}

int hdg_set_precision(hdg_ctx *C, int precision)
{
  if (!C)
    return fail("hdg_set_precision: null context");
  if (precision < 0 || precision > 2)
    return fail("hdg_set_precision: precision must be 0 (FP64), 1 (FP32 + refinement in the "
                "register kernels) or 2 (in every kernel)", -2);
  C->precision = precision;
  return 0;
}

int hdg_csr_pattern(const hdg_ctx *C, long long *rowptr, int *colind)
{
  if (!C || !rowptr || !colind)
    return fail("hdg_csr_pattern: null argument");
  const long long tdf = (long long)C->ne * C->nf;
  if (C->nfaces * tdf > 2147483647LL)
    return fail("hdg_csr_pattern: trace system too large for int32 columns", -2);
  long long pos = 0;
  for (long long f = 0; f < C->nfaces; ++f)
    for (long long r = 0; r < tdf; ++r)
      {
        rowptr[f * tdf + r] = pos;
        for (long long b = C->topo.blkptr[f]; b < C->topo.blkptr[f + 1]; ++b)
          for (long long c = 0; c < tdf; ++c)
            colind[pos++] = (int)(C->topo.blk_cols[b] * tdf + c);
      }
  rowptr[C->nfaces * tdf] = pos;
  return 0;
}

int hdg_dirichlet_dofs(const hdg_ctx *C, long long *idx)
{
  if (!C)
    return fail("hdg_dirichlet_dofs: null context");
  const long long tdf = (long long)C->ne * C->nf;
  int             cnt = 0;
  for (long long c = 0; c < C->ncells; ++c)
    for (int lf = 0; lf < C->nfc; ++lf)
      {
        const int bid = boundary_id(C, c, lf);
        if (bid < 0)
          continue;
        const long long f = cell_face(C, c, lf);
        for (int ci = 0; ci < C->ne; ++ci)
          if (C->phys.bc_type[bid][C->encomp[ci]] == 1)
            for (int m = 0; m < C->nf; ++m)
              {
                if (idx)
                  idx[cnt] = f * tdf + ci * C->nf + m;
                ++cnt;
              }
      }
  return cnt;
}

int hdg_cell_qpoints(const hdg_ctx *C, double *xyz)
{
  if (!C || !xyz)
    return fail("hdg_cell_qpoints: null argument");
  const int d = C->d, k = C->k;
  for (long long c = 0; c < C->ncells; ++c)
    {
      int ic[3];
      cell_coords(C, c, ic);
      for (int q = 0; q < C->qc; ++q)
        {
          int r = q;
          for (int a = 0; a < d; ++a)
            {
              const int qa = r % (k + 1);
              r /= (k + 1);
              xyz[(c * C->qc + q) * d + a] = C->lo[a] + (ic[a] + C->ht.qx1[qa]) * C->h[a];
            }
        }
    }
  return 0;
}

int hdg_face_qpoints(const hdg_ctx *C, double *xyz)
{
  if (!C || !xyz)
    return fail("hdg_face_qpoints: null argument");
  const int d = C->d, k = C->k;
  for (int a = 0; a < d; ++a)
    {
      const long long *m = C->face_dims[a];
      for (long long fi = 0; fi < C->face_off[a + 1] - C->face_off[a]; ++fi)
        {
          const long long pos[3] = {fi % m[0], (fi / m[0]) % m[1], fi / (m[0] * m[1])};
          const long long f      = C->face_off[a] + fi;
          for (int s = 0; s < C->qf; ++s)
            {
              double pt[3];
              for (int b = 0; b < d; ++b)
                pt[b] = C->lo[b] + pos[b] * C->h[b];
              int r = s;
              for (int mm = 0; mm < d - 1; ++mm)
                {
                  const int b = C->ht.tang[a][mm];
                  pt[b] += C->ht.qx1[r % (k + 1)] * C->h[b];
                  r /= (k + 1);
                }
              for (int b = 0; b < d; ++b)
                xyz[(f * C->qf + s) * d + b] = pt[b];
            }
        }
    }
  return 0;
}

int hdg_set_doping(hdg_ctx *C, const double *doping_qp)
{
  if (!C || !doping_qp)
    return fail("hdg_set_doping: null argument");
  hipError_t e = hipMemcpy(C->d_doping, doping_qp, C->ncells * C->qc * sizeof(double),
                           hipMemcpyHostToDevice);
  return e == hipSuccess ? 0 : hip_fail(e, "hdg_set_doping");
}

int hdg_set_cell_face_tau(hdg_ctx *C, const double *tau)
{
  if (!C)
    return fail("hdg_set_cell_face_tau: null context");
  if (!tau)
    {
      if (C->d_taucf)
        hipFree(C->d_taucf);
      C->d_taucf = nullptr;
      return 0;
    }
  const long long n = C->ncells * C->nfc * 3;
  for (long long i = 0; i < n; ++i)
    if (!(tau[i] >= 0) || tau[i] > 1e300)
      return fail("hdg_set_cell_face_tau: tau values must be finite and >= 0 (cell " +
                      std::to_string(i / (C->nfc * 3)) + ", face " +
                      std::to_string((i / 3) % C->nfc) + ")",
                  -2);
  // values change in place: launches recorded in a graph keep the address
  if (!C->d_taucf)
    {
      hipError_t e = hipMalloc((void **)&C->d_taucf, n * sizeof(double));
      if (e != hipSuccess)
        return hip_fail(e, "hdg_set_cell_face_tau");
    }
  hipError_t e = hipMemcpy(C->d_taucf, tau, n * sizeof(double), hipMemcpyHostToDevice);
  return e == hipSuccess ? 0 : hip_fail(e, "hdg_set_cell_face_tau");
}

int hdg_set_bc_values(hdg_ctx *C, const double *dir, const double *neu)
{
  if (!C)
    return fail("hdg_set_bc_values: null context");
  const size_t n = C->nfaces * 3 * C->qf * sizeof(double);
  if (dir)
    {
      hipError_t e = hipMemcpy(C->d_dirv, dir, n, hipMemcpyHostToDevice);
      if (e != hipSuccess)
        return hip_fail(e, "hdg_set_bc_values");
    }
  if (neu)
    {
      hipError_t e = hipMemcpy(C->d_neuv, neu, n, hipMemcpyHostToDevice);
      if (e != hipSuccess)
        return hip_fail(e, "hdg_set_bc_values");
    }
  return 0;
}

int hdg_condense(hdg_ctx *C, const double *x_cell, const double *x_trace, int thermodyn,
                 double *csr_vals, double *rhs, double *dir_vals, void *stream)
{
  if (!C || !x_cell || !x_trace || !csr_vals || !rhs || !dir_vals)
    return fail("hdg_condense: null argument");
  if (thermodyn && (C->en[1] || C->en[2] || !C->en[0]))
    return fail("hdg_condense: thermodynamic equilibrium runs with V only", -2);
  hdg::KParams p = params(C, thermodyn);
  p.xc     = x_cell;
  p.xt     = x_trace;
  p.vals   = csr_vals;
  p.rhs    = rhs;
  p.dirout = dir_vals;
  hipStream_t s    = (hipStream_t)stream;
  const bool  fast = use_fast(C, thermodyn);
  for (int col = 0; col < 2; ++col)
    {
      const int  nc = (int)C->color_cells[col].size();
      hipError_t e  = launch_pass1(C, fast, p, C->d_cells[col], nc, col, 0, s);
      if (e != hipSuccess)
        return hip_fail(e, "hdg_condense launch");
    }
  return 0;
}

int hdg_reconstruct(hdg_ctx *C, const double *x_cell, const double *x_trace,
                    const double *dx_trace, int thermodyn, double *dx_cell, void *stream)
{
  if (!C || !x_cell || !x_trace || !dx_trace || !dx_cell)
    return fail("hdg_reconstruct: null argument");
  if (thermodyn && (C->en[1] || C->en[2] || !C->en[0]))
    return fail("hdg_reconstruct: thermodynamic equilibrium runs with V only", -2);
  hdg::KParams p = params(C, thermodyn);
  p.xc  = x_cell;
  p.xt  = x_trace;
  p.dxt = dx_trace;
  p.dxc = dx_cell;
  hipStream_t s = (hipStream_t)stream;
  hipError_t  e = !use_fast_pass2(C, thermodyn)
                    ? hdg::launch_reconstruct(C->d, C->k, p, (int)C->ncells, s)
                  : C->d == 3 ? hdg::launch_fast3_reconstruct(p, (int)C->ncells, s)
                  : C->k == 3 ? hdg::launch_fast2p_reconstruct(p, (int)C->ncells, s)
                              : hdg::launch_fast_reconstruct(C->k, p, (int)C->ncells, s);
  return e == hipSuccess ? 0 : hip_fail(e, "hdg_reconstruct launch");
}

int hdg_cell_systems(hdg_ctx *C, const double *x_cell, const double *x_trace, int thermodyn,
                     const int *cells, int ncell, double *S, double *g, void *stream)
{
  if (!C || !x_cell || !x_trace || !cells || !S || !g)
    return fail("hdg_cell_systems: null argument");
  if (thermodyn && (C->en[1] || C->en[2] || !C->en[0]))
    return fail("hdg_cell_systems: thermodynamic equilibrium runs with V only", -2);
  hdg::KParams p = params(C, thermodyn);
  p.xc   = x_cell;
  p.xt   = x_trace;
  p.Sout = S;
  p.gout = g;
  hipError_t e = launch_pass1(C, use_fast(C, thermodyn), p, cells, ncell, 0, 1,
                              (hipStream_t)stream);
  return e == hipSuccess ? 0 : hip_fail(e, "hdg_cell_systems launch");
}

int hdg_local_system(hdg_ctx *C, const double *x_cell, const double *x_trace, int thermodyn,
                     const int *cells, int ncell, double *A, double *CT, double *TC, double *TT,
                     double *c, double *t, void *stream)
{
  if (!C || !x_cell || !x_trace || !cells || !A || !CT || !TC || !TT || !c || !t || ncell < 0)
    return fail("hdg_local_system: null argument");
  if (thermodyn && (C->en[1] || C->en[2] || !C->en[0]))
    return fail("hdg_local_system: thermodynamic equilibrium runs with V only", -2);
  hdg::KParams p = params(C, thermodyn);
  p.xc           = x_cell;
  p.xt           = x_trace;
  hipStream_t s  = (hipStream_t)stream;
  // the column forms go through a workspace: cells in chunks of at most
  // 256 MB of it
  const size_t per   = hdg::local_system_workspace(C->d, C->k, C->ne);
  const int    chunk = (int)std::max<size_t>(1, std::min<size_t>(ncell, (256u << 20) / per));
  double      *G     = nullptr;
  hipError_t   e     = hipMalloc((void **)&G, per * std::max(chunk, 1));
  if (e != hipSuccess)
    return hip_fail(e, "hdg_local_system workspace");
  const long long N = (long long)C->ne * (C->d + 1) * C->nb, T = (long long)C->nfc * C->ne * C->nf;
  for (int a = 0; a < ncell && e == hipSuccess; a += chunk)
    {
      const int m = std::min(chunk, ncell - a);
      e = hdg::launch_local_system(C->d, C->k, p, cells + a, m, G, A + a * N * N, CT + a * N * T,
                                   TC + a * T * N, TT + a * T * T, c + a * N, t + a * T, s);
    }
  const hipError_t e2 = hipStreamSynchronize(s);
  (void)hipFree(G);
  if (e == hipSuccess)
    e = e2;
  return e == hipSuccess ? 0 : hip_fail(e, "hdg_local_system launch");
}

int hdg_terminal_current(hdg_ctx *C, const double *x_cell, const double *x_trace,
                         int boundary_id, double *out, void *stream)
{
  if (!C || !x_cell || !x_trace || !out)
    return fail("hdg_terminal_current: null argument");
  if (boundary_id < 0 || boundary_id >= 2 * C->d)
    return fail("hdg_terminal_current: boundary id out of range", -2);
  hdg::KParams p = params(C, 0);
  p.xc           = x_cell;
  p.xt           = x_trace;
  hipError_t e   = hdg::launch_current(C->d, C->nb, C->nf, C->qf, p, boundary_id, out,
                                       (hipStream_t)stream);
  return e == hipSuccess ? 0 : hip_fail(e, "hdg_terminal_current launch");
}

} // extern "C"

// several contexts condensed by one persistent launch per colour
struct hdg_batch
{
  std::vector<hdg_ctx *> ctx;
  int                    ne = 0, thermo_ok = 0, full_ok = 0, wpb = 4, ncu = 256;
  int                    en[3] = {0, 0, 0}; // enabled components at creation
  int                    deg3 = 0; // degree-3 meshes: the pair kernel
  int                   *d_chunk_dev[2] = {nullptr, nullptr}, *d_items[2] = {nullptr, nullptr};
  int                    nchunks[2] = {0, 0};
  long long              ncells = 0;
  hdg::KParams          *d_params = nullptr, *h_params = nullptr; // h_params pinned
  hipEvent_t             copied   = nullptr;
  // the parameter blocks of the last upload: an unchanged call (the same
  // buffers and data) skips the upload, which a stream capture cannot hold
  std::vector<hdg::KParams> last;
  int                    grid     = 0; // WGs per launch; 0 until tuned
  int                    tile = 4; // waves per workgroup (register batches)
  // the autotuner's candidates: (tile, grid, ms of a two-colour pass)
  int                    ncand = 0, cand_tile[8] = {}, cand_grid[8] = {};
  float                  cand_ms[8] = {};
};

namespace
{
  int batch_launch(hdg_batch *B, int thermodyn, int mixed, int grid, int tile, hipStream_t s)
  {
    for (int col = 0; col < 2; ++col)
      {
        hipError_t e =
          B->deg3 ? hdg::launch_batch_pair(B->d_params, B->d_chunk_dev[col], B->d_items[col],
                                           B->nchunks[col], col, grid, mixed, s)
                  : hdg::launch_batch_condense(B->ne, thermodyn, mixed, B->d_params,
                                               B->d_chunk_dev[col], B->d_items[col],
                                               B->nchunks[col], col, grid, tile, s);
        if (e != hipSuccess)
          return hip_fail(e, "hdg_batch_condense launch");
      }
    return 0;
  }
} // namespace

extern "C" {

int hdg_batch_create(hdg_ctx *const *ctxs, int n, hdg_batch **out)
{
  if (!ctxs || n < 1 || !out)
    return fail("hdg_batch_create: invalid argument");
  hdg_batch *B = new hdg_batch();
  B->ne        = ctxs[0] ? ctxs[0]->ne : 0;
  for (int c = 0; c < 3; ++c)
    B->en[c] = ctxs[0] ? ctxs[0]->en[c] : 0;
  B->deg3      = ctxs[0] && ctxs[0]->k == 3;
  for (int i = 0; i < n; ++i)
    {
      hdg_ctx   *C  = ctxs[i];
      const bool ok = C && C->d == 2 && C->ne == B->ne && C->en[0] == ctxs[0]->en[0] &&
                      C->en[1] == ctxs[0]->en[1] && C->en[2] == ctxs[0]->en[2] &&
                      (B->deg3 ? (C->k == 3 && C->ne == 3) : (C->k >= 1 && C->k <= 2));
      if (!ok)
        {
          delete B;
          return fail("hdg_batch_create: contexts must be 2D with the same enabled components, "
                      "all of degree 1 or 2, or all of degree 3 with V, n and p", -2);
        }
      B->ctx.push_back(C);
      B->ncells += C->ncells;
    }
  B->wpb = B->deg3 ? hdg::batch_pairs_per_group() : hdg::batch_waves_per_group();
  int rc = 0;
  for (int col = 0; col < 2 && rc == 0; ++col)
    {
      std::vector<int> cdev, items;
      for (int i = 0; i < n; ++i)
        {
          const auto &cl = B->ctx[i]->color_cells[col];
          for (size_t a = 0; a < cl.size(); a += B->wpb)
            {
              cdev.push_back(i);
              for (int w = 0; w < B->wpb; ++w)
                items.push_back(a + w < cl.size() ? cl[a + w] : -1);
            }
        }
      B->nchunks[col] = (int)cdev.size();
      rc              = upload(&B->d_chunk_dev[col], cdev);
      if (rc == 0)
        rc = upload(&B->d_items[col], items);
    }
  hipError_t e = hipSuccess;
  if (rc == 0)
    e = hipMalloc((void **)&B->d_params, n * sizeof(hdg::KParams));
  if (rc == 0 && e == hipSuccess)
    e = hipHostMalloc((void **)&B->h_params, n * sizeof(hdg::KParams), hipHostMallocDefault);
  if (rc == 0 && e == hipSuccess)
    e = hipEventCreateWithFlags(&B->copied, hipEventDisableTiming);
  if (rc == 0 && e == hipSuccess)
    {
      int dev = 0;
      if (hipGetDevice(&dev) == hipSuccess)
        hipDeviceGetAttribute(&B->ncu, hipDeviceAttributeMultiprocessorCount, dev);
    }
  if (rc == 0 && e != hipSuccess)
    rc = hip_fail(e, "hdg_batch_create");
  if (rc != 0)
    {
      hdg_batch_destroy(B);
      return rc;
    }
  *out = B;
  return 0;
}

void hdg_batch_destroy(hdg_batch *B)
{
  if (!B)
    return;
  for (int col = 0; col < 2; ++col)
    for (void *p : {(void *)B->d_chunk_dev[col], (void *)B->d_items[col]})
      if (p)
        (void)hipFree(p);
  if (B->d_params)
    (void)hipFree(B->d_params);
  if (B->h_params)
    (void)hipHostFree(B->h_params);
  if (B->copied)
    (void)hipEventDestroy(B->copied);
  delete B;
}

int hdg_batch_condense(hdg_batch *B, int thermodyn, const double *const *x_cell,
                       const double *const *x_trace, double *const *csr_vals,
                       double *const *rhs, double *const *dir_vals, void *stream)
{
  if (!B || !x_cell || !x_trace || !csr_vals || !rhs || !dir_vals)
    return fail("hdg_batch_condense: null argument");
  if (thermodyn && (B->ne != 1 || !B->en[0]))
    return fail("hdg_batch_condense: thermodynamic equilibrium runs with V only", -2);
  if (!thermodyn && B->ne != 3)
    return fail("hdg_batch_condense: the register kernel needs V, n and p or V alone", -2);
  if (thermodyn && B->deg3)
    return fail("hdg_batch_condense: degree-3 batches run the full operator only", -2);
  hipStream_t s = (hipStream_t)stream;
  const int   n     = (int)B->ctx.size();
  int         mixed = B->ctx[0]->precision;
  std::vector<hdg::KParams> np(n);
  for (int i = 0; i < n; ++i)
    {
      if (!x_cell[i] || !x_trace[i] || !csr_vals[i] || !rhs[i] || !dir_vals[i])
        return fail("hdg_batch_condense: null device pointer");
      // a context changed after hdg_batch_create would be run with the
      // batch's component template (out of bounds in its smaller buffers)
      // or in context 0's precision
      const hdg_ctx *C = B->ctx[i];
      if (C->ne != B->ne || C->en[0] != B->en[0] || C->en[1] != B->en[1] || C->en[2] != B->en[2])
        return fail("hdg_batch_condense: context " + std::to_string(i) +
                        " changed its enabled components since hdg_batch_create",
                    -2);
      if (C->precision != mixed)
        return fail("hdg_batch_condense: context " + std::to_string(i) +
                        " has another precision than context 0",
                    -2);
      if (!default_variant(B->ctx[i], thermodyn))
        return fail("hdg_batch_condense: context " + std::to_string(i) +
                        " uses a non-default flux, tau or phi_linearize; condense it alone",
                    -2);
      hdg::KParams p = params(B->ctx[i], thermodyn);
      p.xc           = x_cell[i];
      p.xt           = x_trace[i];
      p.vals         = csr_vals[i];
      p.rhs          = rhs[i];
      p.dirout       = dir_vals[i];
      std::memcpy((void *)&np[i], (const void *)&p, sizeof(p));
    }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  hipError_t             e   = hipStreamIsCapturing(s, &cap);
  if (e != hipSuccess)
    return hip_fail(e, "hdg_batch_condense");
  const bool capturing = cap != hipStreamCaptureStatusNone;
  const bool same      = B->last.size() == (size_t)n &&
                    std::memcmp((const void *)B->last.data(), (const void *)np.data(),
                                n * sizeof(hdg::KParams)) == 0;
  if (!same)
    {
      if (capturing)
        return fail("hdg_batch_condense: a captured call must use the buffers and data of the "
                    "previous (eager) call",
                    -2);
      // the pinned parameter blocks may still feed the previous call's copy
      e = hipEventSynchronize(B->copied);
      if (e != hipSuccess)
        return hip_fail(e, "hdg_batch_condense");
      std::memcpy((void *)B->h_params, (const void *)np.data(), n * sizeof(hdg::KParams));
      e = hipMemcpyAsync(B->d_params, B->h_params, n * sizeof(hdg::KParams),
                         hipMemcpyHostToDevice, s);
      if (e == hipSuccess)
        e = hipEventRecord(B->copied, s);
      if (e != hipSuccess)
        return hip_fail(e, "hdg_batch_condense copy");
      B->last = np;
    }
  if (B->grid == 0 && capturing)
    return fail("hdg_batch_condense: the first call tunes the grid and cannot be captured", -2);
  if (B->grid == 0)
    {
      // autotuner: tiles (waves per workgroup of the register kernel, 4 or
      // 8; the pair kernel has one) and persistent grids (workgroups per
      // CU), each timed on this batch over a complete two-colour pass (the
      // outputs are the same for every candidate); the fastest is kept
      const int tiles4[2] = {4, 8}, per_cu4[2][5] = {{1, 2, 3, 4, 6}, {1, 2, 0, 0, 0}};
      const int ntile = B->deg3 ? 1 : 2;
      const int mx    = std::max(1, std::max(B->nchunks[0], B->nchunks[1]));
      B->ncand        = 0;
      for (int ti = 0; ti < ntile; ++ti)
        for (int k = 0; k < 5; ++k)
          if (per_cu4[ti][k] > 0)
            {
              B->cand_tile[B->ncand] = B->deg3 ? hdg::batch_pairs_per_group() : tiles4[ti];
              B->cand_grid[B->ncand] = std::max(1, std::min(per_cu4[ti][k] * B->ncu, mx));
              ++B->ncand;
            }
      hipEvent_t a, b;
      if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
        return fail("hdg_batch_condense: event creation failed", -3);
      float best = 1e30f;
      for (int k = 0; k < B->ncand; ++k)
        {
          const int g = B->cand_grid[k], tl = B->cand_tile[k];
          int       rc = batch_launch(B, thermodyn, mixed, g, tl, s);
          if (rc == 0)
            {
              (void)hipEventRecord(a, s);
              rc = batch_launch(B, thermodyn, mixed, g, tl, s);
              (void)hipEventRecord(b, s);
              (void)hipEventSynchronize(b);
            }
          if (rc != 0)
            {
              (void)hipEventDestroy(a);
              (void)hipEventDestroy(b);
              B->grid = 0;
              return rc;
            }
          float ms = 0;
          (void)hipEventElapsedTime(&ms, a, b);
          B->cand_ms[k] = ms;
          if (ms < best)
            {
              best    = ms;
              B->grid = g;
              B->tile = tl;
            }
        }
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  return batch_launch(B, thermodyn, mixed, B->grid, B->tile, s);
}

int hdg_batch_info(const hdg_batch *B, double out[6])
{
  if (!B || !out)
    return fail("hdg_batch_info: null argument");
  out[0] = (double)B->ncells;
  out[1] = (double)B->grid;
  out[2] = (double)B->ncu;
  out[3] = (double)B->tile;
  out[4] = (double)B->ncand;
  float best = 0;
  for (int k = 0; k < B->ncand; ++k)
    if (B->cand_tile[k] == B->tile && B->cand_grid[k] == B->grid)
      best = B->cand_ms[k];
  out[5] = best;
  return 0;
}

int hdg_batch_tuning(const hdg_batch *B, double *out, int cap)
{
  if (!B || (cap > 0 && !out) || cap < 0)
    return fail("hdg_batch_tuning: invalid argument");
  for (int k = 0; k < B->ncand && k < cap; ++k)
    {
      out[3 * k]     = (double)B->cand_tile[k];
      out[3 * k + 1] = (double)B->cand_grid[k];
      out[3 * k + 2] = (double)B->cand_ms[k];
    }
  return B->ncand;
}

int hdg_max_abs4(const double *const *vectors, const long long *lengths, double *scratch,
                 double *out, void *stream)
{
  if (!vectors || !lengths || !scratch || !out)
    return fail("hdg_max_abs4: null argument");
  for (int v = 0; v < 4; ++v)
    if (!vectors[v] || lengths[v] < 0)
      return fail("hdg_max_abs4: invalid vector");
  hipError_t e = hdg::launch_norms(vectors, lengths, scratch, out, (hipStream_t)stream);
  return e == hipSuccess ? 0 : hip_fail(e, "hdg_max_abs4 launch");
}

int hdg_max_abs4_noalloc(const double *const *vectors, const long long *lengths, double *out,
                         void *stream)
{
  if (!vectors || !lengths || !out)
    return fail("hdg_max_abs4_noalloc: null argument");
  for (int v = 0; v < 4; ++v)
    if (!vectors[v] || lengths[v] < 0)
      return fail("hdg_max_abs4_noalloc: invalid vector");
  hipError_t e = hdg::launch_norms_static(vectors, lengths, out, (hipStream_t)stream);
  return e == hipSuccess ? 0 : hip_fail(e, "hdg_max_abs4_noalloc launch");
}

int hdg_trace_plan_size(int dim, const int *n_cells, int dofs_per_face, double out[6])
{
  if (!n_cells || !out || (dim != 2 && dim != 3) || dofs_per_face < 1)
    return fail("hdg_trace_plan_size: invalid argument");
  int n[3] = {1, 1, 1};
  for (int a = 0; a < dim; ++a)
    {
      if (n_cells[a] < 1)
        return fail("hdg_trace_plan_size: invalid mesh", -2);
      n[a] = n_cells[a];
    }
  hdg::Topology T;
  hdg::build_topology(dim, n, T);
  hdg::NDPlan P;
  std::string err;
  if (!hdg::build_nd_plan(T, dofs_per_face, P, err, true))
    return fail("hdg_trace_plan_size: " + err, -4);
  out[0] = (double)P.Ftot;
  out[1] = (double)P.wtot;
  out[2] = (double)P.ptot;
  out[3] = P.flops;
  out[4] = (double)P.max_front;
  out[5] = (double)P.nnodes;
  return 0;
}

int hdg_gmres_solve(const long long *rowptr, const int *colind, const double *vals, long long n,
                    const double *rhs, double *x, int restart, double tol, long long max_it,
                    long long *iterations, double *residual, void *stream)
{
  if (!iterations || !residual || n < 0 || (n > 0 && (!rowptr || !colind || !vals || !rhs || !x)))
    return fail("hdg_gmres_solve: invalid argument");
  if (restart < 1 || max_it < 0 || !(tol >= 0))
    return fail("hdg_gmres_solve: restart >= 1, max_it >= 0 and tol >= 0 are required", -2);
  std::string err;
  const int   rc = hdg::gmres_solve(rowptr, colind, vals, n, rhs, x, restart, tol, max_it,
                                    iterations, residual, (hipStream_t)stream, err);
  return rc ? fail(err, rc) : 0;
}

int hdg_trace_plan_selftest(int dim, const int *n_cells, int dofs_per_face, unsigned seed,
                            const double *csr_vals, double out[3])
{
  if (!n_cells || !out || (dim != 2 && dim != 3) || dofs_per_face < 1)
    return fail("hdg_trace_plan_selftest: invalid argument");
  for (int a = 0; a < dim; ++a)
    if (n_cells[a] < 1)
      return fail("hdg_trace_plan_selftest: invalid mesh", -2);
  int n[3] = {1, 1, 1};
  for (int a = 0; a < dim; ++a)
    n[a] = n_cells[a];
  std::string err;
  if (!hdg::nd_plan_selftest(dim, n, dofs_per_face, seed, csr_vals, out, err))
    return fail("hdg_trace_plan_selftest: " + err, -4);
  return 0;
}

} // extern "C"
