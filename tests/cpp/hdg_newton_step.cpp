// hdg_newton_step.cpp -- a C++ host driving one Newton step of the 2D/3D
// silicon diode through the C ABI of libddhdg_hip.so (include/ddhdg_hip.h),
// with device buffers from hipMalloc: the call sequence a C++ NPSolver
// binding makes (reference src/np_solver.cpp:1483-1561, private_run).
//
//   hdg_newton_step <in.bin> <out.bin>
//
// in.bin (little-endian):
//   int32  dim, degree, nx, ny, nz, bc_from_file
//   double bias, nd, na, junction
//   int64  ncell_values, ntrace_values
//   double x_cell[ncell_values], x_trace[ntrace_values]
//   if bc_from_file: double doping[ncells*QC], dirichlet[nfaces*3*QF],
//                    neumann[nfaces*3*QF]   (rescaled, as uploaded)
// out.bin: double x_cell[...], x_trace[...] after the step, then the
//   update norms |dx_cell|_inf, |dx_trace|_inf.
//
// Physics: silicon at 300 K (eps_r 11.7, mu_n 0.14, mu_p 0.045 m^2/Vs,
// Nc = Nv = n_i = 1.45e16 m^-3, SRH with tau_n = tau_p = 1e-6 s),
// Adimensionalizer(1e-6 m, q/k_B, 1e23 m^-3, 1), HDG tau 1 in rescaled
// units, step doping nd | na at x = junction, ohmic contacts on ids 0 and
// 1 (bias on id 1).  Without bc_from_file the doping and the contact
// values are evaluated here, at the quadrature points the library reports.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "ddhdg_hip.h"

namespace
{
  // include/constants.h:12-15
  constexpr double EPSILON0 = 8.85418781762e-12;
  constexpr double Q        = 1.602176634e-19;
  constexpr double KB       = 1.380649e-23;
  constexpr double EV       = 1.602176565e-19;
  constexpr double NI       = 1.45e16;
  constexpr double TEMP     = 300.0;

  void check(int rc, const char *what)
  {
    if (rc < 0)
      throw std::runtime_error(std::string(what) + ": " + hdg_last_error());
  }
  void hcheck(hipError_t e, const char *what)
  {
    if (e != hipSuccess)
      throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
  }

  // Adimensionalizer (include/adimensionalizer.h:28-262)
  struct Adim
  {
    double l = 1e-6, T = Q / KB, N = 1e23, mu = 1.0;
    double V() const { return T * KB / Q; }
    double perm() const { return l * l * Q * Q * N / (T * KB); }
    double rec() const { return KB * N * T * mu / (l * l * Q); }
    double tau(int c) const
    {
      const double Tv = T * KB / Q;
      return c == 0 ? Q * N / Tv : Tv * mu / (l * l);
    }
  };

  // the reference ScratchData set-up, folded once
  hdg_phys fold_physics(const Adim &a)
  {
    hdg_phys ph{};
    ph.eps  = 11.7 * EPSILON0 / a.perm();
    ph.mu_n = 0.14 / a.mu;
    ph.mu_p = 0.045 / a.mu;
    ph.U_T  = (KB * TEMP / Q) / a.V();
    for (int c = 0; c < 3; ++c)
      ph.tau[c] = a.tau(c) / a.tau(c);   // tau = 1 in rescaled units
    ph.recombination = 1;
    ph.rec[0]        = NI;
    ph.rec[1]        = 1e-6;
    ph.rec[2]        = 1e-6;
    ph.n_rescale     = a.N;
    ph.rec_rescale   = a.rec();
    ph.kT_over_q     = KB * TEMP / (Q * a.V());
    ph.ec_r          = 0.0 * EV / (Q * a.V());
    ph.ev_r          = 0.0 * EV / (Q * a.V());
    ph.nc            = NI;
    ph.nv            = NI;
    ph.flux_type     = 0;
    ph.phi_linearize = 0;
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 3; ++c)
        ph.bc_type[b][c] = 1;            // ohmic contacts: Dirichlet V, n, p
    ph.drift_jacobian = 1;
    return ph;
  }

  // compute_local_charge_neutrality_on_a_point
  // (src/thermodynamic_equilibrium.cpp:9-43) with Nc = Nv, Ec = Ev = 0
  double neutral_potential(double doping)
  {
    const double ut  = TEMP * KB / Q;
    const double ni2 = NI * NI * std::exp((0.0 * EV - 0.0 * EV) / (TEMP * KB));
    return (0.0 * EV + 0.0 * EV) / (2 * Q) + ut / 2 * std::log(NI / NI) +
           ut * std::asinh(doping / (2 * std::sqrt(ni2)));
  }
  // Solver::compute_density (include/solver.h:378-412), quasi-Fermi 0
  double density(int c, double V)
  {
    const double kbt = KB * TEMP;
    const double e   = c == 1 ? (Q * (V - 0.0) - 0.0 * EV) / kbt : (Q * (0.0 - V) + 0.0 * EV) / kbt;
    return NI * std::exp(e);
  }

  template <class T>
  T read1(FILE *f)
  {
    T v;
    if (std::fread(&v, sizeof(T), 1, f) != 1)
      throw std::runtime_error("short input file");
    return v;
  }
  void readn(FILE *f, double *p, size_t n)
  {
    if (std::fread(p, sizeof(double), n, f) != n)
      throw std::runtime_error("short input file");
  }

  template <class T>
  T *dalloc(size_t n)
  {
    T *p = nullptr;
    hcheck(hipMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
    return p;
  }
} // namespace

int main(int argc, char **argv)
{
  if (argc != 3)
    {
      std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
      return 2;
    }
  try
    {
      FILE *fin = std::fopen(argv[1], "rb");
      if (!fin)
        throw std::runtime_error("cannot open input");
      hdg_mesh_desc m{};
      m.dim          = read1<int32_t>(fin);
      m.degree       = read1<int32_t>(fin);
      m.n_cells[0]   = read1<int32_t>(fin);
      m.n_cells[1]   = read1<int32_t>(fin);
      m.n_cells[2]   = read1<int32_t>(fin);
      const int from_file = read1<int32_t>(fin);
      const double bias = read1<double>(fin), nd = read1<double>(fin), na = read1<double>(fin),
                   junction = read1<double>(fin);
      for (int a = 0; a < 3; ++a)
        {
          m.lower[a] = 0.0;
          m.upper[a] = 1.0;
        }
      const Adim     adim;
      const hdg_phys ph = fold_physics(adim);
      hdg_ctx       *ctx = nullptr;
      check(hdg_ctx_create(&m, &ph, &ctx), "hdg_ctx_create");
      long long sz[HDG_NSIZES];
      check(hdg_sizes(ctx, sz), "hdg_sizes");
      const long long ncells = sz[HDG_NCELLS], nfaces = sz[HDG_NFACES], NC = sz[HDG_NC];
      const long long QC = sz[HDG_QC], QF = sz[HDG_QF], nf = sz[HDG_NF];
      const long long ntrace = sz[HDG_NTRACE], nnz = sz[HDG_NNZ];
      const long long ncv = read1<int64_t>(fin), ntv = read1<int64_t>(fin);
      if (ncv != ncells * NC || ntv != nfaces * 3 * nf)
        throw std::runtime_error("state sizes do not match the mesh");
      std::vector<double> xc(ncv), xt(ntv);
      readn(fin, xc.data(), ncv);
      readn(fin, xt.data(), ntv);

      // problem data at the quadrature points
      std::vector<double> dop(ncells * QC), dirv(nfaces * 3 * QF, 0.0), neuv(nfaces * 3 * QF, 0.0);
      if (from_file)
        {
          readn(fin, dop.data(), dop.size());
          readn(fin, dirv.data(), dirv.size());
          readn(fin, neuv.data(), neuv.size());
        }
      else
        {
          const int d = m.dim;
          std::vector<double> cq(ncells * QC * d), fq(nfaces * QF * d);
          check(hdg_cell_qpoints(ctx, cq.data()), "hdg_cell_qpoints");
          check(hdg_face_qpoints(ctx, fq.data()), "hdg_face_qpoints");
          auto doping = [&](double x) { return x < junction ? nd : -na; };
          for (long long i = 0; i < ncells * QC; ++i)
            dop[i] = doping(cq[i * d]) / adim.N;
          // contacts: the faces with normal axis 0 at x = 0 (id 0) and at
          // x = 1 (id 1), the first (nx+1)*ny*nz faces, x-index fastest
          const long long nx1 = m.n_cells[0] + 1;
          const long long nax = nx1 * m.n_cells[1] * (d == 3 ? m.n_cells[2] : 1);
          for (long long f = 0; f < nax; ++f)
            {
              const long long ix = f % nx1;
              if (ix != 0 && ix != nx1 - 1)
                continue;
              const double b = ix == 0 ? 0.0 : bias;
              for (long long s = 0; s < QF; ++s)
                {
                  const double V0 = neutral_potential(doping(fq[(f * QF + s) * d]));
                  dirv[(f * 3 + 0) * QF + s] = (V0 + b) / adim.V();
                  dirv[(f * 3 + 1) * QF + s] = density(1, V0) / adim.N;
                  dirv[(f * 3 + 2) * QF + s] = density(2, V0) / adim.N;
                }
            }
        }
      std::fclose(fin);
      check(hdg_set_doping(ctx, dop.data()), "hdg_set_doping");
      check(hdg_set_bc_values(ctx, dirv.data(), neuv.data()), "hdg_set_bc_values");

      // device state and buffers
      double *d_xc = dalloc<double>(ncv), *d_xt = dalloc<double>(ntv);
      double *d_vals = dalloc<double>(nnz), *d_rhs = dalloc<double>(ntrace);
      double *d_dir = dalloc<double>(ntrace), *d_sol = dalloc<double>(ntrace);
      double *d_dt = dalloc<double>(ntv), *d_dc = dalloc<double>(ncv);
      hcheck(hipMemcpy(d_xc, xc.data(), ncv * 8, hipMemcpyHostToDevice), "copy");
      hcheck(hipMemcpy(d_xt, xt.data(), ntv * 8, hipMemcpyHostToDevice), "copy");
      hcheck(hipMemset(d_dir, 0, ntrace * 8), "memset");

      // pass 1: assembly, condensation, scatter (assemble_system(false, ...))
      check(hdg_condense(ctx, d_xc, d_xt, 0, d_vals, d_rhs, d_dir, nullptr), "hdg_condense");
      // the trace solve (SparseDirectUMFPACK in the reference)
      hdg_trace_solver *ts = nullptr;
      check(hdg_trace_solver_create(ctx, &ts), "hdg_trace_solver_create");
      check(hdg_trace_factor(ts, d_vals, nullptr), "hdg_trace_factor");
      check(hdg_trace_factor_status(ts, nullptr), "hdg_trace_factor_status");
      check(hdg_trace_solve(ts, d_rhs, d_sol, nullptr), "hdg_trace_solve");
      // Dirichlet dofs take their projected values (src/np_solver.cpp:
      // 1503-1513); restricted -> full trace (all components enabled: the
      // same order)
      std::vector<double> sol(ntrace), dv(ntrace);
      hcheck(hipMemcpy(sol.data(), d_sol, ntrace * 8, hipMemcpyDeviceToHost), "copy");
      hcheck(hipMemcpy(dv.data(), d_dir, ntrace * 8, hipMemcpyDeviceToHost), "copy");
      const int ndir = hdg_dirichlet_dofs(ctx, nullptr);
      check(ndir, "hdg_dirichlet_dofs");
      std::vector<long long> dirs(std::max(ndir, 1));
      check(hdg_dirichlet_dofs(ctx, dirs.data()), "hdg_dirichlet_dofs");
      for (int i = 0; i < ndir; ++i)
        sol[dirs[i]] = dv[dirs[i]];
      if (ntrace != ntv)
        throw std::runtime_error("all three components must be enabled");
      hcheck(hipMemcpy(d_dt, sol.data(), ntv * 8, hipMemcpyHostToDevice), "copy");
      // pass 2: the cell update (assemble_system(true, ...))
      check(hdg_reconstruct(ctx, d_xc, d_xt, d_dt, 0, d_dc, nullptr), "hdg_reconstruct");
      std::vector<double> dc(ncv);
      hcheck(hipMemcpy(dc.data(), d_dc, ncv * 8, hipMemcpyDeviceToHost), "copy");
      // x += alpha dx with alpha = 1, and the update norms
      double nc_ = 0, nt_ = 0;
      for (long long i = 0; i < ncv; ++i)
        {
          nc_ = std::fmax(nc_, std::fabs(dc[i]));
          xc[i] += dc[i];
        }
      for (long long i = 0; i < ntv; ++i)
        {
          nt_ = std::fmax(nt_, std::fabs(sol[i]));
          xt[i] += sol[i];
        }
      hdg_trace_solver_destroy(ts);
      for (double *p : {d_xc, d_xt, d_vals, d_rhs, d_dir, d_sol, d_dt, d_dc})
        (void)hipFree(p);
      hdg_ctx_destroy(ctx);

      FILE *fout = std::fopen(argv[2], "wb");
      if (!fout)
        throw std::runtime_error("cannot open output");
      std::fwrite(xc.data(), 8, ncv, fout);
      std::fwrite(xt.data(), 8, ntv, fout);
      std::fwrite(&nc_, 8, 1, fout);
      std::fwrite(&nt_, 8, 1, fout);
      std::fclose(fout);
      std::printf("newton step: %lld cells, %lld trace dofs, |dx_cell| %.6e |dx_trace| %.6e\n",
                  ncells, ntrace, nc_, nt_);
      return 0;
    }
  catch (const std::exception &e)
    {
      std::fprintf(stderr, "hdg_newton_step: %s\n", e.what());
      return 1;
    }
}
