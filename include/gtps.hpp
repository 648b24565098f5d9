// gtps.hpp - header-only C++ facade over the C ABI in gtps.h, with the class
// and method names of the reference's hot-path API (namespace PSMF):
//   PSMF::LaplaceOperator   include/laplace_operator.cuh:71-205
//   PSMF::PatchSmoother     include/patch_smoother.cuh:86-124, 313-335
//   PSMF::MGTransferCUDA    include/cuda_mg_transfer.cuh:57-358
//   PSMF::MultigridSolvers  include/solver.cuh:1541-2210
//   PSMF::LaplaceDGOperator, PSMF::DGMultigridSolvers  tests/mf_dg.cu:209-274
//   gtps::rhs_sine / error_norms  apps/poisson.cu:44-87, 353-382
// Vectors are raw device pointers (what deal.II's
// LinearAlgebra::distributed::Vector<Number, MemorySpace::CUDA>::get_values()
// returns), lexicographic with boundary DoFs.  Errors throw gtps::Error, as
// the reference throws deal.II exceptions; the C ABI underneath never throws.
#ifndef GTPS_HPP
#define GTPS_HPP

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "gtps.h"

namespace gtps
{
  class Error : public std::runtime_error
  {
  public:
    Error(int code, const std::string &what)
      : std::runtime_error("gtps error " + std::to_string(code) + ": " + what)
      , code(code)
    {}
    int code;
  };

  inline void
  check(int rc)
  {
    if (rc != GTPS_OK)
      throw Error(rc, gtps_last_error());
  }

  template <typename Number>
  constexpr int precision_of()
  {
    static_assert(sizeof(Number) == 4 || sizeof(Number) == 8,
                  "float or double");
    return sizeof(Number) == 4 ? GTPS_FP32 : GTPS_FP64;
  }

  class Context
  {
  public:
    explicit Context(int device = 0, void *stream = nullptr)
    {
      gtps_ctx *c = nullptr;
      check(gtps_ctx_create(device, stream, &c));
      h.reset(c, [](gtps_ctx *p) { gtps_ctx_destroy(p); });
    }
    gtps_ctx *
    get() const
    {
      return h.get();
    }
    void
    synchronize() const
    {
      check(gtps_ctx_synchronize(h.get()));
    }

  private:
    std::shared_ptr<gtps_ctx> h;
  };
} // namespace gtps

namespace PSMF
{
  template <int dim, int fe_degree, typename Number>
  class LaplaceOperator
  {
  public:
    using value_type = Number;

    LaplaceOperator(const gtps::Context &ctx, unsigned int level)
      : ctx(ctx)
    {
      gtps_level *l = nullptr;
      gtps::check(gtps_level_create(ctx.get(), dim, fe_degree, (int)level,
                                    gtps::precision_of<Number>(), &l));
      h.reset(l, [](gtps_level *p) { gtps_level_destroy(p); });
    }
    long long
    n_dofs() const
    {
      return gtps_level_n_dofs(h.get());
    }
    void
    vmult(Number *dst, const Number *src) const
    {
      gtps::check(gtps_laplace_vmult(h.get(), dst, src));
    }
    void
    Tvmult(Number *dst, const Number *src) const
    {
      vmult(dst, src);
    }
    gtps_level *
    get() const
    {
      return h.get();
    }

  private:
    gtps::Context               ctx;
    std::shared_ptr<gtps_level> h;
  };

  template <int dim, int fe_degree, typename Number>
  class PatchSmoother
  {
  public:
    struct AdditionalData
    {
      double relaxation = 1.;
      int    variant    = GTPS_MULT_FWD;
    };
    void
    initialize(const LaplaceOperator<dim, fe_degree, Number> &op,
               const AdditionalData &data = AdditionalData())
    {
      gtps_smoother *s = nullptr;
      gtps::check(
        gtps_smoother_create(op.get(), data.variant, data.relaxation, &s));
      h.reset(s, [](gtps_smoother *p) { gtps_smoother_destroy(p); });
    }
    void
    vmult(Number *dst, const Number *src) const
    {
      gtps::check(gtps_smoother_vmult(h.get(), dst, src));
    }
    void
    step(Number *dst, const Number *src) const
    {
      gtps::check(gtps_smoother_step(h.get(), dst, src));
    }

  private:
    std::shared_ptr<gtps_smoother> h;
  };

  template <int dim, typename Number>
  class MGTransferCUDA
  {
  public:
    MGTransferCUDA(const gtps::Context &ctx, int fe_degree,
                   unsigned int min_level, unsigned int max_level)
    {
      gtps_transfer *t = nullptr;
      gtps::check(gtps_transfer_create(ctx.get(), dim, fe_degree,
                                       (int)min_level, (int)max_level,
                                       gtps::precision_of<Number>(), &t));
      h.reset(t, [](gtps_transfer *p) { gtps_transfer_destroy(p); });
    }
    void
    prolongate(unsigned int to_level, Number *dst, const Number *src) const
    {
      gtps::check(gtps_transfer_prolongate(h.get(), (int)to_level, dst, src));
    }
    void
    prolongate_and_add(unsigned int to_level, Number *dst,
                       const Number *src) const
    {
      gtps::check(
        gtps_transfer_prolongate_add(h.get(), (int)to_level, dst, src));
    }
    void
    restrict_and_add(unsigned int from_level, Number *dst,
                     const Number *src) const
    {
      gtps::check(
        gtps_transfer_restrict_add(h.get(), (int)from_level, dst, src));
    }

  private:
    std::shared_ptr<gtps_transfer> h;
  };

  // Number: V-cycle precision (float for the mixed form); the outer CG and
  // the vectors handed to solve/vmult are double.
  template <int dim, int fe_degree, typename Number>
  class MultigridSolvers
  {
  public:
    MultigridSolvers(const gtps::Context &ctx, unsigned int max_level,
                     int    schedule   = GTPS_MG_SYMMETRIC,
                     double relaxation = 1.0)
    {
      gtps_mg *m = nullptr;
      gtps::check(gtps_mg_create(ctx.get(), dim, fe_degree, (int)max_level,
                                 gtps::precision_of<Number>(), schedule,
                                 relaxation, &m));
      h.reset(m, [](gtps_mg *p) { gtps_mg_destroy(p); });
    }
    // preconditioner application z = V(r)
    void
    vmult(double *dst, const double *src) const
    {
      gtps::check(gtps_mg_vcycle(h.get(), dst, src));
    }
    struct SolveInfo
    {
      int                 iterations = 0;
      double              reduction  = 0.;
      std::vector<double> history;
    };
    // MG-preconditioned CG (replaces solve_gmres); throws on non-convergence
    SolveInfo
    solve_cg(double *x, const double *b, double rtol = 1e-8,
             int maxit = 100) const
    {
      SolveInfo info;
      info.history.resize(maxit + 1);
      gtps::check(gtps_mg_pcg_solve(h.get(), x, b, rtol, maxit,
                                    &info.iterations, &info.reduction,
                                    info.history.data()));
      info.history.resize(info.iterations + 1);
      return info;
    }
    void
    do_matvec(double *dst, const double *src) const
    {
      gtps::check(gtps_mg_matvec(h.get(), dst, src));
    }
    // full multigrid (solver.cuh:1784-1860): rhs[l-1] is level l's FP64
    // right-hand side, l = 1..max_level
    void
    solve(double *x, const std::vector<const double *> &rhs, int n_cycles = 1) const
    {
      gtps::check(gtps_mg_fmg_solve(h.get(), x, rhs.data(), n_cycles));
    }
    // -div(c grad u) with one FP64 device value per finest cell (NULL: c = 1)
    void
    set_coefficient(const double *coef) const
    {
      gtps::check(gtps_mg_set_coefficient(h.get(), coef));
    }
    void
    matvec_transpose(double *dst, const double *src) const
    {
      gtps::check(gtps_mg_matvec_transpose(h.get(), dst, src));
    }
    // grad[K] = -lambda_K^T A^1_K u_K (dJ/dc for the adjoint lambda)
    void
    coefficient_gradient(double *grad, const double *lambda, const double *u) const
    {
      gtps::check(gtps_mg_coefficient_gradient(h.get(), grad, lambda, u));
    }

  private:
    std::shared_ptr<gtps_mg> h;
  };

  // SIPG-DG Laplacian of one level (the operator of tests/mf_dg.cu:209-274
  // on a uniform mesh; penalty = sigma h, k (k+1) there)
  template <int dim, int fe_degree, typename Number>
  class LaplaceDGOperator
  {
  public:
    LaplaceDGOperator(const gtps::Context &ctx, int n_cells, double penalty)
    {
      gtps_dg_level *l = nullptr;
      gtps::check(gtps_dg_level_create(ctx.get(), dim, fe_degree, n_cells,
                                       gtps::precision_of<Number>(), penalty, &l));
      h.reset(l, [](gtps_dg_level *p) { gtps_dg_level_destroy(p); });
    }
    long long
    n_dofs() const
    {
      return gtps_dg_level_n_dofs(h.get());
    }
    void
    vmult(Number *dst, const Number *src) const
    {
      gtps::check(gtps_dg_vmult(h.get(), dst, src));
    }
    void
    smoother_step(Number *u, const Number *b, bool reverse = false,
                  double relaxation = 1.0) const
    {
      gtps::check(gtps_dg_smoother_step(h.get(), u, b,
                                        reverse ? GTPS_MULT_REV : GTPS_MULT_FWD,
                                        relaxation));
    }

  private:
    std::shared_ptr<gtps_dg_level> h;
  };

  // DG multigrid (n_coarse 2^l cells, l = 0..max_level) with FP64 MG-CG
  template <int dim, int fe_degree, typename Number>
  class DGMultigridSolvers
  {
  public:
    DGMultigridSolvers(const gtps::Context &ctx, int n_coarse, int max_level,
                       double penalty, int coarse_sweeps = 2)
    {
      gtps_dg_mg *m = nullptr;
      gtps::check(gtps_dg_mg_create(ctx.get(), dim, fe_degree, n_coarse, max_level,
                                    gtps::precision_of<Number>(), penalty,
                                    coarse_sweeps, &m));
      h.reset(m, [](gtps_dg_mg *p) { gtps_dg_mg_destroy(p); });
    }
    long long
    n_dofs() const
    {
      return gtps_dg_mg_n_dofs(h.get());
    }
    int
    solve_cg(double *x, const double *b, double rtol = 1e-8, int maxit = 200) const
    {
      int    it  = 0;
      double rel = 0.;
      gtps::check(gtps_dg_mg_pcg_solve(h.get(), x, b, rtol, maxit, &it, &rel, nullptr));
      return it;
    }

  private:
    std::shared_ptr<gtps_dg_mg> h;
  };
} // namespace PSMF

namespace gtps
{
  // the reference's Poisson problem (apps/poisson.cu:44-87): load vector of
  // f = dim (w pi)^2 prod sin(w pi x_d) on 2^level cells, and the L2 / H1
  // errors against prod sin(w pi x_d) (compute_error, apps/poisson.cu:353-382)
  inline void
  rhs_sine(const Context &ctx, int dim, int degree, int level, double *b,
           double wave = 3.0)
  {
    check(gtps_rhs_sine(ctx.get(), dim, degree, level, wave, b));
  }
  struct ErrorNorms
  {
    double l2 = 0., h1 = 0.;
  };
  inline ErrorNorms
  error_norms(const Context &ctx, int dim, int degree, int level, const double *u,
              double wave = 3.0, int n_q = 0)
  {
    ErrorNorms e;
    check(gtps_error_norms(ctx.get(), dim, degree, level, wave,
                           n_q > 0 ? n_q : degree + 1, u, &e.l2, &e.h1));
    return e;
  }
  // solution files (64-byte header + values, FNV-1a checked on load)
  inline void
  save_vector(const Context &ctx, const std::string &path, const double *x, int dim,
              int degree, int n_cells, int layout = GTPS_LAYOUT_CG)
  {
    check(gtps_vector_save(ctx.get(), path.c_str(), x, GTPS_FP64, dim, degree, n_cells,
                           layout));
  }
  inline void
  load_vector(const Context &ctx, const std::string &path, double *x, long long n)
  {
    check(gtps_vector_load(ctx.get(), path.c_str(), x, n, GTPS_FP64));
  }
} // namespace gtps

#endif // GTPS_HPP
