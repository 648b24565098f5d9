// Compiles the C++ facade (include/gtps.hpp) against libgtps.so the way a
// caller of the reference's PSMF classes would, and runs it: with a GPU it
// solves the reference's Poisson problem (u = prod sin(3 pi x_d)) with
// 3D Q3 on 8^3 and 16^3 cells by MG-CG and by full multigrid, checks the
// L2 error order, writes and re-reads the solution file, and solves a DG
// problem (argv[1]: where to write the solution file); without a GPU it
// must fail cleanly through gtps::Error (exit code 3), never crash.
#include <cmath>
#include <cstdio>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "gtps.hpp"

namespace
{
  double *
  dalloc(long long n)
  {
    double *p = nullptr;
    if (hipMalloc(&p, n * sizeof(double)) != hipSuccess)
      throw gtps::Error(GTPS_ENOMEM, "hipMalloc");
    return p;
  }
} // namespace

int
main(int argc, char **argv)
{
  const char *path = argc > 1 ? argv[1] : "facade_solution.gtps";
  try
    {
      gtps::Context ctx(0);
      double        err[2] = {0., 0.};
      for (int level = 3; level <= 4; ++level)
        {
          PSMF::MultigridSolvers<3, 3, float> mg(ctx, level);
          const long long n1 = 3LL * (1 << level) + 1, n = n1 * n1 * n1;
          double *b = dalloc(n), *x = dalloc(n);
          gtps::rhs_sine(ctx, 3, 3, level, b);
          auto info = mg.solve_cg(x, b, 1e-10, 50);
          const auto e = gtps::error_norms(ctx, 3, 3, level, x);
          err[level - 3] = e.l2;
          std::printf("level %d iterations %d reduction %.3e L2 %.3e H1 %.3e\n", level,
                      info.iterations, info.reduction, e.l2, e.h1);
          if (info.reduction > 1e-10)
            return 5;
          if (level == 4)
            {
              // full multigrid with the per-level load vectors
              std::vector<double *>       rhs;
              std::vector<const double *> crhs;
              for (int l = 1; l <= level; ++l)
                {
                  const long long m1 = 3LL * (1 << l) + 1;
                  rhs.push_back(dalloc(m1 * m1 * m1));
                  gtps::rhs_sine(ctx, 3, 3, l, rhs.back());
                  crhs.push_back(rhs.back());
                }
              double *xf = dalloc(n);
              mg.solve(xf, crhs, 2);
              const auto ef = gtps::error_norms(ctx, 3, 3, level, xf);
              std::printf("fmg L2 %.3e\n", ef.l2);
              if (!(ef.l2 < 1.5 * e.l2))
                return 6;
              // solution file round trip
              gtps::save_vector(ctx, path, x, 3, 3, 1 << level);
              gtps::load_vector(ctx, path, xf, n);
              const auto er = gtps::error_norms(ctx, 3, 3, level, xf);
              if (er.l2 != e.l2)
                return 7;
              for (double *p : rhs)
                (void)hipFree(p);
              (void)hipFree(xf);
            }
          (void)hipFree(b);
          (void)hipFree(x);
        }
      const double order = std::log2(err[0] / err[1]);
      std::printf("L2 order %.2f\n", order);
      if (order < 3.3)
        return 8;
      // SIPG-DG Q2 on 6^3 cells (3^3 coarse), penalty 2.5 (k+1)^2
      PSMF::DGMultigridSolvers<3, 2, float> dg(ctx, 3, 1, 2.5 * 9);
      const long long nd = dg.n_dofs();
      double *bd = dalloc(nd), *xd = dalloc(nd);
      std::vector<double> hb(nd, 1.0);
      (void)hipMemcpy(bd, hb.data(), nd * sizeof(double), hipMemcpyHostToDevice);
      const int itd = dg.solve_cg(xd, bd, 1e-8, 100);
      std::printf("dg iterations %d\n", itd);
      (void)hipFree(bd);
      (void)hipFree(xd);
      return 0;
    }
  catch (const gtps::Error &e)
    {
      std::printf("gtps::Error %d: %s\n", e.code, e.what());
      return 3;
    }
}
