// Compiles the C++ facade (include/gtps.hpp) against libgtps.so the way a
// caller of the reference's PSMF classes would, and runs it: with a GPU it
// solves 3D Q3 on 8^3 cells; without one it must fail cleanly through
// gtps::Error (exit code 3), never crash.
#include <cstdio>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "gtps.hpp"

int
main()
{
  try
    {
      gtps::Context ctx(0);
      PSMF::MultigridSolvers<3, 3, float> mg(ctx, 3);
      const long long n = 25LL * 25 * 25;
      std::vector<double> hb(n, 0.0);
      for (int z = 1; z < 24; ++z)
        for (int y = 1; y < 24; ++y)
          for (int x = 1; x < 24; ++x)
            hb[x + 25 * (y + 25 * z)] = 1.0;
      double *b = nullptr, *x = nullptr;
      if (hipMalloc(&b, n * sizeof(double)) != hipSuccess ||
          hipMalloc(&x, n * sizeof(double)) != hipSuccess)
        return 4;
      (void)hipMemcpy(b, hb.data(), n * sizeof(double), hipMemcpyHostToDevice);
      auto info = mg.solve_cg(x, b, 1e-8, 50);
      std::printf("iterations %d reduction %.3e\n", info.iterations, info.reduction);
      (void)hipFree(b);
      (void)hipFree(x);
      return info.reduction <= 1e-8 ? 0 : 5;
    }
  catch (const gtps::Error &e)
    {
      std::printf("gtps::Error %d: %s\n", e.code, e.what());
      return 3;
    }
}
