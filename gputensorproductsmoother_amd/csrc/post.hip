// post.hip - load-vector assembly and error norms on the GPU, for the
// continuous Q_k layout (lexicographic, shared cell-end nodes, Dirichlet
// rows) and the DG layout (X = c (k+1) + i, no constrained rows).
//
// * assembly: b_i = sum_cells int f phi_i with QGauss(k+1) per cell, f
//   given at the quadrature points, as the closed form of the reference's
//   RightHandSide (f = dim (w pi)^2 prod sin(w pi x_d), poisson.cu:76-87) or
//   as a constant; the reference assembles this on the host cell by cell
//   (laplace_operator.cuh:140-173).  Here one thread owns one DoF and
//   gathers the (at most 2^dim) cells that contain it in a fixed order, so
//   the result is deterministic without atomics.  Dirichlet rows of the CG
//   layout are zeroed (the reference leaves them; a CG residual needs them
//   zero).
// * error norms: L2 norm and H1 seminorm of u_h - prod sin(w pi x_d) with
//   QGauss(n_q) per cell (poisson.cu:353-382 with n_q = k+1 there), one
//   thread per quadrature point, per-block partial sums reduced on the host
//   in a fixed order.
#include "gtps_internal.h"

namespace gtps
{
  namespace
  {
    constexpr int PNT = 256;

    // 1D data at the quadrature points, copied to LDS at kernel start
    struct QTab
    {
      double phi[MAX_QUAD][8];   // phi_i(xi_q)
      double dphi[MAX_QUAD][8];  // phi_i'(xi_q) / h
      double qp[MAX_QUAD], qw[MAX_QUAD];
    };

    __device__ void
    load_qtab(const QTab &src, QTab &dst)
    {
      const double *s = reinterpret_cast<const double *>(&src);
      double       *d = reinterpret_cast<double *>(&dst);
      for (int e = threadIdx.x; e < (int)(sizeof(QTab) / sizeof(double)); e += PNT)
        d[e] = s[e];
      __syncthreads();
    }

    template <int K, int DIM, bool DG>
    __global__ void __launch_bounds__(PNT)
    error_norms_kernel(const double *__restrict__ u, const int n, const int nq,
                       const double a, const double vol, const QTab tab_arg,
                       double *__restrict__ partials)
    {
      __shared__ QTab   tab;
      __shared__ double red[PNT / 64];
      load_qtab(tab_arg, tab);
      constexpr int   NN = K + 1, ST = DG ? K + 1 : K, NL = DIM == 3 ? NN : 1;
      const long long N1    = DG ? (long long)n * NN : (long long)K * n + 1;
      const long long nqd   = DIM == 3 ? (long long)nq * nq * nq : (long long)nq * nq;
      const long long ncell = DIM == 3 ? (long long)n * n * n : (long long)n * n;
      const long long total = ncell * nqd;
      double          e0 = 0.0, e1 = 0.0;
      for (long long g = (long long)blockIdx.x * PNT + threadIdx.x; g < total;
           g += (long long)gridDim.x * PNT)
        {
          const long long cell = g / nqd;
          const int       q    = (int)(g - cell * nqd);
          const int qx = q % nq, qy = (q / nq) % nq, qz = DIM == 3 ? q / (nq * nq) : 0;
          const int cx = (int)(cell % n), cy = (int)((cell / n) % n),
                    cz = DIM == 3 ? (int)(cell / ((long long)n * n)) : 0;
          const double *u0 = u + (long long)ST * cx +
                             N1 * ((long long)ST * cy + N1 * ((long long)ST * cz));
          double uh = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll 1
          for (int l = 0; l < NL; ++l)
            {
              const double pz = DIM == 3 ? tab.phi[qz][l] : 1.0;
              const double dz = DIM == 3 ? tab.dphi[qz][l] : 0.0;
#pragma unroll
              for (int j = 0; j < NN; ++j)
                {
                  const double *row = u0 + N1 * (j + N1 * l);
                  double        s0 = 0.0, s1 = 0.0;
#pragma unroll
                  for (int i = 0; i < NN; ++i)
                    {
                      const double v = row[i];
                      s0 += v * tab.phi[qx][i];
                      s1 += v * tab.dphi[qx][i];
                    }
                  const double py = tab.phi[qy][j], dy = tab.dphi[qy][j];
                  uh += s0 * py * pz;
                  gx += s1 * py * pz;
                  gy += s0 * dy * pz;
                  gz += s0 * py * dz;
                }
            }
          const double h  = 1.0 / n;
          const double x  = (cx + tab.qp[qx]) * h, y = (cy + tab.qp[qy]) * h,
                       z  = (cz + tab.qp[qz]) * h;
          const double sx = sin(a * x), sy = sin(a * y), sz = DIM == 3 ? sin(a * z) : 1.0;
          const double cxv = a * cos(a * x), cyv = a * cos(a * y),
                       czv = DIM == 3 ? a * cos(a * z) : 0.0;
          const double w = tab.qw[qx] * tab.qw[qy] * (DIM == 3 ? tab.qw[qz] : 1.0) * vol;
          const double du = uh - sx * sy * sz;
          const double dx = gx - cxv * sy * sz, dy = gy - sx * cyv * sz,
                       dzv = gz - sx * sy * czv;
          e0 += w * du * du;
          e1 += w * (dx * dx + dy * dy + dzv * dzv);
        }
      e0 = block_sum<PNT>(e0, red);
      __syncthreads();
      e1 = block_sum<PNT>(e1, red);
      if (threadIdx.x == 0)
        {
          partials[2 * blockIdx.x]     = e0;
          partials[2 * blockIdx.x + 1] = e1;
        }
    }

    // MODE 0: f from fq[cell][q]; 1: f = DIM a^2 prod sin(a x_d); 2: f = fc
    template <int K, int DIM, bool DG, int MODE>
    __global__ void __launch_bounds__(PNT)
    assemble_rhs_kernel(const double *__restrict__ fq, double *__restrict__ b,
                        const int n, const double a, const double fc,
                        const double vol, const QTab tab_arg)
    {
      __shared__ QTab tab;
      load_qtab(tab_arg, tab);
      constexpr int   NN = K + 1, NQ = K + 1, ST = DG ? K + 1 : K;
      constexpr int   NQD = DIM == 3 ? NQ * NQ * NQ : NQ * NQ;
      const long long N1  = DG ? (long long)n * NN : (long long)K * n + 1;
      const long long N   = DIM == 3 ? N1 * N1 * N1 : N1 * N1;
      const double    h   = 1.0 / n;
      for (long long g = (long long)blockIdx.x * PNT + threadIdx.x; g < N;
           g += (long long)gridDim.x * PNT)
        {
          const long long X = g % N1, Y = (g / N1) % N1, Z = DIM == 3 ? g / (N1 * N1) : 0;
          if (!DG)
            {
              bool bd = X == 0 || X == N1 - 1 || Y == 0 || Y == N1 - 1;
              if (DIM == 3)
                bd = bd || Z == 0 || Z == N1 - 1;
              if (bd)
                {
                  b[g] = 0.0;
                  continue;
                }
            }
          // the cells containing the DoF along each direction: (cell, local)
          int c[3][2], li[3][2], nc[3];
          const long long XYZ[3] = {X, Y, Z};
#pragma unroll
          for (int d = 0; d < 3; ++d)
            {
              nc[d] = 0;
              if (d == 2 && DIM == 2)
                {
                  c[d][0] = 0, li[d][0] = 0, nc[d] = 1;
                  continue;
                }
              const int cc = (int)(XYZ[d] / ST), ii = (int)(XYZ[d] - (long long)cc * ST);
              if (!DG && ii == 0 && cc > 0)
                {
                  c[d][nc[d]] = cc - 1, li[d][nc[d]] = K, ++nc[d];
                }
              if (cc < n)
                {
                  c[d][nc[d]] = cc, li[d][nc[d]] = ii, ++nc[d];
                }
            }
          double v = 0.0;
          for (int tz = 0; tz < nc[2]; ++tz)
            for (int ty = 0; ty < nc[1]; ++ty)
              for (int tx = 0; tx < nc[0]; ++tx)
                {
                  const int cx = c[0][tx], cy = c[1][ty], cz = c[2][tz];
                  const int ix = li[0][tx], iy = li[1][ty], iz = li[2][tz];
                  const long long fo =
                    MODE == 0 ? (cx + (long long)n * (cy + (long long)n * cz)) * NQD : 0;
                  double s = 0.0;
#pragma unroll 1
                  for (int qz = 0; qz < (DIM == 3 ? NQ : 1); ++qz)
                    {
                      const double wz = DIM == 3 ? tab.qw[qz] * tab.phi[qz][iz] : 1.0;
                      const double fz =
                        (DIM == 3 && MODE == 1) ? sin(a * (cz + tab.qp[qz]) * h) : 1.0;
#pragma unroll
                      for (int qy = 0; qy < NQ; ++qy)
                        {
                          const double wy = tab.qw[qy] * tab.phi[qy][iy];
                          const double fy = MODE == 1 ? sin(a * (cy + tab.qp[qy]) * h) : 1.0;
                          double       sx = 0.0;
#pragma unroll
                          for (int qx = 0; qx < NQ; ++qx)
                            {
                              double fv;
                              if (MODE == 0)
                                fv = fq[fo + qx + NQ * (qy + NQ * qz)];
                              else if (MODE == 1)
                                fv = sin(a * (cx + tab.qp[qx]) * h);
                              else
                                fv = 1.0;
                              sx += tab.qw[qx] * tab.phi[qx][ix] * fv;
                            }
                          s += wz * wy * sx * (MODE == 1 ? fy * fz : 1.0);
                        }
                    }
                  v += s;
                }
          const double fac = MODE == 1 ? DIM * a * a : (MODE == 2 ? fc : 1.0);
          b[g] = v * vol * fac;
        }
    }

    template <typename F>
    auto
    dispatch_degree(int k, F &&f)
    {
      switch (k)
        {
          case 1: return f(std::integral_constant<int, 1>{});
          case 2: return f(std::integral_constant<int, 2>{});
          case 3: return f(std::integral_constant<int, 3>{});
          case 4: return f(std::integral_constant<int, 4>{});
          case 5: return f(std::integral_constant<int, 5>{});
          case 6: return f(std::integral_constant<int, 6>{});
          case 7: return f(std::integral_constant<int, 7>{});
          default: throw Error(GTPS_EINVAL, "degree must be in 1..7");
        }
    }

    QTab
    make_qtab(gtps_ctx *ctx, int k, int nq, double h)
    {
      QTab t{};
      const Basis1D &b = get_tables(ctx, k)->basis;
      std::vector<double> p, w;
      gauss_rule(nq, p, w);
      for (int q = 0; q < nq; ++q)
        {
          double val[8], der[8];
          lagrange_values(b.nodes, p[q], val, der);
          for (int i = 0; i <= k; ++i)
            {
              t.phi[q][i]  = val[i];
              t.dphi[q][i] = der[i] / h;
            }
          t.qp[q] = p[q];
          t.qw[q] = w[q];
        }
      return t;
    }

    int
    pgrid(long long work, int cap)
    {
      return (int)std::max<long long>(1, std::min<long long>((work + PNT - 1) / PNT, cap));
    }
  } // namespace

  void
  compute_error_norms(gtps_ctx *ctx, int dim, int k, int n, bool dg, double wave,
                      int nq, const double *u, double *l2, double *h1)
  {
    GTPS_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
    GTPS_REQUIRE(k >= 1 && k <= MAX_DEGREE, "degree must be in 1..7");
    GTPS_REQUIRE(n >= 1 && n <= 4096, "cells per direction must be in 1..4096");
    GTPS_REQUIRE(nq >= 1 && nq <= MAX_QUAD, "quadrature points must be in 1..16");
    GTPS_REQUIRE(std::isfinite(wave), "wave number must be finite");
    const double    h     = 1.0 / n;
    const QTab      tab   = make_qtab(ctx, k, nq, h);
    const long long nqd   = dim == 3 ? (long long)nq * nq * nq : (long long)nq * nq;
    const long long cells = dim == 3 ? (long long)n * n * n : (long long)n * n;
    const int       nb    = pgrid(cells * nqd, std::min(ctx->ws.capacity / 2, 4096));
    const double    vol   = dim == 3 ? h * h * h : h * h;
    const double    a     = wave * M_PI;
    hipStream_t     s     = ctx->stream;
    double         *pt    = ctx->ws.partials;
    dispatch_degree(k, [&](auto kk) {
      constexpr int K = decltype(kk)::value;
#define GTPS_EN(D, G)                                                             \
  hipLaunchKernelGGL((error_norms_kernel<K, D, G>), dim3(nb), dim3(PNT), 0, s, u, \
                     n, nq, a, vol, tab, pt)
      if (dim == 3)
        {
          if (dg)
            GTPS_EN(3, true);
          else
            GTPS_EN(3, false);
        }
      else
        {
          if (dg)
            GTPS_EN(2, true);
          else
            GTPS_EN(2, false);
        }
#undef GTPS_EN
      GTPS_LAUNCH_CHECK();
      return 0;
    });
    std::vector<double> host(2 * nb);
    GTPS_HIP(hipMemcpyAsync(host.data(), pt, sizeof(double) * 2 * nb,
                            hipMemcpyDeviceToHost, s));
    GTPS_HIP(hipStreamSynchronize(s));
    double e0 = 0.0, e1 = 0.0;
    for (int i = 0; i < nb; ++i)
      {
        e0 += host[2 * i];
        e1 += host[2 * i + 1];
      }
    *l2 = std::sqrt(e0);
    *h1 = std::sqrt(e1);
  }

  void
  assemble_rhs(gtps_ctx *ctx, int dim, int k, int n, bool dg, int mode,
               const double *fq, double wave, double fc, double *b)
  {
    GTPS_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
    GTPS_REQUIRE(k >= 1 && k <= MAX_DEGREE, "degree must be in 1..7");
    GTPS_REQUIRE(n >= 1 && n <= 4096, "cells per direction must be in 1..4096");
    GTPS_REQUIRE(mode >= 0 && mode <= 2, "bad right-hand-side mode");
    GTPS_REQUIRE(mode != 0 || fq != nullptr, "f values are NULL");
    const double    h   = 1.0 / n;
    const QTab      tab = make_qtab(ctx, k, k + 1, h);
    const long long N1  = dg ? (long long)n * (k + 1) : (long long)k * n + 1;
    const long long N   = dim == 3 ? N1 * N1 * N1 : N1 * N1;
    const int       nb  = pgrid(N, 16384);
    const double    vol = dim == 3 ? h * h * h : h * h;
    const double    a   = wave * M_PI;
    hipStream_t     s   = ctx->stream;
    dispatch_degree(k, [&](auto kk) {
      constexpr int K = decltype(kk)::value;
      auto go = [&](auto dd, auto gg, auto mm) {
        constexpr int  D  = decltype(dd)::value;
        constexpr bool G  = decltype(gg)::value;
        constexpr int  MO = decltype(mm)::value;
        hipLaunchKernelGGL((assemble_rhs_kernel<K, D, G, MO>), dim3(nb), dim3(PNT), 0, s,
                           fq, b, n, a, fc, vol, tab);
      };
      auto by_mode = [&](auto dd, auto gg) {
        if (mode == 0)
          go(dd, gg, std::integral_constant<int, 0>{});
        else if (mode == 1)
          go(dd, gg, std::integral_constant<int, 1>{});
        else
          go(dd, gg, std::integral_constant<int, 2>{});
      };
      auto by_layout = [&](auto dd) {
        if (dg)
          by_mode(dd, std::true_type{});
        else
          by_mode(dd, std::false_type{});
      };
      if (dim == 3)
        by_layout(std::integral_constant<int, 3>{});
      else
        by_layout(std::integral_constant<int, 2>{});
      GTPS_LAUNCH_CHECK();
      return 0;
    });
  }
} // namespace gtps
