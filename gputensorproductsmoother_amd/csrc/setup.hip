// setup.hip - per-degree 1D data: Gauss-Lobatto Lagrange basis, Gauss(k+1)
// cell matrices, 2-cell patch matrices and embedding matrix on the host in
// FP64, and the generalized eigenproblem K_I s = lambda M_I s of the patch
// interior solved on the GPU in FP64 (one wavefront, Cholesky + cyclic
// Jacobi).  The results are packed in the Tab<K> layout for FP64 and FP32.
//
// Reference counterparts: reinit_tensor_product
// (include/patch_base.template.cuh:709-1205) and TensorProductData
// (include/tensor_product.h:60-284), which assemble the same 1D matrices
// with deal.II FE_DGQ<1>/FE_Q<1> + QGauss(k+1) and call deal.II's LAPACK
// based TensorProductMatrixSymmetricSum::reinit in the smoother's precision.
#include <cmath>

#include "gtps_internal.h"

namespace gtps
{
  namespace
  {
    // Legendre P_n and P_{n-1} at x by the three-term recurrence.
    void
    legendre_pair(int n, long double x, long double &pn, long double &pn1)
    {
      long double a = 1.0L, b = x;
      if (n == 0)
        {
          pn  = 1.0L;
          pn1 = 0.0L;
          return;
        }
      for (int m = 1; m < n; ++m)
        {
          const long double c = ((2 * m + 1) * x * b - m * a) / (m + 1);
          a                   = b;
          b                   = c;
        }
      pn  = b;
      pn1 = a;
    }

    // Interior Gauss-Lobatto nodes are the roots of P_k' on (-1, 1); Newton
    // on P_k' with P_k'' from Legendre's equation, started from the
    // Chebyshev-Lobatto points.
    std::vector<double>
    lobatto_nodes(int k)
    {
      std::vector<double> x(k + 1);
      x[0] = 0.0;
      x[k] = 1.0;
      for (int i = 1; i < k; ++i)
        {
          long double t = -std::cos(M_PI * (long double)i / k);
          for (int it = 0; it < 200; ++it)
            {
              long double pk, pk1;
              legendre_pair(k, t, pk, pk1);
              // P_k' = k (P_{k-1} - t P_k) / (1 - t^2)
              const long double dpk = k * (pk1 - t * pk) / (1 - t * t);
              // P_k'' from Legendre's ODE
              const long double d2pk =
                (2 * t * dpk - k * (k + 1) * pk) / (1 - t * t);
              const long double dt = dpk / d2pk;
              t -= dt;
              if (std::fabs((double)dt) < 1e-20)
                break;
            }
          x[i] = (double)((t + 1.0L) * 0.5L);
        }
      return x;
    }

    void
    gauss_points(int n, std::vector<double> &p, std::vector<double> &w)
    {
      p.resize(n);
      w.resize(n);
      for (int i = 0; i < n; ++i)
        {
          long double t = -std::cos(M_PI * (i + 0.75L) / (n + 0.5L));
          long double pn = 0, pn1 = 0, dpn = 0;
          for (int it = 0; it < 200; ++it)
            {
              legendre_pair(n, t, pn, pn1);
              dpn                  = n * (pn1 - t * pn) / (1 - t * t);
              const long double dt = pn / dpn;
              t -= dt;
              if (std::fabs((double)dt) < 1e-20)
                break;
            }
          legendre_pair(n, t, pn, pn1);
          dpn  = n * (pn1 - t * pn) / (1 - t * t);
          p[i] = (double)((t + 1.0L) * 0.5L);
          w[i] = (double)(1.0L / ((1 - t * t) * dpn * dpn));
        }
    }

    // Lagrange basis values and derivatives at x through barycentric
    // weights (one pass over the nodes per basis function).
    void
    lagrange_eval(const std::vector<double> &xn, double x, double *val,
                  double *der)
    {
      const int n = (int)xn.size();
      for (int i = 0; i < n; ++i)
        {
          double v = 1.0, d = 0.0, denom = 1.0;
          for (int j = 0; j < n; ++j)
            {
              if (j == i)
                continue;
              denom *= xn[i] - xn[j];
              // product rule: (v * (x - xj))' = d * (x - xj) + v
              d = d * (x - xn[j]) + v;
              v *= x - xn[j];
            }
          val[i] = v / denom;
          if (der)
            der[i] = d / denom;
        }
    }

    // One wavefront: generalized eigenproblem of an n x n SPD pencil (n <= 16)
    // K s = lam M s by M = L L^T, C = L^-1 K L^-T, cyclic Jacobi on C,
    // S = L^-T Q.  Eigenvalues ascending; S^T M S = I.  Work-group b solves
    // the b-th of a batch of pencils stored back to back.
    __global__ void __launch_bounds__(64)
    gen_eig_kernel(int n, const double *__restrict__ Kin,
                   const double *__restrict__ Min, double *__restrict__ lam,
                   double *__restrict__ Sout)
    {
      Kin += (size_t)blockIdx.x * n * n;
      Min += (size_t)blockIdx.x * n * n;
      Sout += (size_t)blockIdx.x * n * n;
      lam += (size_t)blockIdx.x * n;
      __shared__ double L[16 * 16], Li[16 * 16], C[16 * 16], V[16 * 16],
        T[16 * 16];
      __shared__ double cs[2];
      __shared__ int    order[16];
      __shared__ int    done;
      const int t = threadIdx.x;
      for (int e = t; e < n * n; e += 64)
        {
          L[e]  = 0.0;
          Li[e] = 0.0;
          V[e]  = (e / n == e % n) ? 1.0 : 0.0;
        }
      __syncthreads();
      // Cholesky, column by column
      for (int j = 0; j < n; ++j)
        {
          if (t == 0)
            {
              double s = Min[j * n + j];
              for (int m = 0; m < j; ++m)
                s -= L[j * n + m] * L[j * n + m];
              L[j * n + j] = sqrt(s);
            }
          __syncthreads();
          if (t > j && t < n)
            {
              double s = Min[t * n + j];
              for (int m = 0; m < j; ++m)
                s -= L[t * n + m] * L[j * n + m];
              L[t * n + j] = s / L[j * n + j];
            }
          __syncthreads();
        }
      // inverse of L: lane j solves column j
      if (t < n)
        {
          const int j   = t;
          Li[j * n + j] = 1.0 / L[j * n + j];
          for (int i = j + 1; i < n; ++i)
            {
              double s = 0.0;
              for (int m = j; m < i; ++m)
                s -= L[i * n + m] * Li[m * n + j];
              Li[i * n + j] = s / L[i * n + i];
            }
        }
      __syncthreads();
      for (int e = t; e < n * n; e += 64)
        {
          const int i = e / n, j = e % n;
          double    s = 0.0;
          for (int m = 0; m < n; ++m)
            s += Li[i * n + m] * Kin[m * n + j];
          T[e] = s;
        }
      __syncthreads();
      for (int e = t; e < n * n; e += 64)
        {
          const int i = e / n, j = e % n;
          double    s = 0.0;
          for (int m = 0; m < n; ++m)
            s += T[i * n + m] * Li[j * n + m];
          C[e] = s;
        }
      __syncthreads();
      for (int e = t; e < n * n; e += 64)
        {
          const int i = e / n, j = e % n;
          if (i > j)
            {
              const double s = 0.5 * (C[i * n + j] + C[j * n + i]);
              C[i * n + j]   = s;
              C[j * n + i]   = s;
            }
        }
      __syncthreads();
      for (int sweep = 0; sweep < 64; ++sweep)
        {
          if (t == 0)
            {
              double off = 0.0, diag = 0.0;
              for (int e = 0; e < n * n; ++e)
                {
                  if (e / n == e % n)
                    diag += C[e] * C[e];
                  else
                    off += C[e] * C[e];
                }
              done = off <= 1e-32 * diag;
            }
          __syncthreads();
          if (done)
            break;
          for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q)
              {
                if (t == 0)
                  {
                    const double apq = C[p * n + q];
                    double       c = 1.0, s = 0.0;
                    if (fabs(apq) > 1e-300)
                      {
                        const double theta =
                          (C[q * n + q] - C[p * n + p]) / (2.0 * apq);
                        const double tt =
                          (theta >= 0 ? 1.0 : -1.0) /
                          (fabs(theta) + sqrt(theta * theta + 1.0));
                        c = 1.0 / sqrt(tt * tt + 1.0);
                        s = tt * c;
                      }
                    cs[0] = c;
                    cs[1] = s;
                  }
                __syncthreads();
                const double c = cs[0], s = cs[1];
                if (t < n)
                  {
                    const double ap = C[t * n + p], aq = C[t * n + q];
                    C[t * n + p]    = c * ap - s * aq;
                    C[t * n + q]    = s * ap + c * aq;
                    const double vp = V[t * n + p], vq = V[t * n + q];
                    V[t * n + p]    = c * vp - s * vq;
                    V[t * n + q]    = s * vp + c * vq;
                  }
                __syncthreads();
                if (t < n)
                  {
                    const double ap = C[p * n + t], aq = C[q * n + t];
                    C[p * n + t]    = c * ap - s * aq;
                    C[q * n + t]    = s * ap + c * aq;
                  }
                __syncthreads();
              }
        }
      if (t == 0)
        {
          for (int i = 0; i < n; ++i)
            order[i] = i;
          for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j)
              if (C[order[j] * n + order[j]] < C[order[i] * n + order[i]])
                {
                  const int tmp = order[i];
                  order[i]      = order[j];
                  order[j]      = tmp;
                }
        }
      __syncthreads();
      for (int e = t; e < n * n; e += 64)
        {
          const int i = e / n, jj = e % n, j = order[jj];
          double    s = 0.0;
          for (int m = 0; m < n; ++m)
            s += Li[m * n + i] * V[m * n + j];
          Sout[i * n + jj] = s;
        }
      if (t < n)
        lam[t] = C[order[t] * n + order[t]];
    }

    template <int K>
    std::vector<double>
    pack_tables(const Basis1D &b)
    {
      using TB = Tab<K>;
      std::vector<double> tab(TB::size, 0.0);
      std::copy(b.cellM.begin(), b.cellM.end(), tab.begin() + TB::cellM);
      std::copy(b.cellK.begin(), b.cellK.end(), tab.begin() + TB::cellK);
      std::copy(b.patchM.begin(), b.patchM.end(), tab.begin() + TB::patchM);
      std::copy(b.patchK.begin(), b.patchK.end(), tab.begin() + TB::patchK);
      std::copy(b.S.begin(), b.S.end(), tab.begin() + TB::S);
      std::copy(b.lam.begin(), b.lam.end(), tab.begin() + TB::lam);
      std::copy(b.E.begin(), b.E.end(), tab.begin() + TB::E);
      const int N = K + 1, W = 2 * K + 1;
      for (int which = 0; which < 2; ++which)
        {
          const std::vector<double> &C   = which == 0 ? b.cellM : b.cellK;
          const int                  off = which == 0 ? TB::bandM : TB::bandK;
          for (int p = 0; p < K; ++p)
            for (int o = -K; o <= K; ++o)
              {
                double v = 0.0;
                if (p == 0)
                  {
                    // vertex row: last row of the left cell + first row of
                    // the right cell
                    if (o <= 0)
                      v += C[K * N + K + o];
                    if (o >= 0)
                      v += C[0 * N + o];
                  }
                else if (o >= -p && o <= K - p)
                  v = C[p * N + p + o];
                tab[off + p * W + o + K] = v;
              }
        }
      const int NI = 2 * K - 1;
      for (int c = 0; c < NI; ++c)
        for (int b2 = 0; b2 < NI; ++b2)
          {
            tab[TB::lamInv2 + c * NI + b2] = 1.0 / (b.lam[c] + b.lam[b2]);
            for (int a = 0; a < NI; ++a)
              tab[TB::lamInv3 + (c * NI + b2) * NI + a] =
                1.0 / (b.lam[c] + b.lam[b2] + b.lam[a]);
          }
      // even/odd classification of the eigenvectors (column j of S):
      // even if S[i][j] = S[NI-1-i][j], odd if S[i][j] = -S[NI-1-i][j]
      std::vector<int> even, odd;
      for (int j = 0; j < NI; ++j)
        {
          double de = 0.0, dodd = 0.0, nrm = 0.0;
          for (int i = 0; i < NI; ++i)
            {
              const double a = b.S[i * NI + j], m = b.S[(NI - 1 - i) * NI + j];
              de   = std::max(de, std::fabs(a - m));
              dodd = std::max(dodd, std::fabs(a + m));
              nrm  = std::max(nrm, std::fabs(a));
            }
          if (de <= 1e-10 * nrm)
            even.push_back(j);
          else if (dodd <= 1e-10 * nrm)
            odd.push_back(j);
          else
            throw Error(GTPS_EHIP, "patch eigenvector " + std::to_string(j) +
                                     " is neither even nor odd");
        }
      GTPS_REQUIRE((int)even.size() == TB::NE && (int)odd.size() == TB::NO,
                   "unexpected even/odd split of the patch eigenbasis");
      std::vector<int> perm(even);
      perm.insert(perm.end(), odd.begin(), odd.end());
      for (int i = 0; i <= TB::MH; ++i)
        for (int e = 0; e < TB::NE; ++e)
          tab[TB::SE + i * TB::NE + e] = b.S[i * NI + even[e]];
      for (int i = 0; i < TB::MH; ++i)
        for (int o = 0; o < TB::NO; ++o)
          tab[TB::SO + i * TB::NO + o] = b.S[i * NI + odd[o]];
      for (int c = 0; c < NI; ++c)
        for (int b2 = 0; b2 < NI; ++b2)
          for (int a = 0; a < NI; ++a)
            tab[TB::lamInv3p + (c * NI + b2) * NI + a] =
              1.0 / (b.lam[perm[c]] + b.lam[perm[b2]] + b.lam[perm[a]]);
      return tab;
    }

    std::vector<double>
    pack_any(int k, const Basis1D &b)
    {
      switch (k)
        {
          case 1: return pack_tables<1>(b);
          case 2: return pack_tables<2>(b);
          case 3: return pack_tables<3>(b);
          case 4: return pack_tables<4>(b);
          case 5: return pack_tables<5>(b);
          case 6: return pack_tables<6>(b);
          case 7: return pack_tables<7>(b);
          default: throw Error(GTPS_EINVAL, "degree must be in 1..7");
        }
    }
  } // namespace

  Basis1D
  make_basis_host(int k)
  {
    GTPS_REQUIRE(k >= 1 && k <= MAX_DEGREE, "degree must be in 1..7");
    Basis1D b;
    b.k     = k;
    b.nodes = lobatto_nodes(k);
    gauss_points(k + 1, b.qp, b.qw);
    const int           n = k + 1;
    std::vector<double> val(n), der(n);
    b.cellM.assign(n * n, 0.0);
    b.cellK.assign(n * n, 0.0);
    for (int q = 0; q < n; ++q)
      {
        lagrange_eval(b.nodes, b.qp[q], val.data(), der.data());
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j)
            {
              b.cellM[i * n + j] += b.qw[q] * val[i] * val[j];
              b.cellK[i * n + j] += b.qw[q] * der[i] * der[j];
            }
      }
    const int np = 2 * k + 1;
    b.E.resize(np * n);
    for (int q = 0; q < np; ++q)
      {
        // fine node q of a coarse cell: child 0 for q <= k, child 1 after
        const double x = q <= k ? 0.5 * b.nodes[q] : 0.5 + 0.5 * b.nodes[q - k];
        lagrange_eval(b.nodes, x, val.data(), nullptr);
        for (int a = 0; a < n; ++a)
          b.E[q * n + a] = val[a];
      }
    b.patchM.assign(np * np, 0.0);
    b.patchK.assign(np * np, 0.0);
    for (int c = 0; c < 2; ++c)
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
          {
            b.patchM[(c * k + i) * np + c * k + j] += b.cellM[i * n + j];
            b.patchK[(c * k + i) * np + c * k + j] += b.cellK[i * n + j];
          }
    return b;
  }

  void
  lagrange_values(const std::vector<double> &nodes, double x, double *val,
                  double *der)
  {
    lagrange_eval(nodes, x, val, der);
  }

  void
  gauss_rule(int n, std::vector<double> &p, std::vector<double> &w)
  {
    gauss_points(n, p, w);
  }

  void
  gpu_gen_eig(gtps_ctx *ctx, int n, const std::vector<double> &K,
              const std::vector<double> &M, std::vector<double> &lam,
              std::vector<double> &S)
  {
    GTPS_REQUIRE(n >= 1 && n <= 16, "eigenproblem size must be in 1..16");
    double *dbuf = nullptr;
    GTPS_HIP(hipMalloc(&dbuf, sizeof(double) * (3 * n * n + n)));
    double *dK = dbuf, *dM = dbuf + n * n, *dS = dbuf + 2 * n * n,
           *dl = dbuf + 3 * n * n;
    try
      {
        GTPS_HIP(hipMemcpyAsync(dK, K.data(), sizeof(double) * n * n,
                                hipMemcpyHostToDevice, ctx->stream));
        GTPS_HIP(hipMemcpyAsync(dM, M.data(), sizeof(double) * n * n,
                                hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(gen_eig_kernel, dim3(1), dim3(64), 0, ctx->stream,
                           n, dK, dM, dl, dS);
        GTPS_LAUNCH_CHECK();
        lam.resize(n);
        S.resize(n * n);
        GTPS_HIP(hipMemcpyAsync(lam.data(), dl, sizeof(double) * n,
                                hipMemcpyDeviceToHost, ctx->stream));
        GTPS_HIP(hipMemcpyAsync(S.data(), dS, sizeof(double) * n * n,
                                hipMemcpyDeviceToHost, ctx->stream));
        GTPS_HIP(hipStreamSynchronize(ctx->stream));
      }
    catch (...)
      {
        (void)hipFree(dbuf);
        throw;
      }
    GTPS_HIP(hipFree(dbuf));
    for (double v : lam)
      GTPS_REQUIRE(std::isfinite(v) && v > 0.0,
                   "patch eigenproblem produced a non-positive eigenvalue");
  }

  const void *
  get_sep_tables(gtps_ctx *ctx, int k, int precision)
  {
    DegreeTables *dt = get_tables(ctx, k);
    if (!dt->sep64)
      {
        const Basis1D &b  = dt->basis;
        const int      n  = k + 1, np = 2 * k + 1, ni = 2 * k - 1, nn = ni * ni;
        std::vector<double> KI((size_t)SEP_BINS * nn), MI((size_t)SEP_BINS * nn);
        for (int bin = 0; bin < SEP_BINS; ++bin)
          {
            const double        rho = std::exp((bin - SEP_HALF) * SEP_STEP);
            std::vector<double> Mp(np * np, 0.0), Kp(np * np, 0.0);
            for (int c = 0; c < 2; ++c)
              for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j)
                  {
                    const double w = c ? rho : 1.0;
                    Mp[(c * k + i) * np + c * k + j] += w * b.cellM[i * n + j];
                    Kp[(c * k + i) * np + c * k + j] += w * b.cellK[i * n + j];
                  }
            for (int i = 0; i < ni; ++i)
              for (int j = 0; j < ni; ++j)
                {
                  MI[(size_t)bin * nn + i * ni + j] = Mp[(i + 1) * np + j + 1];
                  KI[(size_t)bin * nn + i * ni + j] = Kp[(i + 1) * np + j + 1];
                }
          }
        const size_t        nb = SEP_BINS;
        std::vector<double> lam(nb * ni), S(nb * nn);
        double             *dbuf = nullptr;
        GTPS_HIP(hipMalloc(&dbuf, sizeof(double) * nb * (3 * nn + ni)));
        double *dK = dbuf, *dM = dbuf + nb * nn, *dS = dbuf + 2 * nb * nn,
               *dl = dbuf + 3 * nb * nn;
        try
          {
            GTPS_HIP(hipMemcpyAsync(dK, KI.data(), sizeof(double) * nb * nn,
                                    hipMemcpyHostToDevice, ctx->stream));
            GTPS_HIP(hipMemcpyAsync(dM, MI.data(), sizeof(double) * nb * nn,
                                    hipMemcpyHostToDevice, ctx->stream));
            hipLaunchKernelGGL(gen_eig_kernel, dim3((unsigned)nb), dim3(64), 0,
                               ctx->stream, ni, dK, dM, dl, dS);
            GTPS_LAUNCH_CHECK();
            GTPS_HIP(hipMemcpyAsync(lam.data(), dl, sizeof(double) * nb * ni,
                                    hipMemcpyDeviceToHost, ctx->stream));
            GTPS_HIP(hipMemcpyAsync(S.data(), dS, sizeof(double) * nb * nn,
                                    hipMemcpyDeviceToHost, ctx->stream));
            GTPS_HIP(hipStreamSynchronize(ctx->stream));
          }
        catch (...)
          {
            (void)hipFree(dbuf);
            throw;
          }
        GTPS_HIP(hipFree(dbuf));
        for (double v : lam)
          GTPS_REQUIRE(std::isfinite(v) && v > 0.0,
                       "ratio-grid eigenproblem produced a non-positive eigenvalue");
        const int           w = SEP_W(k);
        std::vector<double> t64(nb * w);
        for (size_t bin = 0; bin < nb; ++bin)
          {
            std::copy(S.begin() + bin * nn, S.begin() + (bin + 1) * nn, t64.begin() + bin * w);
            std::copy(lam.begin() + bin * ni, lam.begin() + (bin + 1) * ni,
                      t64.begin() + bin * w + nn);
          }
        std::vector<float> t32(t64.begin(), t64.end());
        double *d64 = nullptr;
        float  *d32 = nullptr;
        GTPS_HIP(hipMalloc(&d64, sizeof(double) * t64.size()));
        if (hipMalloc(&d32, sizeof(float) * t32.size()) != hipSuccess)
          {
            (void)hipGetLastError();
            (void)hipFree(d64);
            throw Error(GTPS_ENOMEM, "hipMalloc of the ratio-grid tables failed");
          }
        GTPS_HIP(hipMemcpyAsync(d64, t64.data(), sizeof(double) * t64.size(),
                                hipMemcpyHostToDevice, ctx->stream));
        GTPS_HIP(hipMemcpyAsync(d32, t32.data(), sizeof(float) * t32.size(),
                                hipMemcpyHostToDevice, ctx->stream));
        GTPS_HIP(hipStreamSynchronize(ctx->stream));
        dt->sep64 = d64;
        dt->sep32 = d32;
      }
    return precision == GTPS_FP32 ? (const void *)dt->sep32 : (const void *)dt->sep64;
  }

  DegreeTables *
  get_tables(gtps_ctx *ctx, int k)
  {
    auto it = ctx->tables.find(k);
    if (it != ctx->tables.end())
      return it->second.get();
    auto dt   = std::make_unique<DegreeTables>();
    dt->basis = make_basis_host(k);
    Basis1D  &b  = dt->basis;
    const int ni = 2 * k - 1, np = 2 * k + 1;
    std::vector<double> MI(ni * ni), KI(ni * ni);
    for (int i = 0; i < ni; ++i)
      for (int j = 0; j < ni; ++j)
        {
          MI[i * ni + j] = b.patchM[(i + 1) * np + j + 1];
          KI[i * ni + j] = b.patchK[(i + 1) * np + j + 1];
        }
    double *dbuf = nullptr;
    GTPS_HIP(hipMalloc(&dbuf, sizeof(double) * (4 * ni * ni + ni)));
    double *dK = dbuf, *dM = dbuf + ni * ni, *dS = dbuf + 2 * ni * ni,
           *dl = dbuf + 3 * ni * ni;
    try
      {
        GTPS_HIP(hipMemcpyAsync(dK, KI.data(), sizeof(double) * ni * ni,
                                hipMemcpyHostToDevice, ctx->stream));
        GTPS_HIP(hipMemcpyAsync(dM, MI.data(), sizeof(double) * ni * ni,
                                hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(gen_eig_kernel, dim3(1), dim3(64), 0, ctx->stream,
                           ni, dK, dM, dl, dS);
        GTPS_LAUNCH_CHECK();
        b.lam.resize(ni);
        b.S.resize(ni * ni);
        GTPS_HIP(hipMemcpyAsync(b.lam.data(), dl, sizeof(double) * ni,
                                hipMemcpyDeviceToHost, ctx->stream));
        GTPS_HIP(hipMemcpyAsync(b.S.data(), dS, sizeof(double) * ni * ni,
                                hipMemcpyDeviceToHost, ctx->stream));
        GTPS_HIP(hipStreamSynchronize(ctx->stream));
      }
    catch (...)
      {
        (void)hipFree(dbuf);
        throw;
      }
    GTPS_HIP(hipFree(dbuf));
    for (double v : b.lam)
      GTPS_REQUIRE(std::isfinite(v) && v > 0.0,
                   "patch eigenproblem produced a non-positive eigenvalue");

    const std::vector<double> t64 = pack_any(k, b);
    std::vector<float>        t32(t64.begin(), t64.end());
    GTPS_HIP(hipMalloc(&dt->d64, sizeof(double) * t64.size()));
    GTPS_HIP(hipMalloc(&dt->d32, sizeof(float) * t32.size()));
    GTPS_HIP(hipMemcpyAsync(dt->d64, t64.data(), sizeof(double) * t64.size(),
                            hipMemcpyHostToDevice, ctx->stream));
    GTPS_HIP(hipMemcpyAsync(dt->d32, t32.data(), sizeof(float) * t32.size(),
                            hipMemcpyHostToDevice, ctx->stream));
    GTPS_HIP(hipStreamSynchronize(ctx->stream));
    DegreeTables *raw = dt.get();
    ctx->tables[k]    = std::move(dt);
    return raw;
  }
} // namespace gtps
