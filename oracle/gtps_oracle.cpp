// gtps_oracle.cpp - FP64 CPU oracle for the geometric-multigrid hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker: tests/, the smoke()
// entry and bench.py's CPU-baseline leg load it through ctypes. The product
// library (gputensorproductsmoother_amd/csrc) never links or calls it.
//
// It restates, in plain FP64 C++ (OpenMP over independent cells/patches),
// the algorithm of the reference's hot path on a uniform Cartesian mesh of
// the unit square/cube with Q_k Lagrange elements (Gauss-Lobatto nodes,
// lexicographic numbering, boundary DoFs stored with identity rows):
//
//   * 1D basis: GL nodes, Gauss(k+1) quadrature, cell mass/stiffness
//     (patch_base.template.cuh:709-760, tensor_product.h:107-201)
//   * generalized eigenpairs of the 2-cell patch interior pencil
//     K v = lambda M v (tensor_product.h:205-211 -> deal.II
//     TensorProductMatrixSymmetricSum), solved here by Cholesky + cyclic
//     Jacobi in FP64
//   * Laplace vmult as a cell loop that evaluates gradients at Gauss points,
//     integrates back and scatters (laplace_operator.cuh:34-101,
//     dealii/cuda_fee.cuh:273-529), then copies the Dirichlet rows
//     (dealii/cuda_mf.template.cuh:494-508)
//   * vertex-patch Schwarz smoother, coloured by patch-index parity
//     (patch_base.template.cuh:1208-1285), with the full-box local residual
//     r = b - A_P u and the fast-diagonalisation inverse
//     S (x) S (x) S diag(1/sum lambda) (S (x) S (x) S)^T written back as
//     relaxation * (u_I + A_II^{-1} r_I)
//     (loop_kernel.cuh:969-1060, evaluate_kernel.cuh:2058-2163, 501-590,
//     1327-1434); additive variant accumulates patch solves of one global
//     residual colour by colour
//   * MG transfer as a coarse-cell loop with the 1D embedding matrix and
//     valence weights (cuda_mg_transfer.template.cuh:48-369, 561-671)
//   * V(1,1)-cycle (solver.cuh:1952-2001) with optional reversed colour order
//     for post-smoothing, and MG-preconditioned CG.
//   * variable-coefficient diffusion (no reference counterpart: the
//     reference's MultigridSolvers ignores its coefficient, solver.cuh:1548):
//     per-cell coefficient in the same cell loop, the transpose of the
//     constrained operator, a vertex-patch smoother whose inexact local
//     solve is the FD inverse of a separable (rank-1) fit of the patch's
//     cell coefficients, coefficient coarsening by averaging, and
//     dJ/dc_K = -lambda_K^T A^1_K u_K.
//   * post-processing: load-vector assembly from f at the Gauss points
//     (laplace_operator.cuh:140-173) and the L2 / H1-seminorm error against
//     prod sin(w pi x_d) (poisson.cu:44-74, 353-382), CG and DG layouts.
//
// Every function is pinned by tests/test_oracle_*.py against closed forms,
// high-precision golden fixtures (tests/golden) and brute-force dense
// restatements on tiny grids.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <map>

#include <omp.h>

namespace
{
  constexpr int MAXK = 8;

  // --------------------------------------------------------------- 1D basis
  void
  legendre(int n, long double x, long double &p, long double &dp)
  {
    long double p0 = 1.0L, p1 = x;
    if (n == 0)
      {
        p  = 1.0L;
        dp = 0.0L;
        return;
      }
    for (int m = 1; m < n; ++m)
      {
        long double p2 = ((2 * m + 1) * x * p1 - m * p0) / (m + 1);
        p0             = p1;
        p1             = p2;
      }
    p = p1;
    // P_n' = n (x P_n - P_{n-1}) / (x^2 - 1), valid away from +-1
    dp = n * (x * p1 - p0) / (x * x - 1.0L);
  }

  // Gauss-Lobatto nodes on [0,1]: endpoints and the roots of P_k'.
  std::vector<double>
  gl_nodes(int k)
  {
    std::vector<double> x(k + 1);
    x[0] = 0.0;
    x[k] = 1.0;
    for (int i = 1; i < k; ++i)
      {
        long double t = -std::cos(M_PI * i / k);
        for (int it = 0; it < 100; ++it)
          {
            long double p, dp;
            legendre(k, t, p, dp);
            // P'' from Legendre's equation
            long double d2p = (2 * t * dp - k * (k + 1) * p) / (1 - t * t);
            long double dt  = dp / d2p;
            t -= dt;
            if (std::fabs((double)dt) < 1e-19)
              break;
          }
        x[i] = (double)((t + 1.0L) / 2.0L);
      }
    return x;
  }

  // Gauss-Legendre rule with n points on [0,1].
  void
  gauss_rule(int n, std::vector<double> &pts, std::vector<double> &wts)
  {
    pts.resize(n);
    wts.resize(n);
    for (int i = 0; i < n; ++i)
      {
        long double t = -std::cos(M_PI * (i + 0.75) / (n + 0.5));
        long double p = 0, dp = 0;
        for (int it = 0; it < 100; ++it)
          {
            legendre(n, t, p, dp);
            long double dt = p / dp;
            t -= dt;
            if (std::fabs((double)dt) < 1e-19)
              break;
          }
        legendre(n, t, p, dp);
        pts[i] = (double)((t + 1.0L) / 2.0L);
        wts[i] = (double)(1.0L / ((1 - t * t) * dp * dp));
      }
  }

  double
  lagrange(const std::vector<double> &nodes, int i, double x)
  {
    double v = 1.0;
    for (size_t j = 0; j < nodes.size(); ++j)
      if ((int)j != i)
        v *= (x - nodes[j]) / (nodes[i] - nodes[j]);
    return v;
  }

  double
  lagrange_deriv(const std::vector<double> &nodes, int i, double x)
  {
    double s = 0.0;
    for (size_t m = 0; m < nodes.size(); ++m)
      {
        if ((int)m == i)
          continue;
        double v = 1.0 / (nodes[i] - nodes[m]);
        for (size_t j = 0; j < nodes.size(); ++j)
          if ((int)j != i && j != m)
            v *= (x - nodes[j]) / (nodes[i] - nodes[j]);
        s += v;
      }
    return s;
  }

  // Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (row-major).
  void
  jacobi_eig(int n, std::vector<double> A, std::vector<double> &w,
             std::vector<double> &V)
  {
    V.assign(n * n, 0.0);
    for (int i = 0; i < n; ++i)
      V[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep)
      {
        double off = 0.0, diag = 0.0;
        for (int p = 0; p < n; ++p)
          for (int q = 0; q < n; ++q)
            (p == q ? diag : off) += A[p * n + q] * A[p * n + q];
        if (off <= 1e-32 * diag)
          break;
        for (int p = 0; p < n; ++p)
          for (int q = p + 1; q < n; ++q)
            {
              const double apq = A[p * n + q];
              if (std::fabs(apq) < 1e-300)
                continue;
              const double theta = (A[q * n + q] - A[p * n + p]) / (2 * apq);
              const double t =
                (theta >= 0 ? 1.0 : -1.0) /
                (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
              const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
              for (int r = 0; r < n; ++r)
                {
                  const double ap = A[r * n + p], aq = A[r * n + q];
                  A[r * n + p]    = c * ap - s * aq;
                  A[r * n + q]    = s * ap + c * aq;
                }
              for (int r = 0; r < n; ++r)
                {
                  const double ap = A[p * n + r], aq = A[q * n + r];
                  A[p * n + r]    = c * ap - s * aq;
                  A[q * n + r]    = s * ap + c * aq;
                }
              for (int r = 0; r < n; ++r)
                {
                  const double vp = V[r * n + p], vq = V[r * n + q];
                  V[r * n + p]    = c * vp - s * vq;
                  V[r * n + q]    = s * vp + c * vq;
                }
            }
      }
    w.resize(n);
    for (int i = 0; i < n; ++i)
      w[i] = A[i * n + i];
  }

  // Generalized symmetric-definite eigenproblem K s = lambda M s, with
  // S^T M S = I, eigenvalues ascending. S[i*n+j] = component i of vector j.
  void
  gen_eig(int n, const double *K, const double *M, double *lam, double *S)
  {
    std::vector<double> L(n * n, 0.0);
    for (int j = 0; j < n; ++j)
      {
        double s = M[j * n + j];
        for (int m = 0; m < j; ++m)
          s -= L[j * n + m] * L[j * n + m];
        L[j * n + j] = std::sqrt(s);
        for (int i = j + 1; i < n; ++i)
          {
            double t = M[i * n + j];
            for (int m = 0; m < j; ++m)
              t -= L[i * n + m] * L[j * n + m];
            L[i * n + j] = t / L[j * n + j];
          }
      }
    // Linv (lower triangular)
    std::vector<double> Li(n * n, 0.0);
    for (int j = 0; j < n; ++j)
      {
        Li[j * n + j] = 1.0 / L[j * n + j];
        for (int i = j + 1; i < n; ++i)
          {
            double t = 0.0;
            for (int m = j; m < i; ++m)
              t -= L[i * n + m] * Li[m * n + j];
            Li[i * n + j] = t / L[i * n + i];
          }
      }
    // C = Li K Li^T
    std::vector<double> T(n * n, 0.0), C(n * n, 0.0);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        {
          double t = 0.0;
          for (int m = 0; m < n; ++m)
            t += Li[i * n + m] * K[m * n + j];
          T[i * n + j] = t;
        }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        {
          double t = 0.0;
          for (int m = 0; m < n; ++m)
            t += T[i * n + m] * Li[j * n + m];
          C[i * n + j] = t;
        }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < i; ++j)
        C[i * n + j] = C[j * n + i] = 0.5 * (C[i * n + j] + C[j * n + i]);
    std::vector<double> w, Q;
    jacobi_eig(n, C, w, Q);
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i)
      order[i] = i;
    std::sort(order.begin(), order.end(),
              [&](int a, int b) { return w[a] < w[b]; });
    for (int jj = 0; jj < n; ++jj)
      {
        const int j = order[jj];
        lam[jj]     = w[j];
        for (int i = 0; i < n; ++i)
          {
            // S = Li^T Q
            double t = 0.0;
            for (int m = 0; m < n; ++m)
              t += Li[m * n + i] * Q[m * n + j];
            S[i * n + jj] = t;
          }
      }
  }

  struct Basis
  {
    int                 k = 0;
    std::vector<double> nodes, qp, qw;
    std::vector<double> N, D;    // [q*(k+1)+i] values/derivatives at Gauss points
    std::vector<double> M, K;    // cell (k+1)^2, reference cell [0,1]
    std::vector<double> E;       // embedding (2k+1) x (k+1)
    std::vector<double> MP, KP;  // 2-cell patch (2k+1)^2
    std::vector<double> lam, S;  // interior (2k-1) eigenpairs
  };

  Basis
  make_basis(int k)
  {
    Basis b;
    b.k     = k;
    b.nodes = gl_nodes(k);
    gauss_rule(k + 1, b.qp, b.qw);
    const int n = k + 1;
    b.N.resize(n * n);
    b.D.resize(n * n);
    for (int q = 0; q < n; ++q)
      for (int i = 0; i < n; ++i)
        {
          b.N[q * n + i] = lagrange(b.nodes, i, b.qp[q]);
          b.D[q * n + i] = lagrange_deriv(b.nodes, i, b.qp[q]);
        }
    b.M.assign(n * n, 0.0);
    b.K.assign(n * n, 0.0);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        for (int q = 0; q < n; ++q)
          {
            b.M[i * n + j] += b.qw[q] * b.N[q * n + i] * b.N[q * n + j];
            b.K[i * n + j] += b.qw[q] * b.D[q * n + i] * b.D[q * n + j];
          }
    const int np = 2 * k + 1;
    b.E.resize(np * n);
    for (int q = 0; q < np; ++q)
      {
        const double x =
          q <= k ? 0.5 * b.nodes[q] : 0.5 + 0.5 * b.nodes[q - k];
        for (int a = 0; a < n; ++a)
          b.E[q * n + a] = lagrange(b.nodes, a, x);
      }
    b.MP.assign(np * np, 0.0);
    b.KP.assign(np * np, 0.0);
    for (int c = 0; c < 2; ++c)
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
          {
            b.MP[(c * k + i) * np + c * k + j] += b.M[i * n + j];
            b.KP[(c * k + i) * np + c * k + j] += b.K[i * n + j];
          }
    const int ni = 2 * k - 1;
    std::vector<double> MI(ni * ni), KI(ni * ni);
    for (int i = 0; i < ni; ++i)
      for (int j = 0; j < ni; ++j)
        {
          MI[i * ni + j] = b.MP[(i + 1) * np + j + 1];
          KI[i * ni + j] = b.KP[(i + 1) * np + j + 1];
        }
    b.lam.resize(ni);
    b.S.resize(ni * ni);
    gen_eig(ni, KI.data(), MI.data(), b.lam.data(), b.S.data());
    return b;
  }

  const Basis &
  basis(int k)
  {
    static Basis cache[MAXK + 1];
    static bool  have[MAXK + 1] = {};
#pragma omp critical(gtps_oracle_basis)
    {
      if (!have[k])
        {
          cache[k] = make_basis(k);
          have[k]  = true;
        }
    }
    return cache[k];
  }

  // ------------------------------------------------- small tensor utilities
  // Arrays are (n2, n1, n0) with axis 0 fastest. Apply A (rows x cols) along
  // `axis`: out[.., r, ..] (+)= sum_c A[r*cols+c] in[.., c, ..].
  void
  apply_axis(const double *A, int rows, int cols, const double *in,
             double *out, const int *shape, int axis, bool add,
             bool transpose = false)
  {
    int s_out[3] = {shape[0], shape[1], shape[2]};
    s_out[axis]  = rows;
    const int stride_in[3]  = {1, shape[0], shape[0] * shape[1]};
    const int stride_out[3] = {1, s_out[0], s_out[0] * s_out[1]};
    for (int i2 = 0; i2 < s_out[2]; ++i2)
      for (int i1 = 0; i1 < s_out[1]; ++i1)
        for (int i0 = 0; i0 < s_out[0]; ++i0)
          {
            int idx[3] = {i0, i1, i2};
            const int r = idx[axis];
            idx[axis]   = 0;
            const int base_in =
              idx[0] * stride_in[0] + idx[1] * stride_in[1] + idx[2] * stride_in[2];
            double s = 0.0;
            for (int c = 0; c < cols; ++c)
              {
                const double a = transpose ? A[c * rows + r] : A[r * cols + c];
                s += a * in[base_in + c * stride_in[axis]];
              }
            const int o = i0 * stride_out[0] + i1 * stride_out[1] +
                          i2 * stride_out[2];
            out[o] = add ? out[o] + s : s;
          }
  }

  struct Grid
  {
    int       dim, k, n;  // n cells per direction
    long long N1, N;      // DoFs per direction, total
    double    h;
    Grid(int dim_, int k_, int level)
      : dim(dim_), k(k_), n(1 << level)
    {
      N1 = (long long)k * n + 1;
      N  = dim == 2 ? N1 * N1 : N1 * N1 * N1;
      h  = 1.0 / n;
    }
    bool
    on_boundary(long long i) const
    {
      const long long x = i % N1, y = (i / N1) % N1, z = i / (N1 * N1);
      if (x == 0 || x == N1 - 1 || y == 0 || y == N1 - 1)
        return true;
      return dim == 3 && (z == 0 || z == N1 - 1);
    }
  };

  // ------------------------------------------------------------- operator
  // Cell (cx, cy[, cz]) has the lexicographic cell index cx + n cy + n^2 cz.
  inline long long
  cell_index(const Grid &g, int cx, int cy, int cz)
  {
    return cx + (long long)g.n * cy + (long long)g.n * g.n * cz * (g.dim == 3);
  }

  // Cell-wise quadrature form of the reference-cell Laplacian: out =
  // sum_d (D_d^T W N^T..)(D_d N..) u with JxW = h^dim w_q, J^{-1} = I/h.
  struct CellOp
  {
    const Basis        &b;
    int                 dim, n1d, nl;
    std::vector<double> wq, t1, t2, gq, out;
    int                 shape[3];
    CellOp(const Grid &g)
      : b(basis(g.k)), dim(g.dim), n1d(g.k + 1),
        nl(g.dim == 2 ? (g.k + 1) * (g.k + 1) : (g.k + 1) * (g.k + 1) * (g.k + 1))
    {
      const double scale = std::pow(g.h, dim - 2);
      wq.resize(nl);
      t1.resize(nl);
      t2.resize(nl);
      gq.resize(nl);
      out.resize(nl);
      shape[0] = shape[1] = n1d;
      shape[2]            = dim == 3 ? n1d : 1;
      for (int l = 0; l < shape[2]; ++l)
        for (int j = 0; j < n1d; ++j)
          for (int i = 0; i < n1d; ++i)
            wq[(l * n1d + j) * n1d + i] =
              b.qw[i] * b.qw[j] * (dim == 3 ? b.qw[l] : 1.0) * scale;
    }
    // acc (+)= coef * A_K u
    void
    apply(const double *u, double *acc, double coef, bool add)
    {
      if (!add)
        std::fill(acc, acc + nl, 0.0);
      for (int d = 0; d < dim; ++d)
        {
          const double *src     = u;
          double       *bufs[2] = {t1.data(), t2.data()};
          int           nb      = 0;
          for (int a = 0; a < dim; ++a)
            {
              double *dst = (a == dim - 1) ? gq.data() : bufs[nb++ & 1];
              apply_axis(a == d ? b.D.data() : b.N.data(), n1d, n1d, src, dst,
                         shape, a, false);
              src = dst;
            }
          for (int q = 0; q < nl; ++q)
            gq[q] *= wq[q];
          src = gq.data();
          nb  = 0;
          for (int a = 0; a < dim; ++a)
            {
              double *dst = (a == dim - 1) ? out.data() : bufs[nb++ & 1];
              apply_axis(a == d ? b.D.data() : b.N.data(), n1d, n1d, src, dst,
                         shape, a, false, true);
              src = dst;
            }
          for (int q = 0; q < nl; ++q)
            acc[q] += coef * out[q];
        }
    }
  };

  // y = A x, A = sum_K coef_K A_K (coef == nullptr: all 1) scattered from
  // the cells, then the Dirichlet rows. mode 0: identity rows y_B = x_B
  // (laplace_operator.cuh:86-101, dealii/cuda_mf.template.cuh:494-508);
  // mode 1: the transpose of that matrix, A^T x = [A_II x_I ; A_IB^T x_I +
  // x_B] (boundary inputs masked in the cell products, full boundary rows).
  void
  laplace_vmult(const Grid &g, const double *x, double *y,
                const double *coef = nullptr, int mode = 0)
  {
    const int n1d = g.k + 1;
    const int dim = g.dim;
    std::fill(y, y + g.N, 0.0);
    const int ncol = 1 << dim;
    for (int col = 0; col < ncol; ++col)
      {
        const int px = col & 1, py = (col >> 1) & 1, pz = (col >> 2) & 1;
        const int nz = dim == 3 ? g.n : 1;
#pragma omp parallel
        {
          CellOp              op(g);
          std::vector<double> u(op.nl), acc(op.nl);
#pragma omp for collapse(2) schedule(static)
          for (int cz = pz * (dim == 3); cz < nz; cz += (dim == 3 ? 2 : 1))
            for (int cy = py; cy < g.n; cy += 2)
              {
                if (dim == 2 && pz == 1)
                  continue;
                for (int cx = px; cx < g.n; cx += 2)
                  {
                    const long long base = (long long)g.k * cx +
                                           g.N1 * ((long long)g.k * cy) +
                                           g.N1 * g.N1 *
                                             ((long long)g.k * cz * (dim == 3));
                    for (int l = 0; l < op.shape[2]; ++l)
                      for (int j = 0; j < n1d; ++j)
                        for (int i = 0; i < n1d; ++i)
                          {
                            const long long gi = base + i + g.N1 * j + g.N1 * g.N1 * l;
                            u[(l * n1d + j) * n1d + i] =
                              (mode == 1 && g.on_boundary(gi)) ? 0.0 : x[gi];
                          }
                    op.apply(u.data(), acc.data(),
                             coef ? coef[cell_index(g, cx, cy, cz)] : 1.0, false);
                    for (int l = 0; l < op.shape[2]; ++l)
                      for (int j = 0; j < n1d; ++j)
                        for (int i = 0; i < n1d; ++i)
                          y[base + i + g.N1 * j + g.N1 * g.N1 * l] +=
                            acc[(l * n1d + j) * n1d + i];
                  }
              }
        }
      }
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < g.N; ++i)
      if (g.on_boundary(i))
        y[i] = mode == 1 ? y[i] + x[i] : x[i];
  }

  // ------------------------------------------------------------- smoother
  enum Variant
  {
    MULT_FWD = 0,
    MULT_REV = 1,
    ADDITIVE = 2
  };

  // One patch: box values ub, bb (full box), returns interior correction in
  // `delta` (interior, lexicographic) for r = rb_I (if residual_given) or
  // r = b - A_P u otherwise.
  struct PatchWork
  {
    std::vector<double> box, t1, t2, r, t;
    explicit PatchWork(int np)
      : box(np * np * np), t1(np * np * np), t2(np * np * np),
        r(np * np * np), t(np * np * np)
    {}
  };

  void
  fd_inverse(const Grid &g, const Basis &b, double *r, double *tmp)
  {
    // r: interior (ni^dim), in place -> A_II^{-1} r
    const int    ni     = 2 * g.k - 1;
    const int    shape[3] = {ni, ni, g.dim == 3 ? ni : 1};
    const double scale  = std::pow(g.h, g.dim - 2);
    double      *bufs[2] = {r, tmp};
    int          cur     = 0;
    for (int a = 0; a < g.dim; ++a)
      {
        apply_axis(b.S.data(), ni, ni, bufs[cur], bufs[cur ^ 1], shape, a,
                   false, true);  // S^T
        cur ^= 1;
      }
    for (int l = 0; l < shape[2]; ++l)
      for (int j = 0; j < ni; ++j)
        for (int i = 0; i < ni; ++i)
          {
            const double lam =
              b.lam[i] + b.lam[j] + (g.dim == 3 ? b.lam[l] : 0.0);
            bufs[cur][(l * ni + j) * ni + i] /= scale * lam;
          }
    for (int a = 0; a < g.dim; ++a)
      {
        apply_axis(b.S.data(), ni, ni, bufs[cur], bufs[cur ^ 1], shape, a,
                   false, false);  // S
        cur ^= 1;
      }
    if (cur != 0)
      std::memcpy(r, bufs[cur], sizeof(double) * shape[0] * shape[1] * shape[2]);
  }

  void
  smoother_step(const Grid &g, int variant, double relax, const double *b_,
                double *u)
  {
    const Basis &b   = basis(g.k);
    const int    k   = g.k;
    const int    np  = 2 * k + 1;
    const int    ni  = 2 * k - 1;
    const int    dim = g.dim;
    const int    npatch = g.n - 1;  // patches per direction
    const double scale  = std::pow(g.h, dim - 2);
    std::vector<double> res;
    if (variant == ADDITIVE)
      {
        res.resize(g.N);
        laplace_vmult(g, u, res.data());
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < g.N; ++i)
          res[i] = b_[i] - res[i];
      }
    const int ncol = 1 << dim;
    for (int cc = 0; cc < ncol; ++cc)
      {
        const int col = variant == MULT_REV ? ncol - 1 - cc : cc;
        const int px = col & 1, py = (col >> 1) & 1, pz = (col >> 2) & 1;
        const int nz = dim == 3 ? npatch : 1;
#pragma omp parallel
        {
          PatchWork   w(np);
          const int   shape[3]  = {np, np, dim == 3 ? np : 1};
          const int   shi[3]    = {ni, ni, dim == 3 ? ni : 1};
          const int   nbox      = shape[0] * shape[1] * shape[2];
          const int   nint      = shi[0] * shi[1] * shi[2];
          std::vector<double> bb(nbox), delta(nint);
#pragma omp for collapse(2) schedule(static)
          for (int iz = (dim == 3 ? pz : 0); iz < nz; iz += (dim == 3 ? 2 : 1))
            for (int iy = py; iy < npatch; iy += 2)
              {
                if (dim == 2 && pz == 1)
                  continue;
                for (int ix = px; ix < npatch; ix += 2)
                  {
                    const long long first =
                      (long long)k * ix + g.N1 * ((long long)k * iy) +
                      g.N1 * g.N1 * ((long long)k * iz * (dim == 3));
                    auto gidx = [&](int i, int j, int l) {
                      return first + i + g.N1 * j + g.N1 * g.N1 * l;
                    };
                    if (variant == ADDITIVE)
                      {
                        for (int l = 0; l < shi[2]; ++l)
                          for (int j = 0; j < ni; ++j)
                            for (int i = 0; i < ni; ++i)
                              delta[(l * ni + j) * ni + i] =
                                res[gidx(i + 1, j + 1, l + (dim == 3))];
                      }
                    else
                      {
                        for (int l = 0; l < shape[2]; ++l)
                          for (int j = 0; j < np; ++j)
                            for (int i = 0; i < np; ++i)
                              {
                                w.box[(l * np + j) * np + i] = u[gidx(i, j, l)];
                                bb[(l * np + j) * np + i]    = b_[gidx(i, j, l)];
                              }
                        // A_P u = scale * sum_d KP_d (x) MP_others
                        std::fill(w.r.begin(), w.r.begin() + nbox, 0.0);
                        for (int d = 0; d < dim; ++d)
                          {
                            const double *src = w.box.data();
                            double       *bufs[2] = {w.t1.data(), w.t2.data()};
                            int           nb      = 0;
                            for (int a = 0; a < dim; ++a)
                              {
                                double *dst = bufs[nb++ & 1];
                                apply_axis(a == d ? b.KP.data() : b.MP.data(),
                                           np, np, src, dst, shape, a, false);
                                src = dst;
                              }
                            for (int q = 0; q < nbox; ++q)
                              w.r[q] += src[q];
                          }
                        for (int l = 0; l < shi[2]; ++l)
                          for (int j = 0; j < ni; ++j)
                            for (int i = 0; i < ni; ++i)
                              {
                                const int q = ((l + (dim == 3)) * np + j + 1) * np + i + 1;
                                delta[(l * ni + j) * ni + i] =
                                  bb[q] - scale * w.r[q];
                              }
                      }
                    fd_inverse(g, b, delta.data(), w.t.data());
                    for (int l = 0; l < shi[2]; ++l)
                      for (int j = 0; j < ni; ++j)
                        for (int i = 0; i < ni; ++i)
                          {
                            const long long gi = gidx(i + 1, j + 1, l + (dim == 3));
                            const double    dv = delta[(l * ni + j) * ni + i];
                            if (variant == ADDITIVE)
                              u[gi] += relax * dv;
                            else
                              u[gi] = relax * (u[gi] + dv);
                          }
                  }
              }
        }
      }
  }

  // ------------------------ separable approximation of a patch operator
  // The patch operator sum_K c_K A_K over the 2^d cells of a vertex patch is
  // a Kronecker sum exactly when the coefficient is separable,
  // c_abc = s p_a q_b r_c (a, b, c the cell offsets along x, y, z): then
  //   sum c_abc A_abc = s (K_p (x) M_q (x) M_r + M_p (x) K_q (x) M_r + ...)
  // with K_p = K_cell0 + (p_1/p_0) K_cell1 the 2-cell 1D patch matrix whose
  // second cell is weighted by the ratio rho = p_1/p_0 (p_0 = 1), and the
  // fast-diagonalisation inverse applies with per-direction generalized
  // eigenpairs of (K_rho, M_rho).  For a general coefficient the local solve
  // uses the log-space least-squares fit: log rho_d = mean log c over the
  // upper cells along d minus mean over the lower ones, rounded to a grid of
  // step SEP_STEP (|log rho| <= SEP_HALF SEP_STEP; the eigenpairs of every
  // grid point are tabulated once per degree), fit f_abc = prod_d rho_d^bit,
  // and the scale
  //   sigma = max( max_K c_K / f_K / SEP_THETA, exp(mean_K log(c_K / f_K)) ).
  // A_P = sum c_K A_K <= max_K(c_K/f_K) sum f_K A_K (every A_K is positive
  // semi-definite), so sigma A~ >= A_P / SEP_THETA: the local correction has
  // lambda_max(A~^{-1} A_P,II) <= SEP_THETA < 2 and every patch update is an
  // A-norm contraction (the symmetric V-cycle stays SPD); SEP_THETA = 1.5
  // was the best of 1.0-1.95 in the oracle's MG-PCG iteration counts.  A
  // separable coefficient (c = 1 in particular) gives the exact local solve.
  constexpr int    SEP_HALF  = 64;
  constexpr double SEP_STEP  = 1.0 / 16.0;
  constexpr double SEP_THETA = 1.5;

  struct SepPencil
  {
    std::vector<double> lam, S;
  };
  const SepPencil &
  sep_pencil(int k, int bin)
  {
    static std::map<std::pair<int, int>, SepPencil> cache;
    const SepPencil *res = nullptr;
#pragma omp critical(gtps_oracle_sep)
    {
      auto it = cache.find({k, bin});
      if (it == cache.end())
        {
          const Basis &b   = basis(k);
          const int    n   = k + 1, np = 2 * k + 1, ni = 2 * k - 1;
          const double rho = std::exp(bin * SEP_STEP);
          std::vector<double> MPr(np * np, 0.0), KPr(np * np, 0.0);
          for (int c = 0; c < 2; ++c)
            for (int i = 0; i < n; ++i)
              for (int j = 0; j < n; ++j)
                {
                  const double w = c ? rho : 1.0;
                  MPr[(c * k + i) * np + c * k + j] += w * b.M[i * n + j];
                  KPr[(c * k + i) * np + c * k + j] += w * b.K[i * n + j];
                }
          std::vector<double> MI(ni * ni), KI(ni * ni);
          for (int i = 0; i < ni; ++i)
            for (int j = 0; j < ni; ++j)
              {
                MI[i * ni + j] = MPr[(i + 1) * np + j + 1];
                KI[i * ni + j] = KPr[(i + 1) * np + j + 1];
              }
          SepPencil p;
          p.lam.resize(ni);
          p.S.resize(ni * ni);
          gen_eig(ni, KI.data(), MI.data(), p.lam.data(), p.S.data());
          it = cache.emplace(std::make_pair(k, bin), std::move(p)).first;
        }
      res = &it->second;
    }
    return *res;
  }

  // grid bins of the per-direction log ratios and the scale sigma of the
  // patch whose lower-left cell is (ix, iy, iz)
  void
  separable_fit(const Grid &g, const double *coef, int ix, int iy, int iz, int *bins,
                double *sigma)
  {
    const int dim = g.dim, nc = 1 << dim;
    double    lc[8];
    for (int c = 0; c < nc; ++c)
      lc[c] = std::log(coef[cell_index(g, ix + (c & 1), iy + ((c >> 1) & 1),
                                       iz + ((c >> 2) & 1))]);
    bins[0] = bins[1] = bins[2] = 0;
    for (int a = 0; a < dim; ++a)
      {
        double d = 0.0;
        for (int c = 0; c < nc; ++c)
          d += ((c >> a) & 1) ? lc[c] : -lc[c];
        d /= nc / 2;
        bins[a] = (int)std::max<long>(-SEP_HALF, std::min<long>(SEP_HALF, std::lround(d / SEP_STEP)));
      }
    double rmax = -1e300, rsum = 0.0;
    for (int c = 0; c < nc; ++c)
      {
        double lf = 0.0;
        for (int a = 0; a < dim; ++a)
          if ((c >> a) & 1)
            lf += bins[a] * SEP_STEP;
        rmax = std::max(rmax, lc[c] - lf);
        rsum += lc[c] - lf;
      }
    *sigma = std::max(std::exp(rmax) / SEP_THETA, std::exp(rsum / nc));
  }

  // r <- (sigma A~)^{-1} r with A~ = scale sum_d K_rho_d (x) M_rho_others
  void
  fd_inverse_sep(const Grid &g, const int *bins, double sigma, double *r, double *tmp)
  {
    const int    ni       = 2 * g.k - 1;
    const int    shape[3] = {ni, ni, g.dim == 3 ? ni : 1};
    const double scale    = std::pow(g.h, g.dim - 2);
    const SepPencil *p[3];
    for (int a = 0; a < g.dim; ++a)
      p[a] = &sep_pencil(g.k, bins[a]);
    double *bufs[2] = {r, tmp};
    int     cur     = 0;
    for (int a = 0; a < g.dim; ++a)
      {
        apply_axis(p[a]->S.data(), ni, ni, bufs[cur], bufs[cur ^ 1], shape, a, false, true);
        cur ^= 1;
      }
    for (int l = 0; l < shape[2]; ++l)
      for (int j = 0; j < ni; ++j)
        for (int i = 0; i < ni; ++i)
          {
            const double lam =
              p[0]->lam[i] + p[1]->lam[j] + (g.dim == 3 ? p[2]->lam[l] : 0.0);
            bufs[cur][(l * ni + j) * ni + i] /= sigma * scale * lam;
          }
    for (int a = 0; a < g.dim; ++a)
      {
        apply_axis(p[a]->S.data(), ni, ni, bufs[cur], bufs[cur ^ 1], shape, a, false, false);
        cur ^= 1;
      }
    if (cur != 0)
      std::memcpy(r, bufs[cur], sizeof(double) * shape[0] * shape[1] * shape[2]);
  }

  // ------------------------------------------- variable-coefficient smoother
  // Multiplicative vertex-patch Schwarz for A = sum_K c_K A_K (no reference
  // counterpart: the reference ignores its coefficient, solver.cuh:1548).
  // The patch operator is no longer a Kronecker sum, so the local solve is
  // the fast-diagonalisation inverse of the separable approximation
  // sigma_P A~_P (separable_fit above), applied in correction form
  //   u_I <- u_I + relax (sigma_P A~_P)^{-1} (b - A u)_I,
  // with the local residual from the 2^d cells of the patch (the only cells
  // that touch its interior DoFs).
  void
  smoother_step_coeff(const Grid &g, const double *coef, int variant,
                      double relax, const double *b_, double *u)
  {
    const int    k      = g.k;
    const int    n1d    = k + 1;
    const int    np     = 2 * k + 1;
    const int    ni     = 2 * k - 1;
    const int    dim    = g.dim;
    const int    npatch = g.n - 1;
    const int    ncol   = 1 << dim;
    for (int cc = 0; cc < ncol; ++cc)
      {
        const int col = variant == MULT_REV ? ncol - 1 - cc : cc;
        const int px = col & 1, py = (col >> 1) & 1, pz = (col >> 2) & 1;
        const int nz = dim == 3 ? npatch : 1;
        if (dim == 2 && pz == 1)
          continue;
#pragma omp parallel
        {
          CellOp              op(g);
          const int           shi[3] = {ni, ni, dim == 3 ? ni : 1};
          const int           nbox   = np * np * (dim == 3 ? np : 1);
          std::vector<double> box(nbox), acc(nbox), ul(op.nl), al(op.nl),
            delta(shi[0] * shi[1] * shi[2]), tmp(delta.size());
#pragma omp for collapse(2) schedule(static)
          for (int iz = (dim == 3 ? pz : 0); iz < nz; iz += (dim == 3 ? 2 : 1))
            for (int iy = py; iy < npatch; iy += 2)
              for (int ix = px; ix < npatch; ix += 2)
                {
                  const long long first =
                    (long long)k * ix + g.N1 * ((long long)k * iy) +
                    g.N1 * g.N1 * ((long long)k * iz * (dim == 3));
                  auto gidx = [&](int i, int j, int l) {
                    return first + i + g.N1 * j + g.N1 * g.N1 * l;
                  };
                  for (int l = 0; l < (dim == 3 ? np : 1); ++l)
                    for (int j = 0; j < np; ++j)
                      for (int i = 0; i < np; ++i)
                        box[(l * np + j) * np + i] = u[gidx(i, j, l)];
                  std::fill(acc.begin(), acc.end(), 0.0);
                  for (int c = 0; c < (1 << dim); ++c)
                    {
                      const int ox = c & 1, oy = (c >> 1) & 1, oz = (c >> 2) & 1;
                      const double cK =
                        coef[cell_index(g, ix + ox, iy + oy, iz + oz)];
                      for (int l = 0; l < op.shape[2]; ++l)
                        for (int j = 0; j < n1d; ++j)
                          for (int i = 0; i < n1d; ++i)
                            ul[(l * n1d + j) * n1d + i] =
                              box[((oz * k + l) * np + oy * k + j) * np + ox * k + i];
                      op.apply(ul.data(), al.data(), cK, false);
                      for (int l = 0; l < op.shape[2]; ++l)
                        for (int j = 0; j < n1d; ++j)
                          for (int i = 0; i < n1d; ++i)
                            acc[((oz * k + l) * np + oy * k + j) * np + ox * k + i] +=
                              al[(l * n1d + j) * n1d + i];
                    }
                  for (int l = 0; l < shi[2]; ++l)
                    for (int j = 0; j < ni; ++j)
                      for (int i = 0; i < ni; ++i)
                        {
                          const int lz = l + (dim == 3);
                          delta[(l * ni + j) * ni + i] =
                            b_[gidx(i + 1, j + 1, lz)] -
                            acc[(lz * np + j + 1) * np + i + 1];
                        }
                  int    bins[3];
                  double sigma;
                  separable_fit(g, coef, ix, iy, iz, bins, &sigma);
                  fd_inverse_sep(g, bins, sigma, delta.data(), tmp.data());
                  for (int l = 0; l < shi[2]; ++l)
                    for (int j = 0; j < ni; ++j)
                      for (int i = 0; i < ni; ++i)
                        u[gidx(i + 1, j + 1, l + (dim == 3))] +=
                          relax * delta[(l * ni + j) * ni + i];
                }
        }
      }
  }

  // Coarse-level coefficient: mean of the 2^d children.
  void
  coarsen_coefficient(int dim, int nc, const double *fine, double *coarse)
  {
    const int nf = 2 * nc;
    for (int cz = 0; cz < (dim == 3 ? nc : 1); ++cz)
      for (int cy = 0; cy < nc; ++cy)
        for (int cx = 0; cx < nc; ++cx)
          {
            double s = 0.0;
            for (int c = 0; c < (1 << dim); ++c)
              {
                const int fx = 2 * cx + (c & 1), fy = 2 * cy + ((c >> 1) & 1),
                          fz = 2 * cz + ((c >> 2) & 1);
                s += fine[fx + (long long)nf * fy + (long long)nf * nf * fz * (dim == 3)];
              }
            coarse[cx + (long long)nc * cy + (long long)nc * nc * cz * (dim == 3)] =
              s / (1 << dim);
          }
  }

  // dJ/dc_K = -lambda~_K^T A^1_K u_K with lambda~ = lambda with its Dirichlet
  // entries zeroed (the derivative of the interior rows of A(c) u).
  void
  coefficient_gradient(const Grid &g, const double *lambda, const double *u,
                       double *grad)
  {
    const int n1d = g.k + 1;
    const int dim = g.dim;
    const int nz  = dim == 3 ? g.n : 1;
#pragma omp parallel
    {
      CellOp              op(g);
      std::vector<double> ul(op.nl), al(op.nl);
#pragma omp for collapse(2) schedule(static)
      for (int cz = 0; cz < nz; ++cz)
        for (int cy = 0; cy < g.n; ++cy)
          for (int cx = 0; cx < g.n; ++cx)
            {
              const long long base = (long long)g.k * cx + g.N1 * ((long long)g.k * cy) +
                                     g.N1 * g.N1 * ((long long)g.k * cz * (dim == 3));
              for (int l = 0; l < op.shape[2]; ++l)
                for (int j = 0; j < n1d; ++j)
                  for (int i = 0; i < n1d; ++i)
                    ul[(l * n1d + j) * n1d + i] = u[base + i + g.N1 * j + g.N1 * g.N1 * l];
              op.apply(ul.data(), al.data(), 1.0, false);
              double s = 0.0;
              for (int l = 0; l < op.shape[2]; ++l)
                for (int j = 0; j < n1d; ++j)
                  for (int i = 0; i < n1d; ++i)
                    {
                      const long long gi = base + i + g.N1 * j + g.N1 * g.N1 * l;
                      if (!g.on_boundary(gi))
                        s += lambda[gi] * al[(l * n1d + j) * n1d + i];
                    }
              grad[cell_index(g, cx, cy, cz)] = -s;
            }
    }
  }

  // ------------------------------------------------------------- transfer
  // weight of fine local node q (0..2k) of coarse cell c (0..nc-1): 1/2 on
  // coarse-cell faces shared with a neighbour cell, else 1 (valence weights).
  inline double
  vweight(int q, int c, int nc, int k)
  {
    if (q == 0 && c > 0)
      return 0.5;
    if (q == 2 * k && c < nc - 1)
      return 0.5;
    return 1.0;
  }

  void
  transfer(const Grid &gf, bool prolongate, double *coarse_, double *fine_,
           const double *coarse_in, const double *fine_in)
  {
    // gf: fine grid; coarse grid has gf.n/2 cells
    const Basis &b   = basis(gf.k);
    const int    k   = gf.k;
    const int    dim = gf.dim;
    const int    nc  = gf.n / 2;
    const long long Nc1 = (long long)k * nc + 1;
    const int    n1  = k + 1, np = 2 * k + 1;
    const int    shc[3] = {n1, n1, dim == 3 ? n1 : 1};
    const int    ncoarse_loc = shc[0] * shc[1] * shc[2];
    const int    shf[3] = {np, np, dim == 3 ? np : 1};
    const int    nfine_loc = shf[0] * shf[1] * shf[2];
    const int    ncol = 1 << dim;
    for (int col = 0; col < ncol; ++col)
      {
        const int px = col & 1, py = (col >> 1) & 1, pz = (col >> 2) & 1;
        if (dim == 2 && pz)
          continue;
        const int nz = dim == 3 ? nc : 1;
#pragma omp parallel
        {
          std::vector<double> cv(ncoarse_loc), fv(nfine_loc), t1(nfine_loc),
            t2(nfine_loc);
#pragma omp for collapse(2) schedule(static)
          for (int cz = (dim == 3 ? pz : 0); cz < nz; cz += (dim == 3 ? 2 : 1))
            for (int cy = py; cy < nc; cy += 2)
              for (int cx = px; cx < nc; cx += 2)
                {
                  const int cc[3] = {cx, cy, cz};
                  auto cidx = [&](int i, int j, int l) {
                    return (long long)k * cx + i +
                           Nc1 * ((long long)k * cy + j) +
                           Nc1 * Nc1 * ((long long)k * cz * (dim == 3) + l);
                  };
                  auto fidx = [&](int i, int j, int l) {
                    return (long long)2 * k * cx + i +
                           gf.N1 * ((long long)2 * k * cy + j) +
                           gf.N1 * gf.N1 *
                             ((long long)2 * k * cz * (dim == 3) + l);
                  };
                  auto c_on_bdry = [&](int i, int j, int l) {
                    const long long gx = (long long)k * cx + i,
                                    gy = (long long)k * cy + j,
                                    gz = (long long)k * cz + l;
                    if (gx == 0 || gx == Nc1 - 1 || gy == 0 || gy == Nc1 - 1)
                      return true;
                    return dim == 3 && (gz == 0 || gz == Nc1 - 1);
                  };
                  auto fw = [&](int i, int j, int l) {
                    double w = vweight(i, cc[0], nc, k) * vweight(j, cc[1], nc, k);
                    if (dim == 3)
                      w *= vweight(l, cc[2], nc, k);
                    return w;
                  };
                  if (prolongate)
                    {
                      for (int l = 0; l < shc[2]; ++l)
                        for (int j = 0; j < n1; ++j)
                          for (int i = 0; i < n1; ++i)
                            cv[(l * n1 + j) * n1 + i] =
                              c_on_bdry(i, j, l) ? 0.0 : coarse_in[cidx(i, j, l)];
                      int sh[3] = {shc[0], shc[1], shc[2]};
                      const double *src = cv.data();
                      double       *bufs[2] = {t1.data(), t2.data()};
                      for (int a = 0; a < dim; ++a)
                        {
                          double *dst = (a == dim - 1) ? fv.data() : bufs[a & 1];
                          apply_axis(b.E.data(), np, n1, src, dst, sh, a, false);
                          sh[a] = np;
                          src   = dst;
                        }
                      for (int l = 0; l < shf[2]; ++l)
                        for (int j = 0; j < np; ++j)
                          for (int i = 0; i < np; ++i)
                            fine_[fidx(i, j, l)] +=
                              fw(i, j, l) * fv[(l * np + j) * np + i];
                    }
                  else
                    {
                      for (int l = 0; l < shf[2]; ++l)
                        for (int j = 0; j < np; ++j)
                          for (int i = 0; i < np; ++i)
                            fv[(l * np + j) * np + i] =
                              fw(i, j, l) * fine_in[fidx(i, j, l)];
                      int sh[3] = {shf[0], shf[1], shf[2]};
                      const double *src = fv.data();
                      double       *bufs[2] = {t1.data(), t2.data()};
                      for (int a = 0; a < dim; ++a)
                        {
                          double *dst = (a == dim - 1) ? cv.data() : bufs[a & 1];
                          apply_axis(b.E.data(), n1, np, src, dst, sh, a,
                                     false, true);
                          sh[a] = n1;
                          src   = dst;
                        }
                      for (int l = 0; l < shc[2]; ++l)
                        for (int j = 0; j < n1; ++j)
                          for (int i = 0; i < n1; ++i)
                            if (!c_on_bdry(i, j, l))
                              coarse_[cidx(i, j, l)] += cv[(l * n1 + j) * n1 + i];
                    }
                }
        }
      }
  }

  // ------------------------------------------------------------ multigrid
  struct MG
  {
    int dim, k, minlevel, maxlevel, variant;  // variant: 0 sym, 1 ref-order, 2 additive
    double relax;
    std::vector<Grid> grids;
    mutable std::vector<std::vector<double>> d, u, t;
    std::vector<std::vector<double>> coef;  // per level; empty: constant 1
    MG(int dim_, int k_, int maxl, int var, double rel,
       const double *fine_coef = nullptr)
      : dim(dim_), k(k_), minlevel(1), maxlevel(maxl), variant(var), relax(rel)
    {
      for (int l = 0; l <= maxlevel; ++l)
        grids.emplace_back(dim, k, std::max(l, 1));
      d.resize(maxlevel + 1);
      u.resize(maxlevel + 1);
      t.resize(maxlevel + 1);
      for (int l = minlevel; l <= maxlevel; ++l)
        {
          d[l].assign(grids[l].N, 0.0);
          u[l].assign(grids[l].N, 0.0);
          t[l].assign(grids[l].N, 0.0);
        }
      if (fine_coef)
        {
          coef.resize(maxlevel + 1);
          const long long nf = dim == 3 ? (1LL << (3 * maxlevel)) : (1LL << (2 * maxlevel));
          coef[maxlevel].assign(fine_coef, fine_coef + nf);
          for (int l = maxlevel - 1; l >= minlevel; --l)
            {
              const int nc = 1 << l;
              coef[l].assign(dim == 3 ? (long long)nc * nc * nc : (long long)nc * nc, 0.0);
              coarsen_coefficient(dim, nc, coef[l + 1].data(), coef[l].data());
            }
        }
    }
    const double *level_coef(int l) const
    {
      return coef.empty() ? nullptr : coef[l].data();
    }
    void smooth(int l, int var, double rel, const double *b, double *x) const
    {
      if (coef.empty())
        smoother_step(grids[l], var, rel, b, x);
      else
        smoother_step_coeff(grids[l], coef[l].data(), var, rel, b, x);
    }
    int pre_variant() const { return variant == 2 ? ADDITIVE : MULT_FWD; }
    int post_variant() const
    {
      return variant == 2 ? ADDITIVE : (variant == 0 ? MULT_REV : MULT_FWD);
    }
    // n_cycles > 1 repeats smoothing + coarse correction on the same defect
    // (MultigridSolvers::v_cycle(level, my_n_cycles), solver.cuh:1950-2001)
    void
    vcycle(int l, int n_cycles = 1) const
    {
      const Grid &g = grids[l];
      std::fill(u[l].begin(), u[l].end(), 0.0);
      if (l == minlevel)
        {
          // single vertex patch: exact solve (one inexact local solve with a
          // variable coefficient)
          smooth(l, MULT_FWD, 1.0, d[l].data(), u[l].data());
          return;
        }
      for (int c = 0; c < n_cycles; ++c)
        {
          smooth(l, pre_variant(), relax, d[l].data(), u[l].data());
          laplace_vmult(g, u[l].data(), t[l].data(), level_coef(l));
          for (long long i = 0; i < g.N; ++i)
            t[l][i] = d[l][i] - t[l][i];
          std::fill(d[l - 1].begin(), d[l - 1].end(), 0.0);
          transfer(g, false, d[l - 1].data(), nullptr, nullptr, t[l].data());
          vcycle(l - 1);
          transfer(g, true, nullptr, u[l].data(), u[l - 1].data(), nullptr);
          smooth(l, post_variant(), relax, d[l].data(), u[l].data());
        }
    }
    // Full multigrid (MultigridSolvers::solve, solver.cuh:1784-1860): exact
    // coarse solve plus one more smoothing step on level 1, then per finer
    // level: prolongate the coarse solution, impose the Dirichlet values
    // (boundary entries of that level's right-hand side), form the residual
    // and add the correction of n_cycles V-cycles.  rhs[l] for l = 1..max.
    void
    fmg(const std::vector<const double *> &rhs, int n_cycles, double *x) const
    {
      std::vector<std::vector<double>> sol(maxlevel + 1);
      const Grid &g1 = grids[minlevel];
      std::copy(rhs[minlevel], rhs[minlevel] + g1.N, d[minlevel].begin());
      vcycle(minlevel);
      smooth(minlevel, MULT_FWD, 1.0, d[minlevel].data(), u[minlevel].data());
      sol[minlevel] = u[minlevel];
      for (int l = minlevel + 1; l <= maxlevel; ++l)
        {
          const Grid &g = grids[l];
          sol[l].assign(g.N, 0.0);
          transfer(g, true, nullptr, sol[l].data(), sol[l - 1].data(), nullptr);
          for (long long i = 0; i < g.N; ++i)
            if (g.on_boundary(i))
              sol[l][i] = rhs[l][i];
          laplace_vmult(g, sol[l].data(), t[l].data(), level_coef(l));
          for (long long i = 0; i < g.N; ++i)
            d[l][i] = rhs[l][i] - t[l][i];
          vcycle(l, n_cycles);
          for (long long i = 0; i < g.N; ++i)
            sol[l][i] += u[l][i];
        }
      std::copy(sol[maxlevel].begin(), sol[maxlevel].end(), x);
    }
    void
    apply(const double *r, double *z) const
    {
      std::copy(r, r + grids[maxlevel].N, d[maxlevel].begin());
      vcycle(maxlevel);
      std::copy(u[maxlevel].begin(), u[maxlevel].end(), z);
    }
  };

  double
  dot(long long n, const double *a, const double *b)
  {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long long i = 0; i < n; ++i)
      s += a[i] * b[i];
    return s;
  }

  uint64_t
  splitmix64(uint64_t x)
  {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
  }
} // namespace

namespace
{
  // ============================================================== SIPG-DG
  // Discontinuous Galerkin Q_k (FE_DGQ on Gauss-Lobatto points) on a
  // uniform mesh of n cells per direction, DoFs stored as one tensor grid
  // X = cx (k+1) + ix, index X + N1 Y + N1^2 Z with N1 = n (k+1).  The
  // symmetric interior penalty form follows the reference's matrix-free DG
  // test operator (tests/mf_dg.cu:63-274): cell term grad u . grad v; on an
  // interior face (sigma [u] - {du/dn}) [v] - [u] {dv/dn} / 2 * 2; on a
  // boundary face (2 sigma u - du/dn) v - u dv/dn, with sigma = pf / h.
  // Here it is evaluated the way that test does it: cell and face
  // quadrature loops with Gauss(k+1) points.
  struct DGGrid
  {
    int       dim, k, n;
    long long N1, N;
    double    h, pf;
    DGGrid(int dim_, int k_, int n_, double pf_)
      : dim(dim_), k(k_), n(n_), pf(pf_)
    {
      N1 = (long long)n * (k + 1);
      N  = dim == 3 ? N1 * N1 * N1 : N1 * N1;
      h  = 1.0 / n;
    }
    long long
    idx(long long X, long long Y, long long Z) const
    {
      return X + N1 * Y + N1 * N1 * Z;
    }
  };

  // values and reference derivatives of the 1D basis at the cell ends
  struct FaceBasis
  {
    std::vector<double> v0, v1, d0, d1;
  };
  FaceBasis
  face_basis(int k)
  {
    const Basis &b = basis(k);
    FaceBasis    f;
    for (int i = 0; i <= k; ++i)
      {
        f.v0.push_back(lagrange(b.nodes, i, 0.0));
        f.v1.push_back(lagrange(b.nodes, i, 1.0));
        f.d0.push_back(lagrange_deriv(b.nodes, i, 0.0));
        f.d1.push_back(lagrange_deriv(b.nodes, i, 1.0));
      }
    return f;
  }

  // y = A x (cell loop + interior-face loop + boundary-face loop)
  void
  dg_vmult(const DGGrid &g, const double *x, double *y)
  {
    const Basis    &b   = basis(g.k);
    const FaceBasis fb  = face_basis(g.k);
    const int       n1  = g.k + 1;
    const int       dim = g.dim;
    const int       nl  = dim == 3 ? n1 * n1 * n1 : n1 * n1;
    const double    h   = g.h, sigma = g.pf / h;
    const double    scale = std::pow(h, dim - 2);
    std::fill(y, y + g.N, 0.0);
    const int nz = dim == 3 ? g.n : 1;
    auto gather = [&](int cx, int cy, int cz, const double *src, double *loc) {
      for (int l = 0; l < (dim == 3 ? n1 : 1); ++l)
        for (int j = 0; j < n1; ++j)
          for (int i = 0; i < n1; ++i)
            loc[(l * n1 + j) * n1 + i] =
              src[g.idx((long long)cx * n1 + i, (long long)cy * n1 + j,
                        (long long)cz * n1 * (dim == 3) + l)];
    };
    auto scatter = [&](int cx, int cy, int cz, const double *loc) {
      for (int l = 0; l < (dim == 3 ? n1 : 1); ++l)
        for (int j = 0; j < n1; ++j)
          for (int i = 0; i < n1; ++i)
            y[g.idx((long long)cx * n1 + i, (long long)cy * n1 + j,
                    (long long)cz * n1 * (dim == 3) + l)] += loc[(l * n1 + j) * n1 + i];
    };
    // cell terms (quadrature of grad u . grad v), same as the CG cell loop
    {
      std::vector<double> wq(nl);
      for (int l = 0; l < (dim == 3 ? n1 : 1); ++l)
        for (int j = 0; j < n1; ++j)
          for (int i = 0; i < n1; ++i)
            wq[(l * n1 + j) * n1 + i] =
              b.qw[i] * b.qw[j] * (dim == 3 ? b.qw[l] : 1.0) * scale;
      const int shape[3] = {n1, n1, dim == 3 ? n1 : 1};
      std::vector<double> u(nl), t1(nl), t2(nl), gq(nl), out(nl), acc(nl);
      for (int cz = 0; cz < nz; ++cz)
        for (int cy = 0; cy < g.n; ++cy)
          for (int cx = 0; cx < g.n; ++cx)
            {
              gather(cx, cy, cz, x, u.data());
              std::fill(acc.begin(), acc.end(), 0.0);
              for (int d = 0; d < dim; ++d)
                {
                  const double *src     = u.data();
                  double       *bufs[2] = {t1.data(), t2.data()};
                  int           nb      = 0;
                  for (int a = 0; a < dim; ++a)
                    {
                      double *dst = (a == dim - 1) ? gq.data() : bufs[nb++ & 1];
                      apply_axis(a == d ? b.D.data() : b.N.data(), n1, n1, src,
                                 dst, shape, a, false);
                      src = dst;
                    }
                  for (int q = 0; q < nl; ++q)
                    gq[q] *= wq[q];
                  src = gq.data();
                  nb  = 0;
                  for (int a = 0; a < dim; ++a)
                    {
                      double *dst = (a == dim - 1) ? out.data() : bufs[nb++ & 1];
                      apply_axis(a == d ? b.D.data() : b.N.data(), n1, n1, src,
                                 dst, shape, a, false, true);
                      src = dst;
                    }
                  for (int q = 0; q < nl; ++q)
                    acc[q] += out[q];
                }
              scatter(cx, cy, cz, acc.data());
            }
    }
    // face terms: per face quadrature point (tensor Gauss points in the
    // tangential directions), the values and normal derivatives of the two
    // traces; basis function (i,j,l) = phi_i(x) phi_j(y) phi_l(z)
    const int ntan = dim == 3 ? n1 * n1 : n1;  // tangential quad points
    std::vector<double> um(nl), up(nl), om(nl), op(nl);
    // value / normal-derivative of basis function `a` (local multi-index)
    // of a cell at a face point: side s = 0 (face at xi_d = 0) or 1
    auto eval = [&](int d, int s, const int *ai, const int *qt, double &val,
                    double &dn) {
      // ai: local 1D indices per direction; qt: tangential quad indices
      val = 1.0;
      dn  = 1.0;
      int t = 0;
      for (int a = 0; a < dim; ++a)
        {
          if (a == d)
            {
              val *= s ? fb.v1[ai[a]] : fb.v0[ai[a]];
              dn *= (s ? fb.d1[ai[a]] : fb.d0[ai[a]]) / h;
            }
          else
            {
              const double nv = b.N[qt[t] * n1 + ai[a]];
              val *= nv;
              dn *= nv;
              ++t;
            }
        }
    };
    auto tan_weight = [&](const int *qt) {
      double w = 1.0;
      for (int t = 0; t < dim - 1; ++t)
        w *= b.qw[qt[t]] * h;
      return w;
    };
    for (int d = 0; d < dim; ++d)
      for (int cz = 0; cz < nz; ++cz)
        for (int cy = 0; cy < g.n; ++cy)
          for (int cx = 0; cx < g.n; ++cx)
            {
              const int c[3] = {cx, cy, cz};
              // faces at the low side (interior with the lower neighbour or
              // boundary) and at the high side (boundary only)
              for (int side = 0; side < 2; ++side)
                {
                  const bool low   = side == 0;
                  const bool bdry  = low ? c[d] == 0 : c[d] == g.n - 1;
                  if (!low && !bdry)
                    continue;  // interior high faces are the neighbour's low
                  if (bdry)
                    {
                      // boundary face of this cell: outward normal -e_d
                      // (low) or +e_d (high)
                      const double sgn = low ? -1.0 : 1.0;
                      gather(cx, cy, cz, x, um.data());
                      std::fill(om.begin(), om.end(), 0.0);
                      for (int q = 0; q < ntan; ++q)
                        {
                          const int qt[2] = {q % n1, q / n1};
                          double    uval = 0, udn = 0;
                          int       ai[3];
                          for (int a = 0; a < nl; ++a)
                            {
                              ai[0] = a % n1;
                              ai[1] = (a / n1) % n1;
                              ai[2] = a / (n1 * n1);
                              double v, dn;
                              eval(d, low ? 0 : 1, ai, qt, v, dn);
                              uval += um[a] * v;
                              udn += um[a] * dn * sgn;
                            }
                          const double w  = tan_weight(qt);
                          const double tv = (2.0 * sigma * uval - udn) * w;
                          const double td = -uval * w;
                          for (int a = 0; a < nl; ++a)
                            {
                              ai[0] = a % n1;
                              ai[1] = (a / n1) % n1;
                              ai[2] = a / (n1 * n1);
                              double v, dn;
                              eval(d, low ? 0 : 1, ai, qt, v, dn);
                              om[a] += tv * v + td * dn * sgn;
                            }
                        }
                      scatter(cx, cy, cz, om.data());
                      if (low)
                        continue;
                    }
                  if (!low || bdry)
                    continue;
                  // interior face between the lower neighbour (minus side,
                  // its high face) and this cell (plus side, its low face);
                  // normal +e_d
                  int cm[3] = {cx, cy, cz};
                  cm[d] -= 1;
                  gather(cm[0], cm[1], cm[2], x, um.data());
                  gather(cx, cy, cz, x, up.data());
                  std::fill(om.begin(), om.end(), 0.0);
                  std::fill(op.begin(), op.end(), 0.0);
                  for (int q = 0; q < ntan; ++q)
                    {
                      const int qt[2] = {q % n1, q / n1};
                      double    vm = 0, vp = 0, dm = 0, dp = 0;
                      int       ai[3];
                      for (int a = 0; a < nl; ++a)
                        {
                          ai[0] = a % n1;
                          ai[1] = (a / n1) % n1;
                          ai[2] = a / (n1 * n1);
                          double v, dn;
                          eval(d, 1, ai, qt, v, dn);
                          vm += um[a] * v;
                          dm += um[a] * dn;
                          eval(d, 0, ai, qt, v, dn);
                          vp += up[a] * v;
                          dp += up[a] * dn;
                        }
                      const double w    = tan_weight(qt);
                      const double jump = vm - vp;
                      const double tv   = (jump * sigma - 0.5 * (dm + dp)) * w;
                      const double td   = -0.5 * jump * w;
                      for (int a = 0; a < nl; ++a)
                        {
                          ai[0] = a % n1;
                          ai[1] = (a / n1) % n1;
                          ai[2] = a / (n1 * n1);
                          double v, dn;
                          eval(d, 1, ai, qt, v, dn);
                          om[a] += tv * v + td * dn;
                          eval(d, 0, ai, qt, v, dn);
                          op[a] += -tv * v + td * dn;
                        }
                    }
                  scatter(cm[0], cm[1], cm[2], om.data());
                  scatter(cx, cy, cz, op.data());
                }
            }
  }

  // 1D SIPG data of one direction: the global 1D matrix L (size n(k+1))
  // of the unit-scaled form (cell stiffness + face terms with sigma = pf),
  // from the face formulas of dg_vmult restricted to 1D; the cell mass M.
  std::vector<double>
  dg_global_L(int k, int n, double pf)
  {
    const Basis    &b  = basis(k);
    const FaceBasis fb = face_basis(k);
    const int       n1 = k + 1, N = n * n1;
    std::vector<double> L((size_t)N * N, 0.0);
    auto at = [&](int r, int c) -> double & { return L[(size_t)r * N + c]; };
    for (int c = 0; c < n; ++c)
      for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n1; ++j)
          at(c * n1 + i, c * n1 + j) += b.K[i * n1 + j];
    // interior faces between cells c (minus, its x = 1 end) and c + 1
    for (int c = 0; c + 1 < n; ++c)
      for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n1; ++j)
          {
            // test i / trial j on sides (m,m), (m,p), (p,m), (p,p)
            const double tm[2] = {fb.v1[i], -fb.v0[i]};   // [v]
            const double dtm[2] = {0.5 * fb.d1[i], 0.5 * fb.d0[i]};  // {v'}
            const double sm[2] = {fb.v1[j], -fb.v0[j]};   // [u]
            const double dsm[2] = {0.5 * fb.d1[j], 0.5 * fb.d0[j]};  // {u'}
            for (int si = 0; si < 2; ++si)
              for (int sj = 0; sj < 2; ++sj)
                at((c + si) * n1 + i, (c + sj) * n1 + j) +=
                  pf * sm[sj] * tm[si] - dsm[sj] * tm[si] - sm[sj] * dtm[si];
          }
    // boundary faces: left (normal -1) at cell 0, right (normal +1) at n-1
    for (int i = 0; i < n1; ++i)
      for (int j = 0; j < n1; ++j)
        {
          at(i, j) += 2 * pf * fb.v0[i] * fb.v0[j] + fb.d0[j] * fb.v0[i] +
                      fb.v0[j] * fb.d0[i];
          const int o = (n - 1) * n1;
          at(o + i, o + j) += 2 * pf * fb.v1[i] * fb.v1[j] -
                              fb.d1[j] * fb.v1[i] - fb.v1[j] * fb.d1[i];
        }
    return L;
  }

  // Patch data of one direction for the vertex patch on cells (c0, c0+1):
  // restricted 1D matrices and their generalized eigenpairs.
  struct DGPatch1D
  {
    std::vector<double> L, M, lam, S;  // (2(k+1))^2, S column j = vector j
  };
  DGPatch1D
  dg_patch_1d(int k, int n, double pf, int c0)
  {
    const Basis &b  = basis(k);
    const int    n1 = k + 1, np = 2 * n1, N = n * n1;
    const std::vector<double> Lg = dg_global_L(k, n, pf);
    DGPatch1D p;
    p.L.assign(np * np, 0.0);
    p.M.assign(np * np, 0.0);
    for (int i = 0; i < np; ++i)
      for (int j = 0; j < np; ++j)
        p.L[i * np + j] = Lg[(size_t)(c0 * n1 + i) * N + c0 * n1 + j];
    for (int c = 0; c < 2; ++c)
      for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n1; ++j)
          p.M[(c * n1 + i) * np + c * n1 + j] = b.M[i * n1 + j];
    p.lam.resize(np);
    p.S.resize(np * np);
    gen_eig(np, p.L.data(), p.M.data(), p.lam.data(), p.S.data());
    return p;
  }

  // One multiplicative sweep over the 3^d colours of vertex patches
  // (interior vertices v in 1..n-1 per direction, patch index p = v-1 with
  // cells p, p+1; colour = p mod 3 per direction: DG patches two apart would
  // be face neighbours, three apart they share nothing).  Local problem
  // u_P = A_PP^{-1} (b_P - A_P,ext u_ext) with A_PP inverted by fast
  // diagonalisation of the restricted 1D matrices (exact: A_PP is their
  // Kronecker sum); b_P - A_P,ext u_ext comes from one operator application
  // with the colour's patch DoFs zeroed.
  void
  dg_smoother_step(const DGGrid &g, int variant, double relax, const double *b_,
                   double *u)
  {
    const int    k = g.k, n1 = k + 1, np = 2 * n1, dim = g.dim;
    const int    npd = g.n - 1;  // patches per direction
    const double scale = std::pow(g.h, dim - 2);
    std::vector<DGPatch1D> P1(g.n);  // by first cell
    for (int c0 = 0; c0 + 1 < g.n; ++c0)
      P1[c0] = dg_patch_1d(k, g.n, g.pf, c0);
    const int nloc = dim == 3 ? np * np * np : np * np;
    const int ncol = dim == 3 ? 27 : 9;
    std::vector<double> tmp(g.N), ax(g.N);
    for (int cc = 0; cc < ncol; ++cc)
      {
        const int col = variant == MULT_REV ? ncol - 1 - cc : cc;
        const int px = col % 3, py = (col / 3) % 3, pz = col / 9;
        // residual with the patch DoFs of this colour zeroed: the patches
        // of one colour are separated by a cell, so one global product
        // gives every patch's b_P - A_P,ext u_ext
        std::copy(u, u + g.N, tmp.begin());
        std::vector<std::array<int, 3>> plist;
        for (int iz = (dim == 3 ? pz : 0); iz < (dim == 3 ? npd : 1); iz += (dim == 3 ? 3 : 1))
          for (int iy = py; iy < npd; iy += 3)
            for (int ix = px; ix < npd; ix += 3)
              plist.push_back({ix, iy, iz});
        if (plist.empty())
          continue;
        auto pidx = [&](const std::array<int, 3> &p, int i, int j, int l) {
          return g.idx((long long)p[0] * n1 + i, (long long)p[1] * n1 + j,
                       dim == 3 ? (long long)p[2] * n1 + l : 0);
        };
        for (const auto &p : plist)
          for (int l = 0; l < (dim == 3 ? np : 1); ++l)
            for (int j = 0; j < np; ++j)
              for (int i = 0; i < np; ++i)
                tmp[pidx(p, i, j, l)] = 0.0;
        dg_vmult(g, tmp.data(), ax.data());
#pragma omp parallel
        {
          std::vector<double> r(nloc), t(nloc);
#pragma omp for schedule(static)
          for (size_t pi = 0; pi < plist.size(); ++pi)
            {
              const auto &p = plist[pi];
              for (int l = 0; l < (dim == 3 ? np : 1); ++l)
                for (int j = 0; j < np; ++j)
                  for (int i = 0; i < np; ++i)
                    {
                      const long long gi = pidx(p, i, j, l);
                      r[(l * np + j) * np + i] = b_[gi] - ax[gi];
                    }
              // S^T per direction, 1/(scale sum lambda), S per direction
              const DGPatch1D *d1[3] = {&P1[p[0]], &P1[p[1]], &P1[dim == 3 ? p[2] : 0]};
              const int shape[3] = {np, np, dim == 3 ? np : 1};
              double *bufs[2] = {r.data(), t.data()};
              int     cur     = 0;
              for (int a = 0; a < dim; ++a)
                {
                  apply_axis(d1[a]->S.data(), np, np, bufs[cur], bufs[cur ^ 1],
                             shape, a, false, true);
                  cur ^= 1;
                }
              for (int l = 0; l < shape[2]; ++l)
                for (int j = 0; j < np; ++j)
                  for (int i = 0; i < np; ++i)
                    {
                      double lam = d1[0]->lam[i] + d1[1]->lam[j];
                      if (dim == 3)
                        lam += d1[2]->lam[l];
                      bufs[cur][(l * np + j) * np + i] /= scale * lam;
                    }
              for (int a = 0; a < dim; ++a)
                {
                  apply_axis(d1[a]->S.data(), np, np, bufs[cur], bufs[cur ^ 1],
                             shape, a, false, false);
                  cur ^= 1;
                }
              for (int l = 0; l < shape[2]; ++l)
                for (int j = 0; j < np; ++j)
                  for (int i = 0; i < np; ++i)
                    {
                      const long long gi = pidx(p, i, j, l);
                      u[gi] = (1.0 - relax) * u[gi] + relax * bufs[cur][(l * np + j) * np + i];
                    }
            }
        }
      }
  }

  // DG transfer: fine cell f of coarse cell f/2, child c = f % 2; the
  // coarse polynomial interpolated at the child's nodes, x = (c + xi) / 2.
  std::vector<double>
  dg_child_embedding(int k)
  {
    const Basis        &b  = basis(k);
    const int           n1 = k + 1;
    std::vector<double> E(2 * n1 * n1);
    for (int c = 0; c < 2; ++c)
      for (int q = 0; q < n1; ++q)
        for (int a = 0; a < n1; ++a)
          E[(c * n1 + q) * n1 + a] = lagrange(b.nodes, a, 0.5 * (c + b.nodes[q]));
    return E;
  }

  void
  dg_transfer(const DGGrid &gf, bool prolongate, double *coarse, double *fine)
  {
    const int    k = gf.k, n1 = k + 1, dim = gf.dim;
    const DGGrid gc(dim, k, gf.n / 2, gf.pf);
    const std::vector<double> E = dg_child_embedding(k);
    const int nz = dim == 3 ? gf.n : 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int fz = 0; fz < nz; ++fz)
      for (int fy = 0; fy < gf.n; ++fy)
        for (int fx = 0; fx < gf.n; ++fx)
          {
            const int f[3] = {fx, fy, fz};
            for (int l = 0; l < (dim == 3 ? n1 : 1); ++l)
              for (int j = 0; j < n1; ++j)
                for (int i = 0; i < n1; ++i)
                  {
                    const long long gfi = gf.idx((long long)fx * n1 + i, (long long)fy * n1 + j,
                                                 dim == 3 ? (long long)fz * n1 + l : 0);
                    const int q[3] = {i, j, l};
                    double    acc = 0.0;
                    for (int cl = 0; cl < (dim == 3 ? n1 : 1); ++cl)
                      for (int cb = 0; cb < n1; ++cb)
                        for (int ca = 0; ca < n1; ++ca)
                          {
                            const int a[3] = {ca, cb, cl};
                            double    w   = 1.0;
                            for (int d = 0; d < dim; ++d)
                              w *= E[((f[d] & 1) * n1 + q[d]) * n1 + a[d]];
                            const long long gci =
                              gc.idx((long long)(fx / 2) * n1 + ca, (long long)(fy / 2) * n1 + cb,
                                     dim == 3 ? (long long)(fz / 2) * n1 + cl : 0);
                            if (prolongate)
                              acc += w * coarse[gci];
                            else
                              {
#pragma omp atomic
                                coarse[gci] += w * fine[gfi];
                              }
                          }
                    if (prolongate)
                      fine[gfi] += acc;
                  }
          }
  }

  struct DGMG
  {
    int dim, k, n0, maxl, variant, coarse_sweeps;
    double pf, relax;
    std::vector<DGGrid> grids;
    mutable std::vector<std::vector<double>> d, u, t;
    DGMG(int dim_, int k_, int n0_, int maxl_, double pf_, int var, double rel, int cs)
      : dim(dim_), k(k_), n0(n0_), maxl(maxl_), variant(var), coarse_sweeps(cs),
        pf(pf_), relax(rel)
    {
      for (int l = 0; l <= maxl; ++l)
        grids.emplace_back(dim, k, n0 << l, pf);
      d.resize(maxl + 1);
      u.resize(maxl + 1);
      t.resize(maxl + 1);
      for (int l = 0; l <= maxl; ++l)
        {
          d[l].assign(grids[l].N, 0.0);
          u[l].assign(grids[l].N, 0.0);
          t[l].assign(grids[l].N, 0.0);
        }
    }
    void
    vcycle(int l) const
    {
      const DGGrid &g = grids[l];
      std::fill(u[l].begin(), u[l].end(), 0.0);
      if (l == 0)
        {
          // coarsest level: symmetric sweep pairs from zero
          for (int s = 0; s < coarse_sweeps; ++s)
            {
              dg_smoother_step(g, MULT_FWD, 1.0, d[l].data(), u[l].data());
              dg_smoother_step(g, MULT_REV, 1.0, d[l].data(), u[l].data());
            }
          return;
        }
      dg_smoother_step(g, MULT_FWD, relax, d[l].data(), u[l].data());
      dg_vmult(g, u[l].data(), t[l].data());
      for (long long i = 0; i < g.N; ++i)
        t[l][i] = d[l][i] - t[l][i];
      std::fill(d[l - 1].begin(), d[l - 1].end(), 0.0);
      dg_transfer(g, false, d[l - 1].data(), t[l].data());
      vcycle(l - 1);
      dg_transfer(g, true, u[l - 1].data(), u[l].data());
      dg_smoother_step(g, variant == 0 ? MULT_REV : MULT_FWD, relax, d[l].data(),
                       u[l].data());
    }
    void
    apply(const double *r, double *z) const
    {
      std::copy(r, r + grids[maxl].N, d[maxl].begin());
      vcycle(maxl);
      std::copy(u[maxl].begin(), u[maxl].end(), z);
    }
  };
} // namespace

extern "C"
{
  int
  oracle_max_degree()
  {
    return MAXK;
  }

  // 1D data for degree k: nodes[k+1], cell M,K [(k+1)^2], embedding
  // E[(2k+1)(k+1)], patch MP,KP [(2k+1)^2], interior eigenpairs lam[2k-1],
  // S[(2k-1)^2] (column j = eigenvector j).
  int
  oracle_basis_1d(int k, double *nodes, double *M, double *K, double *E,
                  double *MP, double *KP, double *lam, double *S)
  {
    if (k < 1 || k > MAXK)
      return 1;
    const Basis &b = basis(k);
    if (nodes) std::copy(b.nodes.begin(), b.nodes.end(), nodes);
    if (M) std::copy(b.M.begin(), b.M.end(), M);
    if (K) std::copy(b.K.begin(), b.K.end(), K);
    if (E) std::copy(b.E.begin(), b.E.end(), E);
    if (MP) std::copy(b.MP.begin(), b.MP.end(), MP);
    if (KP) std::copy(b.KP.begin(), b.KP.end(), KP);
    if (lam) std::copy(b.lam.begin(), b.lam.end(), lam);
    if (S) std::copy(b.S.begin(), b.S.end(), S);
    return 0;
  }

  int
  oracle_gen_eig(int n, const double *K, const double *M, double *lam,
                 double *S)
  {
    if (n < 1 || n > 64)
      return 1;
    gen_eig(n, K, M, lam, S);
    return 0;
  }

  long long
  oracle_n_dofs(int dim, int k, int level)
  {
    return Grid(dim, k, level).N;
  }

  int
  oracle_laplace_vmult(int dim, int k, int level, const double *x, double *y)
  {
    if ((dim != 2 && dim != 3) || k < 1 || k > MAXK || level < 1)
      return 1;
    laplace_vmult(Grid(dim, k, level), x, y);
    return 0;
  }

  // variant: 0 forward colours, 1 reverse colours, 2 additive
  int
  oracle_smoother_step(int dim, int k, int level, int variant, double relax,
                       const double *b, double *u)
  {
    if ((dim != 2 && dim != 3) || k < 1 || k > MAXK || level < 1 ||
        variant < 0 || variant > 2)
      return 1;
    smoother_step(Grid(dim, k, level), variant, relax, b, u);
    return 0;
  }

  int
  oracle_prolongate_add(int dim, int k, int fine_level, const double *coarse,
                        double *fine)
  {
    if ((dim != 2 && dim != 3) || k < 1 || k > MAXK || fine_level < 2)
      return 1;
    transfer(Grid(dim, k, fine_level), true, nullptr, fine, coarse, nullptr);
    return 0;
  }

  int
  oracle_restrict_add(int dim, int k, int fine_level, const double *fine,
                      double *coarse)
  {
    if ((dim != 2 && dim != 3) || k < 1 || k > MAXK || fine_level < 2)
      return 1;
    transfer(Grid(dim, k, fine_level), false, coarse, nullptr, nullptr, fine);
    return 0;
  }

  // mg_variant: 0 multiplicative with reversed post-smoothing (symmetric),
  // 1 multiplicative with the reference's colour order, 2 additive.
  int
  oracle_vcycle(int dim, int k, int level, int mg_variant, double relax,
                const double *r, double *z)
  {
    if ((dim != 2 && dim != 3) || k < 1 || k > MAXK || level < 1)
      return 1;
    MG mg(dim, k, level, mg_variant, relax);
    mg.apply(r, z);
    return 0;
  }

  // Preconditioned CG on the zero-Dirichlet system. Boundary entries of b are
  // taken as the boundary values of x (identity rows); stops when
  // ||r_k|| <= rtol ||r_0||. history (optional, maxit+1) gets ||r_k||.
  int
  oracle_coeff_pcg(int dim, int k, int level, const double *coef,
                   int mg_variant, double relax, const double *b, double *x,
                   double rtol, int maxit, int *iters, double *relres,
                   double *history)
  {
    if ((dim != 2 && dim != 3) || k < 1 || k > MAXK || level < 1)
      return 1;
    const Grid g(dim, k, level);
    MG         mg(dim, k, level, mg_variant, relax, coef);
    const long long N = g.N;
    std::vector<double> r(N), z(N), p(N), q(N);
    for (long long i = 0; i < N; ++i)
      x[i] = g.on_boundary(i) ? b[i] : 0.0;
    laplace_vmult(g, x, q.data(), coef);
    for (long long i = 0; i < N; ++i)
      r[i] = b[i] - q[i];
    const double r0 = std::sqrt(dot(N, r.data(), r.data()));
    if (history)
      history[0] = r0;
    int    it  = 0;
    double rn  = r0;
    if (r0 == 0.0)
      {
        *iters  = 0;
        *relres = 0.0;
        return 0;
      }
    mg.apply(r.data(), z.data());
    std::copy(z.begin(), z.end(), p.begin());
    double rz = dot(N, r.data(), z.data());
    while (it < maxit)
      {
        laplace_vmult(g, p.data(), q.data(), coef);
        const double alpha = rz / dot(N, p.data(), q.data());
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < N; ++i)
          {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
          }
        ++it;
        rn = std::sqrt(dot(N, r.data(), r.data()));
        if (history)
          history[it] = rn;
        if (!(rn > rtol * r0))
          break;
        mg.apply(r.data(), z.data());
        const double rz_new = dot(N, r.data(), z.data());
        const double beta   = rz_new / rz;
        rz                  = rz_new;
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < N; ++i)
          p[i] = z[i] + beta * p[i];
      }
    *iters  = it;
    *relres = rn / r0;
    return std::isfinite(rn) ? 0 : 2;
  }

  int
  oracle_pcg(int dim, int k, int level, int mg_variant, double relax,
             const double *b, double *x, double rtol, int maxit, int *iters,
             double *relres, double *history)
  {
    return oracle_coeff_pcg(dim, k, level, nullptr, mg_variant, relax, b, x,
                            rtol, maxit, iters, relres, history);
  }

  // Variable-coefficient operator (coef per cell, lexicographic cells; NULL
  // = 1). mode 0: Dirichlet identity rows; mode 1: transpose.
  int
  oracle_coeff_vmult(int dim, int k, int level, const double *coef,
                     const double *x, double *y, int mode)
  {
    if ((dim != 2 && dim != 3) || k < 1 || k > MAXK || level < 1 ||
        mode < 0 || mode > 1)
      return 1;
    laplace_vmult(Grid(dim, k, level), x, y, coef, mode);
    return 0;
  }

  // variant: 0 forward colours, 1 reverse colours
  int
  oracle_coeff_smoother_step(int dim, int k, int level, const double *coef,
                             int variant, double relax, const double *b,
                             double *u)
  {
    if ((dim != 2 && dim != 3) || k < 1 || k > MAXK || level < 1 ||
        variant < 0 || variant > 1 || !coef)
      return 1;
    smoother_step_coeff(Grid(dim, k, level), coef, variant, relax, b, u);
    return 0;
  }

  int
  oracle_coeff_coarsen(int dim, int coarse_level, const double *fine,
                       double *coarse)
  {
    if ((dim != 2 && dim != 3) || coarse_level < 0)
      return 1;
    coarsen_coefficient(dim, 1 << coarse_level, fine, coarse);
    return 0;
  }

  int
  oracle_coeff_vcycle(int dim, int k, int level, const double *coef,
                      int mg_variant, double relax, const double *r, double *z)
  {
    if ((dim != 2 && dim != 3) || k < 1 || k > MAXK || level < 1 ||
        mg_variant < 0 || mg_variant > 1)
      return 1;
    MG mg(dim, k, level, mg_variant, relax, coef);
    mg.apply(r, z);
    return 0;
  }

  // rhs_all: the right-hand sides of levels 1..level, concatenated
  int
  oracle_fmg(int dim, int k, int level, const double *coef, int mg_variant,
             double relax, const double *rhs_all, int n_cycles, double *x)
  {
    if ((dim != 2 && dim != 3) || k < 1 || k > MAXK || level < 1 ||
        mg_variant < 0 || mg_variant > 2 || n_cycles < 1)
      return 1;
    MG mg(dim, k, level, mg_variant, relax, coef);
    std::vector<const double *> rhs(level + 1, nullptr);
    const double *p = rhs_all;
    for (int l = 1; l <= level; ++l)
      {
        rhs[l] = p;
        p += mg.grids[l].N;
      }
    mg.fmg(rhs, n_cycles, x);
    return 0;
  }

  int
  oracle_coeff_gradient(int dim, int k, int level, const double *lambda,
                        const double *u, double *grad)
  {
    if ((dim != 2 && dim != 3) || k < 1 || k > MAXK || level < 1)
      return 1;
    coefficient_gradient(Grid(dim, k, level), lambda, u, grad);
    return 0;
  }

  // Variable-coefficient problem data from the SplitMix64 generator
  // u01(j) = (splitmix64(seed ^ j) >> 11) 2^-53:
  //   coef_K = exp(sigma g_K), g_K = sqrt(-2 ln(1 - u01(2^32 + 2K)))
  //            cos(2 pi u01(2^32 + 2K + 1))   (Box-Muller)
  //   ud_i   = 0.01 u01(2^33 + i) (interior DoFs; 0 on the boundary)
  //   cells  = the first n_cells distinct floor(u01(2^34 + j) n_cells_total)
  int
  oracle_coeff_problem(int dim, int k, int level, unsigned long long seed,
                       double sigma, double *coef, double *ud, int n_cells,
                       long long *cells)
  {
    const Grid g(dim, k, level);
    const long long nc = dim == 3 ? (long long)g.n * g.n * g.n : (long long)g.n * g.n;
    auto u01 = [&](uint64_t j) {
      return (double)(splitmix64(seed ^ j) >> 11) * 0x1.0p-53;
    };
    if (coef)
      for (long long K = 0; K < nc; ++K)
        {
          const double a = u01((1ULL << 32) + 2 * (uint64_t)K);
          const double c = u01((1ULL << 32) + 2 * (uint64_t)K + 1);
          const double gK = std::sqrt(-2.0 * std::log(1.0 - a)) * std::cos(2.0 * M_PI * c);
          coef[K] = std::exp(sigma * gK);
        }
    if (ud)
      for (long long i = 0; i < g.N; ++i)
        ud[i] = g.on_boundary(i) ? 0.0 : 0.01 * u01((1ULL << 33) + (uint64_t)i);
    if (cells)
      {
        int got = 0;
        for (uint64_t j = 0; got < n_cells && j < (1ULL << 20); ++j)
          {
            const long long K = (long long)(u01((1ULL << 34) + j) * (double)nc);
            bool dup = false;
            for (int q = 0; q < got; ++q)
              dup = dup || cells[q] == K;
            if (!dup)
              cells[got++] = K;
          }
        if (got < n_cells)
          return 1;
      }
    return 0;
  }

  // ---------------------------------------------------------- SIPG-DG
  // n = cells per direction; pf = penalty factor (sigma = pf / h).
  long long
  oracle_dg_n_dofs(int dim, int k, int n)
  {
    return DGGrid(dim, k, n, 1.0).N;
  }

  int
  oracle_dg_vmult(int dim, int k, int n, double pf, const double *x, double *y)
  {
    if ((dim != 2 && dim != 3) || k < 1 || k > MAXK || n < 1)
      return 1;
    dg_vmult(DGGrid(dim, k, n, pf), x, y);
    return 0;
  }

  // global 1D SIPG matrix (n(k+1))^2 of the unit-scaled form
  int
  oracle_dg_global_1d(int k, int n, double pf, double *L)
  {
    if (k < 1 || k > MAXK || n < 1)
      return 1;
    const std::vector<double> v = dg_global_L(k, n, pf);
    std::copy(v.begin(), v.end(), L);
    return 0;
  }

  int
  oracle_dg_smoother_step(int dim, int k, int n, double pf, int variant,
                          double relax, const double *b, double *u)
  {
    if ((dim != 2 && dim != 3) || k < 1 || k > MAXK || n < 2 || variant < 0 ||
        variant > 1)
      return 1;
    dg_smoother_step(DGGrid(dim, k, n, pf), variant, relax, b, u);
    return 0;
  }

  // fine mesh of n_fine cells (even); coarse has n_fine/2
  int
  oracle_dg_prolongate_add(int dim, int k, int n_fine, const double *coarse,
                           double *fine)
  {
    if ((dim != 2 && dim != 3) || k < 1 || k > MAXK || n_fine < 2 || n_fine % 2)
      return 1;
    dg_transfer(DGGrid(dim, k, n_fine, 1.0), true, const_cast<double *>(coarse), fine);
    return 0;
  }

  int
  oracle_dg_restrict_add(int dim, int k, int n_fine, const double *fine,
                         double *coarse)
  {
    if ((dim != 2 && dim != 3) || k < 1 || k > MAXK || n_fine < 2 || n_fine % 2)
      return 1;
    dg_transfer(DGGrid(dim, k, n_fine, 1.0), false, coarse, const_cast<double *>(fine));
    return 0;
  }

  int
  oracle_dg_vcycle(int dim, int k, int n0, int max_level, double pf,
                   int mg_variant, double relax, int coarse_sweeps,
                   const double *r, double *z)
  {
    if ((dim != 2 && dim != 3) || k < 1 || k > MAXK || n0 < 2 || max_level < 0)
      return 1;
    DGMG mg(dim, k, n0, max_level, pf, mg_variant, relax, coarse_sweeps);
    mg.apply(r, z);
    return 0;
  }

  // PCG (no Dirichlet rows: boundary conditions are weak)
  int
  oracle_dg_pcg(int dim, int k, int n0, int max_level, double pf,
                int mg_variant, double relax, int coarse_sweeps,
                const double *b, double *x, double rtol, int maxit, int *iters,
                double *relres, double *history)
  {
    if ((dim != 2 && dim != 3) || k < 1 || k > MAXK || n0 < 2 || max_level < 0)
      return 1;
    DGMG mg(dim, k, n0, max_level, pf, mg_variant, relax, coarse_sweeps);
    const DGGrid &g = mg.grids[max_level];
    const long long N = g.N;
    std::vector<double> r(b, b + N), z(N), p(N), q(N);
    std::fill(x, x + N, 0.0);
    const double r0 = std::sqrt(dot(N, r.data(), r.data()));
    if (history)
      history[0] = r0;
    int    it = 0;
    double rn = r0;
    if (r0 == 0.0)
      {
        *iters  = 0;
        *relres = 0.0;
        return 0;
      }
    mg.apply(r.data(), z.data());
    std::copy(z.begin(), z.end(), p.begin());
    double rz = dot(N, r.data(), z.data());
    while (it < maxit)
      {
        dg_vmult(g, p.data(), q.data());
        const double alpha = rz / dot(N, p.data(), q.data());
        for (long long i = 0; i < N; ++i)
          {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
          }
        ++it;
        rn = std::sqrt(dot(N, r.data(), r.data()));
        if (history)
          history[it] = rn;
        if (!(rn > rtol * r0))
          break;
        mg.apply(r.data(), z.data());
        const double rz_new = dot(N, r.data(), z.data());
        const double beta   = rz_new / rz;
        rz                  = rz_new;
        for (long long i = 0; i < N; ++i)
          p[i] = z[i] + beta * p[i];
      }
    *iters  = it;
    *relres = rn / r0;
    return std::isfinite(rn) ? 0 : 2;
  }

  // load vector of f = c prod_d sin(pi x_d) (c = dim pi^2: u = prod sin(pi
  // x_d)), kind 1; or f = 1, kind 0; Gauss(k+3) per cell
  int
  oracle_dg_rhs(int dim, int k, int n, int kind, double *b)
  {
    const DGGrid g(dim, k, n, 1.0);
    const Basis &bs = basis(k);
    std::vector<double> qp, qw;
    gauss_rule(k + 3, qp, qw);
    const int n1 = k + 1, nq = k + 3;
    std::vector<double> m(g.N1, 0.0);
    for (int c = 0; c < n; ++c)
      for (int i = 0; i < n1; ++i)
        for (int q = 0; q < nq; ++q)
          {
            const double xq = (c + qp[q]) * g.h;
            const double f1 = kind == 0 ? 1.0 : std::sin(M_PI * xq);
            m[(long long)c * n1 + i] += qw[q] * g.h * f1 * lagrange(bs.nodes, i, qp[q]);
          }
    const double cst = kind == 0 ? 1.0 : dim * M_PI * M_PI;
    for (long long i = 0; i < g.N; ++i)
      {
        const long long X = i % g.N1, Y = (i / g.N1) % g.N1, Z = i / (g.N1 * g.N1);
        double v = cst * m[X] * m[Y];
        if (dim == 3)
          v *= m[Z];
        b[i] = v;
      }
    return 0;
  }

  // L2 error of the DG function u against prod_d sin(pi x_d), Gauss(k+3)
  double
  oracle_dg_l2_error_sine(int dim, int k, int n, const double *u)
  {
    const DGGrid g(dim, k, n, 1.0);
    const Basis &bs = basis(k);
    std::vector<double> qp, qw;
    gauss_rule(k + 3, qp, qw);
    const int nq = k + 3, n1 = k + 1;
    std::vector<double> phi(nq * n1);
    for (int q = 0; q < nq; ++q)
      for (int i = 0; i < n1; ++i)
        phi[q * n1 + i] = lagrange(bs.nodes, i, qp[q]);
    double err = 0.0;
    const int nz = dim == 3 ? n : 1;
#pragma omp parallel for collapse(2) reduction(+ : err) schedule(static)
    for (int cz = 0; cz < nz; ++cz)
      for (int cy = 0; cy < n; ++cy)
        for (int cx = 0; cx < n; ++cx)
          for (int qz = 0; qz < (dim == 3 ? nq : 1); ++qz)
            for (int qy = 0; qy < nq; ++qy)
              for (int qx = 0; qx < nq; ++qx)
                {
                  double uh = 0.0;
                  for (int l = 0; l < (dim == 3 ? n1 : 1); ++l)
                    for (int j = 0; j < n1; ++j)
                      for (int i = 0; i < n1; ++i)
                        {
                          double w = phi[qx * n1 + i] * phi[qy * n1 + j];
                          if (dim == 3)
                            w *= phi[qz * n1 + l];
                          uh += w * u[g.idx((long long)cx * n1 + i, (long long)cy * n1 + j,
                                            dim == 3 ? (long long)cz * n1 + l : 0)];
                        }
                  const double x = (cx + qp[qx]) * g.h, y = (cy + qp[qy]) * g.h,
                               z = (cz + qp[qz]) * g.h;
                  double ue = std::sin(M_PI * x) * std::sin(M_PI * y);
                  double w  = qw[qx] * qw[qy] * g.h * g.h;
                  if (dim == 3)
                    {
                      ue *= std::sin(M_PI * z);
                      w *= qw[qz] * g.h;
                    }
                  err += w * (uh - ue) * (uh - ue);
                }
    return std::sqrt(err);
  }

  // b_i = 2*u01(i) - 1 with u01(i) = (splitmix64(seed ^ i) >> 11) 2^-53;
  // Dirichlet entries set to zero.
  int
  oracle_rhs_random(int dim, int k, int level, unsigned long long seed,
                    double *b)
  {
    const Grid g(dim, k, level);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < g.N; ++i)
      {
        const double u01 =
          (double)(splitmix64(seed ^ (uint64_t)i) >> 11) * 0x1.0p-53;
        b[i] = g.on_boundary(i) ? 0.0 : 2.0 * u01 - 1.0;
      }
    return 0;
  }

  // Assembled load vector b_i = int f phi_i for a separable f(x) =
  // c * prod_d f1(x_d) with f1 = 1 (kind 0) or f1 = sin(3 pi x) and
  // c = dim (3 pi)^2 (kind 1, the reference's Poisson problem), Gauss(k+1)
  // quadrature; Dirichlet entries zero.
  int
  oracle_rhs_assembled(int dim, int k, int level, int kind, double *b)
  {
    const Grid   g(dim, k, level);
    const Basis &bs = basis(k);
    std::vector<double> m(g.N1, 0.0);
    for (int c = 0; c < g.n; ++c)
      for (int i = 0; i <= k; ++i)
        for (int q = 0; q <= k; ++q)
          {
            const double xq = (c + bs.qp[q]) * g.h;
            const double f1 = kind == 0 ? 1.0 : std::sin(3.0 * M_PI * xq);
            m[(long long)k * c + i] += bs.qw[q] * g.h * f1 * bs.N[q * (k + 1) + i];
          }
    const double cst = kind == 0 ? 1.0 : dim * 9.0 * M_PI * M_PI;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < g.N; ++i)
      {
        const long long x = i % g.N1, y = (i / g.N1) % g.N1, z = i / (g.N1 * g.N1);
        double v = cst * m[x] * m[y];
        if (dim == 3)
          v *= m[z];
        b[i] = g.on_boundary(i) ? 0.0 : v;
      }
    return 0;
  }

  // L2 error of the FE function u against prod_d sin(3 pi x_d), with a
  // Gauss(k+3) rule per cell.
  double
  oracle_l2_error_sine(int dim, int k, int level, const double *u)
  {
    const Grid   g(dim, k, level);
    const Basis &bs = basis(k);
    std::vector<double> qp, qw;
    gauss_rule(k + 3, qp, qw);
    const int nq = k + 3, n1 = k + 1;
    std::vector<double> phi(nq * n1);
    for (int q = 0; q < nq; ++q)
      for (int i = 0; i < n1; ++i)
        phi[q * n1 + i] = lagrange(bs.nodes, i, qp[q]);
    double err = 0.0;
    const int nz = dim == 3 ? g.n : 1;
#pragma omp parallel for collapse(2) reduction(+ : err) schedule(static)
    for (int cz = 0; cz < nz; ++cz)
      for (int cy = 0; cy < g.n; ++cy)
        for (int cx = 0; cx < g.n; ++cx)
          for (int qz = 0; qz < (dim == 3 ? nq : 1); ++qz)
            for (int qy = 0; qy < nq; ++qy)
              for (int qx = 0; qx < nq; ++qx)
                {
                  double uh = 0.0;
                  for (int l = 0; l < (dim == 3 ? n1 : 1); ++l)
                    for (int j = 0; j < n1; ++j)
                      for (int i = 0; i < n1; ++i)
                        {
                          double w = phi[qx * n1 + i] * phi[qy * n1 + j];
                          if (dim == 3)
                            w *= phi[qz * n1 + l];
                          const long long gi =
                            (long long)k * cx + i + g.N1 * ((long long)k * cy + j) +
                            g.N1 * g.N1 * ((long long)k * cz * (dim == 3) + l);
                          uh += w * u[gi];
                        }
                  const double x = (cx + qp[qx]) * g.h, y = (cy + qp[qy]) * g.h,
                               z = (cz + qp[qz]) * g.h;
                  double ue = std::sin(3 * M_PI * x) * std::sin(3 * M_PI * y);
                  double w  = qw[qx] * qw[qy] * g.h * g.h;
                  if (dim == 3)
                    {
                      ue *= std::sin(3 * M_PI * z);
                      w *= qw[qz] * g.h;
                    }
                  err += w * (uh - ue) * (uh - ue);
                }
    return std::sqrt(err);
  }

  // Load vector b_i = sum_cells int f phi_i with QGauss(nq = k+1) per cell
  // (laplace_operator.cuh:140-173, the reference's compute_residual
  // assembly), from f given at the quadrature points: fq[cell][q] with cell
  // = cx + n (cy + n cz) and q = qx + nq (qy + nq qz), points
  // ((c + xi_q) h).  CG (dg = 0): lexicographic Q_k DoFs, Dirichlet rows
  // zeroed; DG (dg = 1): X = cx (k+1) + ix, no constrained rows.
  int
  oracle_assemble_rhs(int dim, int k, int n, int dg, const double *fq, double *b)
  {
    if (k < 1 || k > MAXK || n < 1 || (dim != 2 && dim != 3))
      return 1;
    const Basis &bs = basis(k);
    const int    nq = k + 1, n1 = k + 1;
    const long long N1 = dg ? (long long)n * n1 : (long long)k * n + 1;
    const long long N  = dim == 3 ? N1 * N1 * N1 : N1 * N1;
    const double    h  = 1.0 / n;
    std::vector<double> phi(nq * n1);
    for (int q = 0; q < nq; ++q)
      for (int i = 0; i < n1; ++i)
        phi[q * n1 + i] = lagrange(bs.nodes, i, bs.qp[q]);
    std::fill(b, b + N, 0.0);
    const int nz = dim == 3 ? n : 1, lz = dim == 3 ? n1 : 1, qz_n = dim == 3 ? nq : 1;
    const long long nqd = dim == 3 ? (long long)nq * nq * nq : (long long)nq * nq;
    // serial over cells: the scatter adds into shared DoFs
    for (int cz = 0; cz < nz; ++cz)
      for (int cy = 0; cy < n; ++cy)
        for (int cx = 0; cx < n; ++cx)
          {
            const long long cell = cx + (long long)n * (cy + (long long)n * cz);
            const double   *f    = fq + cell * nqd;
            const int st = dg ? n1 : k;
            for (int l = 0; l < lz; ++l)
              for (int j = 0; j < n1; ++j)
                for (int i = 0; i < n1; ++i)
                  {
                    double v = 0.0;
                    for (int qz = 0; qz < qz_n; ++qz)
                      for (int qy = 0; qy < nq; ++qy)
                        for (int qx = 0; qx < nq; ++qx)
                          {
                            double w = bs.qw[qx] * bs.qw[qy] * h * h * phi[qx * n1 + i] *
                                       phi[qy * n1 + j];
                            if (dim == 3)
                              w *= bs.qw[qz] * h * phi[qz * n1 + l];
                            v += w * f[qx + nq * (qy + (long long)nq * qz)];
                          }
                    const long long X = (long long)st * cx + i, Y = (long long)st * cy + j,
                                    Z = dim == 3 ? (long long)st * cz + l : 0;
                    b[X + N1 * (Y + N1 * Z)] += v;
                  }
          }
    if (!dg)
      for (long long g = 0; g < N; ++g)
        {
          const long long X = g % N1, Y = (g / N1) % N1, Z = g / (N1 * N1);
          bool bd = X == 0 || X == N1 - 1 || Y == 0 || Y == N1 - 1;
          if (dim == 3)
            bd = bd || Z == 0 || Z == N1 - 1;
          if (bd)
            b[g] = 0.0;
        }
    return 0;
  }

  // L2 norm and H1 seminorm (broken, cell by cell, for DG) of u_h - u with
  // u = prod_d sin(wave pi x_d), QGauss(nq) per cell
  // (poisson.cu:353-382: integrate_difference with QGauss(k+1) there).
  int
  oracle_error_norms(int dim, int k, int n, int dg, double wave, int nq, const double *u,
                     double *l2, double *h1)
  {
    if (k < 1 || k > MAXK || n < 1 || nq < 1 || nq > 32 || (dim != 2 && dim != 3))
      return 1;
    const Basis &bs = basis(k);
    const int    n1 = k + 1;
    const long long N1 = dg ? (long long)n * n1 : (long long)k * n + 1;
    const double    h  = 1.0 / n, a = wave * M_PI;
    std::vector<double> qp, qw;
    gauss_rule(nq, qp, qw);
    std::vector<double> phi(nq * n1), dphi(nq * n1);
    for (int q = 0; q < nq; ++q)
      for (int i = 0; i < n1; ++i)
        {
          phi[q * n1 + i]  = lagrange(bs.nodes, i, qp[q]);
          dphi[q * n1 + i] = lagrange_deriv(bs.nodes, i, qp[q]) / h;
        }
    const int nz = dim == 3 ? n : 1, lz = dim == 3 ? n1 : 1, qz_n = dim == 3 ? nq : 1;
    const int st = dg ? n1 : k;
    double e0 = 0.0, e1 = 0.0;
#pragma omp parallel for collapse(2) reduction(+ : e0, e1) schedule(static)
    for (int cz = 0; cz < nz; ++cz)
      for (int cy = 0; cy < n; ++cy)
        for (int cx = 0; cx < n; ++cx)
          for (int qz = 0; qz < qz_n; ++qz)
            for (int qy = 0; qy < nq; ++qy)
              for (int qx = 0; qx < nq; ++qx)
                {
                  double uh = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
                  for (int l = 0; l < lz; ++l)
                    for (int j = 0; j < n1; ++j)
                      for (int i = 0; i < n1; ++i)
                        {
                          const long long X = (long long)st * cx + i,
                                          Y = (long long)st * cy + j,
                                          Z = dim == 3 ? (long long)st * cz + l : 0;
                          const double v  = u[X + N1 * (Y + N1 * Z)];
                          const double pz = dim == 3 ? phi[qz * n1 + l] : 1.0;
                          const double dz = dim == 3 ? dphi[qz * n1 + l] : 0.0;
                          uh += v * phi[qx * n1 + i] * phi[qy * n1 + j] * pz;
                          gx += v * dphi[qx * n1 + i] * phi[qy * n1 + j] * pz;
                          gy += v * phi[qx * n1 + i] * dphi[qy * n1 + j] * pz;
                          gz += v * phi[qx * n1 + i] * phi[qy * n1 + j] * dz;
                        }
                  const double x = (cx + qp[qx]) * h, y = (cy + qp[qy]) * h,
                               z = (cz + qp[qz]) * h;
                  const double sx = std::sin(a * x), sy = std::sin(a * y),
                               sz = dim == 3 ? std::sin(a * z) : 1.0;
                  const double cxv = a * std::cos(a * x), cyv = a * std::cos(a * y),
                               czv = dim == 3 ? a * std::cos(a * z) : 0.0;
                  double w = qw[qx] * qw[qy] * h * h;
                  if (dim == 3)
                    w *= qw[qz] * h;
                  const double du = uh - sx * sy * sz;
                  const double dx = gx - cxv * sy * sz, dy = gy - sx * cyv * sz,
                               dzv = gz - sx * sy * czv;
                  e0 += w * du * du;
                  e1 += w * (dx * dx + dy * dy + dzv * dzv);
                }
    *l2 = std::sqrt(e0);
    *h1 = std::sqrt(e1);
    return 0;
  }

  int
  oracle_set_threads(int n)
  {
    if (n > 0)
      omp_set_num_threads(n);
    return omp_get_max_threads();
  }
}
