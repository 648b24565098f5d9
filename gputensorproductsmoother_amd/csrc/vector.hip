// vector.hip - FP64/FP32 vector kernels and the fused CG building blocks.
//
// Reductions are two-stage and deterministic: a kernel writes one partial
// sum per work-group (wavefront shuffles, then LDS across the waves), and a
// single work-group adds the partials in a fixed order into a device scalar
// slot.  CG scalars (alpha, beta) are formed on the device, so an iteration
// needs no host round trip except the convergence check.
//
// Reference counterparts: deal.II Vector<double,CUDA> sadd/add/l2_norm/dot
// used by the solver (include/solver.cuh:1712,1830,1853,1977, SolverGMRES
// internals), vec_equ/plain_copy (include/cuda_vector.cuh:79-118) and
// copy_vector<is_d2f> (include/solver.cuh:64-78).
#include "gtps_internal.h"

namespace gtps
{
  namespace
  {
    constexpr int VNT = 256;

    int
    grid_for(long long n, int cap = 8192)
    {
      long long nb = (n + VNT - 1) / VNT;
      return (int)std::max<long long>(1, std::min<long long>(nb, cap));
    }

    template <typename T>
    __global__ void __launch_bounds__(VNT)
    fill_kernel(long long n, T *__restrict__ x, T v)
    {
      for (long long i = (long long)blockIdx.x * VNT + threadIdx.x; i < n;
           i += (long long)gridDim.x * VNT)
        x[i] = v;
    }

    __global__ void __launch_bounds__(VNT)
    axpby_kernel(long long n, double a, const double *__restrict__ x, double b,
                 double *__restrict__ y)
    {
      for (long long i = (long long)blockIdx.x * VNT + threadIdx.x; i < n;
           i += (long long)gridDim.x * VNT)
        y[i] = a * x[i] + b * y[i];
    }

    __global__ void __launch_bounds__(VNT)
    d2f_kernel(long long n, const double *__restrict__ x, float *__restrict__ y)
    {
      for (long long i = (long long)blockIdx.x * VNT + threadIdx.x; i < n;
           i += (long long)gridDim.x * VNT)
        y[i] = (float)x[i];
    }

    __global__ void __launch_bounds__(VNT)
    f2d_kernel(long long n, const float *__restrict__ x, double *__restrict__ y)
    {
      for (long long i = (long long)blockIdx.x * VNT + threadIdx.x; i < n;
           i += (long long)gridDim.x * VNT)
        y[i] = (double)x[i];
    }

    __global__ void __launch_bounds__(VNT)
    dot_kernel(long long n, const double *__restrict__ x,
               const double *__restrict__ y, const float *__restrict__ y32,
               double *__restrict__ partials)
    {
      __shared__ double red[VNT / 64];
      double            s = 0.0;
      for (long long i = (long long)blockIdx.x * VNT + threadIdx.x; i < n;
           i += (long long)gridDim.x * VNT)
        s += x[i] * (y32 ? (double)y32[i] : y[i]);
      s = block_sum<VNT>(s, red);
      if (threadIdx.x == 0)
        partials[blockIdx.x] = s;
    }

    // Block sum of partials[0..nb) (result in thread 0).  Each thread
    // issues its loads 8 at a time into separate accumulators, so the load
    // latencies overlap; the summation order is fixed for a given nb.
    __device__ __forceinline__ double
    sum_partials(const double *__restrict__ partials, int nb, double *red)
    {
      constexpr int U = 8;
      double        acc[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        acc[u] = 0.0;
      for (int i0 = threadIdx.x; i0 < nb; i0 += VNT * U)
        {
#pragma unroll
          for (int u = 0; u < U; ++u)
            {
              const int i = i0 + u * VNT;
              acc[u] += i < nb ? partials[i] : 0.0;
            }
        }
      double s = 0.0;
#pragma unroll
      for (int u = 0; u < U; ++u)
        s += acc[u];
      return block_sum<VNT>(s, red);
    }

    // CG step quotients: a zero denominator only arises when the residual
    // (and with it z, p and q) is exactly zero, where the step is 0; without
    // the guard 0/0 would turn a zero right-hand side of a batch into NaN.
    __device__ __forceinline__ double
    quotient(double num, double den)
    {
      return den != 0.0 ? num / den : 0.0;
    }

    __global__ void __launch_bounds__(VNT)
    finalize_kernel(const double *__restrict__ partials, int nb,
                    double *__restrict__ scalars, int slot,
                    double *__restrict__ host_out)
    {
      __shared__ double red[VNT / 64];
      const double      s = sum_partials(partials, nb, red);
      if (threadIdx.x == 0)
        {
          scalars[slot] = s;
          if (host_out)
            *host_out = s;
        }
    }

    __global__ void __launch_bounds__(VNT)
    boundary_sumsq_kernel(int dim, int N1, const double *__restrict__ b,
                          double *__restrict__ partials)
    {
      __shared__ double red[VNT / 64];
      const long long   face  = dim == 3 ? (long long)N1 * N1 : N1;
      const long long   total = 2 * dim * face;
      double            s     = 0.0;
      for (long long e = (long long)blockIdx.x * VNT + threadIdx.x; e < total;
           e += (long long)gridDim.x * VNT)
        {
          const int       f  = (int)(e / face);
          const long long w  = e - f * face;
          const int       a  = (int)(w % N1), c = (int)(w / N1);
          const int       hi = (f & 1) ? N1 - 1 : 0;
          long long       i;
          if (dim == 3)
            {
              const long long P = (long long)N1 * N1;
              i = f < 2 ? a + (long long)c * N1 + hi * P        // z face
                : f < 4 ? a + (long long)hi * N1 + c * P        // y face
                        : hi + (long long)a * N1 + c * P;       // x face
            }
          else
            i = f < 2 ? a + (long long)hi * N1 : hi + (long long)a * N1;
          const double v = b[i];
          s += v * v;
        }
      s = block_sum<VNT>(s, red);
      if (threadIdx.x == 0)
        partials[blockIdx.x] = s;
    }

    __global__ void __launch_bounds__(VNT)
    cg_init_kernel(long long n, const double *__restrict__ b, double *__restrict__ x,
                   double *__restrict__ r, float *__restrict__ r32,
                   double *__restrict__ p, double *__restrict__ sc,
                   double *__restrict__ partials)
    {
      __shared__ double red[VNT / 64];
      double            s = 0.0;
      for (long long i = (long long)blockIdx.x * VNT + threadIdx.x; i < n;
           i += (long long)gridDim.x * VNT)
        {
          const double bi = b[i];
          x[i]            = 0.0;
          p[i]            = 0.0;
          r[i]            = bi;
          if (r32)
            r32[i] = (float)bi;
          s += bi * bi;
        }
      s = block_sum<VNT>(s, red);
      if (threadIdx.x == 0)
        {
          partials[blockIdx.x] = s;
          if (blockIdx.x == 0)
            {
              sc[cg::RZ]    = 1.0;
              sc[cg::ALPHA] = 0.0;
            }
        }
    }

    // sum of the partials into scalars[slot], then one scalar op; op < 0:
    // none.  OP_SHIFT_BETA reads the old rz before it is replaced.
    __global__ void __launch_bounds__(VNT)
    finalize_op_kernel(const double *__restrict__ partials, int nb,
                       double *__restrict__ sc, int slot, int op)
    {
      __shared__ double red[VNT / 64];
      const double      s = sum_partials(partials, nb, red);
      if (threadIdx.x != 0)
        return;
      if (op == cg::OP_SHIFT_BETA)
        {
          const double old = sc[cg::RZ];
          sc[cg::RZ_OLD]   = old;
          sc[cg::RZ]       = s;
          sc[cg::BETA]     = quotient(s, old);
          return;
        }
      sc[slot] = s;
      if (op == cg::OP_ALPHA)
        sc[cg::ALPHA] = quotient(sc[cg::RZ], s);
    }

    __global__ void
    cg_scalar_kernel(double *sc, int op)
    {
      if (threadIdx.x != 0)
        return;
      switch (op)
        {
          case cg::OP_ALPHA:
            sc[cg::ALPHA] = quotient(sc[cg::RZ], sc[cg::PQ]);
            break;
          case cg::OP_BETA:
            sc[cg::BETA] = quotient(sc[cg::RZ], sc[cg::RZ_OLD]);
            break;
          case cg::OP_SHIFT:
            sc[cg::RZ_OLD] = sc[cg::RZ];
            break;
          case cg::OP_INIT:
            sc[cg::RZ]    = 1.0;
            sc[cg::ALPHA] = 0.0;
            break;
          default:
            break;
        }
    }

    // x += alpha p (x == nullptr: deferred to the next p update); r -=
    // alpha q; partial ||r||^2; r32 = (float) r
    __global__ void __launch_bounds__(VNT)
    cg_update_xr_kernel(long long n, double *__restrict__ x,
                        double *__restrict__ r, const double *__restrict__ p,
                        const double *__restrict__ q, float *__restrict__ r32,
                        const double *__restrict__ sc,
                        double *__restrict__ partials)
    {
      __shared__ double red[VNT / 64];
      const double      alpha = sc[cg::ALPHA];
      double            s     = 0.0;
      for (long long i = (long long)blockIdx.x * VNT + threadIdx.x; i < n;
           i += (long long)gridDim.x * VNT)
        {
          if (x)
            x[i] += alpha * p[i];
          const double ri = r[i] - alpha * q[i];
          r[i]            = ri;
          if (r32)
            r32[i] = (float)ri;
          s += ri * ri;
        }
      s = block_sum<VNT>(s, red);
      if (threadIdx.x == 0)
        partials[blockIdx.x] = s;
    }

    // p = z + beta p, z given in FP64 or FP32; x != nullptr: first the
    // deferred x += alpha p of the last step with the old p (read once for
    // both)
    __global__ void __launch_bounds__(VNT)
    cg_update_p_kernel(long long n, double *__restrict__ p,
                       const double *__restrict__ z,
                       const float *__restrict__ z32,
                       const double *__restrict__ sc, double *__restrict__ x)
    {
      const double beta = sc[cg::BETA], alpha = sc[cg::ALPHA];
      for (long long i = (long long)blockIdx.x * VNT + threadIdx.x; i < n;
           i += (long long)gridDim.x * VNT)
        {
          const double zi = z32 ? (double)z32[i] : z[i];
          const double pi = p[i];
          if (x)
            x[i] += alpha * pi;
          p[i] = zi + beta * pi;
        }
    }

    // x += alpha p: the deferred update of the last step
    __global__ void __launch_bounds__(VNT)
    cg_add_alpha_p_kernel(long long n, double *__restrict__ x,
                          const double *__restrict__ p, const double *__restrict__ sc)
    {
      const double alpha = sc[cg::ALPHA];
      for (long long i = (long long)blockIdx.x * VNT + threadIdx.x; i < n;
           i += (long long)gridDim.x * VNT)
        x[i] += alpha * p[i];
    }

    template <typename T>
    __global__ void __launch_bounds__(VNT)
    set_boundary_kernel(int dim, int N1, long long n, T *__restrict__ v, T val)
    {
      for (long long i = (long long)blockIdx.x * VNT + threadIdx.x; i < n;
           i += (long long)gridDim.x * VNT)
        {
          const int  x  = (int)(i % N1);
          const int  y  = (int)((i / N1) % N1);
          const int  z  = (int)(i / ((long long)N1 * N1));
          const bool bd = x == 0 || x == N1 - 1 || y == 0 || y == N1 - 1 ||
                          (dim == 3 && (z == 0 || z == N1 - 1));
          if (bd)
            v[i] = val;
        }
    }

    template <typename T>
    __global__ void __launch_bounds__(VNT)
    copy_boundary_kernel(int dim, int N1, long long n, const T *__restrict__ src,
                         T *__restrict__ dst)
    {
      for (long long i = (long long)blockIdx.x * VNT + threadIdx.x; i < n;
           i += (long long)gridDim.x * VNT)
        {
          const int  x  = (int)(i % N1);
          const int  y  = (int)((i / N1) % N1);
          const int  z  = (int)(i / ((long long)N1 * N1));
          const bool bd = x == 0 || x == N1 - 1 || y == 0 || y == N1 - 1 ||
                          (dim == 3 && (z == 0 || z == N1 - 1));
          if (bd)
            dst[i] = src[i];
        }
    }

    __global__ void __launch_bounds__(VNT)
    add_f32_to_f64_kernel(long long n, const float *__restrict__ x,
                          double *__restrict__ y)
    {
      for (long long i = (long long)blockIdx.x * VNT + threadIdx.x; i < n;
           i += (long long)gridDim.x * VNT)
        y[i] += (double)x[i];
    }

    __global__ void __launch_bounds__(VNT)
    lift_boundary_kernel(int dim, int N1, long long n,
                         const double *__restrict__ b, double *__restrict__ x)
    {
      for (long long i = (long long)blockIdx.x * VNT + threadIdx.x; i < n;
           i += (long long)gridDim.x * VNT)
        {
          const int  xx = (int)(i % N1);
          const int  y  = (int)((i / N1) % N1);
          const int  z  = (int)(i / ((long long)N1 * N1));
          const bool bd = xx == 0 || xx == N1 - 1 || y == 0 || y == N1 - 1 ||
                          (dim == 3 && (z == 0 || z == N1 - 1));
          x[i] = bd ? b[i] : 0.0;
        }
    }
  } // namespace

  void
  launch_fill(hipStream_t s, long long n, int prec, void *x, double v)
  {
    if (n <= 0)
      return;
    if (prec == GTPS_FP32)
      hipLaunchKernelGGL(fill_kernel<float>, dim3(grid_for(n)), dim3(VNT), 0,
                         s, n, (float *)x, (float)v);
    else
      hipLaunchKernelGGL(fill_kernel<double>, dim3(grid_for(n)), dim3(VNT), 0,
                         s, n, (double *)x, v);
    GTPS_LAUNCH_CHECK();
  }

  void
  launch_axpby(hipStream_t s, long long n, double a, const double *x, double b,
               double *y)
  {
    if (n <= 0)
      return;
    hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(VNT), 0, s, n, a,
                       x, b, y);
    GTPS_LAUNCH_CHECK();
  }

  void
  launch_convert_f64_f32(hipStream_t s, long long n, const double *x, float *y)
  {
    if (n <= 0)
      return;
    hipLaunchKernelGGL(d2f_kernel, dim3(grid_for(n)), dim3(VNT), 0, s, n, x, y);
    GTPS_LAUNCH_CHECK();
  }

  void
  launch_convert_f32_f64(hipStream_t s, long long n, const float *x, double *y)
  {
    if (n <= 0)
      return;
    hipLaunchKernelGGL(f2d_kernel, dim3(grid_for(n)), dim3(VNT), 0, s, n, x, y);
    GTPS_LAUNCH_CHECK();
  }

  int
  launch_dot_partials(hipStream_t s, long long n, const double *x,
                      const double *y, const float *y32, double *partials,
                      int capacity)
  {
    const int nb = grid_for(n, std::min(capacity, 2048));
    hipLaunchKernelGGL(dot_kernel, dim3(nb), dim3(VNT), 0, s, n, x, y, y32,
                       partials);
    GTPS_LAUNCH_CHECK();
    return nb;
  }

  void
  launch_finalize_sum(hipStream_t s, const double *partials, int nb,
                      double *scalars, int slot, double *host_out)
  {
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(VNT), 0, s, partials, nb,
                       scalars, slot, host_out);
    GTPS_LAUNCH_CHECK();
  }

  int
  launch_boundary_sumsq(const gtps_level *lvl, const double *b, double *partials,
                        int capacity)
  {
    const long long face = lvl->dim == 3 ? (long long)lvl->N1 * lvl->N1 : lvl->N1;
    const int       nb   = grid_for(2 * lvl->dim * face, std::min(capacity, 2048));
    hipLaunchKernelGGL(boundary_sumsq_kernel, dim3(nb), dim3(VNT), 0, lvl->ctx->stream,
                       lvl->dim, lvl->N1, b, partials);
    GTPS_LAUNCH_CHECK();
    return nb;
  }

  int
  launch_cg_init(hipStream_t s, long long n, const double *b, double *x, double *r,
                 float *r32, double *p, double *scalars, double *partials, int capacity)
  {
    const int nb = grid_for(n, std::min(capacity, 2048));
    hipLaunchKernelGGL(cg_init_kernel, dim3(nb), dim3(VNT), 0, s, n, b, x, r, r32, p,
                       scalars, partials);
    GTPS_LAUNCH_CHECK();
    return nb;
  }

  void
  launch_finalize_op(hipStream_t s, const double *partials, int nb,
                     double *scalars, int slot, int op)
  {
    hipLaunchKernelGGL(finalize_op_kernel, dim3(1), dim3(VNT), 0, s, partials,
                       nb, scalars, slot, op);
    GTPS_LAUNCH_CHECK();
  }

  void
  launch_cg_scalar(hipStream_t s, double *scalars, int op)
  {
    hipLaunchKernelGGL(cg_scalar_kernel, dim3(1), dim3(64), 0, s, scalars, op);
    GTPS_LAUNCH_CHECK();
  }

  int
  launch_cg_update_xr(hipStream_t s, long long n, double *x, double *r,
                      const double *p, const double *q, float *r32,
                      const double *scalars, double *partials, int capacity)
  {
    const int nb = grid_for(n, std::min(capacity, 2048));
    hipLaunchKernelGGL(cg_update_xr_kernel, dim3(nb), dim3(VNT), 0, s, n, x, r,
                       p, q, r32, scalars, partials);
    GTPS_LAUNCH_CHECK();
    return nb;
  }

  void
  launch_cg_update_p(hipStream_t s, long long n, double *p, const double *z,
                     const float *z32, const double *scalars, double *x)
  {
    hipLaunchKernelGGL(cg_update_p_kernel, dim3(grid_for(n)), dim3(VNT), 0, s,
                       n, p, z, z32, scalars, x);
    GTPS_LAUNCH_CHECK();
  }

  void
  launch_cg_add_alpha_p(hipStream_t s, long long n, double *x, const double *p,
                        const double *scalars)
  {
    hipLaunchKernelGGL(cg_add_alpha_p_kernel, dim3(grid_for(n)), dim3(VNT), 0, s,
                       n, x, p, scalars);
    GTPS_LAUNCH_CHECK();
  }

  void
  launch_set_boundary(const gtps_level *lvl, void *v, double value)
  {
    hipStream_t s = lvl->ctx->stream;
    if (lvl->prec == GTPS_FP32)
      hipLaunchKernelGGL(set_boundary_kernel<float>, dim3(grid_for(lvl->N)),
                         dim3(VNT), 0, s, lvl->dim, lvl->N1, lvl->N, (float *)v,
                         (float)value);
    else
      hipLaunchKernelGGL(set_boundary_kernel<double>, dim3(grid_for(lvl->N)),
                         dim3(VNT), 0, s, lvl->dim, lvl->N1, lvl->N,
                         (double *)v, value);
    GTPS_LAUNCH_CHECK();
  }

  void
  launch_copy_boundary(const gtps_level *lvl, const void *src, void *dst)
  {
    hipStream_t s = lvl->ctx->stream;
    if (lvl->prec == GTPS_FP32)
      hipLaunchKernelGGL(copy_boundary_kernel<float>, dim3(grid_for(lvl->N)),
                         dim3(VNT), 0, s, lvl->dim, lvl->N1, lvl->N,
                         (const float *)src, (float *)dst);
    else
      hipLaunchKernelGGL(copy_boundary_kernel<double>, dim3(grid_for(lvl->N)),
                         dim3(VNT), 0, s, lvl->dim, lvl->N1, lvl->N,
                         (const double *)src, (double *)dst);
    GTPS_LAUNCH_CHECK();
  }

  void
  launch_add_f32_to_f64(hipStream_t s, long long n, const float *x, double *y)
  {
    if (n <= 0)
      return;
    hipLaunchKernelGGL(add_f32_to_f64_kernel, dim3(grid_for(n)), dim3(VNT), 0,
                       s, n, x, y);
    GTPS_LAUNCH_CHECK();
  }

  void
  launch_lift_boundary(const gtps_level *lvl, const double *b, double *x)
  {
    hipLaunchKernelGGL(lift_boundary_kernel, dim3(grid_for(lvl->N)), dim3(VNT),
                       0, lvl->ctx->stream, lvl->dim, lvl->N1, lvl->N, b, x);
    GTPS_LAUNCH_CHECK();
  }
} // namespace gtps
