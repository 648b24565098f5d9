// coarse.hip - the V-cycle from level 2 down as one dense matrix.
//
// Below level 3 the V-cycle (pre-smoothing from zero, residual,
// restriction, the exact level-1 solve, prolongation, post-smoothing) is a
// fixed linear map of the level-2 defect, zero on the Dirichlet DoFs.  On
// these levels every kernel is latency-bound (a few microseconds each for a
// few hundred to a few thousand DoFs, about twenty launches per V-cycle), so
// the map is tabulated once at setup, column by column with the ordinary
// V-cycle kernels applied to unit defects, and each V-cycle then replaces
// the whole sub-cycle by one matrix-vector product over the interior DoFs.
// The result is the same linear map, rounded differently.  The DG
// hierarchy tabulates its coarsest solve (symmetric sweep pairs from zero on
// a mesh of at most 4 cells per direction) the same way, with the plain
// forms below (no constrained rows), one column per work-group.
#include "gtps_internal.h"

namespace gtps
{
  namespace
  {
    constexpr int CNT = 256;
    // loads in flight per lane in the row loop of dense_apply_kernel
    constexpr int ROW_BATCH = 8;

    // interior DoF j (x fastest over the (N1-2)^dim interior) -> level index;
    // DIM = 0: no constrained rows, the identity
    template <int DIM>
    __device__ __forceinline__ long long
    interior_index(int j, int N1)
    {
      if (DIM == 0)
        return j;
      const int ni1 = N1 - 2;
      const int x = j % ni1 + 1, y = (j / ni1) % ni1 + 1;
      const int z = DIM == 3 ? j / (ni1 * ni1) + 1 : 0;
      return x + (long long)N1 * (y + (long long)N1 * z);
    }

    // interior_index with host-computed divisors (ni1 = N1 - 2 and ni1^2;
    // the dense form only serves levels of a few thousand DoFs: 32-bit)
    template <int DIM>
    __device__ __forceinline__ int
    interior_index_fast(int j, int N1, const FastDiv &da, const FastDiv &db)
    {
      if (DIM == 0)
        return j;
      const int q = (int)fastdiv((unsigned)j, da), x = j - q * (int)da.d + 1;
      if (DIM == 2)
        return x + N1 * (q + 1);
      const int z = (int)fastdiv((unsigned)j, db), y = q - z * (int)da.d + 1;
      return x + N1 * (y + N1 * (z + 1));
    }

    // D[j][*] = e_{interior j}: column j's unit defect
    template <typename T, int DIM>
    __global__ void __launch_bounds__(CNT)
    unit_defects_kernel(T *__restrict__ D, const int N1, const long long N, const int ni)
    {
      for (long long e = (long long)blockIdx.x * CNT + threadIdx.x; e < (long long)ni * N;
           e += (long long)gridDim.x * CNT)
        {
          const long long j = e / N, g = e - j * N;
          D[e] = g == interior_index<DIM>((int)j, N1) ? T(1) : T(0);
        }
    }

    // B[i][j] = U[j][interior i] for j < ni, 0 in the padding columns
    // ni <= j < ld (rows of ld entries, 16-byte aligned)
    template <typename T, int DIM>
    __global__ void __launch_bounds__(CNT)
    gather_columns_kernel(const T *__restrict__ U, T *__restrict__ B, const int N1,
                          const long long N, const int ni, const int ld)
    {
      for (long long e = (long long)blockIdx.x * CNT + threadIdx.x; e < (long long)ni * ld;
           e += (long long)gridDim.x * CNT)
        {
          const int i = (int)(e / ld), j = (int)(e - (long long)i * ld);
          B[e]        = j < ni ? U[(long long)j * N + interior_index<DIM>(i, N1)] : T(0);
        }
    }

    template <typename T>
    struct Vec16;
    template <>
    struct Vec16<float>
    {
      using type = float4;
      __device__ static float
      dot(const float4 a, const float4 b)
      {
        return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
      }
    };
    template <>
    struct Vec16<double>
    {
      using type = double2;
      __device__ static double
      dot(const double2 a, const double2 b)
      {
        return a.x * b.x + a.y * b.y;
      }
    };

    // u = B d on the interior, 0 on the Dirichlet DoFs.  One wavefront per
    // row: the lanes read the row in 16-byte pieces, ROW_BATCH of them in
    // flight per lane before the first product, so a row of up to
    // 64 * ROW_BATCH pieces costs one memory round trip; d (gathered to the
    // interior numbering) sits in LDS.
    template <typename T, int DIM>
    __global__ void __launch_bounds__(CNT)
    dense_apply_kernel(const T *__restrict__ B, const T *__restrict__ d, T *__restrict__ u,
                       const int N1, const long long N, const int ni, const int ld,
                       const FastDiv da, const FastDiv db)
    {
      using V  = typename Vec16<T>::type;
      constexpr int PER = 16 / sizeof(T);
      __shared__ __align__(16) T sd[MAX_DENSE_COARSE];
      const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
      // batch: blockIdx.y selects one of several contiguous vectors
      d += (long long)blockIdx.y * N;
      u += (long long)blockIdx.y * N;
      {
        // the defect's interior values: every load issued before the first
        // wait (a load-store loop waited for each element in turn)
        constexpr int ITD = (MAX_DENSE_COARSE + CNT - 1) / CNT;
        T             v[ITD];
#pragma unroll
        for (int q = 0; q < ITD; ++q)
          {
            const int j = tid + q * CNT;
            v[q]        = j < ni ? d[interior_index_fast<DIM>(j, N1, da, db)] : T(0);
          }
#pragma unroll
        for (int q = 0; q < ITD; ++q)
          if (tid + q * CNT < ld)
            sd[tid + q * CNT] = v[q];
      }
      // Dirichlet rows: zero (spread over the grid)
      // (the dense form only serves levels of a few thousand DoFs: 32-bit)
      for (int g = blockIdx.x * CNT + tid; DIM != 0 && g < (int)N; g += gridDim.x * CNT)
        {
          const int x = g % N1, y = (g / N1) % N1, z = DIM == 3 ? g / (N1 * N1) : 1;
          bool bd = x == 0 || x == N1 - 1 || y == 0 || y == N1 - 1;
          if (DIM == 3)
            bd = bd || z == 0 || z == N1 - 1;
          if (bd)
            u[g] = T(0);
        }
      __syncthreads();
      const int i = blockIdx.x * (CNT / 64) + wave;
      if (i >= ni)
        return;
      const V  *row = reinterpret_cast<const V *>(B + (long long)i * ld);
      const V  *sv  = reinterpret_cast<const V *>(sd);
      const int nv  = ld / PER;
      T         acc = 0;
      for (int base = 0; base < nv; base += 64 * ROW_BATCH)
        {
          V r[ROW_BATCH];
#pragma unroll
          for (int q = 0; q < ROW_BATCH; ++q)
            {
              const int j = base + q * 64 + lane;
              if (j < nv)
                r[q] = row[j];
            }
#pragma unroll
          for (int q = 0; q < ROW_BATCH; ++q)
            {
              const int j = base + q * 64 + lane;
              if (j < nv)
                acc += Vec16<T>::dot(r[q], sv[j]);
            }
        }
      const double s = wave_sum((double)acc);
      if (lane == 0)
        u[interior_index_fast<DIM>(i, N1, da, db)] = (T)s;
    }

    int
    cgrid(long long work, int cap = 16384)
    {
      return (int)std::max<long long>(1, std::min<long long>((work + CNT - 1) / CNT, cap));
    }

    template <typename F>
    void
    by_type(int dim, int prec, F &&f)
    {
      if (prec == GTPS_FP32)
        {
          if (dim == 3)
            f((float *)nullptr, std::integral_constant<int, 3>{});
          else
            f((float *)nullptr, std::integral_constant<int, 2>{});
        }
      else
        {
          if (dim == 3)
            f((double *)nullptr, std::integral_constant<int, 3>{});
          else
            f((double *)nullptr, std::integral_constant<int, 2>{});
        }
    }
  } // namespace

  int
  dense_ld(int n)
  {
    return (n + 3) & ~3;
  }

  void
  launch_unit_defects_plain(hipStream_t s, int n, int prec, void *D)
  {
    if (prec == GTPS_FP32)
      hipLaunchKernelGGL((unit_defects_kernel<float, 0>), dim3(cgrid((long long)n * n)),
                         dim3(CNT), 0, s, (float *)D, 0, (long long)n, n);
    else
      hipLaunchKernelGGL((unit_defects_kernel<double, 0>), dim3(cgrid((long long)n * n)),
                         dim3(CNT), 0, s, (double *)D, 0, (long long)n, n);
    GTPS_LAUNCH_CHECK();
  }

  void
  launch_transpose_plain(hipStream_t s, int n, int prec, const void *U, void *B)
  {
    const int ld = dense_ld(n);
    if (prec == GTPS_FP32)
      hipLaunchKernelGGL((gather_columns_kernel<float, 0>), dim3(cgrid((long long)n * ld)),
                         dim3(CNT), 0, s, (const float *)U, (float *)B, 0, (long long)n, n,
                         ld);
    else
      hipLaunchKernelGGL((gather_columns_kernel<double, 0>), dim3(cgrid((long long)n * ld)),
                         dim3(CNT), 0, s, (const double *)U, (double *)B, 0, (long long)n, n,
                         ld);
    GTPS_LAUNCH_CHECK();
  }

  void
  launch_dense_apply_plain(hipStream_t s, int n, int prec, const void *B, const void *d,
                           void *u, int nvec)
  {
    GTPS_REQUIRE(n >= 1 && dense_ld(n) <= MAX_DENSE_COARSE, "dense coarse operator too large");
    GTPS_REQUIRE(nvec >= 1, "bad batch");
    const int  ld = dense_ld(n);
    const dim3 grid((n + CNT / 64 - 1) / (CNT / 64), nvec);
    if (prec == GTPS_FP32)
      hipLaunchKernelGGL((dense_apply_kernel<float, 0>), grid, dim3(CNT), 0, s,
                         (const float *)B, (const float *)d, (float *)u, 0, (long long)n, n,
                         ld, make_fastdiv(1), make_fastdiv(1));
    else
      hipLaunchKernelGGL((dense_apply_kernel<double, 0>), grid, dim3(CNT), 0, s,
                         (const double *)B, (const double *)d, (double *)u, 0, (long long)n,
                         n, ld, make_fastdiv(1), make_fastdiv(1));
    GTPS_LAUNCH_CHECK();
  }

  int
  dense_coarse_size(int dim, int N1)
  {
    const int ni1 = N1 - 2;
    return dim == 3 ? ni1 * ni1 * ni1 : ni1 * ni1;
  }

  void
  launch_unit_defects(hipStream_t s, int dim, int N1, int prec, void *D)
  {
    const int       ni = dense_coarse_size(dim, N1);
    const long long N  = dim == 3 ? (long long)N1 * N1 * N1 : (long long)N1 * N1;
    by_type(dim, prec, [&](auto tp, auto dd) {
      using T = std::remove_pointer_t<decltype(tp)>;
      hipLaunchKernelGGL((unit_defects_kernel<T, decltype(dd)::value>),
                         dim3(cgrid((long long)ni * N)), dim3(CNT), 0, s, (T *)D, N1, N, ni);
    });
    GTPS_LAUNCH_CHECK();
  }

  void
  launch_gather_columns(hipStream_t s, int dim, int N1, int prec, const void *U, void *B)
  {
    const int       ni = dense_coarse_size(dim, N1), ld = dense_ld(ni);
    const long long N  = dim == 3 ? (long long)N1 * N1 * N1 : (long long)N1 * N1;
    by_type(dim, prec, [&](auto tp, auto dd) {
      using T = std::remove_pointer_t<decltype(tp)>;
      hipLaunchKernelGGL((gather_columns_kernel<T, decltype(dd)::value>),
                         dim3(cgrid((long long)ni * ld)), dim3(CNT), 0, s, (const T *)U,
                         (T *)B, N1, N, ni, ld);
    });
    GTPS_LAUNCH_CHECK();
  }

  void
  launch_dense_apply(hipStream_t s, int dim, int N1, int prec, const void *B, const void *d,
                     void *u)
  {
    const int ni = dense_coarse_size(dim, N1), ld = dense_ld(ni);
    GTPS_REQUIRE(ni >= 1 && ld <= MAX_DENSE_COARSE, "dense coarse operator too large");
    const long long N  = dim == 3 ? (long long)N1 * N1 * N1 : (long long)N1 * N1;
    const int       nb = (ni + CNT / 64 - 1) / (CNT / 64);
    const FastDiv   da = make_fastdiv((unsigned)(N1 - 2)),
                  db = make_fastdiv((unsigned)((N1 - 2) * (N1 - 2)));
    by_type(dim, prec, [&](auto tp, auto dd) {
      using T = std::remove_pointer_t<decltype(tp)>;
      hipLaunchKernelGGL((dense_apply_kernel<T, decltype(dd)::value>), dim3(nb), dim3(CNT), 0,
                         s, (const T *)B, (const T *)d, (T *)u, N1, N, ni, ld, da, db);
    });
    GTPS_LAUNCH_CHECK();
  }
} // namespace gtps
