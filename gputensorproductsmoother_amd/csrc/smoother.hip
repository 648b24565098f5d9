// smoother.hip - vertex-patch Schwarz smoother, one colour per launch.
//
// A vertex patch is the 2^d cells around an interior mesh vertex; its DoF
// box has (2k+1)^d nodes and (2k-1)^d interior nodes.  Patches are coloured
// by the parity of their lower-left cell index in each direction, so the
// interiors of one colour are disjoint and no patch of a colour reads
// another's interior (patch_base.template.cuh:1208-1285).
//
// Multiplicative pass (reference FUSED_CF, loop_kernel.cuh:969-1060 with
// evaluate_kernel.cuh:2058-2163): u_I <- w (u_I + A_II^{-1} (b - A_P u)_I).
// This kernel evaluates the algebraically identical form
// u_I <- w A_II^{-1} (b_I - A_IB u_B) of the reference's FUSED_BD kernel
// (loop_kernel.cuh:674-783): it reads u only on the patch skeleton and b
// only on the interior, and writes the interior.  A_P is the Kronecker sum
// of the 2-cell patch 1D matrices (block-sparse rows), evaluated by
// sum factorisation z -> y -> x with only interior output rows kept, and
//   A_II^{-1} = (S (x) S (x) S) diag(1 / (h^(d-2) (l_i + l_j + l_k))) (S (x) S (x) S)^T
// with M-orthonormal eigenvectors S of the interior pencil (fast
// diagonalisation, evaluate_kernel.cuh:1327-1434).
//
// Additive pass: u_I += w A_II^{-1} r_I for a given global residual r.
//
// Layout: NPB same-colour patches per work-group, stages separated by
// barriers, every stage a loop of independent 1D line tasks over LDS.
#include "gtps_internal.h"

namespace gtps
{
  namespace
  {
    constexpr int SM_NT = 256;

    constexpr int
    cmax(int a, int b)
    {
      return a > b ? a : b;
    }
    constexpr int
    cmin(int a, int b)
    {
      return a < b ? a : b;
    }

    template <int K, typename T>
    struct SM3
    {
      static constexpr int NP  = 2 * K + 1, NI = 2 * K - 1;
      static constexpr int R1  = cmax(NP * NP * NP, 2 * NI * NI * NP);
      static constexpr int R2  = cmax(2 * NI * NP * NP, NI * NI * NI);
      static constexpr int SB  = NI * NI * NI;
      static constexpr int PER = R1 + R2 + SB;
      static constexpr int NPB_T = cmax(1, cmin(64, SM_NT / (NI * NI)));
      // keep the dynamic LDS at or below 64 KiB when more than one patch
      static constexpr int NPB_L = cmax(1, (int)(65536 / (PER * sizeof(T))));
      static constexpr int NPB   = cmin(NPB_T, NPB_L);
      static constexpr size_t LDS = (size_t)NPB * PER * sizeof(T);
    };

    template <int K, typename T>
    struct SM2
    {
      static constexpr int NP  = 2 * K + 1, NI = 2 * K - 1;
      static constexpr int R1  = NP * NP;
      static constexpr int R2  = cmax(2 * NI * NP, NI * NI);
      static constexpr int SB  = NI * NI;
      static constexpr int PER = R1 + R2 + SB;
      static constexpr int NPB_T = cmax(1, cmin(64, SM_NT / NP));
      static constexpr int NPB_L = cmax(1, (int)(65536 / (PER * sizeof(T))));
      static constexpr int NPB   = cmin(NPB_T, NPB_L);
      static constexpr size_t LDS = (size_t)NPB * PER * sizeof(T);
    };

    // patches of colour c along one direction: lower-left cell indices
    // 2i + parity in [0, n-2]
    __device__ __forceinline__ int
    n_patches_1d(int n, int parity)
    {
      return (n - parity) / 2;
    }

    template <int K, typename T, bool ADD>
    __global__ void __launch_bounds__(SM_NT)
    patch3d_kernel(T *__restrict__ u, const T *__restrict__ src,
                   const T *__restrict__ tab, const int n, const int colour,
                   const T scale, const T relax)
    {
      using TB             = Tab<K>;
      using C              = SM3<K, T>;
      constexpr int NP     = C::NP, NI = C::NI, NPB = C::NPB;
      constexpr int R1     = C::R1, R2 = C::R2, SB = C::SB, NT = SM_NT;
      extern __shared__ double smem_raw[];
      T *sR1 = reinterpret_cast<T *>(smem_raw);
      T *sR2 = sR1 + NPB * R1;
      T *sSB = sR2 + NPB * R2;
      __shared__ long long sfirst[NPB];

      const int       N1  = K * n + 1;
      const long long P   = (long long)N1 * N1;
      const int       tid = threadIdx.x;
      const int px = colour & 1, py = (colour >> 1) & 1, pz = (colour >> 2) & 1;
      const int npx = n_patches_1d(n, px), npy = n_patches_1d(n, py),
                npz    = n_patches_1d(n, pz);
      const int npatch = npx * npy * npz;
      const int blk = xcd_logical_id(blockIdx.x, gridDim.x);
      if (tid < NPB)
        {
          const int q = blk * NPB + tid;
          if (q < npatch)
            {
              const int qx = q % npx, qy = (q / npx) % npy, qz = q / (npx * npy);
              sfirst[tid]  = (long long)K * (2 * qx + px) +
                            (long long)K * (2 * qy + py) * N1 +
                            (long long)K * (2 * qz + pz) * P;
            }
          else
            sfirst[tid] = -1;
        }
      __syncthreads();
      const T *pM = tab + TB::patchM;
      const T *pK = tab + TB::patchK;
      const T *S  = tab + TB::S;
      const T *lam = tab + TB::lam;

      if (!ADD)
        {
          // stage 0: skeleton of u (interior zeroed) and interior of b; all
          // loads of a thread issued before the first wait
          constexpr int IT0 = (NPB * NP * NP * NP + NT - 1) / NT;
          T             v0[IT0];
#pragma unroll
          for (int it = 0; it < IT0; ++it)
            {
              const int t  = tid + it * NT;
              const int lp = t / (NP * NP * NP), r = t - lp * (NP * NP * NP);
              const long long f = t < NPB * NP * NP * NP ? sfirst[lp] : -1;
              const int ix = r % NP, iy = (r / NP) % NP, iz = r / (NP * NP);
              const bool inner = ix > 0 && ix < NP - 1 && iy > 0 && iy < NP - 1 &&
                                 iz > 0 && iz < NP - 1;
              v0[it] = f >= 0 ? (inner ? src : u)[f + ix + (long long)iy * N1 + iz * P] : T(0);
            }
#pragma unroll
          for (int it = 0; it < IT0; ++it)
            {
              const int t = tid + it * NT;
              if (t >= NPB * NP * NP * NP)
                continue;
              const int lp = t / (NP * NP * NP), r = t - lp * (NP * NP * NP);
              if (sfirst[lp] < 0)
                continue;
              const int ix = r % NP, iy = (r / NP) % NP, iz = r / (NP * NP);
              const bool inner = ix > 0 && ix < NP - 1 && iy > 0 && iy < NP - 1 &&
                                 iz > 0 && iz < NP - 1;
              if (inner)
                {
                  sR1[lp * R1 + r] = T(0);
                  sSB[lp * SB + ((iz - 1) * NI + iy - 1) * NI + ix - 1] = v0[it];
                }
              else
                sR1[lp * R1 + r] = v0[it];
            }
          __syncthreads();
          // stage 1: z-rows (interior rows only) of M and K
          for (int t = tid; t < NPB * NP * NP; t += NT)
            {
              const int lp = t / (NP * NP), r = t - lp * (NP * NP);
              T         v[NP];
#pragma unroll
              for (int iz = 0; iz < NP; ++iz)
                v[iz] = sR1[lp * R1 + iz * NP * NP + r];
#pragma unroll
              for (int zi = 0; zi < NI; ++zi)
                {
                  const int row = zi + 1;
                  T         m = 0, kk = 0;
#pragma unroll
                  for (int c = prow_lo<K>(row); c <= prow_hi<K>(row); ++c)
                    {
                      m += pM[row * NP + c] * v[c];
                      kk += pK[row * NP + c] * v[c];
                    }
                  sR2[lp * R2 + zi * NP * NP + r]           = m;
                  sR2[lp * R2 + (NI + zi) * NP * NP + r]    = kk;
                }
            }
          __syncthreads();
          // stage 2: y-rows; C1 = My Mz u, C2 = Ky Mz u + My Kz u
          for (int t = tid; t < NPB * NI * NP; t += NT)
            {
              const int lp = t / (NI * NP), r = t - lp * (NI * NP);
              const int zi = r / NP, ix = r - zi * NP;
              T         a[NP], b[NP];
#pragma unroll
              for (int iy = 0; iy < NP; ++iy)
                {
                  a[iy] = sR2[lp * R2 + (zi * NP + iy) * NP + ix];
                  b[iy] = sR2[lp * R2 + ((NI + zi) * NP + iy) * NP + ix];
                }
#pragma unroll
              for (int yi = 0; yi < NI; ++yi)
                {
                  const int row = yi + 1;
                  T         c1 = 0, c2 = 0;
#pragma unroll
                  for (int c = prow_lo<K>(row); c <= prow_hi<K>(row); ++c)
                    {
                      c1 += pM[row * NP + c] * a[c];
                      c2 += pK[row * NP + c] * a[c];
                      c2 += pM[row * NP + c] * b[c];
                    }
                  sR1[lp * R1 + (zi * NI + yi) * NP + ix]            = c1;
                  sR1[lp * R1 + NI * NI * NP + (zi * NI + yi) * NP + ix] = c2;
                }
            }
          __syncthreads();
          // stage 3: x-rows, r_I = b_I - h (Kx C1 + Mx C2), then S^T along x
          for (int t = tid; t < NPB * NI * NI; t += NT)
            {
              const int lp = t / (NI * NI), r = t - lp * (NI * NI);
              T         c1[NP], c2[NP], res[NI];
#pragma unroll
              for (int ix = 0; ix < NP; ++ix)
                {
                  c1[ix] = sR1[lp * R1 + r * NP + ix];
                  c2[ix] = sR1[lp * R1 + NI * NI * NP + r * NP + ix];
                }
#pragma unroll
              for (int xi = 0; xi < NI; ++xi)
                {
                  const int row = xi + 1;
                  T         v   = 0;
#pragma unroll
                  for (int c = prow_lo<K>(row); c <= prow_hi<K>(row); ++c)
                    {
                      v += pK[row * NP + c] * c1[c];
                      v += pM[row * NP + c] * c2[c];
                    }
                  res[xi] = sSB[lp * SB + r * NI + xi] - scale * v;
                }
#pragma unroll
              for (int j = 0; j < NI; ++j)
                {
                  T s = 0;
#pragma unroll
                  for (int i = 0; i < NI; ++i)
                    s += S[i * NI + j] * res[i];
                  sR2[lp * R2 + r * NI + j] = s;
                }
            }
        }
      else
        {
          // additive: interior residual, S^T along x
          for (int t = tid; t < NPB * NI * NI; t += NT)
            {
              const int       lp = t / (NI * NI), r = t - lp * (NI * NI);
              const long long f  = sfirst[lp];
              if (f < 0)
                continue;
              const int       zi = r / NI, yi = r - zi * NI;
              const long long g0 = f + 1 + (long long)(yi + 1) * N1 + (zi + 1) * P;
              T               res[NI];
#pragma unroll
              for (int xi = 0; xi < NI; ++xi)
                res[xi] = src[g0 + xi];
#pragma unroll
              for (int j = 0; j < NI; ++j)
                {
                  T s = 0;
#pragma unroll
                  for (int i = 0; i < NI; ++i)
                    s += S[i * NI + j] * res[i];
                  sR2[lp * R2 + r * NI + j] = s;
                }
            }
        }
      __syncthreads();
      // stage 4: S^T along y
      for (int t = tid; t < NPB * NI * NI; t += NT)
        {
          const int lp = t / (NI * NI), r = t - lp * (NI * NI);
          const int zi = r / NI, xi = r - zi * NI;
          T        *base = sR2 + lp * R2 + zi * NI * NI + xi;
          T         v[NI];
#pragma unroll
          for (int i = 0; i < NI; ++i)
            v[i] = base[i * NI];
#pragma unroll
          for (int j = 0; j < NI; ++j)
            {
              T s = 0;
#pragma unroll
              for (int i = 0; i < NI; ++i)
                s += S[i * NI + j] * v[i];
              base[j * NI] = s;
            }
        }
      __syncthreads();
      // stage 5: S^T along z, divide by the eigenvalue sums, S along z
      for (int t = tid; t < NPB * NI * NI; t += NT)
        {
          const int lp = t / (NI * NI), r = t - lp * (NI * NI);
          const int yi = r / NI, xi = r - yi * NI;
          T        *base = sR2 + lp * R2 + yi * NI + xi;
          const T   lxy  = lam[xi] + lam[yi];
          T         v[NI], tt[NI];
#pragma unroll
          for (int i = 0; i < NI; ++i)
            v[i] = base[i * NI * NI];
#pragma unroll
          for (int j = 0; j < NI; ++j)
            {
              T s = 0;
#pragma unroll
              for (int i = 0; i < NI; ++i)
                s += S[i * NI + j] * v[i];
              tt[j] = s / (scale * (lxy + lam[j]));
            }
#pragma unroll
          for (int i = 0; i < NI; ++i)
            {
              T s = 0;
#pragma unroll
              for (int j = 0; j < NI; ++j)
                s += S[i * NI + j] * tt[j];
              base[i * NI * NI] = s;
            }
        }
      __syncthreads();
      // stage 6: S along y
      for (int t = tid; t < NPB * NI * NI; t += NT)
        {
          const int lp = t / (NI * NI), r = t - lp * (NI * NI);
          const int zi = r / NI, xi = r - zi * NI;
          T        *base = sR2 + lp * R2 + zi * NI * NI + xi;
          T         v[NI];
#pragma unroll
          for (int j = 0; j < NI; ++j)
            v[j] = base[j * NI];
#pragma unroll
          for (int i = 0; i < NI; ++i)
            {
              T s = 0;
#pragma unroll
              for (int j = 0; j < NI; ++j)
                s += S[i * NI + j] * v[j];
              base[i * NI] = s;
            }
        }
      __syncthreads();
      // stage 7: S along x and write the interior
      for (int t = tid; t < NPB * NI * NI; t += NT)
        {
          const int       lp = t / (NI * NI), r = t - lp * (NI * NI);
          const long long f  = sfirst[lp];
          if (f < 0)
            continue;
          const int zi = r / NI, yi = r - zi * NI;
          T         v[NI];
#pragma unroll
          for (int j = 0; j < NI; ++j)
            v[j] = sR2[lp * R2 + r * NI + j];
          const long long g0 = f + 1 + (long long)(yi + 1) * N1 + (zi + 1) * P;
#pragma unroll
          for (int i = 0; i < NI; ++i)
            {
              T s = 0;
#pragma unroll
              for (int j = 0; j < NI; ++j)
                s += S[i * NI + j] * v[j];
              if (ADD)
                u[g0 + i] += relax * s;
              else
                u[g0 + i] = relax * s;
            }
        }
    }

    template <int K, typename T, bool ADD>
    __global__ void __launch_bounds__(SM_NT)
    patch2d_kernel(T *__restrict__ u, const T *__restrict__ src,
                   const T *__restrict__ tab, const int n, const int colour,
                   const T scale, const T relax)
    {
      using TB         = Tab<K>;
      using C          = SM2<K, T>;
      constexpr int NP = C::NP, NI = C::NI, NPB = C::NPB;
      constexpr int R1 = C::R1, R2 = C::R2, SB = C::SB, NT = SM_NT;
      extern __shared__ double smem_raw[];
      T *sR1 = reinterpret_cast<T *>(smem_raw);
      T *sR2 = sR1 + NPB * R1;
      T *sSB = sR2 + NPB * R2;
      __shared__ long long sfirst[NPB];
      const int N1  = K * n + 1;
      const int tid = threadIdx.x;
      const int px = colour & 1, py = (colour >> 1) & 1;
      const int npx = n_patches_1d(n, px), npy = n_patches_1d(n, py);
      const int npatch = npx * npy;
      const int blk = xcd_logical_id(blockIdx.x, gridDim.x);
      if (tid < NPB)
        {
          const int q = blk * NPB + tid;
          if (q < npatch)
            {
              const int qx = q % npx, qy = q / npx;
              sfirst[tid]  = (long long)K * (2 * qx + px) +
                            (long long)K * (2 * qy + py) * N1;
            }
          else
            sfirst[tid] = -1;
        }
      __syncthreads();
      const T *pM  = tab + TB::patchM;
      const T *pK  = tab + TB::patchK;
      const T *S   = tab + TB::S;
      const T *lam = tab + TB::lam;
      if (!ADD)
        {
          for (int t = tid; t < NPB * NP * NP; t += NT)
            {
              const int       lp = t / (NP * NP), r = t - lp * (NP * NP);
              const long long f  = sfirst[lp];
              if (f < 0)
                continue;
              const int ix = r % NP, iy = r / NP;
              const long long g = f + ix + (long long)iy * N1;
              const bool inner = ix > 0 && ix < NP - 1 && iy > 0 && iy < NP - 1;
              if (inner)
                {
                  sR1[lp * R1 + r] = T(0);
                  sSB[lp * SB + (iy - 1) * NI + ix - 1] = src[g];
                }
              else
                sR1[lp * R1 + r] = u[g];
            }
          __syncthreads();
          // y-rows of M (for Kx My) and K (for Mx Ky)
          for (int t = tid; t < NPB * NP; t += NT)
            {
              const int lp = t / NP, ix = t - lp * NP;
              T         v[NP];
#pragma unroll
              for (int iy = 0; iy < NP; ++iy)
                v[iy] = sR1[lp * R1 + iy * NP + ix];
#pragma unroll
              for (int yi = 0; yi < NI; ++yi)
                {
                  const int row = yi + 1;
                  T         m = 0, kk = 0;
#pragma unroll
                  for (int c = prow_lo<K>(row); c <= prow_hi<K>(row); ++c)
                    {
                      m += pM[row * NP + c] * v[c];
                      kk += pK[row * NP + c] * v[c];
                    }
                  sR2[lp * R2 + yi * NP + ix]        = m;
                  sR2[lp * R2 + (NI + yi) * NP + ix] = kk;
                }
            }
          __syncthreads();
          for (int t = tid; t < NPB * NI; t += NT)
            {
              const int lp = t / NI, yi = t - lp * NI;
              T         c1[NP], c2[NP], res[NI];
#pragma unroll
              for (int ix = 0; ix < NP; ++ix)
                {
                  c1[ix] = sR2[lp * R2 + yi * NP + ix];
                  c2[ix] = sR2[lp * R2 + (NI + yi) * NP + ix];
                }
#pragma unroll
              for (int xi = 0; xi < NI; ++xi)
                {
                  const int row = xi + 1;
                  T         v   = 0;
#pragma unroll
                  for (int c = prow_lo<K>(row); c <= prow_hi<K>(row); ++c)
                    {
                      v += pK[row * NP + c] * c1[c];
                      v += pM[row * NP + c] * c2[c];
                    }
                  res[xi] = sSB[lp * SB + yi * NI + xi] - scale * v;
                }
#pragma unroll
              for (int j = 0; j < NI; ++j)
                {
                  T s = 0;
#pragma unroll
                  for (int i = 0; i < NI; ++i)
                    s += S[i * NI + j] * res[i];
                  sR1[lp * R1 + yi * NI + j] = s;
                }
            }
        }
      else
        {
          for (int t = tid; t < NPB * NI; t += NT)
            {
              const int       lp = t / NI, yi = t - lp * NI;
              const long long f  = sfirst[lp];
              if (f < 0)
                continue;
              const long long g0 = f + 1 + (long long)(yi + 1) * N1;
              T               res[NI];
#pragma unroll
              for (int xi = 0; xi < NI; ++xi)
                res[xi] = src[g0 + xi];
#pragma unroll
              for (int j = 0; j < NI; ++j)
                {
                  T s = 0;
#pragma unroll
                  for (int i = 0; i < NI; ++i)
                    s += S[i * NI + j] * res[i];
                  sR1[lp * R1 + yi * NI + j] = s;
                }
            }
        }
      __syncthreads();
      // y: S^T, divide, S
      for (int t = tid; t < NPB * NI; t += NT)
        {
          const int lp = t / NI, xi = t - lp * NI;
          T        *base = sR1 + lp * R1 + xi;
          T         v[NI], tt[NI];
#pragma unroll
          for (int i = 0; i < NI; ++i)
            v[i] = base[i * NI];
#pragma unroll
          for (int j = 0; j < NI; ++j)
            {
              T s = 0;
#pragma unroll
              for (int i = 0; i < NI; ++i)
                s += S[i * NI + j] * v[i];
              tt[j] = s / (scale * (lam[xi] + lam[j]));
            }
#pragma unroll
          for (int i = 0; i < NI; ++i)
            {
              T s = 0;
#pragma unroll
              for (int j = 0; j < NI; ++j)
                s += S[i * NI + j] * tt[j];
              base[i * NI] = s;
            }
        }
      __syncthreads();
      for (int t = tid; t < NPB * NI; t += NT)
        {
          const int       lp = t / NI, yi = t - lp * NI;
          const long long f  = sfirst[lp];
          if (f < 0)
            continue;
          T v[NI];
#pragma unroll
          for (int j = 0; j < NI; ++j)
            v[j] = sR1[lp * R1 + yi * NI + j];
          const long long g0 = f + 1 + (long long)(yi + 1) * N1;
#pragma unroll
          for (int i = 0; i < NI; ++i)
            {
              T s = 0;
#pragma unroll
              for (int j = 0; j < NI; ++j)
                s += S[i * NI + j] * v[j];
              if (ADD)
                u[g0 + i] += relax * s;
              else
                u[g0 + i] = relax * s;
            }
        }
    }

    // ---------------------------------------------------------------------
    // Wave kernel (3D, small k): one wavefront per work-group owns NPW
    // x-adjacent patches of one colour.  Phase A: lane (patch, x) streams its
    // (y,z) plane of the skeleton from HBM (lanes of a wave read contiguous
    // x-runs) and applies the z- and y-rows of the patch matrices in
    // registers, leaving C1 = My Mz u and C2 = Ky Mz u + My Kz u on the
    // interior (y,z) rows.  Phase B: lane (patch, interior (z,y) line)
    // applies the x-rows, forms r = b - h^(d-2)(Kx C1 + Mx C2) and S^T along
    // x.  Phase C: lane (patch, x eigen-index) holds its (z,y) plane: S^T y,
    // S^T z, 1/(h^(d-2) sum lambda), S z, S y.  Phase D: S along x and the
    // interior store.  Phases exchange through LDS (C1/C2 rows, later the
    // transformed interior in place).
    //   MODE 0: multiplicative (src = b)
    //   MODE 1: additive, u_I += w A_II^{-1} r_I (src = global residual r)
    //   MODE 2: multiplicative on u = 0 (src = b): u is not read; the patch
    //           skeleton is written with zeros, so a colour whose boxes cover
    //           the mesh (colour 0) replaces "u = 0; first colour".
    template <int K>
    struct SW3
    {
      static constexpr int NP  = 2 * K + 1, NI = 2 * K - 1;
      static constexpr int NPW = 64 / NP;
      // per-patch LDS stride: C1 and C2 rows (2 NI^2 rows of NP), padded so
      // that phase-A lanes (patch, x) hit distinct banks
      static constexpr int RAW = 2 * NI * NI * NP;
      static constexpr int PER = RAW + ((NP - RAW % 32) + 32) % 32;
    };

    template <int K, typename T>
    constexpr bool
    use_wave_kernel()
    {
      // register budget: the phase-A accumulators (2 NI^2) and the plane
      // column must stay well below 128 VGPRs
      return K <= 2 || (K == 3 && sizeof(T) == 4);
    }

    // at most 93 VGPRs (5 waves per SIMD) for the Q3 multiplicative form,
    // without spilling
    template <int K, typename T, int MODE>
    __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5)))
    patch3d_wave_kernel(T *__restrict__ u, const T *__restrict__ src,
                        const T *__restrict__ tab, const int n,
                        const int colour, const T scale, const T inv_scale,
                        const T relax)
    {
      using TB          = Tab<K>;
      using C           = SW3<K>;
      constexpr int NP  = C::NP, NI = C::NI, NPW = C::NPW, PER = C::PER;
      extern __shared__ double smem_raw[];
      T *sC = reinterpret_cast<T *>(smem_raw);
      __shared__ long long sfirst[NPW];

      const int       N1   = K * n + 1;
      const long long P    = (long long)N1 * N1;
      const int       lane = threadIdx.x;
      const int px = colour & 1, py = (colour >> 1) & 1, pz = (colour >> 2) & 1;
      const int npx = n_patches_1d(n, px), npy = n_patches_1d(n, py),
                npz = n_patches_1d(n, pz);
      const int blk = xcd_logical_id(blockIdx.x, gridDim.x);
      if (lane < NPW)
        {
          const int q = blk * NPW + lane;
          if (q < npx * npy * npz)
            {
              const int qx = q % npx, qy = (q / npx) % npy, qz = q / (npx * npy);
              sfirst[lane] = (long long)K * (2 * qx + px) +
                             (long long)K * (2 * qy + py) * N1 +
                             (long long)K * (2 * qz + pz) * P;
            }
          else
            sfirst[lane] = -1;
        }
      __syncthreads();
      const T *pM   = tab + TB::patchM;
      const T *pK   = tab + TB::patchK;
      const T *S    = tab + TB::S;
      const T *linv = tab + TB::lamInv3;

      if (MODE == 2)
        {
          // zero the skeleton of every patch box
          for (int t = lane; t < NPW * NP * NP * NP; t += 64)
            {
              const int       lp = t % NPW, r = t / NPW;
              const long long f  = sfirst[lp];
              const int ix = r % NP, iy = (r / NP) % NP, iz = r / (NP * NP);
              const bool inner = ix > 0 && ix < NP - 1 && iy > 0 && iy < NP - 1 &&
                                 iz > 0 && iz < NP - 1;
              if (f >= 0 && !inner)
                u[f + ix + (long long)iy * N1 + iz * P] = T(0);
            }
        }
      if (MODE == 0)
        {
          // phase A
          const int       lp = lane / NP, ix = lane - lp * NP;
          const long long f  = lp < NPW ? sfirst[lp] : -1;
          if (f >= 0)
            {
              const bool xface = ix == 0 || ix == NP - 1;
              T          c1[NI][NI], c2[NI][NI];
#pragma unroll
              for (int a = 0; a < NI; ++a)
#pragma unroll
                for (int b = 0; b < NI; ++b)
                  {
                    c1[a][b] = 0;
                    c2[a][b] = 0;
                  }
#pragma unroll
              for (int iy = 0; iy < NP; ++iy)
                {
                  const bool yface = iy == 0 || iy == NP - 1;
                  T          v[NP];
#pragma unroll
                  for (int iz = 0; iz < NP; ++iz)
                    {
                      const bool      zface = iz == 0 || iz == NP - 1;
                      const long long g = f + ix + (long long)iy * N1 + iz * P;
                      if (yface || zface)
                        v[iz] = u[g];
                      else
                        v[iz] = xface ? u[g] : T(0);
                    }
                  T mz[NI], kz[NI];
#pragma unroll
                  for (int zi = 0; zi < NI; ++zi)
                    {
                      const int row = zi + 1;
                      T         m = 0, kk = 0;
#pragma unroll
                      for (int c = prow_lo<K>(row); c <= prow_hi<K>(row); ++c)
                        {
                          m += pM[row * NP + c] * v[c];
                          kk += pK[row * NP + c] * v[c];
                        }
                      mz[zi] = m;
                      kz[zi] = kk;
                    }
#pragma unroll
                  for (int yi = 0; yi < NI; ++yi)
                    {
                      const int row = yi + 1;
                      if (iy < prow_lo<K>(row) || iy > prow_hi<K>(row))
                        continue;
                      const T my = pM[row * NP + iy], ky = pK[row * NP + iy];
#pragma unroll
                      for (int zi = 0; zi < NI; ++zi)
                        {
                          c1[zi][yi] += my * mz[zi];
                          c2[zi][yi] += ky * mz[zi] + my * kz[zi];
                        }
                    }
                }
              T *base = sC + lp * PER + ix;
#pragma unroll
              for (int zi = 0; zi < NI; ++zi)
#pragma unroll
                for (int yi = 0; yi < NI; ++yi)
                  {
                    base[(zi * NI + yi) * NP]             = c1[zi][yi];
                    base[((NI + zi) * NI + yi) * NP]      = c2[zi][yi];
                  }
            }
          __syncthreads();
        }
      // phase B: residual on interior x-lines, S^T along x (in place over
      // the C1 row of the line); the fixed trip count is unrolled so that a
      // lane's lines overlap their loads and arithmetic
#pragma unroll
      for (int t = lane; t < NPW * NI * NI; t += 64)
        {
          const int       lp = t % NPW, r = t / NPW;
          const long long f  = sfirst[lp];
          if (f < 0)
            continue;
          const int       zi = r / NI, yi = r - zi * NI;
          const long long g0 = f + 1 + (long long)(yi + 1) * N1 + (zi + 1) * P;
          T               res[NI];
          if (MODE == 0)
            {
              T c1[NP], c2[NP];
#pragma unroll
              for (int ix = 0; ix < NP; ++ix)
                {
                  c1[ix] = sC[lp * PER + r * NP + ix];
                  c2[ix] = sC[lp * PER + (NI * NI + r) * NP + ix];
                }
#pragma unroll
              for (int xi = 0; xi < NI; ++xi)
                {
                  const int row = xi + 1;
                  T         v   = 0;
#pragma unroll
                  for (int c = prow_lo<K>(row); c <= prow_hi<K>(row); ++c)
                    {
                      v += pK[row * NP + c] * c1[c];
                      v += pM[row * NP + c] * c2[c];
                    }
                  res[xi] = src[g0 + xi] - scale * v;
                }
            }
          else
            {
#pragma unroll
              for (int xi = 0; xi < NI; ++xi)
                res[xi] = src[g0 + xi];
            }
#pragma unroll
          for (int j = 0; j < NI; ++j)
            {
              T s = 0;
#pragma unroll
              for (int i = 0; i < NI; ++i)
                s += S[i * NI + j] * res[i];
              sC[lp * PER + r * NP + j] = s;
            }
        }
      __syncthreads();
      // phase C: per x eigen-index, the (z,y) plane through y and z
      for (int t = lane; t < NPW * NI; t += 64)
        {
          const int lp = t % NPW, xi = t / NPW;
          if (sfirst[lp] < 0)
            continue;
          T *base = sC + lp * PER + xi;
          T  V[NI][NI], W[NI][NI];
#pragma unroll
          for (int zi = 0; zi < NI; ++zi)
#pragma unroll
            for (int yi = 0; yi < NI; ++yi)
              V[zi][yi] = base[(zi * NI + yi) * NP];
          // S^T along y
#pragma unroll
          for (int zi = 0; zi < NI; ++zi)
#pragma unroll
            for (int j = 0; j < NI; ++j)
              {
                T s = 0;
#pragma unroll
                for (int i = 0; i < NI; ++i)
                  s += S[i * NI + j] * V[zi][i];
                W[zi][j] = s;
              }
          // S^T along z, eigenvalue scaling
#pragma unroll
          for (int jy = 0; jy < NI; ++jy)
#pragma unroll
            for (int jz = 0; jz < NI; ++jz)
              {
                T s = 0;
#pragma unroll
                for (int i = 0; i < NI; ++i)
                  s += S[i * NI + jz] * W[i][jy];
                V[jz][jy] = s * (linv[(jz * NI + jy) * NI + xi] * inv_scale);
              }
          // S along z
#pragma unroll
          for (int jy = 0; jy < NI; ++jy)
#pragma unroll
            for (int iz = 0; iz < NI; ++iz)
              {
                T s = 0;
#pragma unroll
                for (int j = 0; j < NI; ++j)
                  s += S[iz * NI + j] * V[j][jy];
                W[iz][jy] = s;
              }
          // S along y
#pragma unroll
          for (int iz = 0; iz < NI; ++iz)
#pragma unroll
            for (int iy = 0; iy < NI; ++iy)
              {
                T s = 0;
#pragma unroll
                for (int j = 0; j < NI; ++j)
                  s += S[iy * NI + j] * W[iz][j];
                base[(iz * NI + iy) * NP] = s;
              }
        }
      __syncthreads();
      // phase D: S along x, store the interior
#pragma unroll
      for (int t = lane; t < NPW * NI * NI; t += 64)
        {
          const int       lp = t % NPW, r = t / NPW;
          const long long f  = sfirst[lp];
          if (f < 0)
            continue;
          const int zi = r / NI, yi = r - zi * NI;
          T         v[NI];
#pragma unroll
          for (int j = 0; j < NI; ++j)
            v[j] = sC[lp * PER + r * NP + j];
          const long long g0 = f + 1 + (long long)(yi + 1) * N1 + (zi + 1) * P;
#pragma unroll
          for (int i = 0; i < NI; ++i)
            {
              T s = 0;
#pragma unroll
              for (int j = 0; j < NI; ++j)
                s += S[i * NI + j] * v[j];
              if (MODE == 1)
                u[g0 + i] += relax * s;
              else
                u[g0 + i] = relax * s;
            }
        }
    }

    // ---------------------------------------------------------------------
    // Patch-group kernel: one wavefront per work-group handles NPW patches
    // (NPW = 1 for small colours on coarse levels, where latency dominates;
    // 2-4 on the finest level, where packing the 25-line FD stages of
    // several patches into one wave cuts the issued VALU instructions).
    // Lane tasks per stage: (patch, z-pencil) loads the skeleton column and
    // applies the z-rows; (patch, z-row, x) applies the y-rows; (patch,
    // interior (z,y) line) forms the residual and S^T along x; then S^T y,
    // S^T z + 1/(h^(d-2) sum lambda) + S z, S y and S x with the store.
    // The interior b-rows are loaded at the start so that every global load
    // is in flight at once.  Same MODE semantics as the wave kernel.
    // Needs (2k+1)^2 <= 64.
    // Fast-diagonalisation transforms in the even/odd eigenbasis (tables
    // Tab::SE, Tab::SO): t = S^T v with t numbered even vectors first, and
    // its inverse v = S t.  The reflection symmetry of the eigenvectors
    // halves the products: K(K) + (K-1)^2 multiply-adds and 2(K-1)
    // additions instead of (2K-1)^2.
    template <int K, typename T>
    __device__ __forceinline__ void
    fd_fwd(const T *SE, const T *SO, const T (&v)[2 * K - 1], T (&t)[2 * K - 1])
    {
      constexpr int NI = 2 * K - 1, MH = K - 1, NE = K, NO = K - 1;
      T e[MH + 1], o[MH > 0 ? MH : 1];
#pragma unroll
      for (int i = 0; i < MH; ++i)
        {
          e[i] = v[i] + v[NI - 1 - i];
          o[i] = v[i] - v[NI - 1 - i];
        }
      e[MH] = v[MH];
#pragma unroll
      for (int q = 0; q < NE; ++q)
        {
          T s = 0;
#pragma unroll
          for (int i = 0; i <= MH; ++i)
            s += SE[i * NE + q] * e[i];
          t[q] = s;
        }
#pragma unroll
      for (int q = 0; q < NO; ++q)
        {
          T s = 0;
#pragma unroll
          for (int i = 0; i < MH; ++i)
            s += SO[i * NO + q] * o[i];
          t[NE + q] = s;
        }
    }

    template <int K, typename T>
    __device__ __forceinline__ void
    fd_inv(const T *SE, const T *SO, const T (&t)[2 * K - 1], T (&v)[2 * K - 1])
    {
      constexpr int NI = 2 * K - 1, MH = K - 1, NE = K, NO = K - 1;
#pragma unroll
      for (int i = 0; i < MH; ++i)
        {
          T pe = 0, po = 0;
#pragma unroll
          for (int q = 0; q < NE; ++q)
            pe += SE[i * NE + q] * t[q];
#pragma unroll
          for (int q = 0; q < NO; ++q)
            po += SO[i * NO + q] * t[NE + q];
          v[i]          = pe + po;
          v[NI - 1 - i] = pe - po;
        }
      T pm = 0;
#pragma unroll
      for (int q = 0; q < NE; ++q)
        pm += SE[MH * NE + q] * t[q];
      v[MH] = pm;
    }

    template <int K, typename T>
    struct SG
    {
      static constexpr int NP = 2 * K + 1, NI = 2 * K - 1;
      // z-rows of M and K: 2 NI slabs of NP^2, ZST apart.  Stage 2 reads
      // slab offsets zi ZST + ix across the lanes: k = 3 in FP32 pads the
      // slab from 49 to 57 words (25 mod 32), which spreads the 35 (zi, ix)
      // lanes of a patch over distinct banks but for 3 (49 = 17 mod 32
      // gives 2- and 3-way conflicts); FP64 keeps 49 (its occupancy is set
      // by LDS).
      static constexpr int ZST = NP * NP + (K == 3 && sizeof(T) == 4 ? 8 : 0);
      static constexpr int ZS  = 2 * NI * ZST;
      static constexpr int YS  = 2 * NI * NI * NP;  // y-rows (C1, C2)
      // The y-rows live at the end of the z-row region (stage 2 reads all
      // of its z-row inputs before its first store) and the FD work buffer
      // (NI^3) at its start, clear of the y-rows that stage 3 reads: one
      // region per patch, which keeps LDS from limiting the wavefronts per
      // SIMD.
      static constexpr int YOFF = ZS - YS;
      static constexpr int PER  = ZS | 1;
      static_assert(NI * NI * NI <= YOFF, "FD buffer overlaps the y-rows");
    };

    // Stage boundary of the patch-group kernel.  Every wavefront owns its
    // own LDS region, so only its own lanes' LDS stores have to be seen: a
    // wavefront's LDS operations are performed in issue order, and the
    // fences keep the compiler from moving them across the boundary.
    template <int W>
    __device__ __forceinline__ void
    group_sync()
    {
      if constexpr (W == 1)
        __syncthreads();
      else
        {
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }

    // W wavefronts per work-group, each with its own NPW patches and LDS
    // region.
    template <int K, typename T, int MODE, int NPW, int W>
    __global__ void __launch_bounds__(64 * W)
    patch3d_group_kernel(T *__restrict__ u, const T *__restrict__ src,
                         const T *__restrict__ tab, const int n,
                         const int colour, const T scale, const T inv_scale,
                         const T relax, const FastDiv div_x, const FastDiv div_xy)
    {
      using TB         = Tab<K>;
      using C          = SG<K, T>;
      constexpr int NP = C::NP, NI = C::NI, PER = C::PER, ZST = C::ZST;
      static_assert(NP * NP <= 64, "one lane per z-pencil");
      constexpr int P2 = (NPW * NI * NP + 63) / 64;  // passes per stage
      constexpr int P3 = (NPW * NI * NI + 63) / 64;
      __shared__ T      sm_all[W * NPW * PER];
      const int       N1   = K * n + 1;
      const int       P    = N1 * N1;  // level sizes < 2^31 (host check)
      const int       lane = W > 1 ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
      const int       wv   = W > 1 ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) : 0;
      T              *sm   = sm_all + wv * (NPW * PER);
      const int px = colour & 1, py = (colour >> 1) & 1, pz = (colour >> 2) & 1;
      const int npx = n_patches_1d(n, px), npy = n_patches_1d(n, py),
                npz = n_patches_1d(n, pz);
      const int blk = xcd_logical_id(blockIdx.x, gridDim.x) * W + wv;
      // patch origins: wave-uniform arithmetic, divisions by the patch
      // counts as multiply-high + shift (divisors set up on the host)
      int fst[NPW];
#pragma unroll
      for (int p = 0; p < NPW; ++p)
        {
          const int q  = blk * NPW + p;
          const int qz = (int)fastdiv((unsigned)q, div_xy);
          const int rq = q - qz * npx * npy;
          const int qy = (int)fastdiv((unsigned)rq, div_x);
          const int qx = rq - qy * npx;
          fst[p] = q < npx * npy * npz
                     ? K * (2 * qx + px) + K * (2 * qy + py) * N1 + K * (2 * qz + pz) * P
                     : -1;
        }
      auto sfirst_of = [&](int lp) -> int {
        int f = fst[0];
#pragma unroll
        for (int p = 1; p < NPW; ++p)
          if (lp == p)
            f = fst[p];
        return f;
      };
      const T *pM   = tab + TB::patchM;
      const T *pK   = tab + TB::patchK;
      const T *SE   = tab + TB::SE;
      const T *SO   = tab + TB::SO;
      const T *linv = tab + TB::lamInv3p;

      // interior right-hand side (or residual) rows of this lane's stage-3
      // tasks and the inverse eigenvalue sums of its stage-5 tasks (the same
      // task index r), loaded up front
      T rb[P3][NI], lv[P3][NI];
#pragma unroll
      for (int pp = 0; pp < P3; ++pp)
        {
          const int t = lane + 64 * pp;
          if (t < NPW * NI * NI)
            {
              const int       lp = t / (NI * NI), r = t - lp * (NI * NI);
              const int       f  = sfirst_of(lp);
              const int       zi = r / NI, yi = r - zi * NI;
              if (f >= 0)
                {
                  const int       g0 = f + 1 + (yi + 1) * N1 + (zi + 1) * P;
#pragma unroll
                  for (int xi = 0; xi < NI; ++xi)
                    rb[pp][xi] = src[g0 + xi];
                }
#pragma unroll
              for (int j = 0; j < NI; ++j)
                lv[pp][j] = linv[j * NI * NI + r];
            }
        }
      if (MODE == 0)
        {
          // stage 1: z-pencils of the skeleton, z-rows of M and K.  When the
          // pencils of the group need more than one pass, the 8K face
          // pencils (ix or iy on the box face, all NP values on the
          // skeleton) run first and the NI^2 interior pencils (only their
          // two end values are skeleton) after them with the two-column
          // form of the contraction.
          auto zrows = [&](int lp, int pc, const T (&v)[NP]) {
            T *sz = sm + lp * PER;
#pragma unroll
            for (int zi = 0; zi < NI; ++zi)
              {
                const int row = zi + 1;
                T         m = 0, kk = 0;
#pragma unroll
                for (int c = prow_lo<K>(row); c <= prow_hi<K>(row); ++c)
                  {
                    m += pM[row * NP + c] * v[c];
                    kk += pK[row * NP + c] * v[c];
                  }
                sz[zi * ZST + pc]        = m;
                sz[(NI + zi) * ZST + pc] = kk;
              }
          };
          if constexpr (NPW * NP * NP <= 64)
            {
              const int t = lane;
              if (t < NPW * NP * NP)
                {
                  const int       lp = t / (NP * NP), pc = t - lp * (NP * NP);
                  const int       f  = sfirst_of(lp);
                  if (f >= 0)
                    {
                      const int  iy = pc / NP, ix = pc - iy * NP;
                      const bool face = ix == 0 || ix == NP - 1 || iy == 0 || iy == NP - 1;
                      T          v[NP];
#pragma unroll
                      for (int iz = 0; iz < NP; ++iz)
                        {
                          const int       g = f + ix + iy * N1 + iz * P;
                          v[iz] = (face || iz == 0 || iz == NP - 1) ? u[g] : T(0);
                        }
                      zrows(lp, pc, v);
                    }
                }
            }
          else
            {
              constexpr int NF = NP * NP - NI * NI;
#pragma unroll
              for (int pp = 0; pp < (NPW * NF + 63) / 64; ++pp)
                {
                  const int t = lane + 64 * pp;
                  if (t >= NPW * NF)
                    continue;
                  const int       lp = t / NF, fi = t - lp * NF;
                  const int       f  = sfirst_of(lp);
                  if (f < 0)
                    continue;
                  int ix, iy;
                  if (fi < NP)
                    {
                      ix = fi;
                      iy = 0;
                    }
                  else if (fi < 2 * NP)
                    {
                      ix = fi - NP;
                      iy = NP - 1;
                    }
                  else
                    {
                      const int g2 = fi - 2 * NP;
                      iy           = 1 + (g2 >> 1);
                      ix           = (g2 & 1) ? NP - 1 : 0;
                    }
                  T v[NP];
#pragma unroll
                  for (int iz = 0; iz < NP; ++iz)
                    v[iz] = u[f + ix + iy * N1 + iz * P];
                  zrows(lp, iy * NP + ix, v);
                }
#pragma unroll
              for (int pp = 0; pp < (NPW * NI * NI + 63) / 64; ++pp)
                {
                  const int t = lane + 64 * pp;
                  if (t >= NPW * NI * NI)
                    continue;
                  const int       lp = t / (NI * NI), r = t - lp * (NI * NI);
                  const int       f  = sfirst_of(lp);
                  if (f < 0)
                    continue;
                  const int       iy = 1 + r / NI, ix = 1 + r % NI;
                  const int       g  = f + ix + iy * N1;
                  const T         v0 = u[g], v1 = u[g + (NP - 1) * P];
                  T              *sz = sm + lp * PER;
                  const int       pc = iy * NP + ix;
#pragma unroll
                  for (int zi = 0; zi < NI; ++zi)
                    {
                      const int row = zi + 1;
                      T         m = 0, kk = 0;
                      if (prow_lo<K>(row) == 0)
                        {
                          m += pM[row * NP] * v0;
                          kk += pK[row * NP] * v0;
                        }
                      if (prow_hi<K>(row) == NP - 1)
                        {
                          m += pM[row * NP + NP - 1] * v1;
                          kk += pK[row * NP + NP - 1] * v1;
                        }
                      sz[zi * ZST + pc]        = m;
                      sz[(NI + zi) * ZST + pc] = kk;
                    }
                }
            }
          group_sync<W>();
          // stage 2: y-rows, C1 = My Mz u, C2 = Ky Mz u + My Kz u.  Every
          // pass reads its z-rows before the first y-row is stored (the
          // y-rows overwrite the end of the z-row region).
          T a2[P2][NP], b2[P2][NP];
#pragma unroll
          for (int pp = 0; pp < P2; ++pp)
            {
              const int t = lane + 64 * pp;
              if (t >= NPW * NI * NP)
                continue;
              const int lp = t / (NI * NP), r = t - lp * (NI * NP);
              const int zi = r / NP, ix = r - zi * NP;
              const T  *sz = sm + lp * PER;
#pragma unroll
              for (int iy = 0; iy < NP; ++iy)
                {
                  a2[pp][iy] = sz[zi * ZST + iy * NP + ix];
                  b2[pp][iy] = sz[(NI + zi) * ZST + iy * NP + ix];
                }
            }
          group_sync<W>();
#pragma unroll
          for (int pp = 0; pp < P2; ++pp)
            {
              const int t = lane + 64 * pp;
              if (t >= NPW * NI * NP)
                continue;
              const int lp = t / (NI * NP), r = t - lp * (NI * NP);
              if (sfirst_of(lp) < 0)
                continue;
              const int zi = r / NP, ix = r - zi * NP;
              T        *sy = sm + lp * PER + C::YOFF;
              const T (&a)[NP] = a2[pp];
              const T (&b)[NP] = b2[pp];
#pragma unroll
              for (int yi = 0; yi < NI; ++yi)
                {
                  const int row = yi + 1;
                  T         c1 = 0, c2 = 0;
#pragma unroll
                  for (int c = prow_lo<K>(row); c <= prow_hi<K>(row); ++c)
                    {
                      c1 += pM[row * NP + c] * a[c];
                      c2 += pK[row * NP + c] * a[c];
                      c2 += pM[row * NP + c] * b[c];
                    }
                  sy[(zi * NI + yi) * NP + ix]        = c1;
                  sy[((NI + zi) * NI + yi) * NP + ix] = c2;
                }
            }
          group_sync<W>();
        }
      if (MODE == 2)
        {
          for (int t = lane; t < NPW * NP * NP * NP; t += 64)
            {
              const int       lp = t / (NP * NP * NP), r = t - lp * (NP * NP * NP);
              const int       f  = sfirst_of(lp);
              const int ix = r % NP, iy = (r / NP) % NP, iz = r / (NP * NP);
              const bool inner = ix > 0 && ix < NP - 1 && iy > 0 && iy < NP - 1 &&
                                 iz > 0 && iz < NP - 1;
              if (f >= 0 && !inner)
                u[f + ix + iy * N1 + iz * P] = T(0);
            }
        }
      // stage 3: x-rows, residual, S^T along x -> FD buffer (z-row region)
#pragma unroll
      for (int pp = 0; pp < P3; ++pp)
        {
          const int t = lane + 64 * pp;
          if (t >= NPW * NI * NI)
            continue;
          const int lp = t / (NI * NI), r = t - lp * (NI * NI);
          if (sfirst_of(lp) < 0)
            continue;
          T res[NI];
          if (MODE == 0)
            {
              const T *sy = sm + lp * PER + C::YOFF;
              T        c1[NP], c2[NP];
#pragma unroll
              for (int ix = 0; ix < NP; ++ix)
                {
                  c1[ix] = sy[r * NP + ix];
                  c2[ix] = sy[(NI * NI + r) * NP + ix];
                }
#pragma unroll
              for (int xi = 0; xi < NI; ++xi)
                {
                  const int row = xi + 1;
                  T         v   = 0;
#pragma unroll
                  for (int c = prow_lo<K>(row); c <= prow_hi<K>(row); ++c)
                    {
                      v += pK[row * NP + c] * c1[c];
                      v += pM[row * NP + c] * c2[c];
                    }
                  res[xi] = rb[pp][xi] - scale * v;
                }
            }
          else
            {
#pragma unroll
              for (int xi = 0; xi < NI; ++xi)
                res[xi] = rb[pp][xi];
            }
          T *sr = sm + lp * PER;
          T  tt[NI];
          fd_fwd<K>(SE, SO, res, tt);
#pragma unroll
          for (int j = 0; j < NI; ++j)
            sr[r * NI + j] = tt[j];
        }
      group_sync<W>();
      // stage 4: S^T along y
#pragma unroll
      for (int pp = 0; pp < P3; ++pp)
        {
          const int t = lane + 64 * pp;
          if (t >= NPW * NI * NI)
            continue;
          const int lp = t / (NI * NI), r = t - lp * (NI * NI);
          if (sfirst_of(lp) < 0)
            continue;
          const int zi = r / NI, xi = r - zi * NI;
          T        *base = sm + lp * PER + zi * NI * NI + xi;
          T         v[NI], tt[NI];
#pragma unroll
          for (int i = 0; i < NI; ++i)
            v[i] = base[i * NI];
          fd_fwd<K>(SE, SO, v, tt);
#pragma unroll
          for (int j = 0; j < NI; ++j)
            base[j * NI] = tt[j];
        }
      group_sync<W>();
      // stage 5: S^T along z, eigenvalue scaling, S along z
#pragma unroll
      for (int pp = 0; pp < P3; ++pp)
        {
          const int t = lane + 64 * pp;
          if (t >= NPW * NI * NI)
            continue;
          const int lp = t / (NI * NI), r = t - lp * (NI * NI);
          if (sfirst_of(lp) < 0)
            continue;
          const int yi = r / NI, xi = r - yi * NI;
          T        *base = sm + lp * PER + yi * NI + xi;
          T         v[NI], tt[NI];
#pragma unroll
          for (int i = 0; i < NI; ++i)
            v[i] = base[i * NI * NI];
          fd_fwd<K>(SE, SO, v, tt);
#pragma unroll
          for (int j = 0; j < NI; ++j)
            tt[j] *= lv[pp][j] * inv_scale;
          fd_inv<K>(SE, SO, tt, v);
#pragma unroll
          for (int i = 0; i < NI; ++i)
            base[i * NI * NI] = v[i];
        }
      group_sync<W>();
      // stage 6: S along y
#pragma unroll
      for (int pp = 0; pp < P3; ++pp)
        {
          const int t = lane + 64 * pp;
          if (t >= NPW * NI * NI)
            continue;
          const int lp = t / (NI * NI), r = t - lp * (NI * NI);
          if (sfirst_of(lp) < 0)
            continue;
          const int zi = r / NI, xi = r - zi * NI;
          T        *base = sm + lp * PER + zi * NI * NI + xi;
          T         v[NI], o[NI];
#pragma unroll
          for (int j = 0; j < NI; ++j)
            v[j] = base[j * NI];
          fd_inv<K>(SE, SO, v, o);
#pragma unroll
          for (int i = 0; i < NI; ++i)
            base[i * NI] = o[i];
        }
      group_sync<W>();
      // stage 7: S along x, store
#pragma unroll
      for (int pp = 0; pp < P3; ++pp)
        {
          const int t = lane + 64 * pp;
          if (t >= NPW * NI * NI)
            continue;
          const int       lp = t / (NI * NI), r = t - lp * (NI * NI);
          const int       f  = sfirst_of(lp);
          if (f < 0)
            continue;
          const int zi = r / NI, yi = r - zi * NI;
          T         v[NI], o[NI];
#pragma unroll
          for (int j = 0; j < NI; ++j)
            v[j] = sm[lp * PER + r * NI + j];
          fd_inv<K>(SE, SO, v, o);
          const int       g0 = f + 1 + (yi + 1) * N1 + (zi + 1) * P;
#pragma unroll
          for (int i = 0; i < NI; ++i)
            {
              if (MODE == 1)
                u[g0 + i] += relax * o[i];
              else
                u[g0 + i] = relax * o[i];
            }
        }
    }

    // Row-staged patch-group kernel: the same stages as the patch-group
    // kernel, but the NPW patches of a work-group are x-consecutive patches
    // of one patch row, and every global access goes through contiguous
    // x-runs of their union box: the skeleton box (width 2k NPW + 1), the
    // interior right-hand side rows and the interior results are moved by
    // lanes over (row, x), staged in LDS.  A row of the union box costs one
    // or two cache-line accesses instead of one per lane-pencil.
    template <int K, int NPW>
    struct SR
    {
      static constexpr int NP = 2 * K + 1, NI = 2 * K - 1;
      static constexpr int UX = 2 * K * NPW + 1;   // union box width
      static constexpr int BX = UX - 2;            // interior x-run width
      static constexpr int ZS = 2 * NI * NP * NP;  // z-rows of M and K / FD
      static constexpr int YS = 2 * NI * NI * NP;  // y-rows
      static constexpr int US = NP * NP * UX;      // union skeleton box
      // the box is dead after stage 1 and lives in the y-row region
      static constexpr int YU = NPW * YS > US ? NPW * YS : US;
      static constexpr int BS = NI * NI * BX;      // interior b rows
    };

    template <int K, typename T, int MODE, int NPW>
    __global__ void __launch_bounds__(64)
    patch3d_row_kernel(T *__restrict__ u, const T *__restrict__ src,
                       const T *__restrict__ tab, const int n,
                       const int colour, const T scale, const T inv_scale,
                       const T relax)
    {
      using TB         = Tab<K>;
      using C          = SR<K, NPW>;
      constexpr int NP = C::NP, NI = C::NI, UX = C::UX, BX = C::BX;
      constexpr int ZS = C::ZS, YS = C::YS;
      __shared__ T sZ[NPW * ZS + 1], sYU[C::YU + 1], sB[C::BS + 1];
      const int       N1   = K * n + 1;
      const long long P    = (long long)N1 * N1;
      const int       lane = threadIdx.x;
      const int px = colour & 1, py = (colour >> 1) & 1, pz = (colour >> 2) & 1;
      const int npx = n_patches_1d(n, px), npy = n_patches_1d(n, py);
      const int nbx = (npx + NPW - 1) / NPW;
      const int blk = xcd_logical_id(blockIdx.x, gridDim.x);
      const int bx = blk % nbx, qy = (blk / nbx) % npy, qz = blk / (nbx * npy);
      const int qx0 = bx * NPW;
      const int nph = min(NPW, npx - qx0);         // patches in this group
      const int ux  = 2 * K * nph + 1, bxw = ux - 2;
      const long long f0 = (long long)K * (2 * qx0 + px) +
                           (long long)K * (2 * qy + py) * N1 +
                           (long long)K * (2 * qz + pz) * P;
      const T *pM   = tab + TB::patchM;
      const T *pK   = tab + TB::patchK;
      const T *SE   = tab + TB::SE;
      const T *SO   = tab + TB::SO;
      const T *linv = tab + TB::lamInv3p;
      // position x (0..ux-1) of the union box lies inside some patch
      // interior iff it is not a multiple of 2k
      auto x_inner = [](int x) { return x % (2 * K) != 0; };

      // phase L: interior b rows and (multiplicative) the skeleton box.
      // Lanes form RPI row segments of XS lanes; a lane keeps its x and
      // row-in-block for the whole phase, so every address is a per-lane
      // base plus a compile-time multiple of the row strides.
      constexpr int XS  = UX <= 8 ? 8 : (UX <= 16 ? 16 : (UX <= 32 ? 32 : 64));
      constexpr int RPI = 64 / XS;
      const int     lx  = lane % XS, lr = lane / XS;
      const bool    xin = x_inner(lx);
      {
        constexpr int NBR = (NI + RPI - 1) / RPI;  // y-blocks per z
        T             vb[NI][NBR];
        const bool    bok = lx >= 1 && lx - 1 < bxw;
#pragma unroll
        for (int z = 1; z <= NI; ++z)
#pragma unroll
          for (int yb = 0; yb < NBR; ++yb)
            {
              const int y = 1 + yb * RPI + lr;
              vb[z - 1][yb] = (bok && y <= NI)
                                ? src[f0 + lx + (long long)y * N1 + z * P]
                                : T(0);
            }
#pragma unroll
        for (int z = 1; z <= NI; ++z)
#pragma unroll
          for (int yb = 0; yb < NBR; ++yb)
            {
              const int y = 1 + yb * RPI + lr;
              if (bok && y <= NI)
                sB[((z - 1) * NI + (y - 1)) * BX + lx - 1] = vb[z - 1][yb];
            }
      }
      if (MODE == 0)
        {
          constexpr int NBR = (NP + RPI - 1) / RPI;
          T             vu[NP][NBR];
          const bool    uok = lx < ux;
#pragma unroll
          for (int z = 0; z < NP; ++z)
#pragma unroll
            for (int yb = 0; yb < NBR; ++yb)
              {
                const int  y     = yb * RPI + lr;
                const bool inner = z > 0 && z < NP - 1 && y > 0 && y < NP - 1 && xin;
                vu[z][yb] = (uok && y < NP && !inner)
                              ? u[f0 + lx + (long long)y * N1 + z * P]
                              : T(0);
              }
#pragma unroll
          for (int z = 0; z < NP; ++z)
#pragma unroll
            for (int yb = 0; yb < NBR; ++yb)
              {
                const int y = yb * RPI + lr;
                if (lx < UX && y < NP)
                  sYU[(z * NP + y) * UX + lx] = vu[z][yb];
              }
        }
      __syncthreads();
      if (MODE == 0)
        {
          // stage 1: z-pencils of each patch box from the union box; when
          // one pass does not hold them all, the 8k face pencils first, then
          // the interior pencils in two-value form (their only skeleton
          // values are the two ends)
          auto zrows = [&](int lp, int pc, const T (&v)[NP]) {
            T *sz = sZ + lp * ZS;
#pragma unroll
            for (int zi = 0; zi < NI; ++zi)
              {
                const int row = zi + 1;
                T         m = 0, kk = 0;
#pragma unroll
                for (int c = prow_lo<K>(row); c <= prow_hi<K>(row); ++c)
                  {
                    m += pM[row * NP + c] * v[c];
                    kk += pK[row * NP + c] * v[c];
                  }
                sz[zi * NP * NP + pc]        = m;
                sz[(NI + zi) * NP * NP + pc] = kk;
              }
          };
          if constexpr (NPW * NP * NP <= 64)
            {
              const int t = lane;
              const int lp = t / (NP * NP), pc = t - lp * (NP * NP);
              if (t < NPW * NP * NP && lp < nph)
                {
                  const int iy = pc / NP, ix = pc - iy * NP;
                  const T  *col = sYU + iy * UX + 2 * K * lp + ix;
                  T         v[NP];
#pragma unroll
                  for (int iz = 0; iz < NP; ++iz)
                    v[iz] = col[iz * NP * UX];
                  zrows(lp, pc, v);
                }
            }
          else
            {
              constexpr int NF = NP * NP - NI * NI;
#pragma unroll
              for (int pp = 0; pp < (NPW * NF + 63) / 64; ++pp)
                {
                  const int t = lane + 64 * pp;
                  const int lp = t / NF, fi = t - lp * NF;
                  if (t >= NPW * NF || lp >= nph)
                    continue;
                  int ix, iy;
                  if (fi < NP)
                    {
                      ix = fi;
                      iy = 0;
                    }
                  else if (fi < 2 * NP)
                    {
                      ix = fi - NP;
                      iy = NP - 1;
                    }
                  else
                    {
                      const int g2 = fi - 2 * NP;
                      iy           = 1 + (g2 >> 1);
                      ix           = (g2 & 1) ? NP - 1 : 0;
                    }
                  const T *col = sYU + iy * UX + 2 * K * lp + ix;
                  T        v[NP];
#pragma unroll
                  for (int iz = 0; iz < NP; ++iz)
                    v[iz] = col[iz * NP * UX];
                  zrows(lp, iy * NP + ix, v);
                }
#pragma unroll
              for (int pp = 0; pp < (NPW * NI * NI + 63) / 64; ++pp)
                {
                  const int t = lane + 64 * pp;
                  const int lp = t / (NI * NI), r = t - lp * (NI * NI);
                  if (t >= NPW * NI * NI || lp >= nph)
                    continue;
                  const int iy = 1 + r / NI, ix = 1 + r % NI;
                  const T  *col = sYU + iy * UX + 2 * K * lp + ix;
                  const T   v0 = col[0], v1 = col[(NP - 1) * NP * UX];
                  T        *sz = sZ + lp * ZS;
                  const int pc = iy * NP + ix;
#pragma unroll
                  for (int zi = 0; zi < NI; ++zi)
                    {
                      const int row = zi + 1;
                      T         m = 0, kk = 0;
                      if (prow_lo<K>(row) == 0)
                        {
                          m += pM[row * NP] * v0;
                          kk += pK[row * NP] * v0;
                        }
                      if (prow_hi<K>(row) == NP - 1)
                        {
                          m += pM[row * NP + NP - 1] * v1;
                          kk += pK[row * NP + NP - 1] * v1;
                        }
                      sz[zi * NP * NP + pc]        = m;
                      sz[(NI + zi) * NP * NP + pc] = kk;
                    }
                }
            }
          __syncthreads();
          // stage 2: y-rows (the box region is dead now)
#pragma unroll
          for (int pp = 0; pp < (NPW * NI * NP + 63) / 64; ++pp)
            {
              const int t = lane + 64 * pp;
              if (t >= NPW * NI * NP)
                continue;
              const int lp = t / (NI * NP), r = t - lp * (NI * NP);
              if (lp >= nph)
                continue;
              const int zi = r / NP, ix = r - zi * NP;
              const T  *sz = sZ + lp * ZS;
              T        *sy = sYU + lp * YS;
              T         a[NP], b[NP];
#pragma unroll
              for (int iy = 0; iy < NP; ++iy)
                {
                  a[iy] = sz[(zi * NP + iy) * NP + ix];
                  b[iy] = sz[((NI + zi) * NP + iy) * NP + ix];
                }
#pragma unroll
              for (int yi = 0; yi < NI; ++yi)
                {
                  const int row = yi + 1;
                  T         c1 = 0, c2 = 0;
#pragma unroll
                  for (int c = prow_lo<K>(row); c <= prow_hi<K>(row); ++c)
                    {
                      c1 += pM[row * NP + c] * a[c];
                      c2 += pK[row * NP + c] * a[c];
                      c2 += pM[row * NP + c] * b[c];
                    }
                  sy[(zi * NI + yi) * NP + ix]        = c1;
                  sy[((NI + zi) * NI + yi) * NP + ix] = c2;
                }
            }
          __syncthreads();
        }
      // stage 3: x-rows, residual, S^T along x -> FD buffer (z-row region)
#pragma unroll
      for (int pp = 0; pp < (NPW * NI * NI + 63) / 64; ++pp)
        {
          const int t = lane + 64 * pp;
          if (t >= NPW * NI * NI)
            continue;
          const int lp = t / (NI * NI), r = t - lp * (NI * NI);
          if (lp >= nph)
            continue;
          const T *brow = sB + r * BX + 2 * K * lp;
          T        res[NI];
          if (MODE == 0)
            {
              const T *sy = sYU + lp * YS;
              T        c1[NP], c2[NP];
#pragma unroll
              for (int ix = 0; ix < NP; ++ix)
                {
                  c1[ix] = sy[r * NP + ix];
                  c2[ix] = sy[(NI * NI + r) * NP + ix];
                }
#pragma unroll
              for (int xi = 0; xi < NI; ++xi)
                {
                  const int row = xi + 1;
                  T         v   = 0;
#pragma unroll
                  for (int c = prow_lo<K>(row); c <= prow_hi<K>(row); ++c)
                    {
                      v += pK[row * NP + c] * c1[c];
                      v += pM[row * NP + c] * c2[c];
                    }
                  res[xi] = brow[xi] - scale * v;
                }
            }
          else
            {
#pragma unroll
              for (int xi = 0; xi < NI; ++xi)
                res[xi] = brow[xi];
            }
          T tt[NI];
          fd_fwd<K>(SE, SO, res, tt);
          T *sr = sZ + lp * ZS;
#pragma unroll
          for (int j = 0; j < NI; ++j)
            sr[r * NI + j] = tt[j];
        }
      __syncthreads();
      // stage 4: S^T along y
#pragma unroll
      for (int pp = 0; pp < (NPW * NI * NI + 63) / 64; ++pp)
        {
          const int t = lane + 64 * pp;
          if (t >= NPW * NI * NI)
            continue;
          const int lp = t / (NI * NI), r = t - lp * (NI * NI);
          if (lp >= nph)
            continue;
          const int zi = r / NI, xi = r - zi * NI;
          T        *base = sZ + lp * ZS + zi * NI * NI + xi;
          T         v[NI], tt[NI];
#pragma unroll
          for (int i = 0; i < NI; ++i)
            v[i] = base[i * NI];
          fd_fwd<K>(SE, SO, v, tt);
#pragma unroll
          for (int j = 0; j < NI; ++j)
            base[j * NI] = tt[j];
        }
      __syncthreads();
      // stage 5: S^T along z, eigenvalue scaling, S along z
#pragma unroll
      for (int pp = 0; pp < (NPW * NI * NI + 63) / 64; ++pp)
        {
          const int t = lane + 64 * pp;
          if (t >= NPW * NI * NI)
            continue;
          const int lp = t / (NI * NI), r = t - lp * (NI * NI);
          if (lp >= nph)
            continue;
          const int yi = r / NI, xi = r - yi * NI;
          T        *base = sZ + lp * ZS + yi * NI + xi;
          T         v[NI], tt[NI];
#pragma unroll
          for (int i = 0; i < NI; ++i)
            v[i] = base[i * NI * NI];
          fd_fwd<K>(SE, SO, v, tt);
#pragma unroll
          for (int j = 0; j < NI; ++j)
            tt[j] *= linv[(j * NI + yi) * NI + xi] * inv_scale;
          fd_inv<K>(SE, SO, tt, v);
#pragma unroll
          for (int i = 0; i < NI; ++i)
            base[i * NI * NI] = v[i];
        }
      __syncthreads();
      // stage 6: S along y
#pragma unroll
      for (int pp = 0; pp < (NPW * NI * NI + 63) / 64; ++pp)
        {
          const int t = lane + 64 * pp;
          if (t >= NPW * NI * NI)
            continue;
          const int lp = t / (NI * NI), r = t - lp * (NI * NI);
          if (lp >= nph)
            continue;
          const int zi = r / NI, xi = r - zi * NI;
          T        *base = sZ + lp * ZS + zi * NI * NI + xi;
          T         v[NI], o[NI];
#pragma unroll
          for (int j = 0; j < NI; ++j)
            v[j] = base[j * NI];
          fd_inv<K>(SE, SO, v, o);
#pragma unroll
          for (int i = 0; i < NI; ++i)
            base[i * NI] = o[i];
        }
      __syncthreads();
      // stage 7: S along x, scaled result back into the line
#pragma unroll
      for (int pp = 0; pp < (NPW * NI * NI + 63) / 64; ++pp)
        {
          const int t = lane + 64 * pp;
          if (t >= NPW * NI * NI)
            continue;
          const int lp = t / (NI * NI), r = t - lp * (NI * NI);
          if (lp >= nph)
            continue;
          T *line = sZ + lp * ZS + r * NI;
          T  v[NI], o[NI];
#pragma unroll
          for (int j = 0; j < NI; ++j)
            v[j] = line[j];
          fd_inv<K>(SE, SO, v, o);
#pragma unroll
          for (int i = 0; i < NI; ++i)
            line[i] = relax * o[i];
        }
      __syncthreads();
      // phase S: interior results by x-runs (x multiple of 2k: not interior)
      {
        constexpr int NBR = (NI + RPI - 1) / RPI;
        const bool    sok = lx >= 1 && lx - 1 < bxw && xin;
        const int     lps = (lx - 1) / (2 * K), xis = lx - 1 - 2 * K * lps;
        T             old[MODE == 1 ? NI : 1][MODE == 1 ? NBR : 1];
        if constexpr (MODE == 1)
          {
#pragma unroll
            for (int z = 1; z <= NI; ++z)
#pragma unroll
              for (int yb = 0; yb < NBR; ++yb)
                {
                  const int y = 1 + yb * RPI + lr;
                  old[z - 1][yb] = (sok && y <= NI)
                                     ? u[f0 + lx + (long long)y * N1 + z * P]
                                     : T(0);
                }
          }
#pragma unroll
        for (int z = 1; z <= NI; ++z)
#pragma unroll
          for (int yb = 0; yb < NBR; ++yb)
            {
              const int y = 1 + yb * RPI + lr;
              if (!sok || y > NI)
                continue;
              const T val = sZ[lps * ZS + ((z - 1) * NI + (y - 1)) * NI + xis];
              const long long g = f0 + lx + (long long)y * N1 + z * P;
              if constexpr (MODE == 1)
                u[g] = old[z - 1][yb] + val;
              else
                u[g] = val;
            }
      }
      if (MODE == 2)
        {
          // the skeleton of the union box is zero (u = 0 on entry)
          for (int e = lane; e < C::US; e += 64)
            {
              const int x = e % UX, row = e / UX;
              const int y = row % NP, z = row / NP;
              const bool inner = y > 0 && y < NP - 1 && z > 0 && z < NP - 1 &&
                                 x_inner(x);
              if (x < ux && !inner)
                u[f0 + x + (long long)y * N1 + z * P] = T(0);
            }
        }
    }

    long long
    small_colour_limit()
    {
      return tuning().small_colour_patches;
    }

    template <typename KernelPtr>
    void
    ensure_lds(KernelPtr kern, size_t bytes)
    {
      if (bytes > 65536)
        GTPS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)bytes));
    }

    template <int K, typename T>
    void
    launch_k(const gtps_level *lvl, void *u, const void *src, int colour,
             int mode, double relax)
    {
      const T    *tab = reinterpret_cast<const T *>(lvl->dtab());
      const int   n   = lvl->n;
      hipStream_t s   = lvl->ctx->stream;
      const int px = colour & 1, py = (colour >> 1) & 1, pz = (colour >> 2) & 1;
      long long np = (long long)((n - px) / 2) * ((n - py) / 2);
      if (lvl->dim == 3)
        np *= (n - pz) / 2;
      if (np <= 0)
        return;
      T       *up = reinterpret_cast<T *>(u);
      const T *sp = reinterpret_cast<const T *>(src);
      if constexpr (2 * K + 1 <= 8)
        if (lvl->dim == 3 && np <= small_colour_limit() && lvl->N < (1LL << 31))
          {
            const T   scale = (T)lvl->scale, inv = (T)(1.0 / lvl->scale);
            // one patch per wave while a colour has too few patches to fill
            // the GPU several times over; packed groups beyond that
            const int npw = np <= tuning().group_min_patches ? 1 : tuning().patches_per_wave;
            // row-staged form: groups never straddle a patch row
            const int npx_ = (n - (colour & 1)) / 2, npy_ = (n - ((colour >> 1) & 1)) / 2,
                      npz_ = (n - (colour >> 2)) / 2;
            const bool rows = tuning().row_staged != 0;
            const FastDiv dx  = make_fastdiv((unsigned)std::max(npx_, 1));
            const FastDiv dxy = make_fastdiv((unsigned)std::max(npx_ * npy_, 1));
            auto grid_of = [&](int w) {
              return rows ? (unsigned)((long long)((npx_ + w - 1) / w) * npy_ * npz_)
                          : (unsigned)((np + w - 1) / w);
            };
            // patch-group kernel: wavefronts per work-group (automatic: 4 in
            // FP32, 1 in FP64)
            const int gw     = tuning().group_waves != 0 ? tuning().group_waves
                                                          : (sizeof(T) == 4 ? 4 : 1);
            auto      wgs_of = [&](int w, int waves) {
              return (unsigned)((np + (long long)w * waves - 1) / ((long long)w * waves));
            };
#define GTPS_SG(M, W)                                                         \
  do                                                                          \
    {                                                                         \
      if (rows)                                                               \
        hipLaunchKernelGGL((patch3d_row_kernel<K, T, M, W>), dim3(grid_of(W)), \
                           dim3(64), 0, s, up, sp, tab, n, colour, scale, inv,  \
                           (T)relax);                                         \
      else if (gw == 4 && W <= 2)                                             \
        hipLaunchKernelGGL((patch3d_group_kernel<K, T, M, W, W <= 2 ? 4 : 1>), \
                           dim3(wgs_of(W, 4)), dim3(256), 0, s, up, sp, tab,   \
                           n, colour, scale, inv, (T)relax, dx, dxy);          \
      else                                                                    \
        hipLaunchKernelGGL((patch3d_group_kernel<K, T, M, W, 1>),             \
                           dim3(wgs_of(W, 1)), dim3(64), 0, s, up, sp, tab, n, \
                           colour, scale, inv, (T)relax, dx, dxy);             \
    }                                                                         \
  while (0)
#define GTPS_SG_W(M)                                                          \
  do                                                                          \
    {                                                                         \
      if (npw == 2)                                                           \
        GTPS_SG(M, 2);                                                        \
      else if (npw == 3)                                                      \
        GTPS_SG(M, 3);                                                        \
      else if (npw == 4)                                                      \
        GTPS_SG(M, 4);                                                        \
      else if (npw == 5)                                                      \
        GTPS_SG(M, 5);                                                        \
      else                                                                    \
        GTPS_SG(M, 1);                                                        \
    }                                                                         \
  while (0)
            if (mode == 0)
              GTPS_SG_W(0);
            else if (mode == 1)
              GTPS_SG_W(1);
            else
              GTPS_SG_W(2);
#undef GTPS_SG_W
#undef GTPS_SG
            GTPS_LAUNCH_CHECK();
            return;
          }
      if constexpr (use_wave_kernel<K, T>())
        if (lvl->dim == 3)
        {
          using C            = SW3<K>;
          const long long nb = (np + C::NPW - 1) / C::NPW;
          const size_t    lds = sizeof(T) * C::NPW * C::PER;
          const T scale = (T)lvl->scale, inv = (T)(1.0 / lvl->scale);
#define GTPS_SW(M)                                                            \
  hipLaunchKernelGGL((patch3d_wave_kernel<K, T, M>), dim3((unsigned)nb),     \
                     dim3(64), lds, s, up, sp, tab, n, colour, scale, inv,    \
                     (T)relax)
          if (mode == 0)
            GTPS_SW(0);
          else if (mode == 1)
            GTPS_SW(1);
          else
            GTPS_SW(2);
#undef GTPS_SW
          GTPS_LAUNCH_CHECK();
          return;
        }
      if (mode == 2)
        {
          // no zero-skeleton form: clear u, then a multiplicative pass
          launch_fill(s, lvl->N, lvl->prec, u, 0.0);
          mode = 0;
        }
      if (mode == 0 && tuning().mfma_smoother != 0 &&
          mfma_smoother_supported(lvl->dim, K, lvl->prec))
        {
          launch_patch_colour_mfma(lvl, u, src, colour, relax);
          return;
        }
      if (lvl->dim == 3)
        {
          using C          = SM3<K, T>;
          const long long nb = (np + C::NPB - 1) / C::NPB;
          if (mode == 0)
            {
              auto kern = patch3d_kernel<K, T, false>;
              ensure_lds(kern, C::LDS);
              hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(SM_NT), C::LDS,
                                 s, up, sp, tab, n, colour, (T)lvl->scale,
                                 (T)relax);
            }
          else
            {
              auto kern = patch3d_kernel<K, T, true>;
              ensure_lds(kern, C::LDS);
              hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(SM_NT), C::LDS,
                                 s, up, sp, tab, n, colour, (T)lvl->scale,
                                 (T)relax);
            }
        }
      else
        {
          using C          = SM2<K, T>;
          const long long nb = (np + C::NPB - 1) / C::NPB;
          if (mode == 0)
            {
              auto kern = patch2d_kernel<K, T, false>;
              ensure_lds(kern, C::LDS);
              hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(SM_NT), C::LDS,
                                 s, up, sp, tab, n, colour, (T)lvl->scale,
                                 (T)relax);
            }
          else
            {
              auto kern = patch2d_kernel<K, T, true>;
              ensure_lds(kern, C::LDS);
              hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(SM_NT), C::LDS,
                                 s, up, sp, tab, n, colour, (T)lvl->scale,
                                 (T)relax);
            }
        }
      GTPS_LAUNCH_CHECK();
    }

    template <typename T>
    void
    launch_t(const gtps_level *lvl, void *u, const void *src, int colour,
             int mode, double relax)
    {
      switch (lvl->k)
        {
          case 1: return launch_k<1, T>(lvl, u, src, colour, mode, relax);
          case 2: return launch_k<2, T>(lvl, u, src, colour, mode, relax);
          case 3: return launch_k<3, T>(lvl, u, src, colour, mode, relax);
          case 4: return launch_k<4, T>(lvl, u, src, colour, mode, relax);
          case 5: return launch_k<5, T>(lvl, u, src, colour, mode, relax);
          case 6: return launch_k<6, T>(lvl, u, src, colour, mode, relax);
          case 7: return launch_k<7, T>(lvl, u, src, colour, mode, relax);
          default: throw Error(GTPS_EINVAL, "degree must be in 1..7");
        }
    }
  } // namespace

  void
  launch_patch_colour(const gtps_level *lvl, void *u, const void *src,
                      int colour, int mode, double relax)
  {
    if (lvl->prec == GTPS_FP32)
      launch_t<float>(lvl, u, src, colour, mode, relax);
    else
      launch_t<double>(lvl, u, src, colour, mode, relax);
  }
} // namespace gtps
