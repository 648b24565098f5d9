// laplace.hip - matrix-free Laplace vmult on the uniform Q_k mesh.
//
// On a uniform Cartesian mesh with constant coefficient, the cell operator
// the reference evaluates with Gauss(k+1) quadrature
// (laplace_operator.cuh:34-65, dealii/cuda_fee.cuh:273-529) sums to the
// Kronecker sum A = h^(d-2) sum_d K (x) M (x) ... of the assembled 1D
// matrices (Gauss(k+1) is exact for the degree-2k 1D integrands).  The
// kernels therefore work DoF-centred: every output row is owned by exactly
// one work-group, so there are no atomics and results are deterministic.
//
// 3D kernel: a work-group owns an x-y tile of TCX x TCY cells and marches in
// z over a chunk of cells.  Each thread keeps, for the columns of the tile
// plus a halo of k nodes, a window of 2k+1 z-planes in registers; per output
// plane it applies the z-rows of M and K (pass 1), then the y-rows (pass 2)
// and x-rows (pass 3) through LDS.  The node position p inside a cell is a
// compile-time loop index in every pass, so stencil coefficients are uniform
// (scalar) reads.  Dirichlet rows copy the input (y_B = x_B), as
// copy_constrained_values does (dealii/cuda_mf.template.cuh:494-508).
//
// HBM traffic per call: read x once, write y once (2 N sizeof(T)); halo
// columns are re-read from L2 by neighbouring tiles.
#include "gtps_internal.h"

namespace gtps
{
  namespace
  {
    template <int K>
    struct VT3;
    // cells per tile in x and y, threads per work-group
    template <>
    struct VT3<1>
    {
      static constexpr int TCX = 48, TCY = 16, NT = 512;
    };
    template <>
    struct VT3<2>
    {
      static constexpr int TCX = 24, TCY = 10, NT = 512;
    };
    template <>
    struct VT3<3>
    {
      static constexpr int TCX = 16, TCY = 8, NT = 512;
    };
    template <>
    struct VT3<4>
    {
      static constexpr int TCX = 12, TCY = 4, NT = 512;
    };
    template <>
    struct VT3<5>
    {
      static constexpr int TCX = 10, TCY = 3, NT = 512;
    };
    template <>
    struct VT3<6>
    {
      static constexpr int TCX = 8, TCY = 2, NT = 512;
    };
    template <>
    struct VT3<7>
    {
      static constexpr int TCX = 8, TCY = 2, NT = 512;
    };

    // alternative tiles of the metric degree (tuning knob vmult_tile)
    template <int K, int CFG>
    struct VT3C : VT3<K>
    {
    };
    template <>
    struct VT3C<3, 1>
    {
      static constexpr int TCX = 8, TCY = 8, NT = 256;
    };
    template <>
    struct VT3C<3, 2>
    {
      static constexpr int TCX = 16, TCY = 8, NT = 256;
    };
    // k = 7: 4x4 cells (the 8x2 tile of VT3<7> holds 5 columns of a
    // 15-plane window per thread: 256 VGPRs and spills in FP64)
    template <>
    struct VT3C<7, 2>
    {
      static constexpr int TCX = 4, TCY = 4, NT = 512;
    };
    // k = 7: the same 4x4 cells on 1024 threads (2 columns of the window
    // per thread instead of 4: about half the registers)
    template <>
    struct VT3C<7, 1>
    {
      static constexpr int TCX = 4, TCY = 4, NT = 1024;
    };

    template <int K, typename T, bool RESID, bool DOT, int CFG = 0>
    __global__ void __launch_bounds__((VT3C<K, CFG>::NT))
    laplace3d_kernel(const T *__restrict__ x, T *__restrict__ y,
                     const T *__restrict__ bv, const T *__restrict__ tab,
                     const int n, const int zc, const T scale,
                     double *__restrict__ partial)
    {
      using TB               = Tab<K>;
      constexpr int TCX      = VT3C<K, CFG>::TCX;
      constexpr int TCY      = VT3C<K, CFG>::TCY;
      constexpr int NT       = VT3C<K, CFG>::NT;
      constexpr int W        = 2 * K + 1;
      constexpr int WX       = K * (TCX + 2) + 1;
      constexpr int WY       = K * (TCY + 2) + 1;
      constexpr int TY       = K * TCY;
      constexpr int NCOLS    = WX * WY;
      constexpr int NCOL     = (NCOLS + NT - 1) / NT;

      __shared__ T      sA[NCOLS], sB[NCOLS], sC1[TY * WX], sC2[TY * WX];
      __shared__ T      sX0[DOT ? NCOLS : 1];
      __shared__ double sred[NT / 64];

      const int       N1  = K * n + 1;
      const long long P   = (long long)N1 * N1;
      const int       tid = threadIdx.x;
      // logical tile: z-chunk fastest, then x, then y, so that the tiles an
      // XCD works through share halo columns and window planes
      const int       L   = xcd_logical_id(linear_block_id(),
                                           gridDim.x * gridDim.y * gridDim.z);
      const int       bz  = L % gridDim.z, bxy = L / gridDim.z;
      const int       cx0 = (bxy % gridDim.x) * TCX, cy0 = (bxy / gridDim.x) * TCY;
      const int       cz0 = bz * zc;
      const int       cz1 = min(cz0 + zc, n);
      const int       gx0 = K * cx0 - K, gy0 = K * cy0 - K;
      const int       ox0 = K * cx0, oy0 = K * cy0;
      const int       ox1 = (cx0 + TCX >= n) ? N1 : K * (cx0 + TCX);
      const int       oy1 = (cy0 + TCY >= n) ? N1 : K * (cy0 + TCY);
      const T *bM = tab + TB::bandM;
      const T *bK = tab + TB::bandK;

      // register window of 2K+1 z-planes per column (planes K(c-1)..K(c+1)
      // for cell c) and, for K <= 4, the K planes of the next cell loaded one
      // cell ahead (larger K would spill)
      constexpr bool PF = K <= 4;
      T         w[NCOL][W], nx[NCOL][PF ? K : 1];
      long long cidx[NCOL];
      bool      own[NCOL], bcol[NCOL];
#pragma unroll
      for (int c = 0; c < NCOL; ++c)
        {
          const int  col   = tid + c * NT;
          const int  lx    = col % WX, ly = col / WX;
          const int  gx    = gx0 + lx, gy = gy0 + ly;
          const bool valid = col < NCOLS && gx >= 0 && gx < N1 && gy >= 0 &&
                             gy < N1;
          cidx[c] = valid ? gx + (long long)gy * N1 : -1;
          own[c]  = valid && gx >= ox0 && gx < ox1 && gy >= oy0 && gy < oy1;
          bcol[c] = gx == 0 || gx == N1 - 1 || gy == 0 || gy == N1 - 1;
#pragma unroll
          for (int j = 0; j < W; ++j)
            {
              const int gz = K * (cz0 - 1) + j;
              w[c][j] = (valid && gz >= 0 && gz < N1) ? x[cidx[c] + gz * P] : T(0);
            }
          if constexpr (PF)
            {
#pragma unroll
              for (int j = 0; j < K; ++j)
                {
                  const int gz = K * (cz0 + 1) + 1 + j;
                  nx[c][j] = (valid && cz0 + 1 < cz1 && gz < N1) ? x[cidx[c] + gz * P] : T(0);
                }
            }
        }

      double dot = 0.0;
      // residual, k >= 5: b at this thread's Dirichlet (boundary-column or
      // first-plane) DoFs, loaded in one plane and stored at the top of the
      // next, when the load has long completed (stored in pass 1 of its own
      // plane, the wait for it would also wait for the pass-3 prefetch).
      // Smaller k keep the pass-1 store: their kernels stay within 128
      // registers that way.
      constexpr bool DEFER = RESID && K >= 5;
      T              bnd[DEFER ? NCOL : 1];
      bool           bpend[DEFER ? NCOL : 1];
#pragma unroll
      for (int c = 0; c < (DEFER ? NCOL : 1); ++c)
        {
          bnd[c]   = T(0);
          bpend[c] = false;
        }
      for (int cz = cz0; cz < cz1; ++cz)
        {
#pragma unroll
          for (int p = 0; p < K; ++p)
            {
              const int       gz   = K * cz + p;
              const long long zoff = gz * P;
              if constexpr (DEFER)
                {
#pragma unroll
                  for (int c = 0; c < NCOL; ++c)
                    {
                      // previous plane (x in the window just below plane p)
                      if (bpend[c])
                        {
                          const T xv = w[c][K + p - 1];
                          const T yv = bnd[c] - xv;
                          y[cidx[c] + zoff - P] = yv;
                          if (DOT)
                            dot += (double)xv * (double)yv;
                        }
                      // this plane: loaded unconditionally (index 0 where not
                      // needed), so no load sits behind a branch
                      const bool      need = own[c] && (bcol[c] || gz == 0);
                      const long long gi   = need ? cidx[c] + zoff : 0;
                      bnd[c]               = bv[gi];
                      bpend[c]             = need;
                    }
                }
              // right-hand side of this thread's pass-3 outputs, loaded now
              // so that the loads overlap passes 1 and 2
              // (not for k = 7 in FP64, where the registers would spill)
              constexpr int  NT3 = (TCX * TY + NT - 1) / NT;
              constexpr bool PFB = RESID && (K < 7 || sizeof(T) == 4);
              T              rbv[NT3][PFB ? K : 1];
              if constexpr (PFB)
                {
#pragma unroll
                  for (int j3 = 0; j3 < NT3; ++j3)
                    {
                      const int task = tid + j3 * NT;
                      const int cxl = task % TCX, ty = task / TCX;
                      const int gyy = K * cy0 + ty;
#pragma unroll
                      for (int q = 0; q < K; ++q)
                        {
                          const int gxx = K * (cx0 + cxl) + q;
                          const bool in = task < TCX * TY && gz != 0 && gxx >= 1 &&
                                          gxx < N1 - 1 && gyy >= 1 && gyy < N1 - 1;
                          rbv[j3][q] = in ? bv[gxx + (long long)gyy * N1 + zoff] : T(0);
                        }
                    }
                }
              // pass 1: z-rows of M and K on every column of tile + halo
#pragma unroll
              for (int c = 0; c < NCOL; ++c)
                {
                  const int col = tid + c * NT;
                  if (col < NCOLS)
                    {
                      T a = 0, b = 0;
#pragma unroll
                      for (int o = band_lo<K>(p); o <= band_hi<K>(p); ++o)
                        {
                          a += bM[p * W + o + K] * w[c][K + p + o];
                          b += bK[p * W + o + K] * w[c][K + p + o];
                        }
                      sA[col] = a;
                      sB[col] = b;
                      if (DOT)
                        sX0[col] = w[c][K + p];
                      if (!DEFER && own[c] && (bcol[c] || gz == 0))
                        {
                          const long long g  = cidx[c] + zoff;
                          const T         xv = w[c][K + p];
                          const T         yv = RESID ? bv[g] - xv : xv;
                          y[g]               = yv;
                          if (DOT)
                            dot += (double)xv * (double)yv;
                        }
                    }
                }
              __syncthreads();
              if (gz != 0)
                {
                  // pass 2: y-rows; C1 = My (Mz x), C2 = Ky (Mz x) + My (Kz x)
                  for (int task = tid; task < WX * TCY; task += NT)
                    {
                      const int lx = task % WX, cyl = task / WX;
                      T         a[W], b[W];
#pragma unroll
                      for (int j = 0; j < W; ++j)
                        {
                          a[j] = sA[(K * cyl + j) * WX + lx];
                          b[j] = sB[(K * cyl + j) * WX + lx];
                        }
#pragma unroll
                      for (int q = 0; q < K; ++q)
                        {
                          T c1 = 0, c2 = 0;
#pragma unroll
                          for (int o = band_lo<K>(q); o <= band_hi<K>(q); ++o)
                            {
                              c1 += bM[q * W + o + K] * a[K + q + o];
                              c2 += bK[q * W + o + K] * a[K + q + o];
                              c2 += bM[q * W + o + K] * b[K + q + o];
                            }
                          sC1[(K * cyl + q) * WX + lx] = c1;
                          sC2[(K * cyl + q) * WX + lx] = c2;
                        }
                    }
                  __syncthreads();
                  // pass 3: x-rows; y = h^(d-2) (Kx C1 + Mx C2) on interior rows
#pragma unroll
                  for (int j3 = 0; j3 < NT3; ++j3)
                    {
                      const int task = tid + j3 * NT;
                      if (task >= TCX * TY)
                        continue;
                      const int cxl = task % TCX, ty = task / TCX;
                      const int gy  = K * cy0 + ty;
                      T         c1[W], c2[W];
#pragma unroll
                      for (int j = 0; j < W; ++j)
                        {
                          c1[j] = sC1[ty * WX + K * cxl + j];
                          c2[j] = sC2[ty * WX + K * cxl + j];
                        }
#pragma unroll
                      for (int q = 0; q < K; ++q)
                        {
                          const int gx = K * (cx0 + cxl) + q;
                          if (gx >= 1 && gx < N1 - 1 && gy >= 1 && gy < N1 - 1)
                            {
                              T v = 0;
#pragma unroll
                              for (int o = band_lo<K>(q); o <= band_hi<K>(q);
                                   ++o)
                                {
                                  v += bK[q * W + o + K] * c1[K + q + o];
                                  v += bM[q * W + o + K] * c2[K + q + o];
                                }
                              v *= scale;
                              const long long g  = gx + (long long)gy * N1 + zoff;
                              T               yv = v;
                              if constexpr (PFB)
                                yv = rbv[j3][q] - v;
                              else if constexpr (RESID)
                                yv = bv[g] - v;
                              y[g]               = yv;
                              if (DOT)
                                dot += (double)sX0[(ty + K) * WX + K * cxl + q + K] *
                                       (double)yv;
                            }
                        }
                    }
                }
              __syncthreads();
            }
          // shift the window by one cell, move the prefetched planes in and
          // start loading the cell after next
#pragma unroll
          for (int c = 0; c < NCOL; ++c)
            {
#pragma unroll
              for (int j = 0; j <= K; ++j)
                w[c][j] = w[c][j + K];
              if constexpr (PF)
                {
#pragma unroll
                  for (int j = 0; j < K; ++j)
                    w[c][K + 1 + j] = nx[c][j];
                  if (cz + 2 < cz1)
                    {
#pragma unroll
                      for (int j = 0; j < K; ++j)
                        {
                          const int gz = K * (cz + 2) + 1 + j;
                          nx[c][j] = (cidx[c] >= 0 && gz < N1) ? x[cidx[c] + gz * P] : T(0);
                        }
                    }
                }
              else if (cz + 1 < cz1)
                {
#pragma unroll
                  for (int j = K + 1; j < W; ++j)
                    {
                      const int gz = K * cz + j;
                      w[c][j] = (cidx[c] >= 0 && gz < N1) ? x[cidx[c] + gz * P] : T(0);
                    }
                }
            }
        }
      // the chunk's last plane (x now just below the window centre)
      if constexpr (DEFER)
        {
#pragma unroll
          for (int c = 0; c < NCOL; ++c)
            if (bpend[c])
              {
                const T xv = w[c][K - 1];
                const T yv = bnd[c] - xv;
                y[cidx[c] + (long long)(K * cz1 - 1) * P] = yv;
                if (DOT)
                  dot += (double)xv * (double)yv;
              }
        }
      // last plane of the domain: Dirichlet rows
      if (cz1 == n)
        {
          const long long zoff = (long long)(N1 - 1) * P;
#pragma unroll
          for (int c = 0; c < NCOL; ++c)
            if (own[c])
              {
                const long long g  = cidx[c] + zoff;
                const T         xv = w[c][K];
                const T         yv = RESID ? bv[g] - xv : xv;
                y[g]               = yv;
                if (DOT)
                  dot += (double)xv * (double)yv;
              }
        }
      if (DOT)
        {
          const double s = block_sum<NT>(dot, sred);
          if (tid == 0)
            partial[linear_block_id()] = s;
        }
    }

    // Box kernel for small levels: a work-group owns a B^3 block of DoFs,
    // loads it with a halo of k in one round of loads (unrolled, all in
    // flight together) and applies the z-, y- and x-rows of the assembled
    // band matrices through LDS.  Latency-bound levels (a few thousand to a
    // few hundred thousand DoFs) finish in one memory round trip instead of
    // the serial plane chain of the z-marching kernel; the halo re-reads
    // ((B+2k)^3 / B^3) come from L2.
    // B is a multiple of k, so block origins sit on cell boundaries and the
    // node position inside a cell (which selects the band coefficients) is
    // a compile-time function of the unrolled output index.  From k = 5 on
    // a block is a single cell (halo volume 27x the block), so the box form
    // pays only up to 8 cells per direction; k = 6, 7 in FP64 exceed the
    // 64 KiB of static LDS and always march.
    template <int K>
    struct VB
    {
      static constexpr int B = K <= 3 ? K * ((8 + K - 1) / K) : (K == 4 ? 4 : K), NT = 256;
      static constexpr int WB = B + 2 * K;
      // static LDS of the box kernel (box, the two z-contractions)
      template <typename T>
      static constexpr bool fits()
      {
        constexpr int NBOX = WB * WB * WB, NCC = WB * B * B;
        return sizeof(T) * ((NBOX > 2 * NCC ? NBOX : 2 * NCC) + 2 * WB * WB * B) +
                 sizeof(double) * (NT / 64) <=
               65536;
      }
    };

    template <int K, typename T, bool RESID, bool DOT>
    __global__ void __launch_bounds__(256)
    laplace3d_box_kernel(const T *__restrict__ x, T *__restrict__ y,
                         const T *__restrict__ bv, const T *__restrict__ tab,
                         const int n, const T scale, double *__restrict__ partial)
    {
      using TB         = Tab<K>;
      constexpr int B  = VB<K>::B, NT = VB<K>::NT, WB = VB<K>::WB;
      constexpr int W  = 2 * K + 1;
      // box (later reused for C1, C2), then A = Mz x and Bz = Kz x
      constexpr int NBOX = WB * WB * WB;
      constexpr int NCC  = WB * B * B;
      constexpr int NBUF = NBOX > 2 * NCC ? NBOX : 2 * NCC;
      __shared__ T      sBox[NBUF], sA[WB * WB * B], sB[WB * WB * B];
      __shared__ double sred[NT / 64];
      const T  *bM  = tab + TB::bandM;
      const T  *bK  = tab + TB::bandK;
      const int N1  = K * n + 1;
      const long long P = (long long)N1 * N1;
      const int tid = threadIdx.x;
      const int nbx = (N1 + B - 1) / B;
      const int L   = xcd_logical_id(blockIdx.x, gridDim.x);
      const int ox0 = (L % nbx) * B, oy0 = ((L / nbx) % nbx) * B,
                oz0 = (L / (nbx * nbx)) * B;
      constexpr int NRB  = RESID ? B : 1;
      T             xo[B], bo[NRB];
      const int     rly = tid % B, rlz = tid / B;
      const int     rgy = oy0 + rly, rgz = oz0 + rlz;
      const bool    rrow = tid < B * B && rgy < N1 && rgz < N1;
      {
        // one z-column (x, y) of the box per thread: the column's address and
        // x/y bounds set up once, the z bounds uniform
        constexpr int NC = WB * WB, IT = (NC + NT - 1) / NT;
        T             v[IT][WB];
#pragma unroll
        for (int j = 0; j < IT; ++j)
          {
            const int  col = tid + j * NT;
            const int  by = col / WB, bx = col - by * WB;
            const int  gx = ox0 - K + bx, gy = oy0 - K + by;
            const bool in = col < NC && gx >= 0 && gx < N1 && gy >= 0 && gy < N1;
            const T   *xc = x + (in ? gx + (long long)gy * N1 : 0);
#pragma unroll
            for (int bz = 0; bz < WB; ++bz)
              {
                const int gz = oz0 - K + bz;
                v[j][bz]     = in && gz >= 0 && gz < N1 ? xc[gz * P] : T(0);
              }
          }
        // the x-pass outputs of this thread (row (ly, lz) = tid < B*B):
        // their input values and right-hand side, issued with the box loads
#pragma unroll
        for (int lx = 0; lx < B; ++lx)
          {
            const int       gx = ox0 + lx;
            const long long g  = gx + (long long)rgy * N1 + rgz * P;
            xo[lx]             = (rrow && gx < N1) ? x[g] : T(0);
            if constexpr (RESID)
              bo[lx] = (rrow && gx < N1) ? bv[g] : T(0);
          }
#pragma unroll
        for (int j = 0; j < IT; ++j)
          {
            const int col = tid + j * NT;
            if (col < NC)
              {
#pragma unroll
                for (int bz = 0; bz < WB; ++bz)
                  sBox[bz * NC + col] = v[j][bz];
              }
          }
      }
      static_assert(B * B <= NT, "one x-pass row per thread");
      __syncthreads();
      // z-rows on every (x, y) column of the box, B output planes
      for (int t = tid; t < WB * WB; t += NT)
        {
          T c[WB];
#pragma unroll
          for (int z = 0; z < WB; ++z)
            c[z] = sBox[z * WB * WB + t];
#pragma unroll
          for (int lz = 0; lz < B; ++lz)
            {
              const int p = lz % K;
              T         a = 0, b = 0;
              // band entries outside band_lo .. band_hi are zero
#pragma unroll
              for (int o = band_lo<K>(p); o <= band_hi<K>(p); ++o)
                {
                  const T m = bM[p * W + o + K], kk = bK[p * W + o + K];
                  a += m * c[lz + K + o];
                  b += kk * c[lz + K + o];
                }
              sA[lz * WB * WB + t] = a;
              sB[lz * WB * WB + t] = b;
            }
        }
      __syncthreads();
      // y-rows on every (x, z) column: C1 = My A, C2 = Ky A + My Bz
      T *sC1 = sBox, *sC2 = sBox + NCC;
      for (int t = tid; t < WB * B; t += NT)
        {
          const int bx = t % WB, lz = t / WB;
          T         a[WB], b[WB];
#pragma unroll
          for (int j = 0; j < WB; ++j)
            {
              a[j] = sA[(lz * WB + j) * WB + bx];
              b[j] = sB[(lz * WB + j) * WB + bx];
            }
#pragma unroll
          for (int ly = 0; ly < B; ++ly)
            {
              const int p = ly % K;
              T         c1 = 0, c2 = 0;
#pragma unroll
              for (int o = band_lo<K>(p); o <= band_hi<K>(p); ++o)
                {
                  const T m = bM[p * W + o + K], kk = bK[p * W + o + K];
                  c1 += m * a[ly + K + o];
                  c2 += kk * a[ly + K + o];
                  c2 += m * b[ly + K + o];
                }
              sC1[(lz * B + ly) * WB + bx] = c1;
              sC2[(lz * B + ly) * WB + bx] = c2;
            }
        }
      __syncthreads();
      // x-rows per (y, z) row of the block, then the Dirichlet rows
      double dot = 0.0;
      if (rrow)
        {
          const int t  = tid;
          const int gy = rgy, gz = rgz;
          T c1[WB], c2[WB];
#pragma unroll
          for (int j = 0; j < WB; ++j)
            {
              c1[j] = sC1[t * WB + j];
              c2[j] = sC2[t * WB + j];
            }
#pragma unroll
          for (int lx = 0; lx < B; ++lx)
            {
              const int gx = ox0 + lx;
              if (gx >= N1)
                continue;
              const int p = lx % K;
              T         v = 0;
#pragma unroll
              for (int o = band_lo<K>(p); o <= band_hi<K>(p); ++o)
                {
                  v += bK[p * W + o + K] * c1[lx + K + o];
                  v += bM[p * W + o + K] * c2[lx + K + o];
                }
              const bool bd = gx == 0 || gx == N1 - 1 || gy == 0 || gy == N1 - 1 ||
                              gz == 0 || gz == N1 - 1;
              const long long g  = gx + (long long)gy * N1 + gz * P;
              const T         xv = xo[lx];
              v                  = bd ? xv : scale * v;
              T yv               = v;
              if constexpr (RESID)
                yv = bo[lx] - v;
              y[g] = yv;
              if (DOT)
                dot += (double)xv * (double)yv;
            }
        }
      if (DOT)
        {
          const double s = block_sum<NT>(dot, sred);
          if (tid == 0)
            partial[blockIdx.x] = s;
        }
    }

    // 2D: a work-group owns TCX x TCY cells; tile + halo staged in LDS.
    template <int K>
    struct VT2
    {
      static constexpr int TCX = (64 + K - 1) / K, TCY = (16 + K - 1) / K,
                           NT = 256;
    };

    template <int K, typename T, bool RESID, bool DOT>
    __global__ void __launch_bounds__(256)
    laplace2d_kernel(const T *__restrict__ x, T *__restrict__ y,
                     const T *__restrict__ bv, const T *__restrict__ tab,
                     const int n, const T scale, double *__restrict__ partial)
    {
      using TB          = Tab<K>;
      constexpr int TCX = VT2<K>::TCX, TCY = VT2<K>::TCY, NT = VT2<K>::NT;
      constexpr int W   = 2 * K + 1;
      constexpr int WX  = K * (TCX + 2) + 1, WY = K * (TCY + 2) + 1;
      constexpr int TY  = K * TCY;
      __shared__ T      sX[WX * WY], sC1[TY * WX], sC2[TY * WX];
      __shared__ double sred[NT / 64];
      const int         N1  = K * n + 1;
      const int         tid = threadIdx.x;
      const int         L   = xcd_logical_id(linear_block_id(), gridDim.x * gridDim.y);
      const int         cx0 = (L % gridDim.x) * TCX, cy0 = (L / gridDim.x) * TCY;
      const int         gx0 = K * cx0 - K, gy0 = K * cy0 - K;
      const int         ox0 = K * cx0, oy0 = K * cy0;
      const int         ox1 = (cx0 + TCX >= n) ? N1 : K * (cx0 + TCX);
      const int         oy1 = (cy0 + TCY >= n) ? N1 : K * (cy0 + TCY);
      const T *bM = tab + TB::bandM;
      const T *bK = tab + TB::bandK;
      double   dot = 0.0;
      for (int i = tid; i < WX * WY; i += NT)
        {
          const int lx = i % WX, ly = i / WX;
          const int gx = gx0 + lx, gy = gy0 + ly;
          const bool valid = gx >= 0 && gx < N1 && gy >= 0 && gy < N1;
          const T    xv    = valid ? x[gx + (long long)gy * N1] : T(0);
          sX[i]            = xv;
          if (valid && gx >= ox0 && gx < ox1 && gy >= oy0 && gy < oy1 &&
              (gx == 0 || gx == N1 - 1 || gy == 0 || gy == N1 - 1))
            {
              const long long g  = gx + (long long)gy * N1;
              const T         yv = RESID ? bv[g] - xv : xv;
              y[g]               = yv;
              if (DOT)
                dot += (double)xv * (double)yv;
            }
        }
      __syncthreads();
      for (int task = tid; task < WX * TCY; task += NT)
        {
          const int lx = task % WX, cyl = task / WX;
          T         a[W];
#pragma unroll
          for (int j = 0; j < W; ++j)
            a[j] = sX[(K * cyl + j) * WX + lx];
#pragma unroll
          for (int q = 0; q < K; ++q)
            {
              T c1 = 0, c2 = 0;
#pragma unroll
              for (int o = band_lo<K>(q); o <= band_hi<K>(q); ++o)
                {
                  c1 += bM[q * W + o + K] * a[K + q + o];
                  c2 += bK[q * W + o + K] * a[K + q + o];
                }
              sC1[(K * cyl + q) * WX + lx] = c1;
              sC2[(K * cyl + q) * WX + lx] = c2;
            }
        }
      __syncthreads();
      for (int task = tid; task < TCX * TY; task += NT)
        {
          const int cxl = task % TCX, ty = task / TCX;
          const int gy  = K * cy0 + ty;
          T         c1[W], c2[W];
#pragma unroll
          for (int j = 0; j < W; ++j)
            {
              c1[j] = sC1[ty * WX + K * cxl + j];
              c2[j] = sC2[ty * WX + K * cxl + j];
            }
#pragma unroll
          for (int q = 0; q < K; ++q)
            {
              const int gx = K * (cx0 + cxl) + q;
              if (gx >= 1 && gx < N1 - 1 && gy >= 1 && gy < N1 - 1)
                {
                  T v = 0;
#pragma unroll
                  for (int o = band_lo<K>(q); o <= band_hi<K>(q); ++o)
                    {
                      v += bK[q * W + o + K] * c1[K + q + o];
                      v += bM[q * W + o + K] * c2[K + q + o];
                    }
                  v *= scale;
                  const long long g  = gx + (long long)gy * N1;
                  const T         yv = RESID ? bv[g] - v : v;
                  y[g]               = yv;
                  if (DOT)
                    dot += (double)x[g] * (double)yv;
                }
            }
        }
      if (DOT)
        {
          const double s = block_sum<NT>(dot, sred);
          if (tid == 0)
            partial[linear_block_id()] = s;
        }
    }

    template <int K, typename T>
    int
    launch_k(const gtps_level *lvl, void *y, const void *x, const void *b,
             bool fused_dot)
    {
      const T   *tab   = reinterpret_cast<const T *>(lvl->dtab());
      const T    scale = (T)lvl->scale;
      const int  n     = lvl->n;
      double    *part  = lvl->ctx->ws.partials;
      hipStream_t s    = lvl->ctx->stream;
      const T   *xp    = reinterpret_cast<const T *>(x);
      T         *yp    = reinterpret_cast<T *>(y);
      const T   *bp    = reinterpret_cast<const T *>(b);
      if constexpr (VB<K>::template fits<T>())
        if (lvl->dim == 3 && n <= (K >= 5 ? tuning().box_vmult_max_n / 2 : tuning().box_vmult_max_n))
          {
            using C        = VB<K>;
            const int nbx  = (lvl->N1 + C::B - 1) / C::B;
            const int nb   = nbx * nbx * nbx;
            if (fused_dot)
              GTPS_REQUIRE(nb <= lvl->ctx->ws.capacity, "too many blocks for partials");
#define GTPS_LB(R, D)                                                         \
  hipLaunchKernelGGL((laplace3d_box_kernel<K, T, R, D>), dim3(nb), dim3(C::NT), \
                     0, s, xp, yp, bp, tab, n, scale, part)
            if (b)
              {
                if (fused_dot)
                  GTPS_LB(true, true);
                else
                  GTPS_LB(true, false);
              }
            else
              {
                if (fused_dot)
                  GTPS_LB(false, true);
                else
                  GTPS_LB(false, false);
              }
#undef GTPS_LB
            GTPS_LAUNCH_CHECK();
            return nb;
          }
      if (lvl->dim == 3)
        {
          auto go = [&](auto cfg) -> int {
            constexpr int CFG = decltype(cfg)::value;
            using C       = VT3C<K, CFG>;
            const int tx  = (n + C::TCX - 1) / C::TCX;
            const int ty  = (n + C::TCY - 1) / C::TCY;
            // z-chunks: about vmult_zdiv work-groups in total (0: 512 in
            // FP32, 2048 in FP64 from a sweep at 64^3 Q3; 512 and 256 for
            // k = 7 from a sweep at 32^3 Q7)
            const long long zdiv = tuning().vmult_zdiv > 0 ? tuning().vmult_zdiv
                                   : sizeof(T) == 4        ? 512
                                   : K == 7                ? 256
                                                           : 2048;
            int zc = (int)(((long long)n * tx * ty + zdiv - 1) / zdiv);
            zc     = std::max(1, std::min(zc, n));
            const int tz = (n + zc - 1) / zc;
            const dim3 grid(tx, ty, tz);
            const int  nb = tx * ty * tz;
            if (fused_dot)
              GTPS_REQUIRE(nb <= lvl->ctx->ws.capacity, "too many blocks for partials");
#define GTPS_L3(R, D)                                                              \
  hipLaunchKernelGGL((laplace3d_kernel<K, T, R, D, CFG>), grid, dim3(C::NT), 0, s, \
                     xp, yp, bp, tab, n, zc, scale, part)
            if (b)
              {
                if (fused_dot)
                  GTPS_L3(true, true);
                else
                  GTPS_L3(true, false);
              }
            else
              {
                if (fused_dot)
                  GTPS_L3(false, true);
                else
                  GTPS_L3(false, false);
              }
#undef GTPS_L3
            GTPS_LAUNCH_CHECK();
            return nb;
          };
          if constexpr (K == 3)
            {
              // 3: 16x8 cells / 512 threads in FP32, 8x8 / 256 in FP64
              const int tile = tuning().vmult_tile < 3 ? tuning().vmult_tile
                               : sizeof(T) == 4         ? 0
                                                        : 1;
              if (tile == 1)
                return go(std::integral_constant<int, 1>{});
              if (tile == 2)
                return go(std::integral_constant<int, 2>{});
            }
          if constexpr (K == 7)
            {
              // 3 (auto) and 2: 4x4 cells / 512 threads; 1: 4x4 / 1024;
              // 0: 8x2 / 512
              const int tile = tuning().vmult_tile;
              if (tile == 2 || tile == 3)
                return go(std::integral_constant<int, 2>{});
              if (tile == 1)
                return go(std::integral_constant<int, 1>{});
            }
          return go(std::integral_constant<int, 0>{});
        }
      using C      = VT2<K>;
      const dim3 grid((n + C::TCX - 1) / C::TCX, (n + C::TCY - 1) / C::TCY);
      const int  nb = grid.x * grid.y;
      if (fused_dot)
        GTPS_REQUIRE(nb <= lvl->ctx->ws.capacity, "too many blocks for partials");
#define GTPS_L2(R, D)                                                         \
  hipLaunchKernelGGL((laplace2d_kernel<K, T, R, D>), grid, dim3(C::NT), 0, s, \
                     xp, yp, bp, tab, n, scale, part)
      if (b)
        {
          if (fused_dot)
            GTPS_L2(true, true);
          else
            GTPS_L2(true, false);
        }
      else
        {
          if (fused_dot)
            GTPS_L2(false, true);
          else
            GTPS_L2(false, false);
        }
#undef GTPS_L2
      GTPS_LAUNCH_CHECK();
      return nb;
    }

    template <typename T>
    int
    launch_t(const gtps_level *lvl, void *y, const void *x, const void *b,
             bool fused_dot)
    {
      switch (lvl->k)
        {
          case 1: return launch_k<1, T>(lvl, y, x, b, fused_dot);
          case 2: return launch_k<2, T>(lvl, y, x, b, fused_dot);
          case 3: return launch_k<3, T>(lvl, y, x, b, fused_dot);
          case 4: return launch_k<4, T>(lvl, y, x, b, fused_dot);
          case 5: return launch_k<5, T>(lvl, y, x, b, fused_dot);
          case 6: return launch_k<6, T>(lvl, y, x, b, fused_dot);
          case 7: return launch_k<7, T>(lvl, y, x, b, fused_dot);
          default: throw Error(GTPS_EINVAL, "degree must be in 1..7");
        }
    }
  } // namespace

  int
  launch_laplace(const gtps_level *lvl, void *y, const void *x, const void *b,
                 bool fused_dot)
  {
    if (lvl->coef)
      return launch_coeff_apply(lvl, y, x, b, fused_dot, false);
    if (lvl->prec == GTPS_FP32)
      return launch_t<float>(lvl, y, x, b, fused_dot);
    return launch_t<double>(lvl, y, x, b, fused_dot);
  }
} // namespace gtps
