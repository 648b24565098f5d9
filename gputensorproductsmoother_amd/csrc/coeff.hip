// coeff.hip - variable-coefficient diffusion on the uniform mesh (3D):
// A(c) = sum_K c_K A^1_K with c constant per cell, its transpose, the
// gradient dJ/dc_K = -lambda~_K^T A^1_K u_K, coefficient coarsening, and a
// vertex-patch smoother with an inexact fast-diagonalisation local solve.
//
// The reference carries a coefficient through its solver interface but
// never uses it (include/solver.cuh:1548; the matrix-free hook
// evaluate_coeff, include/dealii/cuda_mf.template.cuh:569-581, 815-824, is
// not called on the hot path), so these kernels have no reference
// counterpart; they reuse the layout and conventions of the
// constant-coefficient path (lexicographic DoFs, identity Dirichlet rows,
// boundary columns kept).  Cells are numbered cx + n cy + n^2 cz.
//
// Operator: cell-wise, the reference-cell Laplacian is the Kronecker sum
// h (K^ (x) M^ (x) M^ + M^ (x) K^ (x) M^ + M^ (x) M^ (x) K^) of the 1D cell
// matrices (exact with Gauss(k+1)), applied by sum factorisation: z-rows,
// then y-rows, then x-rows, one thread per cell pencil.  Cell results go to
// an element-vector buffer that a DoF-centred gather adds up in a fixed
// order (deterministic, no atomics).
//
// Smoother: per colour and patch, u_I <- u_I + relax (sigma_P A~_P)^{-1}
// (b - A u)_I, the local residual formed from the 8 cells of the patch in
// LDS.  A~_P is the patch operator of a separable coefficient f_abc =
// rho_x^a rho_y^b rho_z^c (a, b, c the cell offsets), which is a Kronecker
// sum of the 2-cell 1D pencils (K_cell0 + rho K_cell1, M_cell0 + rho
// M_cell1) and is inverted by fast diagonalisation with one eigenbasis per
// direction (exact when c is separable, c = 1 in particular).  rho_d is the
// log-space least-squares ratio of the upper to the lower cells along d,
// rounded to a grid of step SEP_STEP whose eigenpairs are tabulated once per
// degree (FP64 on the GPU, setup.hip); sigma_P = max(max_K c_K/f_K /
// SEP_THETA, exp(mean log(c_K/f_K))).  Every A_K is positive semi-definite,
// so A_P <= max(c/f) sum f_K A_K <= SEP_THETA sigma_P A~_P: the local
// correction's largest eigenvalue is at most SEP_THETA = 1.5 < 2, each
// patch update is an A-norm contraction and the symmetric V-cycle stays
// positive definite.  (The previous local solve, the constant-coefficient
// inverse scaled by the patch's largest coefficient, took 19 CG iterations
// at 32^3 Q4 where exact local solves take 7.)
#include "gtps_internal.h"

namespace gtps
{
  namespace
  {
    constexpr int CNT = 256;  // threads per work-group

    __host__ __device__ constexpr int
    cells_per_block(int K)
    {
      return CNT / ((K + 1) * (K + 1));
    }

    __device__ __forceinline__ bool
    on_bdry(int gx, int gy, int gz, int N1)
    {
      return gx == 0 || gx == N1 - 1 || gy == 0 || gy == N1 - 1 || gz == 0 ||
             gz == N1 - 1;
    }

    // MODE 0: element vectors E[cell][n^3] = c_K scale A^1 x_K (Dirichlet
    //         inputs masked when mask != 0);
    // MODE 1: grad[cell] = -scale lambda~_K . A^1 u_K (lambda's Dirichlet
    //         entries masked).
    template <int K, typename T, int MODE>
    __global__ void __launch_bounds__(CNT)
    cell_kernel(const T *__restrict__ x, const T *__restrict__ lam,
                const T *__restrict__ coef, T *__restrict__ out,
                const T *__restrict__ tab, const int n, const int N1,
                const T scale, const int mask, const FastDiv div_n,
                const FastDiv div_nn)
    {
      using TB          = Tab<K>;
      constexpr int NN  = K + 1, N2 = NN * NN, N3 = N2 * NN;
      constexpr int CPB = cells_per_block(K);
      __shared__ T  sa[CPB * N3], sb[CPB * N3];
      __shared__ double red[MODE == 1 ? CNT : 1];
      const T      *cM   = tab + TB::cellM;
      const T      *cK   = tab + TB::cellK;
      const int     tid  = threadIdx.x;
      const int     lc   = tid / N2, t = tid - lc * N2;
      // level and element-buffer sizes < 2^31 (host check): 32-bit indices,
      // cell coordinates by host-computed multiply-high divisors
      const int  ncells = n * n * n;
      const int  cell   = xcd_logical_id(blockIdx.x, gridDim.x) * CPB + lc;
      const bool active = lc < CPB && cell < ncells;
      const int  P      = N1 * N1;
      int        cx = 0, cy = 0, cz = 0, base = 0;
      if (active)
        {
          cz   = (int)fastdiv((unsigned)cell, div_nn);
          const int rc = cell - cz * n * n;
          cy   = (int)fastdiv((unsigned)rc, div_n);
          cx   = rc - cy * n;
          base = K * cx + K * cy * N1 + K * cz * P;
        }
      T *ca = sa + (active ? lc : 0) * N3;
      T *cb = sb + (active ? lc : 0) * N3;
      // z-rows: thread (ix, iy) owns the z-pencil
      if (active)
        {
          const int ix = t % NN, iy = t / NN;
          T         v[NN];
#pragma unroll
          for (int iz = 0; iz < NN; ++iz)
            {
              const bool m = mask && on_bdry(K * cx + ix, K * cy + iy, K * cz + iz, N1);
              v[iz] = m ? T(0) : x[base + ix + iy * N1 + iz * P];
            }
#pragma unroll
          for (int lz = 0; lz < NN; ++lz)
            {
              T a = 0, b = 0;
#pragma unroll
              for (int iz = 0; iz < NN; ++iz)
                {
                  a += cM[lz * NN + iz] * v[iz];
                  b += cK[lz * NN + iz] * v[iz];
                }
              ca[(lz * NN + iy) * NN + ix] = a;
              cb[(lz * NN + iy) * NN + ix] = b;
            }
        }
      __syncthreads();
      // y-rows, in place per (lz, ix) column: c1 = My a, c2 = Ky a + My b
      if (active)
        {
          const int ix = t % NN, lz = t / NN;
          T         a[NN], b[NN];
#pragma unroll
          for (int iy = 0; iy < NN; ++iy)
            {
              a[iy] = ca[(lz * NN + iy) * NN + ix];
              b[iy] = cb[(lz * NN + iy) * NN + ix];
            }
#pragma unroll
          for (int ly = 0; ly < NN; ++ly)
            {
              T c1 = 0, c2 = 0;
#pragma unroll
              for (int iy = 0; iy < NN; ++iy)
                {
                  c1 += cM[ly * NN + iy] * a[iy];
                  c2 += cK[ly * NN + iy] * a[iy];
                  c2 += cM[ly * NN + iy] * b[iy];
                }
              ca[(lz * NN + ly) * NN + ix] = c1;
              cb[(lz * NN + ly) * NN + ix] = c2;
            }
        }
      __syncthreads();
      // x-rows per (lz, ly) row: y = Kx c1 + Mx c2
      double part = 0.0;
      if (active)
        {
          const int ly = t % NN, lz = t / NN;
          T         c1[NN], c2[NN];
#pragma unroll
          for (int ix = 0; ix < NN; ++ix)
            {
              c1[ix] = ca[(lz * NN + ly) * NN + ix];
              c2[ix] = cb[(lz * NN + ly) * NN + ix];
            }
          const T f = MODE == 0 ? (coef ? coef[cell] : T(1)) * scale : scale;
#pragma unroll
          for (int lx = 0; lx < NN; ++lx)
            {
              T y = 0;
#pragma unroll
              for (int ix = 0; ix < NN; ++ix)
                {
                  y += cK[lx * NN + ix] * c1[ix];
                  y += cM[lx * NN + ix] * c2[ix];
                }
              if (MODE == 0)
                out[cell * N3 + (lz * NN + ly) * NN + lx] = f * y;
              else
                {
                  const int gx = K * cx + lx, gy = K * cy + ly, gz = K * cz + lz;
                  if (!on_bdry(gx, gy, gz, N1))
                    part += (double)lam[base + lx + ly * N1 + lz * P] *
                            (double)(f * y);
                }
            }
        }
      if (MODE == 1)
        {
          red[tid] = part;
          __syncthreads();
          if (active && t == 0)
            {
              double s = 0.0;
              for (int q = 0; q < N2; ++q)
                s += red[lc * N2 + q];
              out[cell] = (T)(-s);
            }
        }
    }

    // DoF-centred sum of the element vectors, fixed order (lower cell
    // first per direction, z outermost).  Per direction the DoF at node g
    // is node g - K c of the lower cell c = (g-1)/K (g >= 1), and, on a cell
    // boundary (g = K c', c' < n), node 0 of the upper cell c'; the up to 8
    // contributions are read at offsets summed per direction, the absent
    // ones masked.  Rows: RESID y = b - A x; TRANSPOSE boundary rows y_B =
    // sum + x_B; otherwise identity rows.  DOT: per-block partial sums of
    // x.y.  Level and element-buffer sizes < 2^31 (host check).
    template <int K, typename T, bool RESID, bool TRANSPOSE, bool DOT>
    __global__ void __launch_bounds__(CNT)
    gather_kernel(const int N, const int N1, const int n,
                  const T *__restrict__ E, const T *__restrict__ x,
                  const T *__restrict__ b, T *__restrict__ y,
                  double *__restrict__ partials, const FastDiv div_n1,
                  const FastDiv div_p)
    {
      constexpr int NN = K + 1, N3 = NN * NN * NN;
      __shared__ double red[CNT / 64];
      double            dsum = 0.0;
      const int         P    = N1 * N1;
      // element-buffer stride of one cell step and one node step per direction
      const int cs[3] = {N3, n * N3, n * n * N3}, ns[3] = {1, NN, NN * NN};
      for (int i = blockIdx.x * CNT + threadIdx.x; i < N; i += gridDim.x * CNT)
        {
          const int gz = (int)fastdiv((unsigned)i, div_p), r = i - gz * P;
          const int gy = (int)fastdiv((unsigned)r, div_n1), gx = r - gy * N1;
          const int g[3] = {gx, gy, gz};
          int       off[3][2];
          bool      on[3][2];
#pragma unroll
          for (int d = 0; d < 3; ++d)
            {
              const int cl = (g[d] - 1) / K;  // lower cell (g >= 1)
              const int cu = g[d] / K;        // upper cell on a cell boundary
              on[d][0]     = g[d] >= 1;
              on[d][1]     = g[d] == K * cu && cu < n;
              off[d][0]    = cl * cs[d] + (g[d] - K * cl) * ns[d];
              off[d][1]    = cu * cs[d];
            }
          T s = 0;
#pragma unroll
          for (int oz = 0; oz < 2; ++oz)
#pragma unroll
            for (int oy = 0; oy < 2; ++oy)
#pragma unroll
              for (int ox = 0; ox < 2; ++ox)
                {
                  const bool ok = on[0][ox] && on[1][oy] && on[2][oz];
                  const T    v  = ok ? E[off[0][ox] + off[1][oy] + off[2][oz]] : T(0);
                  s += v;
                }
          const bool bd = gx == 0 || gx == N1 - 1 || gy == 0 || gy == N1 - 1 ||
                          gz == 0 || gz == N1 - 1;
          T v;
          if (bd)
            v = TRANSPOSE ? s + x[i] : x[i];
          else
            v = s;
          if (RESID)
            v = b[i] - v;
          y[i] = v;
          if (DOT)
            dsum += (double)x[i] * (double)v;
        }
      if (DOT)
        {
          dsum = block_sum<CNT>(dsum, red);
          if (threadIdx.x == 0)
            partials[blockIdx.x] = dsum;
        }
    }

    template <int K, typename T>
    struct PC
    {
      static constexpr int NN = K + 1, N2 = NN * NN, N3 = N2 * NN;
      static constexpr int NP = 2 * K + 1, NI = 2 * K - 1;
      static constexpr size_t lds =
        sizeof(T) * (NP * NP * NP + 2 * 8 * N3 + NI * NI * NI + 8);
      static constexpr bool ok = lds <= 65536;
    };

    // One thread per vertex patch (lower-left cell (px, py, pz), patches
    // lexicographic over (n-1)^3): the log-space least-squares ratios of the
    // 8 cell coefficients per direction, rounded to the SEP_STEP grid, and
    // sigma = max(max_K c_K/f_K / SEP_THETA, exp(mean log(c_K/f_K))), in
    // FP64 from the level-precision coefficient.
    template <typename T>
    __global__ void __launch_bounds__(CNT)
    separable_fit_kernel(const T *__restrict__ coef, const int n,
                         int *__restrict__ bins, T *__restrict__ sigma)
    {
      const int       m  = n - 1;
      const long long np = (long long)m * m * m;
      for (long long q = (long long)blockIdx.x * CNT + threadIdx.x; q < np;
           q += (long long)gridDim.x * CNT)
        {
          const int px = (int)(q % m), py = (int)((q / m) % m), pz = (int)(q / ((long long)m * m));
          double    lc[8], lf[8];
#pragma unroll
          for (int c = 0; c < 8; ++c)
            {
              lc[c] = log((double)coef[(px + (c & 1)) + (long long)n * (py + ((c >> 1) & 1)) +
                                       (long long)n * n * (pz + (c >> 2))]);
              lf[c] = 0.0;
            }
          int packed = 0;
#pragma unroll
          for (int a = 0; a < 3; ++a)
            {
              double d = 0.0;
#pragma unroll
              for (int c = 0; c < 8; ++c)
                d += ((c >> a) & 1) ? lc[c] : -lc[c];
              const double qq = 0.25 * d / SEP_STEP;
              int          bq = (int)(qq >= 0.0 ? floor(qq + 0.5) : -floor(-qq + 0.5));
              bq              = bq < -SEP_HALF ? -SEP_HALF : (bq > SEP_HALF ? SEP_HALF : bq);
              packed |= (bq + SEP_HALF) << (8 * a);
#pragma unroll
              for (int c = 0; c < 8; ++c)
                if ((c >> a) & 1)
                  lf[c] += bq * SEP_STEP;
            }
          double rmax = lc[0] - lf[0], rsum = 0.0;
#pragma unroll
          for (int c = 0; c < 8; ++c)
            {
              rmax = fmax(rmax, lc[c] - lf[c]);
              rsum += lc[c] - lf[c];
            }
          bins[q]  = packed;
          sigma[q] = (T)fmax(exp(rmax) / SEP_THETA, exp(0.125 * rsum));
        }
    }

    // One vertex patch per work-group.  Only the residual at the patch's
    // interior DoFs is needed, and along each direction those are K of a
    // cell's K+1 local nodes: 1..K in the lower cell, 0..K-1 in the upper
    // one.  Each cell pass therefore forms K outputs per line, the y pass
    // runs only over the interior z outputs and the x pass over the
    // interior (y, z) ones.  The lines of a pass are split into two halves
    // of 128 threads by the cell offset along the pass direction, so that
    // offset, and with it the coefficient row, is uniform per wavefront
    // (scalar loads).  The patch's fit (ratio bins, scale) is read with the
    // first loads.
    template <int K, typename T>
    __global__ void __launch_bounds__(CNT)
    patch_coeff_kernel(T *__restrict__ u, const T *__restrict__ bvec,
                       const T *__restrict__ coef, const T *__restrict__ tab,
                       const T *__restrict__ sep, const int *__restrict__ fit_bins,
                       const T *__restrict__ fit_sigma, const int n, const int colour,
                       const T scale, const T relax, const FastDiv div_x,
                       const FastDiv div_xy)
    {
      using TB         = Tab<K>;
      using C          = PC<K, T>;
      constexpr int NN = C::NN, N2 = C::N2, N3 = C::N3, NP = C::NP, NI = C::NI;
      constexpr int NL = NI * NI, SW = SEP_W(K), HALF = CNT / 2;
      __shared__ T sbox[NP * NP * NP];
      __shared__ T sa[8 * N3], sb[8 * N3];
      __shared__ T sr[NI * NI * NI];
      __shared__ T sc[8];
      const T     *cM   = tab + TB::cellM;
      const T     *cK   = tab + TB::cellK;
      const int    tid  = threadIdx.x;
      // level sizes < 2^31 (host check): offsets within the patch box are
      // 32-bit, from the box's first DoF
      const int    N1   = K * n + 1;
      const int    P    = N1 * N1;
      const int px = colour & 1, py = (colour >> 1) & 1, pz = (colour >> 2) & 1;
      // patch coordinates: divisions by the patch counts as multiply-high
      // and shift (divisors set up on the host)
      const int q  = xcd_logical_id(blockIdx.x, gridDim.x);
      const int qz = (int)fastdiv((unsigned)q, div_xy), rq = q - qz * (int)div_xy.d;
      const int qy = (int)fastdiv((unsigned)rq, div_x), qx = rq - qy * (int)div_x.d;
      const int c0x = 2 * qx + px, c0y = 2 * qy + py, c0z = 2 * qz + pz;
      const int first = K * c0x + K * c0y * N1 + K * c0z * P;
      T *__restrict__ ub       = u + first;
      const T *__restrict__ bp = bvec + first;
      // this patch's separable fit (computed with the coefficient): ratio
      // bins and scale, uniform over the work-group
      const long long pq  = c0x + (long long)(n - 1) * (c0y + (long long)(n - 1) * c0z);
      const int       pk  = fit_bins[pq];
      const T         sig = fit_sigma[pq];
      // the 8 cell coefficients: loaded here, stored to LDS with the box
      // (a store now would wait for this load before the box loads issue)
      const int tc = tid < 8 ? tid : 0;
      const T   cv = coef[(c0x + (tc & 1)) + (long long)n * (c0y + ((tc >> 1) & 1)) +
                        (long long)n * n * (c0z + (tc >> 2))];
      // the box of u and this thread's interior b values: every load issued
      // before the first wait
      constexpr int ITU = (NP * NP * NP + CNT - 1) / CNT, ITR = (NI * NI * NI + CNT - 1) / CNT;
      T             bres[ITR];
      // the fit's per-direction eigenbases and eigenvalues (bins uniform
      // over the work-group: scalar table loads in the FD passes)
      const T *Sx = sep + (pk & 255) * SW;
      const T *Sy = sep + ((pk >> 8) & 255) * SW;
      const T *Sz = sep + ((pk >> 16) & 255) * SW;
      // lines of an FD pass per lane, wavefronts, outputs per wavefront
      constexpr int NW = CNT / 64, NCH = (NL + 63) / 64, OPW = (NI + NW - 1) / NW;
      const int     lane = tid & 63;
      const int     wv   = __builtin_amdgcn_readfirstlane(tid >> 6);
      T             lxy[NCH];
      {
        T vu[ITU];
#pragma unroll
        for (int it = 0; it < ITU; ++it)
          {
            const int t  = tid + it * CNT;
            const int ix = t % NP, iy = (t / NP) % NP, iz = t / (NP * NP);
            vu[it] = t < NP * NP * NP ? ub[ix + iy * N1 + iz * P] : T(0);
          }
#pragma unroll
        for (int it = 0; it < ITR; ++it)
          {
            const int t = tid + it * CNT;
            const int i = t % NI, j = (t / NI) % NI, l = t / (NI * NI);
            bres[it]    = t < NI * NI * NI
                            ? bp[(i + 1) + (j + 1) * N1 + (l + 1) * P]
                            : T(0);
          }
        // this lane's x + y eigenvalue sums for the z-line pass (z-line
        // t = (j, i))
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch)
          {
            const int t = ch * 64 + lane;
            lxy[ch]     = t < NL ? Sx[NL + t % NI] + Sy[NL + t / NI] : T(0);
          }
#pragma unroll
        for (int it = 0; it < ITU; ++it)
          if (tid + it * CNT < NP * NP * NP)
            sbox[tid + it * CNT] = vu[it];
        if (tid < 8)
          sc[tid] = cv;
      }
      __syncthreads();
      const int h  = __builtin_amdgcn_readfirstlane(tid / HALF);  // wave-uniform
      const int hr = tid - h * HALF;
      // cell z-rows: task (cell with z offset h, ix, iy), z outputs lz = m + 1 - h
      for (int t = hr; t < 4 * N2; t += HALF)
        {
          const int cl = t / N2, r = t - cl * N2;
          const int ox = cl & 1, oy = cl >> 1, c = cl + 4 * h;
          const int ix = r % NN, iy = r / NN;
          T         v[NN];
#pragma unroll
          for (int iz = 0; iz < NN; ++iz)
            v[iz] = sbox[((h * K + iz) * NP + oy * K + iy) * NP + ox * K + ix];
#pragma unroll
          for (int m = 0; m < K; ++m)
            {
              const int lz = m + 1 - h;
              T         a = 0, b = 0;
#pragma unroll
              for (int iz = 0; iz < NN; ++iz)
                {
                  a += cM[lz * NN + iz] * v[iz];
                  b += cK[lz * NN + iz] * v[iz];
                }
              sa[c * N3 + (lz * NN + iy) * NN + ix] = a;
              sb[c * N3 + (lz * NN + iy) * NN + ix] = b;
            }
        }
      __syncthreads();
      // cell y-rows in place: task (cell with y offset h, ix, interior lz),
      // y outputs ly = m + 1 - h
      for (int t = hr; t < 4 * NN * K; t += HALF)
        {
          const int cl = t / (NN * K), r = t - cl * (NN * K);
          const int ox = cl & 1, oz = cl >> 1, c = ox + 2 * h + 4 * oz;
          const int ix = r % NN, lz = r / NN + 1 - oz;
          T        *ca = sa + c * N3, *cb = sb + c * N3;
          T         a[NN], b[NN];
#pragma unroll
          for (int iy = 0; iy < NN; ++iy)
            {
              a[iy] = ca[(lz * NN + iy) * NN + ix];
              b[iy] = cb[(lz * NN + iy) * NN + ix];
            }
#pragma unroll
          for (int m = 0; m < K; ++m)
            {
              const int ly = m + 1 - h;
              T         c1 = 0, c2 = 0;
#pragma unroll
              for (int iy = 0; iy < NN; ++iy)
                {
                  c1 += cM[ly * NN + iy] * a[iy];
                  c2 += cK[ly * NN + iy] * a[iy];
                  c2 += cM[ly * NN + iy] * b[iy];
                }
              ca[(lz * NN + ly) * NN + ix] = c1;
              cb[(lz * NN + ly) * NN + ix] = c2;
            }
        }
      __syncthreads();
      // cell x-rows in place into sa: task (cell with x offset h, interior
      // ly, lz), x outputs lx = m + 1 - h, times the cell coefficient
      for (int t = hr; t < 4 * K * K; t += HALF)
        {
          const int cl = t / (K * K), r = t - cl * (K * K);
          const int oy = cl & 1, oz = cl >> 1, c = h + 2 * oy + 4 * oz;
          const int ly = r % K + 1 - oy, lz = r / K + 1 - oz;
          T        *ca = sa + c * N3, *cb = sb + c * N3;
          T         c1[NN], c2[NN];
#pragma unroll
          for (int ix = 0; ix < NN; ++ix)
            {
              c1[ix] = ca[(lz * NN + ly) * NN + ix];
              c2[ix] = cb[(lz * NN + ly) * NN + ix];
            }
          const T f = sc[c] * scale;
#pragma unroll
          for (int m = 0; m < K; ++m)
            {
              const int lx = m + 1 - h;
              T         y  = 0;
#pragma unroll
              for (int ix = 0; ix < NN; ++ix)
                {
                  y += cK[lx * NN + ix] * c1[ix];
                  y += cM[lx * NN + ix] * c2[ix];
                }
              ca[(lz * NN + ly) * NN + lx] = f * y;
            }
        }
      __syncthreads();
      // interior residual r = b - sum of the cells containing the DoF.  Per
      // direction the DoF at patch index p is node p of the lower cell
      // (p <= K) and node p - K of the upper one (p >= K); summed over the 8
      // cells in cell order with the absent ones masked, each read at a
      // fixed offset from the DoF's index in the lower cell (x + y NN + z N2,
      // within the 8-cell buffer for every cell).
#pragma unroll
      for (int it = 0; it < ITR; ++it)
        {
          const int t = tid + it * CNT;
          if (t >= NI * NI * NI)
            break;
          const int  i = t % NI + 1, j = (t / NI) % NI + 1, l = t / (NI * NI) + 1;
          const int  base = (l * NN + j) * NN + i;
          const bool lo[3] = {i <= K, j <= K, l <= K}, hi[3] = {i >= K, j >= K, l >= K};
          T          s = 0;
#pragma unroll
          for (int c = 0; c < 8; ++c)
            {
              const int  ox = c & 1, oy = (c >> 1) & 1, oz = c >> 2;
              const bool on = (ox ? hi[0] : lo[0]) && (oy ? hi[1] : lo[1]) && (oz ? hi[2] : lo[2]);
              const T    v  = sa[base + c * N3 - K * ((oz * NN + oy) * NN + ox)];
              s += on ? v : T(0);
            }
          sr[t] = bres[it] - s;
        }
      __syncthreads();
      // fast diagonalisation with the per-direction eigenpairs of the fit:
      // S_x^T, S_y^T, S_z^T with 1/(sigma scale (lx_i + ly_j + lz_l)), then
      // S_z, S_y, S_x.  In every pass the lanes run over the NI^2 lines and
      // wavefront w forms outputs w, w + NW, ... of each line, so all four
      // wavefronts share the pass and the matrix entries a wavefront
      // multiplies by are uniform (scalar loads); the passes ping-pong
      // between sr and sa (free after the gather).
      const T inv = T(1) / (scale * sig);
      // D = 0 (x): line (l, j), D = 1 (y): line (l, i), D = 2 (z): line (j, i)
      auto line_base = [&](int D, int t) {
        return D == 0 ? t * NI : (D == 1 ? (t / NI) * NL + t % NI : t);
      };
      auto fd_pass = [&](const T *in, T *out, auto dirc, auto trc, auto scc, const T *Sm) {
        constexpr int  D     = decltype(dirc)::value;
        constexpr bool TR    = decltype(trc)::value;
        constexpr bool SCALE = decltype(scc)::value;
        constexpr int  str   = D == 0 ? 1 : (D == 1 ? NI : NL);
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch)
          {
            const int t = ch * 64 + lane;
            if (t >= NL)
              continue;
            const int base = line_base(D, t);
            T         v[NI];
#pragma unroll
            for (int i = 0; i < NI; ++i)
              v[i] = in[base + i * str];
#pragma unroll
            for (int m = 0; m < OPW; ++m)
              {
                const int q = wv + NW * m;
                if (q >= NI)
                  break;
                T acc = 0;
#pragma unroll
                for (int i = 0; i < NI; ++i)
                  acc += (TR ? Sm[i * NI + q] : Sm[q * NI + i]) * v[i];
                if constexpr (SCALE)
                  {
                    // FP32: hardware reciprocal (1 ulp); FP64: exact division
                    if constexpr (sizeof(T) == 4)
                      acc = acc * inv * __builtin_amdgcn_rcpf(lxy[ch] + Sm[NL + q]);
                    else
                      acc = acc * inv / (lxy[ch] + Sm[NL + q]);
                  }
                out[base + q * str] = acc;
              }
          }
        __syncthreads();
      };
      using I0 = std::integral_constant<int, 0>;
      using I1 = std::integral_constant<int, 1>;
      using I2 = std::integral_constant<int, 2>;
      using BT = std::true_type;
      using BF = std::false_type;
      fd_pass(sr, sa, I0{}, BT{}, BF{}, Sx);
      fd_pass(sa, sr, I1{}, BT{}, BF{}, Sy);
      fd_pass(sr, sa, I2{}, BT{}, BT{}, Sz);
      fd_pass(sa, sr, I2{}, BF{}, BF{}, Sz);
      fd_pass(sr, sa, I1{}, BF{}, BF{}, Sy);
      // S_x and the store: output i of x-line (l, j) is DoF (i + 1, j + 1,
      // l + 1) of the patch box
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch)
        {
          const int t = ch * 64 + lane;
          if (t >= NL)
            continue;
          const int j = t % NI, l = t / NI;
          T         v[NI];
#pragma unroll
          for (int jj = 0; jj < NI; ++jj)
            v[jj] = sa[t * NI + jj];
          const int g0 = 1 + (j + 1) * N1 + (l + 1) * P;
#pragma unroll
          for (int m = 0; m < OPW; ++m)
            {
              const int i = wv + NW * m;
              if (i >= NI)
                break;
              T s = 0;
#pragma unroll
              for (int jj = 0; jj < NI; ++jj)
                s += Sx[i * NI + jj] * v[jj];
              const T uold = sbox[((l + 1) * NP + j + 1) * NP + i + 1];
              ub[g0 + i]   = uold + relax * s;
            }
        }
    }

    __global__ void __launch_bounds__(CNT)
    coarsen_kernel(const int nc, const double *__restrict__ fine,
                   double *__restrict__ coarse)
    {
      const long long ncc = (long long)nc * nc * nc;
      const int       nf  = 2 * nc;
      for (long long c = (long long)blockIdx.x * CNT + threadIdx.x; c < ncc;
           c += (long long)gridDim.x * CNT)
        {
          const int cx = (int)(c % nc), cy = (int)((c / nc) % nc),
                    cz = (int)(c / ((long long)nc * nc));
          double s = 0.0;
          for (int q = 0; q < 8; ++q)
            {
              const int fx = 2 * cx + (q & 1), fy = 2 * cy + ((q >> 1) & 1),
                        fz = 2 * cz + (q >> 2);
              s += fine[fx + (long long)nf * fy + (long long)nf * nf * fz];
            }
          coarse[c] = 0.125 * s;
        }
    }

    int
    grid_cap(long long work, int cap)
    {
      const long long nb = (work + CNT - 1) / CNT;
      return (int)std::max<long long>(1, std::min<long long>(nb, cap));
    }

    template <int K, typename T>
    void
    cells_k(const gtps_level *lvl, T *out, const T *x, const T *lam, int mode,
            bool mask)
    {
      const long long ncells = (long long)lvl->n * lvl->n * lvl->n;
      constexpr int   CPB    = cells_per_block(K);
      const long long nb     = (ncells + CPB - 1) / CPB;
      GTPS_REQUIRE(lvl->N < (1LL << 31) && ncells * (K + 1) * (K + 1) * (K + 1) < (1LL << 31),
                   "mesh too large for the cell kernel");
      const T      *tab = (const T *)lvl->dtab();
      const FastDiv dn  = make_fastdiv((unsigned)lvl->n);
      const FastDiv dnn = make_fastdiv((unsigned)(lvl->n * lvl->n));
      if (mode == 0)
        hipLaunchKernelGGL((cell_kernel<K, T, 0>), dim3((unsigned)nb), dim3(CNT),
                           0, lvl->ctx->stream, x, nullptr,
                           (const T *)lvl->coef, out, tab, lvl->n, lvl->N1,
                           (T)lvl->scale, mask ? 1 : 0, dn, dnn);
      else
        hipLaunchKernelGGL((cell_kernel<K, T, 1>), dim3((unsigned)nb), dim3(CNT),
                           0, lvl->ctx->stream, x, lam, nullptr, out, tab,
                           lvl->n, lvl->N1, (T)lvl->scale, 0, dn, dnn);
      GTPS_LAUNCH_CHECK();
    }

    template <int K, typename T>
    int
    apply_k(const gtps_level *lvl, T *y, const T *x, const T *b, bool dot,
            bool transpose)
    {
      T *E = (T *)lvl->celem;
      cells_k<K, T>(lvl, E, x, nullptr, 0, transpose);
      hipStream_t s   = lvl->ctx->stream;
      double     *pt  = lvl->ctx->ws.partials;
      const int   nb  = grid_cap(lvl->N, dot ? std::min(lvl->ctx->ws.capacity, 2048) : 8192);
      GTPS_REQUIRE(lvl->N < (1LL << 31) &&
                     (long long)lvl->n * lvl->n * lvl->n * (K + 1) * (K + 1) * (K + 1) < (1LL << 31),
                   "mesh too large for the gather kernel");
      const FastDiv dn1 = make_fastdiv((unsigned)lvl->N1);
      const FastDiv dp  = make_fastdiv((unsigned)(lvl->N1 * lvl->N1));
#define GTPS_G(R, TR, D)                                                      \
  hipLaunchKernelGGL((gather_kernel<K, T, R, TR, D>), dim3(nb), dim3(CNT), 0, \
                     s, (int)lvl->N, lvl->N1, lvl->n, (const T *)E, x, b, y, pt, dn1, dp)
      if (transpose)
        {
          GTPS_REQUIRE(b == nullptr && !dot, "transpose form has no residual/dot");
          GTPS_G(false, true, false);
        }
      else if (b && dot)
        GTPS_G(true, false, true);
      else if (b)
        GTPS_G(true, false, false);
      else if (dot)
        GTPS_G(false, false, true);
      else
        GTPS_G(false, false, false);
#undef GTPS_G
      GTPS_LAUNCH_CHECK();
      return dot ? nb : 0;
    }

    template <int K, typename T>
    void
    patch_k(const gtps_level *lvl, T *u, const T *b, int colour, double relax)
    {
      if constexpr (PC<K, T>::ok)
        {
          const int px = colour & 1, py = (colour >> 1) & 1, pz = colour >> 2;
          const int n  = lvl->n;
          const long long np =
            (long long)((n - px) / 2) * ((n - py) / 2) * ((n - pz) / 2);
          if (np == 0)
            return;
          GTPS_REQUIRE(np < (1LL << 31) && lvl->N < (1LL << 31),
                       "mesh too large for the patch kernel");
          GTPS_REQUIRE(lvl->sep_bins && lvl->sep_sigma, "level has no separable fit");
          const T *sep = (const T *)get_sep_tables(lvl->ctx, K, lvl->prec);
          const int npx = (n - px) / 2, npy = (n - py) / 2;
          hipLaunchKernelGGL((patch_coeff_kernel<K, T>), dim3((unsigned)np),
                             dim3(CNT), 0, lvl->ctx->stream, u, b,
                             (const T *)lvl->coef, (const T *)lvl->dtab(), sep,
                             (const int *)lvl->sep_bins, (const T *)lvl->sep_sigma, n,
                             colour, (T)lvl->scale, (T)relax, make_fastdiv((unsigned)npx),
                             make_fastdiv((unsigned)(npx * npy)));
          GTPS_LAUNCH_CHECK();
        }
      else
        throw Error(GTPS_EINVAL, "the variable-coefficient smoother supports "
                                 "degree <= 5 in FP64 and <= 7 in FP32");
    }

    template <typename T, typename F>
    auto
    dispatch(int k, F &&f)
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
  } // namespace

  bool
  coeff_smoother_supported(int k, int prec)
  {
    return prec == GTPS_FP32 ? k <= 7 : k <= 5;
  }

  int
  launch_coeff_apply(const gtps_level *lvl, void *y, const void *x,
                     const void *b, bool dot, bool transpose)
  {
    GTPS_REQUIRE(lvl->dim == 3, "the variable-coefficient operator is 3D");
    if (lvl->prec == GTPS_FP32)
      return dispatch<float>(lvl->k, [&](auto kk) {
        return apply_k<decltype(kk)::value, float>(lvl, (float *)y, (const float *)x,
                                                   (const float *)b, dot, transpose);
      });
    return dispatch<double>(lvl->k, [&](auto kk) {
      return apply_k<decltype(kk)::value, double>(lvl, (double *)y, (const double *)x,
                                                  (const double *)b, dot, transpose);
    });
  }

  void
  launch_coeff_gradient(const gtps_level *lvl, void *grad, const void *lambda,
                        const void *u)
  {
    GTPS_REQUIRE(lvl->dim == 3, "the variable-coefficient operator is 3D");
    if (lvl->prec == GTPS_FP32)
      dispatch<float>(lvl->k, [&](auto kk) {
        cells_k<decltype(kk)::value, float>(lvl, (float *)grad, (const float *)u,
                                            (const float *)lambda, 1, false);
        return 0;
      });
    else
      dispatch<double>(lvl->k, [&](auto kk) {
        cells_k<decltype(kk)::value, double>(lvl, (double *)grad, (const double *)u,
                                             (const double *)lambda, 1, false);
        return 0;
      });
  }

  void
  launch_patch_coeff_colour(const gtps_level *lvl, void *u, const void *b,
                            int colour, double relax)
  {
    GTPS_REQUIRE(lvl->dim == 3 && lvl->coef, "level has no 3D coefficient");
    if (lvl->prec == GTPS_FP32)
      dispatch<float>(lvl->k, [&](auto kk) {
        patch_k<decltype(kk)::value, float>(lvl, (float *)u, (const float *)b,
                                            colour, relax);
        return 0;
      });
    else
      dispatch<double>(lvl->k, [&](auto kk) {
        patch_k<decltype(kk)::value, double>(lvl, (double *)u, (const double *)b,
                                             colour, relax);
        return 0;
      });
  }

  void
  launch_separable_fit(const gtps_level *lvl)
  {
    const long long m  = lvl->n - 1;
    const long long np = m * m * m;
    if (np <= 0)
      return;
    const unsigned grid = (unsigned)std::min<long long>((np + CNT - 1) / CNT, 4096);
    if (lvl->prec == GTPS_FP32)
      hipLaunchKernelGGL(separable_fit_kernel<float>, dim3(grid), dim3(CNT), 0,
                         lvl->ctx->stream, (const float *)lvl->coef, lvl->n, lvl->sep_bins,
                         (float *)lvl->sep_sigma);
    else
      hipLaunchKernelGGL(separable_fit_kernel<double>, dim3(grid), dim3(CNT), 0,
                         lvl->ctx->stream, (const double *)lvl->coef, lvl->n, lvl->sep_bins,
                         (double *)lvl->sep_sigma);
    GTPS_LAUNCH_CHECK();
  }

  void
  launch_coarsen_coefficient(hipStream_t s, int nc, const double *fine,
                             double *coarse)
  {
    const long long ncc = (long long)nc * nc * nc;
    hipLaunchKernelGGL(coarsen_kernel, dim3(grid_cap(ncc, 8192)), dim3(CNT), 0,
                       s, nc, fine, coarse);
    GTPS_LAUNCH_CHECK();
  }
} // namespace gtps
