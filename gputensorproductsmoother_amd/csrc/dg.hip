// dg.hip - symmetric interior penalty DG (FE_DGQ on Gauss-Lobatto points)
// on a uniform Cartesian mesh of the unit cube, 3D.
//
// The reference has the operator only as a print-only matrix-free test
// (tests/mf_dg.cu:63-274: cell loop, interior-face loop with
// (sigma [u] - {du/dn}) [v] - [u] {dv/dn}, boundary-face loop with
// (2 sigma u - du/dn) v - u dv/dn, sigma = pf / h) and no DG smoother or
// solver.  On a uniform Cartesian mesh every one of those terms is a tensor
// product, so the operator is exactly
//     A = h^(d-2) (L (x) M (x) M + M (x) L (x) M + M (x) M (x) L)
// with M the block-diagonal 1D DG mass and L the block-tridiagonal 1D SIPG
// matrix (cell stiffness + 1D face terms; its diagonal block differs on the
// first and last cell).  DoFs are stored as one tensor grid
// X = cx (k+1) + ix (x fastest), which keeps every 1D operation a
// contiguous or constant-stride line.
//
// Kernels:
// * operator: three axis passes (z: M x, L x; y: M., L. + M.; x: L. + M.)
//   with one thread per output DoF (coalesced along x);
// * vertex-patch smoother: patch = 2x2x2 cells around an interior vertex,
//   all (2k+2)^3 DoFs; colours = patch index mod 3 per direction (27), so
//   patches of one colour share no face and one colour is race-free.  Local
//   solve u_P = A_PP^{-1}(b_P - A_P,ext u_ext): A_PP is the Kronecker sum
//   of the restricted 1D matrices, inverted exactly by fast diagonalisation
//   (FP64 generalized eigenpairs on the GPU, four boundary variants per
//   direction); A_P,ext u_ext couples the patch to the six face-neighbour
//   cell slabs through the off-diagonal blocks of L and the cell-local mass;
// * transfer: cell-wise embedding of the coarse polynomial at the child
//   nodes (no shared DoFs, no weights); restriction is its transpose.
#include "gtps_internal.h"

namespace gtps
{
  namespace
  {
    constexpr int DNT = 256;

    // packed DG table offsets (NN = k+1, PD = 2 NN)
    template <int K>
    struct DTab
    {
      static constexpr int NN = K + 1, PD = 2 * NN;
      static constexpr int M  = 0;              // cell mass [NN][NN]
      static constexpr int Lm = M + NN * NN;    // block (c, c-1)
      static constexpr int Lp = Lm + NN * NN;   // block (c, c+1)
      static constexpr int Lc = Lp + NN * NN;   // 3 diagonal blocks:
                                                // interior, first, last cell
      static constexpr int S   = Lc + 3 * NN * NN;  // 4 patch variants [PD][PD]
      static constexpr int lam = S + 4 * PD * PD;   // 4 x [PD]
      static constexpr int E   = lam + 4 * PD;      // child embedding [2][NN][NN]
      static constexpr int size = E + 2 * NN * NN;
    };

    __host__ __device__ inline int
    dg_cell_variant(int c, int n)
    {
      return c == 0 ? 1 : (c == n - 1 ? 2 : 0);
    }

    // (L in)[X] along a line with stride st (cells of NN DoFs)
    template <int K, typename T>
    __device__ __forceinline__ T
    dg_L_line(const T *tab, const T *in, long long base, long long st, int X,
              int n)
    {
      using TB     = DTab<K>;
      constexpr int NN = TB::NN;
      const int c = X / NN, i = X - c * NN;
      const T  *Lc = tab + TB::Lc + dg_cell_variant(c, n) * NN * NN + i * NN;
      const T  *in0 = in + base + (long long)c * NN * st;
      T         s   = 0;
#pragma unroll
      for (int j = 0; j < NN; ++j)
        s += Lc[j] * in0[j * st];
      if (c > 0)
        {
          const T *Lm = tab + TB::Lm + i * NN;
#pragma unroll
          for (int j = 0; j < NN; ++j)
            s += Lm[j] * in0[(j - NN) * st];
        }
      if (c < n - 1)
        {
          const T *Lp = tab + TB::Lp + i * NN;
#pragma unroll
          for (int j = 0; j < NN; ++j)
            s += Lp[j] * in0[(j + NN) * st];
        }
      return s;
    }

    template <int K, typename T>
    __device__ __forceinline__ T
    dg_M_line(const T *tab, const T *in, long long base, long long st, int X)
    {
      using TB     = DTab<K>;
      constexpr int NN = TB::NN;
      const int c = X / NN, i = X - c * NN;
      const T  *M   = tab + TB::M + i * NN;
      const T  *in0 = in + base + (long long)c * NN * st;
      T         s   = 0;
#pragma unroll
      for (int j = 0; j < NN; ++j)
        s += M[j] * in0[j * st];
      return s;
    }

    // PASS 0 (z): o1 = M x, o2 = L x;  PASS 1 (y): o1 = M i1, o2 = L i1 + M i2
    // PASS 2 (x): y = scale (L i1 + M i2), residual / dot forms
    template <int K, typename T, int PASS, bool RESID, bool DOT>
    __global__ void __launch_bounds__(DNT)
    dg_op_kernel(const T *__restrict__ i1, const T *__restrict__ i2,
                 T *__restrict__ o1, T *__restrict__ o2,
                 const T *__restrict__ tab, const T *__restrict__ b,
                 const T *__restrict__ x0, const int n, const int N1,
                 const long long N, const T scale,
                 double *__restrict__ partials)
    {
      __shared__ double red[DNT / 64];
      double            dot = 0.0;
      const long long   P   = (long long)N1 * N1;
      for (long long g = (long long)blockIdx.x * DNT + threadIdx.x; g < N;
           g += (long long)gridDim.x * DNT)
        {
          const int X = (int)(g % N1), Y = (int)((g / N1) % N1), Z = (int)(g / P);
          if (PASS == 0)
            {
              const long long base = X + (long long)Y * N1;
              o1[g] = dg_M_line<K>(tab, i1, base, P, Z);
              o2[g] = dg_L_line<K>(tab, i1, base, P, Z, n);
            }
          else if (PASS == 1)
            {
              const long long base = X + (long long)Z * P;
              o1[g] = dg_M_line<K>(tab, i1, base, N1, Y);
              o2[g] = dg_L_line<K>(tab, i1, base, N1, Y, n) +
                      dg_M_line<K>(tab, i2, base, N1, Y);
            }
          else
            {
              const long long base = (long long)Y * N1 + (long long)Z * P;
              T v = scale * (dg_L_line<K>(tab, i1, base, 1, X, n) +
                             dg_M_line<K>(tab, i2, base, 1, X));
              if (RESID)
                v = b[g] - v;
              o1[g] = v;
              if (DOT)
                dot += (double)x0[g] * (double)v;
            }
        }
      if (DOT)
        {
          dot = block_sum<DNT>(dot, red);
          if (threadIdx.x == 0)
            partials[blockIdx.x] = dot;
        }
    }

    template <int K, typename T>
    struct DP
    {
      static constexpr int NN = K + 1, PD = 2 * NN, P3 = PD * PD * PD;
      static constexpr int SL = PD * PD * NN;  // one slab
      static constexpr size_t lds = sizeof(T) * (2 * P3 + 2 * SL);
      static constexpr bool ok = lds <= 65536;
    };

    // A_P,ext u_ext for one direction D (0 x, 1 y, 2 z) and side (0: the
    // neighbour cell below the patch, 1: above): r -= scale (M (x) M) (Lx u)
    // on the patch cell next to that neighbour.  Patch arrays are [z][y][x]
    // with PD per direction; the slab s[a][b][j] has j along D and (a, b)
    // the other two directions (a the slower one).
    template <int K, typename T, int D, int SIDE>
    __device__ void
    dg_patch_ext(const T *__restrict__ u, const T *tab, T *r, T *s, T *w,
                 long long X0, long long Y0, long long Z0, int N1, T scale)
    {
      using TB           = DTab<K>;
      constexpr int NN   = TB::NN, PD = TB::PD, SL = PD * PD * NN;
      const long long P  = (long long)N1 * N1;
      const int       tid = threadIdx.x;
      // strides of (a, b, j) in the global vector and the D-offset of the slab
      const long long st[3] = {1, N1, P};
      constexpr int   da = D == 2 ? 1 : 2, db = D == 0 ? 1 : 0;
      const long long sa = st[da], sb = st[db], sj = st[D];
      const long long o0 = X0 + Y0 * N1 + Z0 * P +
                           (SIDE == 0 ? -(long long)NN : (long long)PD) * sj;
      for (int e = tid; e < SL; e += DNT)
        {
          const int j = e % NN, ab = e / NN, bb = ab % PD, aa = ab / PD;
          s[e] = u[o0 + aa * sa + bb * sb + j * sj];
        }
      __syncthreads();
      // w[a][b][i] = sum_j L?[i][j] s[a][b][j]
      const T *Lo = tab + (SIDE == 0 ? TB::Lm : TB::Lp);
      for (int e = tid; e < SL; e += DNT)
        {
          const int i = e % NN, ab = e / NN;
          T         v = 0;
#pragma unroll
          for (int j = 0; j < NN; ++j)
            v += Lo[i * NN + j] * s[ab * NN + j];
          w[e] = v;
        }
      __syncthreads();
      // M along b (cell-local)
      const T *M = tab + TB::M;
      for (int e = tid; e < SL; e += DNT)
        {
          const int i = e % NN, ab = e / NN, bb = ab % PD, aa = ab / PD;
          const int cb = bb / NN, ib = bb - cb * NN;
          T         v  = 0;
#pragma unroll
          for (int j = 0; j < NN; ++j)
            v += M[ib * NN + j] * w[(aa * PD + cb * NN + j) * NN + i];
          s[e] = v;
        }
      __syncthreads();
      // M along a, then subtract into the patch cell next to the slab
      for (int e = tid; e < SL; e += DNT)
        {
          const int i = e % NN, ab = e / NN, bb = ab % PD, aa = ab / PD;
          const int ca = aa / NN, ia = aa - ca * NN;
          T         v  = 0;
#pragma unroll
          for (int j = 0; j < NN; ++j)
            v += M[ia * NN + j] * s[((ca * NN + j) * PD + bb) * NN + i];
          const int pi = SIDE == 0 ? i : NN + i;  // patch index along D
          int       c3[3];
          c3[D]  = pi;
          c3[da] = aa;
          c3[db] = bb;
          r[(c3[2] * PD + c3[1]) * PD + c3[0]] -= scale * v;
        }
      __syncthreads();
    }

    // one pass of the FD transform along direction D: in place on r with
    // scratch t; TR: S^T (forward) or S (inverse)
    template <int K, typename T, int D, bool TR>
    __device__ void
    dg_fd_pass(T *r, T *t, const T *S)
    {
      constexpr int PD = DTab<K>::PD, P3 = PD * PD * PD;
      const int     tid = threadIdx.x;
      constexpr int str = D == 0 ? 1 : (D == 1 ? PD : PD * PD);
      for (int e = tid; e < P3; e += DNT)
        {
          const int c[3] = {e % PD, (e / PD) % PD, e / (PD * PD)};
          const int q    = c[D];
          const int base = e - q * str;
          T         v    = 0;
#pragma unroll
          for (int j = 0; j < PD; ++j)
            v += (TR ? S[j * PD + q] : S[q * PD + j]) * r[base + j * str];
          t[e] = v;
        }
      __syncthreads();
      for (int e = tid; e < P3; e += DNT)
        r[e] = t[e];
      __syncthreads();
    }

    template <int K, typename T>
    __global__ void __launch_bounds__(DNT)
    dg_patch_kernel(T *__restrict__ u, const T *__restrict__ bvec,
                    const T *__restrict__ tab, const int n, const int colour,
                    const T scale, const T relax)
    {
      using TB         = DTab<K>;
      using C          = DP<K, T>;
      constexpr int NN = C::NN, PD = C::PD, P3 = C::P3, SL = C::SL;
      __shared__ T r[P3], t[P3], s[SL], w[SL];
      const int       N1 = n * NN;
      const long long P  = (long long)N1 * N1;
      const int cx = colour % 3, cy = (colour / 3) % 3, cz = colour / 9;
      const int npx = (n - 1 - cx + 2) / 3, npy = (n - 1 - cy + 2) / 3;
      const int q   = xcd_logical_id(blockIdx.x, gridDim.x);
      const int qx = q % npx, qy = (q / npx) % npy, qz = q / (npx * npy);
      const int px = cx + 3 * qx, py = cy + 3 * qy, pz = cz + 3 * qz;
      const long long X0 = (long long)px * NN, Y0 = (long long)py * NN,
                      Z0 = (long long)pz * NN;
      const long long g0 = X0 + Y0 * N1 + Z0 * P;
      const int tid = threadIdx.x;
      for (int e = tid; e < P3; e += DNT)
        {
          const int x = e % PD, y = (e / PD) % PD, z = e / (PD * PD);
          r[e] = bvec[g0 + x + (long long)y * N1 + (long long)z * P];
        }
      __syncthreads();
      if (px > 0)
        dg_patch_ext<K, T, 0, 0>(u, tab, r, s, w, X0, Y0, Z0, N1, scale);
      if (px + 2 < n)
        dg_patch_ext<K, T, 0, 1>(u, tab, r, s, w, X0, Y0, Z0, N1, scale);
      if (py > 0)
        dg_patch_ext<K, T, 1, 0>(u, tab, r, s, w, X0, Y0, Z0, N1, scale);
      if (py + 2 < n)
        dg_patch_ext<K, T, 1, 1>(u, tab, r, s, w, X0, Y0, Z0, N1, scale);
      if (pz > 0)
        dg_patch_ext<K, T, 2, 0>(u, tab, r, s, w, X0, Y0, Z0, N1, scale);
      if (pz + 2 < n)
        dg_patch_ext<K, T, 2, 1>(u, tab, r, s, w, X0, Y0, Z0, N1, scale);
      auto variant = [&](int p) { return (p == 0 ? 1 : 0) + (p + 2 == n ? 2 : 0); };
      const int vx = variant(px), vy = variant(py), vz = variant(pz);
      const T  *Sx = tab + TB::S + vx * PD * PD, *Sy = tab + TB::S + vy * PD * PD,
               *Sz = tab + TB::S + vz * PD * PD;
      const T  *lx = tab + TB::lam + vx * PD, *ly = tab + TB::lam + vy * PD,
               *lz = tab + TB::lam + vz * PD;
      dg_fd_pass<K, T, 0, true>(r, t, Sx);
      dg_fd_pass<K, T, 1, true>(r, t, Sy);
      dg_fd_pass<K, T, 2, true>(r, t, Sz);
      for (int e = tid; e < P3; e += DNT)
        {
          const int x = e % PD, y = (e / PD) % PD, z = e / (PD * PD);
          r[e] /= scale * (lx[x] + ly[y] + lz[z]);
        }
      __syncthreads();
      dg_fd_pass<K, T, 2, false>(r, t, Sz);
      dg_fd_pass<K, T, 1, false>(r, t, Sy);
      dg_fd_pass<K, T, 0, false>(r, t, Sx);
      for (int e = tid; e < P3; e += DNT)
        {
          const int       x = e % PD, y = (e / PD) % PD, z = e / (PD * PD);
          const long long g = g0 + x + (long long)y * N1 + (long long)z * P;
          u[g]              = relax == T(1) ? r[e] : (T(1) - relax) * u[g] + relax * r[e];
        }
    }

    // fine += P coarse: one thread per fine DoF
    template <int K, typename T>
    __global__ void __launch_bounds__(DNT)
    dg_prolongate_kernel(const T *__restrict__ coarse, T *__restrict__ fine,
                         const T *__restrict__ tab, const int nf,
                         const long long Nf)
    {
      using TB          = DTab<K>;
      constexpr int NN  = TB::NN;
      const int       N1f = nf * NN, N1c = (nf / 2) * NN;
      const long long Pf = (long long)N1f * N1f, Pc = (long long)N1c * N1c;
      const T        *E  = tab + TB::E;
      for (long long g = (long long)blockIdx.x * DNT + threadIdx.x; g < Nf;
           g += (long long)gridDim.x * DNT)
        {
          const int X = (int)(g % N1f), Y = (int)((g / N1f) % N1f), Z = (int)(g / Pf);
          const int fx = X / NN, fy = Y / NN, fz = Z / NN;
          const T  *ex = E + ((fx & 1) * NN + X - fx * NN) * NN;
          const T  *ey = E + ((fy & 1) * NN + Y - fy * NN) * NN;
          const T  *ez = E + ((fz & 1) * NN + Z - fz * NN) * NN;
          const T  *c0 = coarse + (long long)(fx / 2) * NN + (long long)(fy / 2) * NN * N1c +
                        (long long)(fz / 2) * NN * Pc;
          T v = 0;
#pragma unroll
          for (int cz = 0; cz < NN; ++cz)
#pragma unroll
            for (int cy = 0; cy < NN; ++cy)
              {
                T a = 0;
#pragma unroll
                for (int cx = 0; cx < NN; ++cx)
                  a += ex[cx] * c0[cx + (long long)cy * N1c + (long long)cz * Pc];
                v += ez[cz] * ey[cy] * a;
              }
          fine[g] += v;
        }
    }

    // coarse (+)= P^T fine: one thread per coarse DoF
    template <int K, typename T, bool STORE>
    __global__ void __launch_bounds__(DNT)
    dg_restrict_kernel(const T *__restrict__ fine, T *__restrict__ coarse,
                       const T *__restrict__ tab, const int nf,
                       const long long Nc)
    {
      using TB          = DTab<K>;
      constexpr int NN  = TB::NN;
      const int       N1f = nf * NN, N1c = (nf / 2) * NN;
      const long long Pf = (long long)N1f * N1f, Pc = (long long)N1c * N1c;
      const T        *E  = tab + TB::E;
      for (long long g = (long long)blockIdx.x * DNT + threadIdx.x; g < Nc;
           g += (long long)gridDim.x * DNT)
        {
          const int X = (int)(g % N1c), Y = (int)((g / N1c) % N1c), Z = (int)(g / Pc);
          const int cxc = X / NN, cyc = Y / NN, czc = Z / NN;
          const int ax = X - cxc * NN, ay = Y - cyc * NN, az = Z - czc * NN;
          T         v  = 0;
          for (int chz = 0; chz < 2; ++chz)
            for (int chy = 0; chy < 2; ++chy)
              for (int chx = 0; chx < 2; ++chx)
                {
                  const T *f0 = fine + (long long)(2 * cxc + chx) * NN +
                                (long long)(2 * cyc + chy) * NN * N1f +
                                (long long)(2 * czc + chz) * NN * Pf;
#pragma unroll
                  for (int qz = 0; qz < NN; ++qz)
#pragma unroll
                    for (int qy = 0; qy < NN; ++qy)
                      {
                        T a = 0;
#pragma unroll
                        for (int qx = 0; qx < NN; ++qx)
                          a += E[(chx * NN + qx) * NN + ax] *
                               f0[qx + (long long)qy * N1f + (long long)qz * Pf];
                        v += E[(chz * NN + qz) * NN + az] * E[(chy * NN + qy) * NN + ay] * a;
                      }
                }
          coarse[g] = STORE ? v : coarse[g] + v;
        }
    }

    int
    dgrid(long long N, int cap = 8192)
    {
      return (int)std::max<long long>(1, std::min<long long>((N + DNT - 1) / DNT, cap));
    }

    template <typename F>
    auto
    dispatch_k(int k, F &&f)
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

    template <int K>
    std::vector<double>
    dg_pack(gtps_ctx *ctx, const Basis1D &b, double pf)
    {
      using TB         = DTab<K>;
      constexpr int NN = TB::NN, PD = TB::PD;
      std::vector<double> tab(TB::size, 0.0);
      std::vector<double> v0(NN), v1(NN), d0(NN), d1(NN);
      lagrange_values(b.nodes, 0.0, v0.data(), d0.data());
      lagrange_values(b.nodes, 1.0, v1.data(), d1.data());
      for (int i = 0; i < NN; ++i)
        for (int j = 0; j < NN; ++j)
          tab[TB::M + i * NN + j] = b.cellM[i * NN + j];
      // interior face between a minus cell (its x = 1 end) and a plus cell
      // (x = 0): jump [u] = u- - u+, average {u'} = (u-' + u+')/2; the face
      // contributes pf [u][v] - {u'}[v] - [u]{v'} (unit-scaled)
      auto face = [&](int si, int i, int sj, int j) {
        const double ti  = si == 0 ? v1[i] : -v0[i];
        const double dti = 0.5 * (si == 0 ? d1[i] : d0[i]);
        const double tj  = sj == 0 ? v1[j] : -v0[j];
        const double dtj = 0.5 * (sj == 0 ? d1[j] : d0[j]);
        return pf * tj * ti - dtj * ti - tj * dti;
      };
      for (int i = 0; i < NN; ++i)
        for (int j = 0; j < NN; ++j)
          {
            const double K0 = b.cellK[i * NN + j];
            // cell c's row i couples to c-1 (c is the plus side) and to c+1
            tab[TB::Lm + i * NN + j] = face(1, i, 0, j);
            tab[TB::Lp + i * NN + j] = face(0, i, 1, j);
            const double lo_int = face(1, i, 1, j);  // low face, interior
            const double hi_int = face(0, i, 0, j);  // high face, interior
            const double lo_bd  = 2 * pf * v0[i] * v0[j] + d0[j] * v0[i] + v0[j] * d0[i];
            const double hi_bd  = 2 * pf * v1[i] * v1[j] - d1[j] * v1[i] - v1[j] * d1[i];
            tab[TB::Lc + 0 * NN * NN + i * NN + j] = K0 + lo_int + hi_int;
            tab[TB::Lc + 1 * NN * NN + i * NN + j] = K0 + lo_bd + hi_int;
            tab[TB::Lc + 2 * NN * NN + i * NN + j] = K0 + lo_int + hi_bd;
          }
      // patch variants: 0 interior, 1 first cell at the low boundary,
      // 2 second cell at the high boundary, 3 both (n = 2)
      for (int v = 0; v < 4; ++v)
        {
          std::vector<double> L(PD * PD, 0.0), M(PD * PD, 0.0), lam, S;
          const int va = (v & 1) ? 1 : 0, vb = (v & 2) ? 2 : 0;
          for (int i = 0; i < NN; ++i)
            for (int j = 0; j < NN; ++j)
              {
                L[i * PD + j]                 = tab[TB::Lc + va * NN * NN + i * NN + j];
                L[(NN + i) * PD + NN + j]     = tab[TB::Lc + vb * NN * NN + i * NN + j];
                L[i * PD + NN + j]            = tab[TB::Lp + i * NN + j];
                L[(NN + i) * PD + j]          = tab[TB::Lm + i * NN + j];
                M[i * PD + j]                 = b.cellM[i * NN + j];
                M[(NN + i) * PD + NN + j]     = b.cellM[i * NN + j];
              }
          gpu_gen_eig(ctx, PD, L, M, lam, S);
          std::copy(S.begin(), S.end(), tab.begin() + TB::S + v * PD * PD);
          std::copy(lam.begin(), lam.end(), tab.begin() + TB::lam + v * PD);
        }
      // child embedding: fine node q of child c at x = (c + xi_q) / 2
      for (int c = 0; c < 2; ++c)
        for (int q = 0; q < NN; ++q)
          lagrange_values(b.nodes, 0.5 * (c + b.nodes[q]),
                          tab.data() + TB::E + (c * NN + q) * NN, nullptr);
      return tab;
    }

    template <int K, typename T>
    int
    apply_k(const gtps_dg_level *lvl, T *y, const T *x, const T *b, bool dot)
    {
      hipStream_t s   = lvl->ctx->stream;
      const T    *tab = (const T *)lvl->tab;
      T *t0 = (T *)lvl->t[0], *t1 = (T *)lvl->t[1], *t2 = (T *)lvl->t[2],
        *t3 = (T *)lvl->t[3];
      const int  nb = dgrid(lvl->N);
      const int  nbd = dgrid(lvl->N, std::min(lvl->ctx->ws.capacity, 2048));
      double    *pt = lvl->ctx->ws.partials;
      const T    sc = (T)lvl->scale;
      hipLaunchKernelGGL((dg_op_kernel<K, T, 0, false, false>), dim3(nb), dim3(DNT), 0,
                         s, x, nullptr, t0, t1, tab, nullptr, nullptr, lvl->n,
                         lvl->N1, lvl->N, sc, pt);
      hipLaunchKernelGGL((dg_op_kernel<K, T, 1, false, false>), dim3(nb), dim3(DNT), 0,
                         s, t0, t1, t2, t3, tab, nullptr, nullptr, lvl->n,
                         lvl->N1, lvl->N, sc, pt);
#define GTPS_DX(R, D, G)                                                      \
  hipLaunchKernelGGL((dg_op_kernel<K, T, 2, R, D>), dim3(G), dim3(DNT), 0, s, \
                     t2, t3, y, nullptr, tab, b, x, lvl->n, lvl->N1, lvl->N,  \
                     sc, pt)
      if (b && dot)
        GTPS_DX(true, true, nbd);
      else if (b)
        GTPS_DX(true, false, nb);
      else if (dot)
        GTPS_DX(false, true, nbd);
      else
        GTPS_DX(false, false, nb);
#undef GTPS_DX
      GTPS_LAUNCH_CHECK();
      return dot ? nbd : 0;
    }

    template <int K, typename T>
    void
    sweep_k(const gtps_dg_level *lvl, T *u, const T *b, int variant, double relax)
    {
      if constexpr (DP<K, T>::ok)
        {
          const int n = lvl->n;
          for (int cc = 0; cc < 27; ++cc)
            {
              const int c = variant == GTPS_MULT_REV ? 26 - cc : cc;
              const int cx = c % 3, cy = (c / 3) % 3, cz = c / 9;
              const long long np = (long long)((n - 1 - cx + 2) / 3) *
                                   ((n - 1 - cy + 2) / 3) * ((n - 1 - cz + 2) / 3);
              if (np <= 0 || cx >= n - 1 || cy >= n - 1 || cz >= n - 1)
                continue;
              hipLaunchKernelGGL((dg_patch_kernel<K, T>), dim3((unsigned)np), dim3(DNT),
                                 0, lvl->ctx->stream, u, b, (const T *)lvl->tab, n, c,
                                 (T)lvl->scale, (T)relax);
              GTPS_LAUNCH_CHECK();
            }
        }
      else
        throw Error(GTPS_EINVAL, "the DG smoother supports degree <= 5 in FP64 "
                                 "and <= 7 in FP32");
    }
  } // namespace

  bool
  dg_supported(int k, int prec)
  {
    return k >= 1 && (prec == GTPS_FP32 ? k <= 7 : k <= 5);
  }

  void
  dg_level_setup(gtps_dg_level *lvl)
  {
    const Basis1D &b   = get_tables(lvl->ctx, lvl->k)->basis;
    std::vector<double> t64 = dispatch_k(lvl->k, [&](auto kk) {
      return dg_pack<decltype(kk)::value>(lvl->ctx, b, lvl->pf);
    });
    hipStream_t s = lvl->ctx->stream;
    GTPS_HIP(hipMalloc(&lvl->tab, lvl->elem() * t64.size()));
    if (lvl->prec == GTPS_FP32)
      {
        std::vector<float> t32(t64.begin(), t64.end());
        GTPS_HIP(hipMemcpyAsync(lvl->tab, t32.data(), sizeof(float) * t32.size(),
                                hipMemcpyHostToDevice, s));
        GTPS_HIP(hipStreamSynchronize(s));
      }
    else
      {
        GTPS_HIP(hipMemcpyAsync(lvl->tab, t64.data(), sizeof(double) * t64.size(),
                                hipMemcpyHostToDevice, s));
        GTPS_HIP(hipStreamSynchronize(s));
      }
    for (void *&p : lvl->t)
      GTPS_HIP(hipMalloc(&p, lvl->elem() * lvl->N));
  }

  int
  launch_dg_apply(const gtps_dg_level *lvl, void *y, const void *x,
                  const void *b, bool fused_dot)
  {
    return dispatch_k(lvl->k, [&](auto kk) {
      constexpr int K = decltype(kk)::value;
      if (lvl->prec == GTPS_FP32)
        return apply_k<K, float>(lvl, (float *)y, (const float *)x,
                                 (const float *)b, fused_dot);
      return apply_k<K, double>(lvl, (double *)y, (const double *)x,
                                (const double *)b, fused_dot);
    });
  }

  void
  launch_dg_sweep(const gtps_dg_level *lvl, void *u, const void *b,
                  int variant, double relax)
  {
    GTPS_REQUIRE(lvl->n >= 2, "the DG smoother needs at least 2 cells per direction");
    dispatch_k(lvl->k, [&](auto kk) {
      constexpr int K = decltype(kk)::value;
      if (lvl->prec == GTPS_FP32)
        sweep_k<K, float>(lvl, (float *)u, (const float *)b, variant, relax);
      else
        sweep_k<K, double>(lvl, (double *)u, (const double *)b, variant, relax);
      return 0;
    });
  }

  void
  launch_dg_prolongate_add(const gtps_dg_level *fl, void *fine, const void *coarse)
  {
    GTPS_REQUIRE(fl->n % 2 == 0, "fine level must have an even cell count");
    dispatch_k(fl->k, [&](auto kk) {
      constexpr int K = decltype(kk)::value;
      if (fl->prec == GTPS_FP32)
        hipLaunchKernelGGL((dg_prolongate_kernel<K, float>), dim3(dgrid(fl->N)),
                           dim3(DNT), 0, fl->ctx->stream, (const float *)coarse,
                           (float *)fine, (const float *)fl->tab, fl->n, fl->N);
      else
        hipLaunchKernelGGL((dg_prolongate_kernel<K, double>), dim3(dgrid(fl->N)),
                           dim3(DNT), 0, fl->ctx->stream, (const double *)coarse,
                           (double *)fine, (const double *)fl->tab, fl->n, fl->N);
      GTPS_LAUNCH_CHECK();
      return 0;
    });
  }

  void
  launch_dg_restrict(const gtps_dg_level *fl, void *coarse, const void *fine,
                     bool store)
  {
    GTPS_REQUIRE(fl->n % 2 == 0, "fine level must have an even cell count");
    const long long N1c = (long long)(fl->n / 2) * (fl->k + 1);
    const long long Nc  = N1c * N1c * N1c;
    dispatch_k(fl->k, [&](auto kk) {
      constexpr int K = decltype(kk)::value;
      if (fl->prec == GTPS_FP32)
        {
          if (store)
            hipLaunchKernelGGL((dg_restrict_kernel<K, float, true>), dim3(dgrid(Nc)),
                               dim3(DNT), 0, fl->ctx->stream, (const float *)fine,
                               (float *)coarse, (const float *)fl->tab, fl->n, Nc);
          else
            hipLaunchKernelGGL((dg_restrict_kernel<K, float, false>), dim3(dgrid(Nc)),
                               dim3(DNT), 0, fl->ctx->stream, (const float *)fine,
                               (float *)coarse, (const float *)fl->tab, fl->n, Nc);
        }
      else
        {
          if (store)
            hipLaunchKernelGGL((dg_restrict_kernel<K, double, true>), dim3(dgrid(Nc)),
                               dim3(DNT), 0, fl->ctx->stream, (const double *)fine,
                               (double *)coarse, (const double *)fl->tab, fl->n, Nc);
          else
            hipLaunchKernelGGL((dg_restrict_kernel<K, double, false>), dim3(dgrid(Nc)),
                               dim3(DNT), 0, fl->ctx->stream, (const double *)fine,
                               (double *)coarse, (const double *)fl->tab, fl->n, Nc);
        }
      GTPS_LAUNCH_CHECK();
      return 0;
    });
  }
} // namespace gtps
