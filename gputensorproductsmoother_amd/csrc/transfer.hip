// transfer.hip - geometric MG transfer between level l-1 (coarse) and l.
//
// For continuous Q_k on the uniform hierarchy the prolongation is
// interpolation, P = P1 (x) P1 (x) P1 with the global 1D embedding P1 whose
// row for fine node j is row q = j - 2k cc of the (2k+1) x (k+1) embedding
// matrix E of coarse cell cc.  The reference assembles the same operator
// cell by cell with valence weights and atomics
// (cuda_mg_transfer.template.cuh:48-369, 573-621); restriction is its
// transpose R = P^T (261-344, 623-671).  Dirichlet DoFs of the coarse level
// are excluded on both sides (set_mg_constrained_dofs, .template.cuh:586,668).
//
// 3D prolongation (all k) and 3D restriction (k <= 3) run as one fused tile
// kernel each (below); 2D and the remaining restrictions run one pass per
// direction (x, then y, then z) through two level-sized temporaries owned by
// the transfer object.  Either way the last step accumulates into the
// destination, so no atomics are needed and results are deterministic.
#include "gtps_internal.h"

namespace gtps
{
  namespace
  {
    constexpr int TR_NT = 256;

    // OUT: 0 store, 1 add, 2 add / 3 store except on coarse Dirichlet nodes
    template <int K, typename T, bool PROLONG, bool MASK_IN, int OUT>
    __global__ void __launch_bounds__(TR_NT)
    transfer_pass_kernel(const T *__restrict__ in, T *__restrict__ out,
                         const T *__restrict__ tab, const int dim,
                         const int axis, const int e0, const int e1,
                         const int e2, const int i0, const int i1,
                         const int nc, const int Nc1)
    {
      using TB          = Tab<K>;
      constexpr int N   = K + 1;
      const T      *E   = tab + TB::E;
      const long long total = (long long)e0 * e1 * e2;
      const long long sin1  = i0, sin2 = (long long)i0 * i1;
      for (long long o = (long long)blockIdx.x * TR_NT + threadIdx.x; o < total;
           o += (long long)gridDim.x * TR_NT)
        {
          int       oc[3];
          oc[0]              = (int)(o % e0);
          const long long r  = o / e0;
          oc[1]              = (int)(r % e1);
          oc[2]              = (int)(r / e1);
          const int j        = oc[axis];
          int       ic[3]    = {oc[0], oc[1], oc[2]};
          ic[axis]           = 0;
          const long long base = ic[0] + ic[1] * sin1 + ic[2] * sin2;
          const long long s    = axis == 0 ? 1 : (axis == 1 ? sin1 : sin2);
          T acc = 0;
          if (PROLONG)
            {
              const int cc = min(j / (2 * K), nc - 1);
              const int q  = j - 2 * K * cc;
              bool      other_bdry = false;
              if (MASK_IN)
                {
                  // input is the coarse vector: nodes on the coarse
                  // Dirichlet boundary count as zero
                  for (int d = 0; d < dim; ++d)
                    if (d != axis && (oc[d] == 0 || oc[d] == Nc1 - 1))
                      other_bdry = true;
                }
              if (!other_bdry)
                {
#pragma unroll
                  for (int a = 0; a < N; ++a)
                    {
                      const int ia = K * cc + a;
                      if (MASK_IN && (ia == 0 || ia == Nc1 - 1))
                        continue;
                      acc += E[q * N + a] * in[base + ia * s];
                    }
                }
            }
          else
            {
              const int cc = j / K, a = j - K * cc;
              if (a != 0)
                {
#pragma unroll
                  for (int q = 1; q < 2 * K; ++q)
                    acc += E[q * N + a] * in[base + (2LL * K * cc + q) * s];
                }
              else
                {
                  acc = in[base + 2LL * K * cc * s];
                  if (cc < nc)
                    {
#pragma unroll
                      for (int q = 1; q < 2 * K; ++q)
                        acc += E[q * N + 0] * in[base + (2LL * K * cc + q) * s];
                    }
                  if (cc > 0)
                    {
#pragma unroll
                      for (int q = 1; q < 2 * K; ++q)
                        acc += E[q * N + K] *
                               in[base + (2LL * K * (cc - 1) + q) * s];
                    }
                }
            }
          if (OUT == 0)
            out[o] = acc;
          else if (OUT == 1)
            out[o] += acc;
          else
            {
              bool bdry = false;
              for (int d = 0; d < dim; ++d)
                if (oc[d] == 0 || oc[d] == Nc1 - 1)
                  bdry = true;
              if (!bdry)
                {
                  if (OUT == 2)
                    out[o] += acc;
                  else
                    out[o] = acc;
                }
            }
        }
    }

    // ------------------------------------------------------------------
    // Fused 3D kernels: one work-group owns a tile of CX x CY x CZ coarse
    // cells.  Prolongation stages the coarse tile (k C + 1 nodes per
    // direction, Dirichlet nodes as zero) in LDS, applies E along x, y, z in
    // LDS and adds the 2k C fine nodes per direction it owns into the fine
    // vector.  Restriction stages the fine window [2k c0 - 2k + 1,
    // 2k (c0 + C) - 1] per direction, applies the transposed 1D rows along
    // x, y, z and adds the k C coarse nodes per direction it owns into the
    // coarse vector (Dirichlet nodes skipped).
    template <int K>
    struct TT
    {
      static constexpr int CX = K <= 3 ? 4 : 2, CY = K <= 3 ? 2 : 1,
                           CZ = K <= 3 ? 2 : 1, NT = 256;
    };

    template <int K, typename T>
    __global__ void __launch_bounds__(256)
    prolongate3d_kernel(const T *__restrict__ coarse, T *__restrict__ fine,
                        const T *__restrict__ tab, const int nc)
    {
      using TB          = Tab<K>;
      constexpr int CX  = TT<K>::CX, CY = TT<K>::CY, CZ = TT<K>::CZ, NT = 256;
      constexpr int N   = K + 1;
      constexpr int EX  = K * CX + 1, EY = K * CY + 1, EZ = K * CZ + 1;
      constexpr int FX  = 2 * K * CX, FY = 2 * K * CY;
      __shared__ T sC[EZ * EY * EX], sX[EZ * EY * FX], sY[EZ * FY * FX];
      const T  *E   = tab + TB::E;
      const int tid = threadIdx.x;
      const int Nc1 = K * nc + 1, Nf1 = 2 * K * nc + 1;
      const long long Pc = (long long)Nc1 * Nc1, Pf = (long long)Nf1 * Nf1;
      const int L   = xcd_logical_id(linear_block_id(),
                                     gridDim.x * gridDim.y * gridDim.z);
      const int c0x = (L % gridDim.x) * CX, c0y = ((L / gridDim.x) % gridDim.y) * CY,
                c0z = (L / (gridDim.x * gridDim.y)) * CZ;
      // the fine values the z pass adds to (its columns (yf, xf), 2k CZ
      // planes each), loaded with the coarse block so that the
      // read-modify-write costs no round trip of its own
      constexpr int ITZ = (FY * FX + NT - 1) / NT, NZO = CZ * 2 * K;
      T             old[ITZ][NZO];
#pragma unroll
      for (int jz = 0; jz < ITZ; ++jz)
        {
          const int  t  = tid + jz * NT;
          const int  yf = t / FX, xf = t - yf * FX;
          const int  gx = 2 * K * c0x + xf, gy = 2 * K * c0y + yf;
          const bool in = t < FY * FX && gx < Nf1 - 1 && gy < Nf1 - 1;
          const T   *col = fine + (in ? gx + (long long)gy * Nf1 : 0);
#pragma unroll
          for (int o = 0; o < NZO; ++o)
            {
              const int gz = 2 * K * c0z + o;
              old[jz][o]   = in && gz < Nf1 - 1 ? col[gz * Pf] : T(0);
            }
        }
      {
        // coarse block in one round of loads (unrolled), then to LDS
        constexpr int TOT = EZ * EY * EX, IT = (TOT + NT - 1) / NT;
        T             v[IT];
#pragma unroll
        for (int j = 0; j < IT; ++j)
          {
            const int i = tid + j * NT;
            const int x = i % EX, y = (i / EX) % EY, z = i / (EX * EY);
            const int gx = K * c0x + x, gy = K * c0y + y, gz = K * c0z + z;
            const bool inner = i < TOT && gx > 0 && gx < Nc1 - 1 && gy > 0 &&
                               gy < Nc1 - 1 && gz > 0 && gz < Nc1 - 1;
            v[j] = inner ? coarse[gx + (long long)gy * Nc1 + gz * Pc] : T(0);
          }
#pragma unroll
        for (int j = 0; j < IT; ++j)
          {
            const int i = tid + j * NT;
            if (i < TOT)
              sC[i] = v[j];
          }
      }
      __syncthreads();
      // x: one coarse row (z, y) per thread, all 2k CX fine outputs unrolled
      for (int t = tid; t < EZ * EY; t += NT)
        {
          T r[EX];
#pragma unroll
          for (int i = 0; i < EX; ++i)
            r[i] = sC[t * EX + i];
#pragma unroll
          for (int cl = 0; cl < CX; ++cl)
#pragma unroll
            for (int q = 0; q < 2 * K; ++q)
              {
                T v = 0;
#pragma unroll
                for (int a = 0; a < N; ++a)
                  v += E[q * N + a] * r[K * cl + a];
                sX[t * FX + 2 * K * cl + q] = v;
              }
        }
      __syncthreads();
      // y: one column (z, xf) per thread
      for (int t = tid; t < EZ * FX; t += NT)
        {
          const int z = t / FX, xf = t - z * FX;
          T         r[EY];
#pragma unroll
          for (int i = 0; i < EY; ++i)
            r[i] = sX[(z * EY + i) * FX + xf];
#pragma unroll
          for (int cl = 0; cl < CY; ++cl)
#pragma unroll
            for (int q = 0; q < 2 * K; ++q)
              {
                T v = 0;
#pragma unroll
                for (int a = 0; a < N; ++a)
                  v += E[q * N + a] * r[K * cl + a];
                sY[(z * FY + 2 * K * cl + q) * FX + xf] = v;
              }
        }
      __syncthreads();
      // z: one column (yf, xf) per thread, accumulated into the fine vector
#pragma unroll
      for (int jz = 0; jz < ITZ; ++jz)
        {
          const int t = tid + jz * NT;
          if (t >= FY * FX)
            continue;
          const int yf = t / FX, xf = t - yf * FX;
          const int gx = 2 * K * c0x + xf, gy = 2 * K * c0y + yf;
          if (gx >= Nf1 - 1 || gy >= Nf1 - 1)
            continue;  // beyond the mesh or the last (Dirichlet) plane: +0
          T r[EZ];
#pragma unroll
          for (int i = 0; i < EZ; ++i)
            r[i] = sY[(i * FY + yf) * FX + xf];
          T *out = fine + gx + (long long)gy * Nf1;
#pragma unroll
          for (int cl = 0; cl < CZ; ++cl)
#pragma unroll
            for (int q = 0; q < 2 * K; ++q)
              {
                const int gz = 2 * K * (c0z + cl) + q;
                if (gz >= Nf1 - 1)
                  continue;
                T v = 0;
#pragma unroll
                for (int a = 0; a < N; ++a)
                  v += E[q * N + a] * r[K * cl + a];
                out[gz * Pf] = old[jz][cl * 2 * K + q] + v;
              }
        }
    }

    // Transposed 1D rows of the prolongation for the k C coarse nodes of a
    // tile, read from a fine window line r[] whose entry 2k cl + 2k - 1 is
    // the fine node on coarse vertex c0 + cl; c0 = first coarse cell.
    template <int K, int C, typename T>
    __device__ __forceinline__ void
    restrict_line(const T *E, const T *r, int c0, T *o)
    {
      constexpr int N = K + 1;
#pragma unroll
      for (int cl = 0; cl < C; ++cl)
#pragma unroll
        for (int a = 0; a < K; ++a)
          {
            const int base = 2 * K * cl + 2 * K - 1;  // fine node 2k(c0+cl)
            T         v    = 0;
            if (a != 0)
              {
#pragma unroll
                for (int q = 1; q < 2 * K; ++q)
                  v += E[q * N + a] * r[base + q];
              }
            else
              {
                v = r[base];
#pragma unroll
                for (int q = 1; q < 2 * K; ++q)
                  v += E[q * N] * r[base + q];
                if (c0 + cl > 0)
                  {
#pragma unroll
                    for (int q = 1; q < 2 * K; ++q)
                      v += E[q * N + K] * r[base + q - 2 * K];
                  }
              }
            o[K * cl + a] = v;
          }
    }

    template <int K>
    struct RT
    {
      // coarse cells per work-group (k = 4: one, the fine window of 15^3
      // nodes then takes 18 KB of LDS in FP32, 36 KB in FP64)
      static constexpr int CX = K <= 2 ? 4 : (K <= 3 ? 2 : 1),
                           CY = K <= 2 ? 2 : (K <= 3 ? 2 : 1),
                           CZ = K <= 2 ? 2 : (K <= 3 ? 2 : 1);
    };

    // STORE: coarse = P^T fine on interior coarse nodes (Dirichlet entries
    // untouched); otherwise coarse += P^T fine
    template <int K, typename T, bool STORE>
    __global__ void __launch_bounds__(256)
    restrict3d_kernel(const T *__restrict__ fine, T *__restrict__ coarse,
                      const T *__restrict__ tab, const int nc)
    {
      using TB         = Tab<K>;
      constexpr int CX = RT<K>::CX, CY = RT<K>::CY, CZ = RT<K>::CZ, NT = 256;
      // fine window per direction and coarse nodes owned per direction
      constexpr int WX = 2 * K * CX + 2 * K - 1, WY = 2 * K * CY + 2 * K - 1,
                    WZ = 2 * K * CZ + 2 * K - 1;
      constexpr int OX = K * CX, OY = K * CY, OZ = K * CZ;
      __shared__ T sF[WZ * WY * WX], sX[WZ * WY * OX], sY[WZ * OY * OX];
      const T  *E   = tab + TB::E;
      const int tid = threadIdx.x;
      const int Nc1 = K * nc + 1, Nf1 = 2 * K * nc + 1;
      const long long Pc = (long long)Nc1 * Nc1, Pf = (long long)Nf1 * Nf1;
      const int L   = xcd_logical_id(linear_block_id(),
                                     gridDim.x * gridDim.y * gridDim.z);
      const int c0x = (L % gridDim.x) * CX, c0y = ((L / gridDim.x) % gridDim.y) * CY,
                c0z = (L / (gridDim.x * gridDim.y)) * CZ;
      const int w0x = 2 * K * c0x - 2 * K + 1, w0y = 2 * K * c0y - 2 * K + 1,
                w0z = 2 * K * c0z - 2 * K + 1;
      {
        // the whole window in one round of loads, one z-column (y, x) of WZ
        // values per thread and pass: the column's address and x/y bounds
        // are set up once, the z bounds are uniform
        constexpr int NC = WY * WX, IT = (NC + NT - 1) / NT;
        T             v[IT][WZ];
#pragma unroll
        for (int j = 0; j < IT; ++j)
          {
            const int  col = tid + j * NT;
            const int  y = col / WX, x = col - y * WX;
            const int  gx = w0x + x, gy = w0y + y;
            const bool in = col < NC && gx >= 0 && gx < Nf1 && gy >= 0 && gy < Nf1;
            const T   *fc = fine + (in ? gx + (long long)gy * Nf1 : 0);
#pragma unroll
            for (int z = 0; z < WZ; ++z)
              {
                const int gz = w0z + z;
                v[j][z]      = in && gz >= 0 && gz < Nf1 ? fc[gz * Pf] : T(0);
              }
          }
#pragma unroll
        for (int j = 0; j < IT; ++j)
          {
            const int col = tid + j * NT;
            if (col < NC)
              {
#pragma unroll
                for (int z = 0; z < WZ; ++z)
                  sF[z * NC + col] = v[j][z];
              }
          }
      }
      __syncthreads();
      for (int t = tid; t < WZ * WY; t += NT)
        {
          T r[WX], o[OX];
#pragma unroll
          for (int i = 0; i < WX; ++i)
            r[i] = sF[t * WX + i];
          restrict_line<K, CX>(E, r, c0x, o);
#pragma unroll
          for (int i = 0; i < OX; ++i)
            sX[t * OX + i] = o[i];
        }
      __syncthreads();
      for (int t = tid; t < WZ * OX; t += NT)
        {
          const int z = t / OX, xo = t - z * OX;
          T         r[WY], o[OY];
#pragma unroll
          for (int i = 0; i < WY; ++i)
            r[i] = sX[(z * WY + i) * OX + xo];
          restrict_line<K, CY>(E, r, c0y, o);
#pragma unroll
          for (int i = 0; i < OY; ++i)
            sY[(z * OY + i) * OX + xo] = o[i];
        }
      __syncthreads();
      for (int t = tid; t < OY * OX; t += NT)
        {
          const int yo = t / OX, xo = t - yo * OX;
          const int gx = K * c0x + xo, gy = K * c0y + yo;
          if (gx <= 0 || gx >= Nc1 - 1 || gy <= 0 || gy >= Nc1 - 1)
            continue;
          T r[WZ], o[OZ];
#pragma unroll
          for (int i = 0; i < WZ; ++i)
            r[i] = sY[(i * OY + yo) * OX + xo];
          restrict_line<K, CZ>(E, r, c0z, o);
          T *out = coarse + gx + (long long)gy * Nc1;
          T  old[OZ];
#pragma unroll
          for (int i = 0; i < OZ; ++i)
            {
              const int gz = K * c0z + i;
              old[i] = (!STORE && gz > 0 && gz < Nc1 - 1) ? out[gz * Pc] : T(0);
            }
#pragma unroll
          for (int i = 0; i < OZ; ++i)
            {
              const int gz = K * c0z + i;
              if (gz <= 0 || gz >= Nc1 - 1)
                continue;
              out[gz * Pc] = old[i] + o[i];
            }
        }
    }

    template <int K, typename T, bool PROLONG, bool MASK_IN, int OUT>
    void
    pass(hipStream_t s, const T *in, T *out, const T *tab, int dim, int axis,
         const int *ein, const int *eout, int nc, int Nc1)
    {
      const long long total = (long long)eout[0] * eout[1] * eout[2];
      const long long nb    = std::min<long long>((total + TR_NT - 1) / TR_NT, 16384);
      hipLaunchKernelGGL((transfer_pass_kernel<K, T, PROLONG, MASK_IN, OUT>),
                         dim3((unsigned)std::max<long long>(nb, 1)),
                         dim3(TR_NT), 0, s, in, out, tab, dim, axis, eout[0],
                         eout[1], eout[2], ein[0], ein[1], nc, Nc1);
      GTPS_LAUNCH_CHECK();
    }

    template <int K, typename T>
    void
    run(const gtps_transfer *tr, int fine_level, T *fine, const T *fine_in,
        T *coarse, const T *coarse_in, bool prolong, bool store)
    {
      const T    *tab = reinterpret_cast<const T *>(
        tr->prec == GTPS_FP32 ? (const void *)tr->tab->d32
                                 : (const void *)tr->tab->d64);
      hipStream_t s   = tr->ctx->stream;
      const int   dim = tr->dim;
      const int   nc  = 1 << (fine_level - 1);
      const int   Nc1 = K * nc + 1, Nf1 = 2 * K * nc + 1;
      T          *t1  = reinterpret_cast<T *>(tr->tmp1);
      T          *t2  = reinterpret_cast<T *>(tr->tmp2);
      if (dim == 3 && prolong)
        {
          using C = TT<K>;
          const dim3 grid((nc + C::CX - 1) / C::CX, (nc + C::CY - 1) / C::CY,
                          (nc + C::CZ - 1) / C::CZ);
          hipLaunchKernelGGL((prolongate3d_kernel<K, T>), grid, dim3(256), 0, s,
                             coarse_in, fine, tab, nc);
          GTPS_LAUNCH_CHECK();
          return;
        }
      // the restriction window outgrows 64 KiB of LDS beyond k = 4: those
      // degrees keep the three separable passes below
      if constexpr (K <= 4)
        if (dim == 3)
          {
            using C = RT<K>;
            const dim3 grid((nc + C::CX - 1) / C::CX, (nc + C::CY - 1) / C::CY,
                            (nc + C::CZ - 1) / C::CZ);
            if (store)
              hipLaunchKernelGGL((restrict3d_kernel<K, T, true>), grid, dim3(256),
                                 0, s, fine_in, coarse, tab, nc);
            else
              hipLaunchKernelGGL((restrict3d_kernel<K, T, false>), grid,
                                 dim3(256), 0, s, fine_in, coarse, tab, nc);
            GTPS_LAUNCH_CHECK();
            return;
          }
      if (prolong)
        {
          int e0[3] = {Nc1, Nc1, dim == 3 ? Nc1 : 1};
          int e1[3] = {Nf1, Nc1, dim == 3 ? Nc1 : 1};
          int e2[3] = {Nf1, Nf1, dim == 3 ? Nc1 : 1};
          int e3[3] = {Nf1, Nf1, Nf1};
          pass<K, T, true, true, 0>(s, coarse_in, t1, tab, dim, 0, e0, e1, nc, Nc1);
          if (dim == 2)
            pass<K, T, true, false, 1>(s, t1, fine, tab, dim, 1, e1, e2, nc, Nc1);
          else
            {
              pass<K, T, true, false, 0>(s, t1, t2, tab, dim, 1, e1, e2, nc, Nc1);
              pass<K, T, true, false, 1>(s, t2, fine, tab, dim, 2, e2, e3, nc, Nc1);
            }
        }
      else
        {
          int e0[3] = {Nf1, Nf1, dim == 3 ? Nf1 : 1};
          int e1[3] = {Nc1, Nf1, dim == 3 ? Nf1 : 1};
          int e2[3] = {Nc1, Nc1, dim == 3 ? Nf1 : 1};
          int e3[3] = {Nc1, Nc1, Nc1};
          if (dim == 2)
            {
              pass<K, T, false, false, 0>(s, fine_in, t1, tab, dim, 0, e0, e1, nc, Nc1);
              if (store)
                pass<K, T, false, false, 3>(s, t1, coarse, tab, dim, 1, e1, e2, nc, Nc1);
              else
                pass<K, T, false, false, 2>(s, t1, coarse, tab, dim, 1, e1, e2, nc, Nc1);
            }
          else
            {
              pass<K, T, false, false, 0>(s, fine_in, t1, tab, dim, 0, e0, e1, nc, Nc1);
              pass<K, T, false, false, 0>(s, t1, t2, tab, dim, 1, e1, e2, nc, Nc1);
              if (store)
                pass<K, T, false, false, 3>(s, t2, coarse, tab, dim, 2, e2, e3, nc, Nc1);
              else
                pass<K, T, false, false, 2>(s, t2, coarse, tab, dim, 2, e2, e3, nc, Nc1);
            }
        }
    }

    template <typename T>
    void
    run_t(const gtps_transfer *tr, int fine_level, void *fine,
          const void *fine_in, void *coarse, const void *coarse_in,
          bool prolong, bool store = false)
    {
      T       *f  = reinterpret_cast<T *>(fine);
      const T *fi = reinterpret_cast<const T *>(fine_in);
      T       *c  = reinterpret_cast<T *>(coarse);
      const T *ci = reinterpret_cast<const T *>(coarse_in);
      switch (tr->k)
        {
          case 1: return run<1, T>(tr, fine_level, f, fi, c, ci, prolong, store);
          case 2: return run<2, T>(tr, fine_level, f, fi, c, ci, prolong, store);
          case 3: return run<3, T>(tr, fine_level, f, fi, c, ci, prolong, store);
          case 4: return run<4, T>(tr, fine_level, f, fi, c, ci, prolong, store);
          case 5: return run<5, T>(tr, fine_level, f, fi, c, ci, prolong, store);
          case 6: return run<6, T>(tr, fine_level, f, fi, c, ci, prolong, store);
          case 7: return run<7, T>(tr, fine_level, f, fi, c, ci, prolong, store);
          default: throw Error(GTPS_EINVAL, "degree must be in 1..7");
        }
    }
  } // namespace

  size_t
  transfer_tmp_elems(int dim, int k, int max_level)
  {
    const long long nc  = 1LL << (max_level - 1);
    const long long Nc1 = k * nc + 1, Nf1 = 2 * k * nc + 1;
    return dim == 3 ? (size_t)(Nf1 * Nf1 * Nc1) : (size_t)(Nf1 * Nc1);
  }

  void
  launch_prolongate_add(const gtps_transfer *tr, int to_level, void *fine,
                        const void *coarse)
  {
    if (tr->prec == GTPS_FP32)
      run_t<float>(tr, to_level, fine, nullptr, nullptr, coarse, true);
    else
      run_t<double>(tr, to_level, fine, nullptr, nullptr, coarse, true);
  }

  void
  launch_restrict_add(const gtps_transfer *tr, int from_level, void *coarse,
                      const void *fine)
  {
    if (tr->prec == GTPS_FP32)
      run_t<float>(tr, from_level, nullptr, fine, coarse, nullptr, false);
    else
      run_t<double>(tr, from_level, nullptr, fine, coarse, nullptr, false);
  }

  void
  launch_restrict_store(const gtps_transfer *tr, int from_level, void *coarse,
                        const void *fine)
  {
    if (tr->prec == GTPS_FP32)
      run_t<float>(tr, from_level, nullptr, fine, coarse, nullptr, false, true);
    else
      run_t<double>(tr, from_level, nullptr, fine, coarse, nullptr, false, true);
  }
} // namespace gtps
