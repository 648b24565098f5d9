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
    // PASS 2 (x): y = scale (L i1 + M i2), residual / dot forms.
    // Passes 0 and 1: one thread per cell line along the pass direction (the
    // NN DoFs of one cell), the cell and its two neighbour cells loaded once
    // for all NN outputs, lanes along x (coalesced; the cell index, and with
    // it the variant of the diagonal block of L, is uniform across a
    // wavefront except where it crosses a plane).  Pass 2 runs along x
    // itself: one thread per DoF with dg_L_line / dg_M_line, so that the
    // lanes stay coalesced (a line per thread there reads strided rows).  An
    // output of L sums the diagonal block, then the block of the previous
    // cell, then that of the next, in both forms.
    template <int K, typename T, int PASS, bool RESID, bool DOT>
    __global__ void __launch_bounds__(DNT)
    dg_op_kernel(const T *__restrict__ i1, const T *__restrict__ i2,
                 T *__restrict__ o1, T *__restrict__ o2,
                 const T *__restrict__ tab, const T *__restrict__ b,
                 const T *__restrict__ x0, const int n, const int N1,
                 const long long N, const T scale,
                 double *__restrict__ partials, const FastDiv div_n1,
                 const FastDiv div_n)
    {
      using TB         = DTab<K>;
      constexpr int NN = TB::NN;
      __shared__ double red[DNT / 64];
      double            dot = 0.0;
      // level sizes < 2^31 (host check): 32-bit indices within a vector,
      // divisions by N1 and n as multiply-high and shift
      const int         P   = N1 * N1;
      const int         Ni  = (int)N;
      // batch: blockIdx.y selects one of several contiguous vectors
      const long long   bo  = (long long)blockIdx.y * N;
      i1 += bo;
      o1 += bo;
      if (PASS != 0)
        i2 += bo;
      if (PASS != 2)
        o2 += bo;
      if (RESID)
        b += bo;
      if (DOT)
        {
          x0 += bo;
          partials += (long long)blockIdx.y * gridDim.x;
        }
      if constexpr (PASS == 2)
        {
          for (int g = blockIdx.x * DNT + threadIdx.x; g < Ni; g += gridDim.x * DNT)
            {
              const int X    = g - (int)fastdiv((unsigned)g, div_n1) * N1;
              const int base = g - X;
              T v = scale * (dg_L_line<K>(tab, i1, base, 1, X, n) +
                             dg_M_line<K>(tab, i2, base, 1, X));
              if (RESID)
                v = b[g] - v;
              o1[g] = v;
              if (DOT)
                dot += (double)x0[g] * (double)v;
            }
        }
      const int st = PASS == 0 ? P : N1;
      const int NL = PASS == 2 ? 0 : Ni / NN;
      for (int q = blockIdx.x * DNT + threadIdx.x; q < NL; q += gridDim.x * DNT)
        {
          int c, base;
          if (PASS == 0)
            {
              // lanes along x, lines along z: q = x + N1 (y + N1 c)
              const int r  = (int)fastdiv((unsigned)q, div_n1);
              const int r2 = (int)fastdiv((unsigned)r, div_n1);
              base = q - r * N1 + (r - r2 * N1) * N1;
              c    = r2;
            }
          else if (PASS == 1)
            {
              // lanes along x, lines along y: q = x + N1 (c + n z)
              const int r  = (int)fastdiv((unsigned)q, div_n1);
              const int z  = (int)fastdiv((unsigned)r, div_n);
              base = q - r * N1 + z * P;
              c    = r - z * n;
            }
          else
            {
              base = 0;
              c    = 0;
            }
          const int c0 = base + c * NN * st;
          T a[NN], pm[NN], pp[NN], bb[NN];
#pragma unroll
          for (int j = 0; j < NN; ++j)
            {
              a[j]  = i1[c0 + j * st];
              pm[j] = c > 0 ? i1[c0 + (j - NN) * st] : T(0);
              pp[j] = c < n - 1 ? i1[c0 + (j + NN) * st] : T(0);
              bb[j] = PASS != 0 ? i2[c0 + j * st] : T(0);
            }
          auto emit = [&](auto variant) {
            constexpr int V  = decltype(variant)::value;
            const T      *Lc = tab + TB::Lc + V * NN * NN;
            const T      *Lm = tab + TB::Lm, *Lp = tab + TB::Lp, *M = tab + TB::M;
#pragma unroll
            for (int i = 0; i < NN; ++i)
              {
                T l = 0;
#pragma unroll
                for (int j = 0; j < NN; ++j)
                  l += Lc[i * NN + j] * a[j];
                if (c > 0)
                  {
#pragma unroll
                    for (int j = 0; j < NN; ++j)
                      l += Lm[i * NN + j] * pm[j];
                  }
                if (c < n - 1)
                  {
#pragma unroll
                    for (int j = 0; j < NN; ++j)
                      l += Lp[i * NN + j] * pp[j];
                  }
                const int g = c0 + i * st;
                if (PASS == 0)
                  {
                    T m = 0;
#pragma unroll
                    for (int j = 0; j < NN; ++j)
                      m += M[i * NN + j] * a[j];
                    o1[g] = m;
                    o2[g] = l;
                  }
                else
                  {
                    T mb = 0, m = 0;
#pragma unroll
                    for (int j = 0; j < NN; ++j)
                      mb += M[i * NN + j] * bb[j];
#pragma unroll
                    for (int j = 0; j < NN; ++j)
                      m += M[i * NN + j] * a[j];
                    o1[g] = m;
                    o2[g] = l + mb;
                  }
              }
          };
          // diagonal block variant as dg_cell_variant: first, last, interior
          if (c == 0)
            emit(std::integral_constant<int, 1>{});
          else if (c == n - 1)
            emit(std::integral_constant<int, 2>{});
          else
            emit(std::integral_constant<int, 0>{});
        }
      if (DOT)
        {
          dot = block_sum<DNT>(dot, red);
          if (threadIdx.x == 0)
            partials[blockIdx.x] = dot;
        }
    }

    // Pass 2 (x) with the rows staged in LDS: a work-group takes RB whole
    // x-rows (contiguous in memory) of i1 and i2 with coalesced loads, then
    // every output reads its cell and neighbour cells, and the 1D blocks of
    // its node, from LDS, in the sums of dg_L_line / dg_M_line.  Rows of at
    // most DG_ROW_CAP values.
    constexpr int DG_ROW_CAP = 2048;

    template <int K, typename T, bool RESID, bool DOT>
    __global__ void __launch_bounds__(DNT)
    dg_op_x_kernel(const T *__restrict__ i1, const T *__restrict__ i2,
                   T *__restrict__ o1, const T *__restrict__ tab,
                   const T *__restrict__ b, const T *__restrict__ x0, const int n,
                   const int N1, const long long N, const int RB, const T scale,
                   double *__restrict__ partials)
    {
      using TB         = DTab<K>;
      constexpr int NN = TB::NN;
      // M, L-, L+ and the three diagonal blocks of L (they lead the table)
      constexpr int NTAB = TB::Lc + 3 * NN * NN;
      __shared__ T      s1[DG_ROW_CAP], s2[DG_ROW_CAP], st[NTAB];
      __shared__ double red[DNT / 64];
      double            dot = 0.0;
      for (int e = threadIdx.x; e < NTAB; e += DNT)
        st[e] = tab[e];
      const long long   bo  = (long long)blockIdx.y * N;
      i1 += bo;
      i2 += bo;
      o1 += bo;
      if (RESID)
        b += bo;
      if (DOT)
        {
          x0 += bo;
          partials += (long long)blockIdx.y * gridDim.x;
        }
      const long long rows = N / N1;
      const int       E    = RB * N1;
      for (long long r0 = (long long)blockIdx.x * RB; r0 < rows;
           r0 += (long long)gridDim.x * RB)
        {
          const long long g0 = r0 * N1;
          const int       ne = N - g0 < E ? (int)(N - g0) : E;
          __syncthreads();  // the previous round's reads are done
          for (int e = threadIdx.x; e < ne; e += DNT)
            {
              s1[e] = i1[g0 + e];
              s2[e] = i2[g0 + e];
            }
          __syncthreads();
          for (int e = threadIdx.x; e < ne; e += DNT)
            {
              const int X = e % N1, c = X / NN, i = X - c * NN;
              const T  *in1 = s1 + (e - X) + c * NN;
              const T  *in2 = s2 + (e - X) + c * NN;
              const T  *Lc  = st + TB::Lc + dg_cell_variant(c, n) * NN * NN + i * NN;
              T         l   = 0;
#pragma unroll
              for (int j = 0; j < NN; ++j)
                l += Lc[j] * in1[j];
              if (c > 0)
                {
                  const T *Lm = st + TB::Lm + i * NN;
#pragma unroll
                  for (int j = 0; j < NN; ++j)
                    l += Lm[j] * in1[j - NN];
                }
              if (c < n - 1)
                {
                  const T *Lp = st + TB::Lp + i * NN;
#pragma unroll
                  for (int j = 0; j < NN; ++j)
                    l += Lp[j] * in1[j + NN];
                }
              const T *M = st + TB::M + i * NN;
              T        m = 0;
#pragma unroll
              for (int j = 0; j < NN; ++j)
                m += M[j] * in2[j];
              const long long g = g0 + e;
              T               v = scale * (l + m);
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

    // LDS plan of the patch kernel: r and t (patch box, ping-pong in the
    // local solve), NS face slabs and their intermediate (NS = 6: all faces
    // in one round; 2: one direction per round; 1: one face per round,
    // whichever fits 64 KB; t shares the slab buffer, which the local solve
    // no longer needs), and the 1D tables of this patch (S and lambda per
    // direction, M, L-, L+).
    template <int K, typename T>
    struct DP
    {
      static constexpr int    NN = K + 1, PD = 2 * NN, P3 = PD * PD * PD;
      static constexpr int    SL   = PD * PD * NN;  // one slab
      static constexpr int    TABN = 3 * PD * PD + 3 * PD + 3 * NN * NN;
      static constexpr int    sa_len(int ns) { return ns * SL > P3 ? ns * SL : P3; }
      static constexpr size_t bytes(int ns)
      {
        return sizeof(T) * (P3 + sa_len(ns) + ns * SL + TABN);
      }
      static constexpr int    NS  = bytes(6) <= 65536 ? 6 : (bytes(2) <= 65536 ? 2 : 1);
      // threads per patch: about one patch DoF per thread (the kernel is
      // latency-bound: every phase is a dependent LDS pass)
      static constexpr int    NT  = P3 >= 1000 ? 1024 : (P3 >= 500 ? 512 : 256);
      static constexpr size_t lds = bytes(NS);
      static constexpr bool   ok  = lds <= 65536;
    };

    // One work-group per patch (2x2x2 cells around an interior vertex, all
    // (2k+2)^3 DoFs), u_P <- (1-w) u_P + w A_PP^{-1} (b_P - A_P,ext u_ext):
    //  1. b_P into registers; the face-neighbour cell slabs of u into LDS;
    //  2. per face, L-/L+ along the face normal, then the cell mass along the
    //     two in-plane directions, each pass one line of k+1 values per
    //     thread; then every patch DoF subtracts the face of each direction
    //     on its side (so the faces of different directions never race);
    //  3. fast diagonalisation, ping-ponging r and t, one task per line and
    //     half of its outputs, the eigenvalue scaling fused into the last
    //     forward pass; the result is stored coalesced along x.
    // Patch arrays are [z][y][x] with PD per direction; slab f (direction
    // D = f / 2, side f % 2: below / above the patch) is [a][b][j] with j
    // along D and (a, b) the other two directions, a the slower one.  Line
    // tasks are counted in whole wavefronts per face (or output half), so
    // the face and its tables are uniform over a wavefront.
    template <int K, typename T>
    struct DPShared
    {
      using C = DP<K, T>;
      T r[C::P3], sa[C::sa_len(C::NS)], sw[C::NS * C::SL];  // t = sa
      T Ssh[3][C::PD * C::PD], lsh[3][C::PD], Msh[C::NN * C::NN], Lsh[2][C::NN * C::NN];
    };

    // the local solve of the patch with lower-left cell pc; ends with the
    // patch written to u (callers separate dependent patches by a barrier)
    template <int K, typename T>
    __device__ void
    dg_patch_solve(T *__restrict__ u, const T *__restrict__ bvec,
                   const T *__restrict__ tab, const int n, const int pc[3],
                   const T scale, const T relax, DPShared<K, T> &sh)
    {
      using C            = DP<K, T>;
      using TB           = DTab<K>;
      constexpr int NN   = C::NN, PD = C::PD, P3 = C::P3, SL = C::SL, NS = C::NS;
      constexpr int NT   = C::NT;
      // lines per slab (or per direction of the box), padded to wavefronts
      constexpr int NL = PD * PD, NLP = (NL + 63) / 64 * 64;
      static_assert(NT % 64 == 0, "whole wavefronts");
      T *r = sh.r, *t = sh.sa, *sa = sh.sa, *sw = sh.sw;
      auto &Ssh = sh.Ssh;
      auto &lsh = sh.lsh;
      auto &Msh = sh.Msh;
      auto &Lsh = sh.Lsh;
      // level sizes < 2^31 (host check): offsets from the patch's first DoF
      // are 32-bit
      const int       N1    = n * NN, P = N1 * N1;
      const int       st[3] = {1, N1, P};
      const int       g0    = pc[0] * NN + pc[1] * NN * N1 + pc[2] * NN * P;
      T *__restrict__ up       = u + g0;
      const T *__restrict__ bp = bvec + g0;
      const int tid            = threadIdx.x;
      auto variant = [&](int p) { return (p == 0 ? 1 : 0) + (p + 2 == n ? 2 : 0); };
      // face f exists if the neighbour cell slab lies inside the mesh
      auto has_face = [&](int f) {
        const int p = pc[f >> 1];
        return (f & 1) ? p + 2 < n : p > 0;
      };
      // patch coordinates (x fastest) of this thread's DoFs tid + it NT,
      // used by the patch loads, the face terms and the store
      constexpr int ITB = (P3 + NT - 1) / NT;
      int           ec[ITB][3];
#pragma unroll
      for (int it = 0; it < ITB; ++it)
        {
          const int e = tid + it * NT < P3 ? tid + it * NT : 0;
          ec[it][0]   = e % PD;
          ec[it][1]   = (e / PD) % PD;
          ec[it][2]   = e / (PD * PD);
        }
      auto box_offset = [&](int it) { return ec[it][0] + ec[it][1] * N1 + ec[it][2] * P; };
      // Coupling tasks (face ff, line ln) and FD tasks (output block fh,
      // line fln): lines per face or block padded to whole wavefronts where
      // that fits the work-group, else packed.
      constexpr int CL = NS * NLP <= NT ? NLP : NL;
      static_assert(NS * CL <= NT, "one coupling task per thread");
      constexpr int OB = (PD / 2) * NLP <= NT ? 2 : NN, NB = PD / OB;
      constexpr int FL = NB * NLP <= NT ? NLP : NL;
      static_assert(NB * FL <= NT, "one FD task per thread");
      // With padded tasks a wavefront's face (block) is uniform, and so are
      // the 1D matrix entries it multiplies by: they are read from the
      // global table at wave-uniform addresses (scalar loads, used as SGPR
      // operands of the multiply-adds).  Packed tasks read LDS copies.
      constexpr bool CU = CL % 64 == 0, FU = FL % 64 == 0;
      const T *Sg[3] = {tab + TB::S + variant(pc[0]) * PD * PD,
                        tab + TB::S + variant(pc[1]) * PD * PD,
                        tab + TB::S + variant(pc[2]) * PD * PD};
      if constexpr (!FU)
        for (int e = tid; e < 3 * PD * PD; e += NT)
          {
            const int d = e / (PD * PD);
            Ssh[d][e - d * PD * PD] = Sg[d][e - d * PD * PD];
          }
      // slab lines of faces f0 .. f0+NS-1: task (face, a, b), at most NTK
      // per thread.  issue_slabs loads a thread's lines into registers,
      // commit_slabs stores them to sa, so that the first round's slab
      // loads go out before any other load (their index arithmetic waits
      // for nothing) and every load is in flight before the first wait.
      constexpr int NTK = (NS * NLP + NT - 1) / NT;
      T             slv[NTK][NN];
      int           slo[NTK];
      auto issue_slabs = [&](int f0) {
#pragma unroll
        for (int j = 0; j < NTK; ++j)
          {
            const int tk = tid + j * NT;
            slo[j]       = -1;
            if (tk >= NS * NLP)
              continue;
            const int ff = __builtin_amdgcn_readfirstlane(tk / NLP), ln = tk - ff * NLP;
            const int f  = f0 + ff;
            if (ln >= NL || !has_face(f))
              continue;
            const int D = f >> 1, da = D == 2 ? 1 : 2, db = D == 0 ? 1 : 0;
            const int aa = ln / PD, bb = ln - aa * PD;
            const int o  = ((f & 1) ? PD : -NN) * st[D] + aa * st[da] + bb * st[db];
#pragma unroll
            for (int i = 0; i < NN; ++i)
              slv[j][i] = up[o + i * st[D]];
            slo[j] = ff * SL + ln * NN;
          }
      };
      auto commit_slabs = [&]() {
#pragma unroll
        for (int j = 0; j < NTK; ++j)
          if (slo[j] >= 0)
            {
#pragma unroll
              for (int i = 0; i < NN; ++i)
                sa[slo[j] + i] = slv[j][i];
            }
      };
      issue_slabs(0);
      // the eigenvalues, b_P and the old patch values (relax != 1)
      static_assert(3 * PD <= NT, "one eigenvalue per thread");
      const bool lam_t = tid < 3 * PD;
      const int  lam_d = lam_t ? tid / PD : 0, lam_i = tid - lam_d * PD;
      const T    lam_v = lam_t ? tab[TB::lam + variant(pc[lam_d]) * PD + lam_i] : T(0);
      if constexpr (!CU)
        for (int e = tid; e < NN * NN; e += NT)
          {
            Msh[e]    = tab[TB::M + e];
            Lsh[0][e] = tab[TB::Lm + e];
            Lsh[1][e] = tab[TB::Lp + e];
          }
      const bool blend = relax != T(1);
      T          vb[ITB], uold[ITB];
#pragma unroll
      for (int it = 0; it < ITB; ++it)
        {
          const int e = tid + it * NT;
          const int o = box_offset(it);
          vb[it]      = e < P3 ? bp[o] : T(0);
          uold[it]    = e < P3 && blend ? up[o] : T(0);
        }
      commit_slabs();
      if (lam_t)
        lsh[lam_d][lam_i] = lam_v;
      // 2. external coupling, NS faces per round; a thread keeps its task
      // (face ff, line ln) in every pass of a round
      const int  cff  = CU ? __builtin_amdgcn_readfirstlane(tid / CL) : tid / CL;
      const int  cln  = tid - cff * CL;
      const bool cact = tid < NS * CL && cln < NL;
      for (int f0 = 0; f0 < 6; f0 += NS)
        {
          if (f0 > 0)
            {
              issue_slabs(f0);
              commit_slabs();
            }
          __syncthreads();
          const int  f  = f0 + cff;
          const bool on = cact && has_face(f);
          const T   *Lo = CU ? tab + ((f & 1) ? TB::Lp : TB::Lm) : Lsh[f & 1];
          const T   *Mc = CU ? tab + TB::M : Msh;
          if (on)
            {
              // w[a][b][i] = sum_j L-/+[i][j] s[a][b][j]: line (a, b)
              const T *sl = sa + cff * SL + cln * NN;
              T        v[NN];
#pragma unroll
              for (int j = 0; j < NN; ++j)
                v[j] = sl[j];
              T *w = sw + cff * SL + cln * NN;
#pragma unroll
              for (int i = 0; i < NN; ++i)
                {
                  T acc = 0;
#pragma unroll
                  for (int j = 0; j < NN; ++j)
                    acc += Lo[i * NN + j] * v[j];
                  w[i] = acc;
                }
            }
          __syncthreads();
          // cell mass along b: line (a, cell cb of b, i)
          if (on)
            {
              const int i = cln % NN, acb = cln / NN, cb = acb & 1, aa = acb >> 1;
              const int o = cff * SL + (aa * PD + cb * NN) * NN + i;
              T         v[NN];
#pragma unroll
              for (int j = 0; j < NN; ++j)
                v[j] = sw[o + j * NN];
#pragma unroll
              for (int ib = 0; ib < NN; ++ib)
                {
                  T acc = 0;
#pragma unroll
                  for (int j = 0; j < NN; ++j)
                    acc += Mc[ib * NN + j] * v[j];
                  sa[o + ib * NN] = acc;
                }
            }
          __syncthreads();
          // cell mass along a: line (cell ca of a, b, i), into sw[a][b][i];
          // an absent face's slab is set to zero, so that the face terms
          // below are read without a mask
          if (cact)
            {
              const int i = cln % NN, cbb = cln / NN, bb = cbb % PD, ca = cbb / PD;
              const int o = cff * SL + (ca * NN * PD + bb) * NN + i;
              if (on)
                {
                  T v[NN];
#pragma unroll
                  for (int j = 0; j < NN; ++j)
                    v[j] = sa[o + j * PD * NN];
#pragma unroll
                  for (int ia = 0; ia < NN; ++ia)
                    {
                      T acc = 0;
#pragma unroll
                      for (int j = 0; j < NN; ++j)
                        acc += Mc[ia * NN + j] * v[j];
                      sw[o + ia * PD * NN] = acc;
                    }
                }
              else
                {
#pragma unroll
                  for (int ia = 0; ia < NN; ++ia)
                    sw[o + ia * PD * NN] = T(0);
                }
            }
          __syncthreads();
          // r -= scale (face terms): per direction, the face on the DoF's
          // side (absent faces hold zeros); all reads issued together, with
          // faces outside this round masked out when the faces take several
          // rounds
#pragma unroll
          for (int it = 0; it < ITB; ++it)
            {
              const int e = tid + it * NT;
              if (e >= P3)
                continue;
              const int(&c3)[3] = ec[it];
              T         v[3];
              bool      on[3];
#pragma unroll
              for (int D = 0; D < 3; ++D)
                {
                  const int pD = c3[D], side = pD >= NN ? 1 : 0, f = 2 * D + side;
                  const int da = D == 2 ? 1 : 2, db = D == 0 ? 1 : 0;
                  on[D] = NS == 6 || (f >= f0 && f < f0 + NS);
                  v[D]  = sw[on[D] ? (f - f0) * SL + (c3[da] * PD + c3[db]) * NN + pD - side * NN
                                   : 0];
                }
              // b_P enters with the first round (r holds the partial
              // residual of the earlier rounds otherwise)
              const T re  = f0 == 0 ? vb[it] : r[e];
              const T acc = (on[0] ? v[0] : T(0)) + (on[1] ? v[1] : T(0)) + (on[2] ? v[2] : T(0));
              r[e]        = re - scale * acc;
            }
          __syncthreads();
        }
      // 3. fast diagonalisation, ping-ponging r and t: S^T along x, y, z,
      // the eigenvalue scaling, S along z, y, x.  Task (output block h,
      // line): the line's PD values, OB outputs; in FP32 blocks of 2 while
      // the tasks fit the work-group once (more wavefronts share a pass,
      // which is what the latency of a lone patch needs), else halves
      // (measured at Q4 24^3: FP32 7.77 against 7.87 ms per solve, FP64
      // 10.10 against 9.76 ms).  Every thread has one task, the same in all
      // passes; with packed tasks the next pass's matrix entries are read
      // from LDS while the current pass computes.
      const int  fh   = FU ? __builtin_amdgcn_readfirstlane(tid / FL) : tid / FL;
      const int  fln  = tid - fh * FL;
      const bool fact = tid < NB * FL && fln < NL;
      // (lp, lq): the other two coordinates of the line, lp the slower
      const int flp = fln / PD, flq = fln - flp * PD;
      // entries (j, m) of the direction's S^T (TR) or S for output block fh
      auto load_coef = [&](auto &cf, int D, bool TR) {
        const T *S = FU ? Sg[D] : Ssh[D];
#pragma unroll
        for (int j = 0; j < PD; ++j)
#pragma unroll
          for (int m = 0; m < OB; ++m)
            cf[j * OB + m] = TR ? S[j * PD + fh * OB + m] : S[(fh * OB + m) * PD + j];
      };
      // small blocks: the next pass's entries are read into registers during
      // this pass (two register sets); large ones are read from LDS as used
      constexpr bool AHEAD = !FU && PD * OB * sizeof(T) <= 80;
      auto pass = [&](const T *in, T *out, auto dirc, auto trc, auto divc, auto &cf,
                      auto &next, int nD, bool nTR, bool has_next) {
        constexpr int  D      = decltype(dirc)::value;
        constexpr bool TR     = decltype(trc)::value;
        constexpr bool DIVIDE = decltype(divc)::value;
        constexpr int  str    = D == 0 ? 1 : (D == 1 ? PD : PD * PD);
        if (fact)
          {
            const int base = D == 0 ? (flp * PD + flq) * PD
                                    : (D == 1 ? flp * PD * PD + flq : flp * PD + flq);
            const T *S0 = FU ? Sg[D] : Ssh[D];
            const T *S  = TR ? S0 + fh * OB : S0 + fh * OB * PD;
            T        v[PD];
#pragma unroll
            for (int j = 0; j < PD; ++j)
              v[j] = in[base + j * str];
            if constexpr (AHEAD)
              if (has_next)
                load_coef(next, nD, nTR);
#pragma unroll
            for (int m = 0; m < OB; ++m)
              {
                T acc = 0;
#pragma unroll
                for (int j = 0; j < PD; ++j)
                  {
                    T c;
                    if constexpr (AHEAD)
                      c = cf[j * OB + m];
                    else
                      c = TR ? S[j * PD + m] : S[m * PD + j];
                    acc += c * v[j];
                  }
                if constexpr (DIVIDE)
                  {
                    const T den = scale * (lsh[0][flq] + lsh[1][flp] + lsh[2][fh * OB + m]);
                    // FP32: hardware reciprocal (1 ulp), FP64: IEEE division
                    if constexpr (sizeof(T) == 4)
                      acc *= __builtin_amdgcn_rcpf(den);
                    else
                      acc /= den;
                  }
                out[base + (fh * OB + m) * str] = acc;
              }
          }
        __syncthreads();
      };
      using I0 = std::integral_constant<int, 0>;
      using I1 = std::integral_constant<int, 1>;
      using I2 = std::integral_constant<int, 2>;
      using BT = std::true_type;
      using BF = std::false_type;
      T ca[AHEAD ? PD * OB : 1], cb[AHEAD ? PD * OB : 1];
      if (AHEAD && fact)
        load_coef(ca, 0, true);
      pass(r, t, I0{}, BT{}, BF{}, ca, cb, 1, true, true);
      pass(t, r, I1{}, BT{}, BF{}, cb, ca, 2, true, true);
      pass(r, t, I2{}, BT{}, BT{}, ca, cb, 2, false, true);
      pass(t, r, I2{}, BF{}, BF{}, cb, ca, 1, false, true);
      pass(r, t, I1{}, BF{}, BF{}, ca, cb, 0, false, true);
      pass(t, r, I0{}, BF{}, BF{}, cb, ca, 0, false, false);
      // the patch values, coalesced along x
#pragma unroll
      for (int it = 0; it < ITB; ++it)
        {
          const int e = tid + it * NT;
          if (e >= P3)
            continue;
          const T v          = r[e];
          up[box_offset(it)] = blend ? (T(1) - relax) * uold[it] + relax * v : v;
        }
    }

    // one colour: one work-group per patch of the colour
    template <int K, typename T>
    __global__ void __launch_bounds__((DP<K, T>::NT))
    dg_patch_kernel(T *__restrict__ u, const T *__restrict__ bvec,
                    const T *__restrict__ tab, const int n, const int colour,
                    const T scale, const T relax, const FastDiv div_x,
                    const FastDiv div_xy)
    {
      __shared__ DPShared<K, T> sh;
      const long long NNN = (long long)n * (K + 1);
      u += (long long)blockIdx.y * NNN * NNN * NNN;
      bvec += (long long)blockIdx.y * NNN * NNN * NNN;
      const unsigned uc = (unsigned)colour;
      const int      cx = (int)(uc % 3), cy = (int)((uc / 3) % 3), cz = (int)(uc / 9);
      // patch coordinates: divisions by the patch counts as multiply-high
      // and shift (divisors set up on the host)
      const int q  = xcd_logical_id(blockIdx.x, gridDim.x);
      const int qz = (int)fastdiv((unsigned)q, div_xy), rq = q - qz * (int)div_xy.d;
      const int qy = (int)fastdiv((unsigned)rq, div_x), qx = rq - qy * (int)div_x.d;
      const int pc[3] = {cx + 3 * qx, cy + 3 * qy, cz + 3 * qz};
      dg_patch_solve<K, T>(u, bvec, tab, n, pc, scale, relax, sh);
    }

    // meshes of at most 4 cells per direction (at most one patch per
    // colour): n_sweeps whole sweeps in one work-group, alternating forward
    // and reverse colour order from `first`, patches one after another;
    // launched with several work-groups, one independent column each
    template <int K, typename T>
    __global__ void __launch_bounds__((DP<K, T>::NT))
    dg_patch_seq_kernel(T *__restrict__ u, const T *__restrict__ bvec,
                        const T *__restrict__ tab, const int n, const int first,
                        const int n_sweeps, const T scale, const T relax,
                        const long long column_stride)
    {
      __shared__ DPShared<K, T> sh;
      // batched form: work-group j smooths column j of (u, b)
      u += (long long)blockIdx.x * column_stride;
      bvec += (long long)blockIdx.x * column_stride;
      for (int s = 0; s < n_sweeps; ++s)
        {
          const bool rev = ((first + s) & 1) != 0;
          for (int cc = 0; cc < 27; ++cc)
            {
              const int c = rev ? 26 - cc : cc;
              const int pc[3] = {c % 3, (c / 3) % 3, c / 9};
              if (pc[0] >= n - 1 || pc[1] >= n - 1 || pc[2] >= n - 1)
                continue;
              dg_patch_solve<K, T>(u, bvec, tab, n, pc, scale, relax, sh);
              __syncthreads();
            }
        }
    }

    // Transfer, one work-group per coarse cell: its 8 children form one
    // (2k+2)^3 brick of fine DoFs, and the embedding is the tensor product of
    // the 1D map Pm[xf][a] = E[xf][a] (fine node xf = child * (k+1) + q of the
    // brick, coarse node a), applied in three LDS passes.
    template <int K, typename T>
    struct DT
    {
      static constexpr int NN = K + 1, FB = 2 * NN;  // brick width
      static constexpr int F3 = FB * FB * FB, C3 = NN * NN * NN;
      static constexpr int S1 = NN * FB * FB, S2 = NN * NN * FB;
    };

    // fine += P coarse
    template <int K, typename T>
    __global__ void __launch_bounds__(DNT)
    dg_prolongate_kernel(const T *__restrict__ coarse, T *__restrict__ fine,
                         const T *__restrict__ tab, const int nf,
                         const long long Nf)
    {
      using TB          = DTab<K>;
      using C           = DT<K, T>;
      constexpr int NN  = C::NN, FB = C::FB;
      __shared__ T    sc[C::C3], s1[C::S2], s2[C::S1], sE[FB * NN];
      const int       tid = threadIdx.x;
      const int       ncd = nf / 2;  // coarse cells per direction
      const int       N1f = nf * NN, N1c = ncd * NN;
      const long long Pf = (long long)N1f * N1f, Pc = (long long)N1c * N1c;
      // level sizes < 2^31 (host check): offsets inside a brick are 32-bit
      const int       Pf32 = (int)Pf, Pc32 = (int)Pc;
      fine += (long long)blockIdx.y * Nf;
      coarse += (long long)blockIdx.y * Pc * N1c;
      const int cc  = blockIdx.x;
      const int cxc = cc % ncd, cyc = (cc / ncd) % ncd, czc = cc / (ncd * ncd);
      const T  *c0  = coarse + (long long)cxc * NN + (long long)cyc * NN * N1c +
                    (long long)czc * NN * Pc;
      T        *f0  = fine + (long long)cxc * FB + (long long)cyc * FB * N1f +
                    (long long)czc * FB * Pf;
      // every global load (embedding table, coarse cell, the fine values
      // the last pass adds to) issued before the first wait
      constexpr int ITE = (FB * NN + DNT - 1) / DNT, ITC = (C::C3 + DNT - 1) / DNT,
                    ITF = (C::F3 + DNT - 1) / DNT;
      T             ve[ITE], vc[ITC], vf[ITF];
#pragma unroll
      for (int it = 0; it < ITE; ++it)
        {
          const int e = tid + it * DNT;
          ve[it]      = e < FB * NN ? tab[TB::E + e] : T(0);
        }
#pragma unroll
      for (int it = 0; it < ITC; ++it)
        {
          const int e = tid + it * DNT;
          const int a = e % NN, b = (e / NN) % NN, c = e / (NN * NN);
          vc[it] = e < C::C3 ? c0[a + b * N1c + c * Pc32] : T(0);
        }
#pragma unroll
      for (int it = 0; it < ITF; ++it)
        {
          const int e  = tid + it * DNT;
          const int xf = e % FB, yf = (e / FB) % FB, zf = e / (FB * FB);
          vf[it] = e < C::F3 ? f0[xf + yf * N1f + zf * Pf32] : T(0);
        }
#pragma unroll
      for (int it = 0; it < ITE; ++it)
        if (tid + it * DNT < FB * NN)
          sE[tid + it * DNT] = ve[it];
#pragma unroll
      for (int it = 0; it < ITC; ++it)
        if (tid + it * DNT < C::C3)
          sc[tid + it * DNT] = vc[it];
      __syncthreads();
      // x: s1[c][b][xf] = sum_a E[xf][a] sc[c][b][a]
      for (int e = tid; e < C::S2; e += DNT)
        {
          const int xf = e % FB, cb = e / FB;
          T         v  = 0;
#pragma unroll
          for (int a = 0; a < NN; ++a)
            v += sE[xf * NN + a] * sc[cb * NN + a];
          s1[e] = v;
        }
      __syncthreads();
      // y: s2[c][yf][xf] = sum_b E[yf][b] s1[c][b][xf]
      for (int e = tid; e < C::S1; e += DNT)
        {
          const int xf = e % FB, yf = (e / FB) % FB, c = e / (FB * FB);
          T         v  = 0;
#pragma unroll
          for (int b = 0; b < NN; ++b)
            v += sE[yf * NN + b] * s1[(c * NN + b) * FB + xf];
          s2[e] = v;
        }
      __syncthreads();
      // z: fine[zf][yf][xf] += sum_c E[zf][c] s2[c][yf][xf]
#pragma unroll
      for (int it = 0; it < ITF; ++it)
        {
          const int e = tid + it * DNT;
          if (e >= C::F3)
            continue;
          const int xf = e % FB, yf = (e / FB) % FB, zf = e / (FB * FB);
          T         v  = vf[it];
#pragma unroll
          for (int c = 0; c < NN; ++c)
            v += sE[zf * NN + c] * s2[(c * FB + yf) * FB + xf];
          f0[xf + yf * N1f + zf * Pf32] = v;
        }
    }

    // coarse (+)= P^T fine
    template <int K, typename T, bool STORE>
    __global__ void __launch_bounds__(DNT)
    dg_restrict_kernel(const T *__restrict__ fine, T *__restrict__ coarse,
                       const T *__restrict__ tab, const int nf,
                       const long long Nc)
    {
      using TB          = DTab<K>;
      using C           = DT<K, T>;
      constexpr int NN  = C::NN, FB = C::FB;
      __shared__ T    sf[C::F3], s1[C::S1], s2[C::S2], sE[FB * NN];
      const int       tid = threadIdx.x;
      const int       ncd = nf / 2;
      const int       N1f = nf * NN, N1c = ncd * NN;
      const long long Pf = (long long)N1f * N1f, Pc = (long long)N1c * N1c;
      // level sizes < 2^31 (host check): offsets inside a brick are 32-bit
      const int       Pf32 = (int)Pf, Pc32 = (int)Pc;
      fine += (long long)blockIdx.y * Pf * N1f;
      coarse += (long long)blockIdx.y * Nc;
      const int cc  = blockIdx.x;
      const int cxc = cc % ncd, cyc = (cc / ncd) % ncd, czc = cc / (ncd * ncd);
      T        *c0  = coarse + (long long)cxc * NN + (long long)cyc * NN * N1c +
                    (long long)czc * NN * Pc;
      const T  *f0  = fine + (long long)cxc * FB + (long long)cyc * FB * N1f +
                     (long long)czc * FB * Pf;
      // every global load (embedding table, fine brick, the coarse values
      // the last pass adds to) issued before the first wait
      constexpr int ITE = (FB * NN + DNT - 1) / DNT, ITF = (C::F3 + DNT - 1) / DNT,
                    ITC = (C::C3 + DNT - 1) / DNT;
      T             ve[ITE], vf[ITF], vc[STORE ? 1 : ITC];
#pragma unroll
      for (int it = 0; it < ITE; ++it)
        {
          const int e = tid + it * DNT;
          ve[it]      = e < FB * NN ? tab[TB::E + e] : T(0);
        }
#pragma unroll
      for (int it = 0; it < ITF; ++it)
        {
          const int e  = tid + it * DNT;
          const int xf = e % FB, yf = (e / FB) % FB, zf = e / (FB * FB);
          vf[it] = e < C::F3 ? f0[xf + yf * N1f + zf * Pf32] : T(0);
        }
      if constexpr (!STORE)
#pragma unroll
        for (int it = 0; it < ITC; ++it)
          {
            const int e = tid + it * DNT;
            const int a = e % NN, b = (e / NN) % NN, c = e / (NN * NN);
            vc[it] = e < C::C3 ? c0[a + b * N1c + c * Pc32] : T(0);
          }
#pragma unroll
      for (int it = 0; it < ITE; ++it)
        if (tid + it * DNT < FB * NN)
          sE[tid + it * DNT] = ve[it];
#pragma unroll
      for (int it = 0; it < ITF; ++it)
        if (tid + it * DNT < C::F3)
          sf[tid + it * DNT] = vf[it];
      __syncthreads();
      // x: s1[zf][yf][a] = sum_xf E[xf][a] sf[zf][yf][xf]
      for (int e = tid; e < C::S1; e += DNT)
        {
          const int a = e % NN, r = e / NN;
          T         v = 0;
#pragma unroll
          for (int xf = 0; xf < FB; ++xf)
            v += sE[xf * NN + a] * sf[r * FB + xf];
          s1[e] = v;
        }
      __syncthreads();
      // y: s2[zf][b][a] = sum_yf E[yf][b] s1[zf][yf][a]
      for (int e = tid; e < C::S2; e += DNT)
        {
          const int a = e % NN, b = (e / NN) % NN, zf = e / (NN * NN);
          T         v = 0;
#pragma unroll
          for (int yf = 0; yf < FB; ++yf)
            v += sE[yf * NN + b] * s1[(zf * FB + yf) * NN + a];
          s2[e] = v;
        }
      __syncthreads();
      // z: coarse[c][b][a] (+)= sum_zf E[zf][c] s2[zf][b][a]
#pragma unroll
      for (int it = 0; it < ITC; ++it)
        {
          const int e = tid + it * DNT;
          if (e >= C::C3)
            continue;
          const int a = e % NN, b = (e / NN) % NN, c = e / (NN * NN);
          T         v = 0;
#pragma unroll
          for (int zf = 0; zf < FB; ++zf)
            v += sE[zf * NN + c] * s2[(zf * NN + b) * NN + a];
          T *o = c0 + (a + b * N1c + c * Pc32);
          if constexpr (STORE)
            *o = v;
          else
            *o = vc[it] + v;
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
    apply_k(const gtps_dg_level *lvl, T *y, const T *x, const T *b, bool dot, int nvec)
    {
      GTPS_REQUIRE(nvec >= 1 && nvec <= lvl->tcap, "DG operator scratch too small for the batch");
      hipStream_t s   = lvl->ctx->stream;
      const T    *tab = (const T *)lvl->tab;
      T *t0 = (T *)lvl->t[0], *t1 = (T *)lvl->t[1], *t2 = (T *)lvl->t[2],
        *t3 = (T *)lvl->t[3];
      // passes 0 and 1: one thread per cell line of K+1 DoFs; pass 2: per DoF
      const dim3 nbl(dgrid(lvl->N / (K + 1)), nvec);
      const dim3 nb(dgrid(lvl->N), nvec);
      const dim3 nbd(dgrid(lvl->N, std::min(lvl->ctx->ws.capacity / nvec, 2048)), nvec);
      double    *pt = lvl->ctx->ws.partials;
      const T    sc = (T)lvl->scale;
      GTPS_REQUIRE(lvl->N < (1LL << 31), "DG level too large for the operator kernels");
      const FastDiv dn1 = make_fastdiv((unsigned)lvl->N1), dn = make_fastdiv((unsigned)lvl->n);
      hipLaunchKernelGGL((dg_op_kernel<K, T, 0, false, false>), nbl, dim3(DNT), 0,
                         s, x, nullptr, t0, t1, tab, nullptr, nullptr, lvl->n,
                         lvl->N1, lvl->N, sc, pt, dn1, dn);
      hipLaunchKernelGGL((dg_op_kernel<K, T, 1, false, false>), nbl, dim3(DNT), 0,
                         s, t0, t1, t2, t3, tab, nullptr, nullptr, lvl->n,
                         lvl->N1, lvl->N, sc, pt, dn1, dn);
      // x pass: whole rows staged in LDS when a row fits, else per DoF
      const bool rows_fit = lvl->N1 <= DG_ROW_CAP;
      const int  RB       = std::max(1, 1024 / lvl->N1);
      const dim3 nbr(dgrid((lvl->N / lvl->N1 + RB - 1) / RB * DNT), nvec);
      const dim3 nbrd(dgrid((lvl->N / lvl->N1 + RB - 1) / RB * DNT,
                            std::min(lvl->ctx->ws.capacity / nvec, 2048)),
                      nvec);
#define GTPS_DX(R, D, G)                                                             \
  do                                                                                 \
    {                                                                                \
      if (rows_fit)                                                                  \
        hipLaunchKernelGGL((dg_op_x_kernel<K, T, R, D>), D ? nbrd : nbr, dim3(DNT), 0, \
                           s, t2, t3, y, tab, b, x, lvl->n, lvl->N1, lvl->N, RB, sc, pt); \
      else                                                                           \
        hipLaunchKernelGGL((dg_op_kernel<K, T, 2, R, D>), G, dim3(DNT), 0, s, t2, t3, \
                           y, nullptr, tab, b, x, lvl->n, lvl->N1, lvl->N, sc, pt,   \
                           dn1, dn);                                                 \
    }                                                                                \
  while (0)
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
      return dot ? (int)(rows_fit ? nbrd.x : nbd.x) : 0;
    }

    // n_sweeps sweeps, alternating colour order from `first`
    // (GTPS_MULT_FWD / _REV)
    template <int K, typename T>
    void
    sweeps_k(const gtps_dg_level *lvl, T *u, const T *b, int first, int n_sweeps,
             double relax, int columns = 1, int nvec = 1)
    {
      if constexpr (DP<K, T>::ok)
        {
          const int   n = lvl->n;
          hipStream_t s = lvl->ctx->stream;
          GTPS_REQUIRE(columns == 1 || n <= 4, "batched sweeps need at most 4 cells per direction");
          GTPS_REQUIRE(lvl->N < (1LL << 31), "DG level too large for the patch kernel");
          if (n <= 4 && nvec > 1)
            columns = nvec;  // independent vectors: one work-group each
          if (n <= 4)
            {
              hipLaunchKernelGGL((dg_patch_seq_kernel<K, T>), dim3(columns),
                                 dim3(DP<K, T>::NT), 0, s, u, b, (const T *)lvl->tab, n,
                                 first == GTPS_MULT_REV ? 1 : 0, n_sweeps,
                                 (T)lvl->scale, (T)relax, lvl->N);
              GTPS_LAUNCH_CHECK();
              return;
            }
          for (int sw = 0; sw < n_sweeps; ++sw)
            {
              const bool rev = ((first == GTPS_MULT_REV ? 1 : 0) + sw) & 1;
              for (int cc = 0; cc < 27; ++cc)
                {
                  const int c  = rev ? 26 - cc : cc;
                  const int cx = c % 3, cy = (c / 3) % 3, cz = c / 9;
                  if (cx >= n - 1 || cy >= n - 1 || cz >= n - 1)
                    continue;
                  const int       npx = (n - 1 - cx + 2) / 3, npy = (n - 1 - cy + 2) / 3;
                  const long long np  = (long long)npx * npy * ((n - 1 - cz + 2) / 3);
                  hipLaunchKernelGGL((dg_patch_kernel<K, T>), dim3((unsigned)np, nvec),
                                     dim3(DP<K, T>::NT),
                                     0, s, u, b, (const T *)lvl->tab, n, c, (T)lvl->scale,
                                     (T)relax, make_fastdiv((unsigned)npx),
                                     make_fastdiv((unsigned)(npx * npy)));
                  GTPS_LAUNCH_CHECK();
                }
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

  void
  dg_level_reserve(gtps_dg_level *lvl, int nb)
  {
    if (nb <= lvl->tcap)
      return;
    GTPS_HIP(hipStreamSynchronize(lvl->ctx->stream));
    for (void *&p : lvl->t)
      {
        (void)hipFree(p);
        p = nullptr;
      }
    lvl->tcap = 0;
    for (void *&p : lvl->t)
      GTPS_HIP(hipMalloc(&p, lvl->elem() * lvl->N * nb));
    lvl->tcap = nb;
  }

  int
  launch_dg_apply(const gtps_dg_level *lvl, void *y, const void *x,
                  const void *b, bool fused_dot, int nvec)
  {
    return dispatch_k(lvl->k, [&](auto kk) {
      constexpr int K = decltype(kk)::value;
      if (lvl->prec == GTPS_FP32)
        return apply_k<K, float>(lvl, (float *)y, (const float *)x,
                                 (const float *)b, fused_dot, nvec);
      return apply_k<K, double>(lvl, (double *)y, (const double *)x,
                                (const double *)b, fused_dot, nvec);
    });
  }

  void
  launch_dg_sweep(const gtps_dg_level *lvl, void *u, const void *b,
                  int variant, double relax, int nvec)
  {
    launch_dg_sweeps(lvl, u, b, variant, 1, relax, nvec);
  }

  void
  launch_dg_sweeps(const gtps_dg_level *lvl, void *u, const void *b, int first,
                   int n_sweeps, double relax, int nvec)
  {
    GTPS_REQUIRE(lvl->n >= 2, "the DG smoother needs at least 2 cells per direction");
    GTPS_REQUIRE(n_sweeps >= 0, "negative sweep count");
    if (n_sweeps == 0)
      return;
    const int variant = first;
    dispatch_k(lvl->k, [&](auto kk) {
      constexpr int K = decltype(kk)::value;
      if (lvl->prec == GTPS_FP32)
        sweeps_k<K, float>(lvl, (float *)u, (const float *)b, variant, n_sweeps, relax, 1,
                           nvec);
      else
        sweeps_k<K, double>(lvl, (double *)u, (const double *)b, variant, n_sweeps, relax, 1,
                            nvec);
      return 0;
    });
  }

  void
  launch_dg_sweeps_batched(const gtps_dg_level *lvl, void *u, const void *b, int first,
                           int n_sweeps, double relax, int columns)
  {
    GTPS_REQUIRE(lvl->n >= 2 && lvl->n <= 4, "batched sweeps need 2..4 cells per direction");
    GTPS_REQUIRE(columns >= 1 && n_sweeps >= 1, "bad batch");
    dispatch_k(lvl->k, [&](auto kk) {
      constexpr int K = decltype(kk)::value;
      if (lvl->prec == GTPS_FP32)
        sweeps_k<K, float>(lvl, (float *)u, (const float *)b, first, n_sweeps, relax, columns);
      else
        sweeps_k<K, double>(lvl, (double *)u, (const double *)b, first, n_sweeps, relax,
                            columns);
      return 0;
    });
  }

  void
  launch_dg_prolongate_add(const gtps_dg_level *fl, void *fine, const void *coarse, int nvec)
  {
    GTPS_REQUIRE(fl->n % 2 == 0, "fine level must have an even cell count");
    GTPS_REQUIRE(fl->N < (1LL << 31), "DG level too large for the transfer kernels");
    // one work-group per coarse cell
    const unsigned ncc = (unsigned)((fl->n / 2) * (fl->n / 2) * (fl->n / 2));
    dispatch_k(fl->k, [&](auto kk) {
      constexpr int K = decltype(kk)::value;
      if (fl->prec == GTPS_FP32)
        hipLaunchKernelGGL((dg_prolongate_kernel<K, float>), dim3(ncc, nvec),
                           dim3(DNT), 0, fl->ctx->stream, (const float *)coarse,
                           (float *)fine, (const float *)fl->tab, fl->n, fl->N);
      else
        hipLaunchKernelGGL((dg_prolongate_kernel<K, double>), dim3(ncc, nvec),
                           dim3(DNT), 0, fl->ctx->stream, (const double *)coarse,
                           (double *)fine, (const double *)fl->tab, fl->n, fl->N);
      GTPS_LAUNCH_CHECK();
      return 0;
    });
  }

  void
  launch_dg_restrict(const gtps_dg_level *fl, void *coarse, const void *fine,
                     bool store, int nvec)
  {
    GTPS_REQUIRE(fl->n % 2 == 0, "fine level must have an even cell count");
    GTPS_REQUIRE(fl->N < (1LL << 31), "DG level too large for the transfer kernels");
    const long long N1c = (long long)(fl->n / 2) * (fl->k + 1);
    const long long Nc  = N1c * N1c * N1c;
    // one work-group per coarse cell
    const unsigned ncc = (unsigned)((fl->n / 2) * (fl->n / 2) * (fl->n / 2));
    dispatch_k(fl->k, [&](auto kk) {
      constexpr int K = decltype(kk)::value;
      if (fl->prec == GTPS_FP32)
        {
          if (store)
            hipLaunchKernelGGL((dg_restrict_kernel<K, float, true>), dim3(ncc, nvec),
                               dim3(DNT), 0, fl->ctx->stream, (const float *)fine,
                               (float *)coarse, (const float *)fl->tab, fl->n, Nc);
          else
            hipLaunchKernelGGL((dg_restrict_kernel<K, float, false>), dim3(ncc, nvec),
                               dim3(DNT), 0, fl->ctx->stream, (const float *)fine,
                               (float *)coarse, (const float *)fl->tab, fl->n, Nc);
        }
      else
        {
          if (store)
            hipLaunchKernelGGL((dg_restrict_kernel<K, double, true>), dim3(ncc, nvec),
                               dim3(DNT), 0, fl->ctx->stream, (const double *)fine,
                               (double *)coarse, (const double *)fl->tab, fl->n, Nc);
          else
            hipLaunchKernelGGL((dg_restrict_kernel<K, double, false>), dim3(ncc, nvec),
                               dim3(DNT), 0, fl->ctx->stream, (const double *)fine,
                               (double *)coarse, (const double *)fl->tab, fl->n, Nc);
        }
      GTPS_LAUNCH_CHECK();
      return 0;
    });
  }
} // namespace gtps
