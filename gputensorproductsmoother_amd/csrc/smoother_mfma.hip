// smoother_mfma.hip - the multiplicative vertex-patch smoother for high
// degrees (3D, FP32, k = 4..7) with every 1D contraction on the matrix
// cores.
//
// A patch box has NP = 2k+1 <= 15 values per direction and NI = 2k-1
// interior ones.  Padded to 16, each step of the local solve is one 16x16
// matrix applied along one direction to the 256 lines of a 16^3 array:
// 16 GEMM tiles of v_mfma_f32_16x16x4f32 (4 k-steps each), 2 tiles per
// wavefront (8 wavefronts per patch; 4 and 16 are tuning options).  The steps are those of the LDS/FMA kernel
// (loop_kernel.cuh:969-1060 in the reference; same BD form here):
//   residual  r_I = b_I - h (Kx My Mz + Mx Ky Mz + Mx My Kz)_IB u_B
//             (z pass: Mz, Kz; y pass: My Mz, Ky Mz + My Kz; x pass),
//   u_I = relax (S S S) diag(1 / h(l_i+l_j+l_k)) (S S S)^T r_I.
// Lane layout of the instruction (checked on the device): lane l supplies
// A[i = l%16][k = l/16] and B[k = l/16][n = l%16] and receives
// D[4(l/16) + r][l%16], r = 0..3.  Lane l feeds inner index j = 4(l/16) + s
// in k-step s, so a contraction's output (rows 4(l/16)..+3 of line l%16)
// is exactly the B operand of a following contraction along the same
// direction: x-residual -> S^T along x and the eigenvalue scaling -> S
// along z are chained in registers without an LDS round trip.
// LDS holds two arrays of the 13 interior z-planes (16 x 16, x fastest).
#include "gtps_internal.h"

namespace gtps
{
  namespace
  {
    typedef float floatx4 __attribute__((ext_vector_type(4)));

    constexpr int NZ    = 13;   // z-planes of an LDS array (interior rows)
    constexpr int ARR   = NZ * 256;

    __device__ __forceinline__ int
    lds_index(int x, int y, int z)
    {
      return x + 16 * y + 256 * z;
    }

    // element (j along D) of line n of tile t; lines of a tile: D = x: y = n,
    // z = t; D = y: x = n, z = t; D = z: x = n, y = t
    template <int D>
    __device__ __forceinline__ int
    line_index(int t, int n, int j)
    {
      if (D == 0)
        return lds_index(j, n, t);
      if (D == 1)
        return lds_index(n, j, t);
      return lds_index(n, t, j);
    }

    // acc += W (16x16, this lane's 4 values w[s] = W[l%16][4(l/16)+s])
    // applied along D to the 16 lines of tile t of `in`; along z only the
    // NZ stored planes are read (the padded weights of the others are 0)
    template <int D>
    __device__ __forceinline__ floatx4
    contract(const float *in, int t, const float (&w)[4], floatx4 acc)
    {
      const int l = threadIdx.x & 63, n = l & 15, kq = l >> 4;
#pragma unroll
      for (int s = 0; s < 4; ++s)
        {
          const int j = 4 * kq + s;
          const float b = (D != 2 || j < NZ) ? in[line_index<D>(t, n, j)] : 0.f;
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(w[s], b, acc, 0, 0, 0);
        }
      return acc;
    }

    // acc += W applied along the direction in which `v` is a contraction
    // output (register chaining)
    __device__ __forceinline__ floatx4
    contract_regs(const floatx4 &v, const float (&w)[4], floatx4 acc)
    {
#pragma unroll
      for (int s = 0; s < 4; ++s)
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(w[s], v[s], acc, 0, 0, 0);
      return acc;
    }

    // rows 4(l/16)..+3 of line l%16; along z only the NZ stored planes
    template <int D>
    __device__ __forceinline__ void
    store_tile(float *out, int t, const floatx4 &v)
    {
      const int l = threadIdx.x & 63, n = l & 15, kq = l >> 4;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (D != 2 || 4 * kq + r < NZ)
          out[line_index<D>(t, n, 4 * kq + r)] = v[r];
    }

    // this lane's A-operand values of a padded 16x16 matrix given by f(i, j)
    template <typename F>
    __device__ __forceinline__ void
    load_weights(float (&w)[4], F &&f)
    {
      const int l = threadIdx.x & 63, i = l & 15, kq = l >> 4;
#pragma unroll
      for (int s = 0; s < 4; ++s)
        w[s] = f(i, 4 * kq + s);
    }

    // NW wavefronts per patch (4, 8 or 16; 8 by default).  The matrix-core
    // work of a patch is the same, but with 8 wavefronts 4 patches (and 8
    // wavefronts per SIMD) are resident per CU, where 4 give 5 patches at 5
    // wavefronts per SIMD (92 registers at k = 7) and 16 give 2: more
    // wavefronts overlap each other's LDS and global latency on the patch
    // chain.  Tiles of a stage: t = wave + NW tt, tt < 16 / NW; a wavefront
    // owns the same z-planes t in every stage.  Every LDS array holds the
    // NZ = 13 interior z-planes (16 x 16 each, x fastest): the z pass reads
    // the box straight from global memory into its B operands, so no array
    // holds the 15 box planes, and 2 arrays (26 KiB) fit 6 work-groups per
    // CU.  Tiles whose lines lie outside the interior are skipped (or
    // stored as zeros where a later stage reads them).
    template <int K, int NW>
    __global__ void __launch_bounds__(64 * NW)
    patch3d_mfma_kernel(float *__restrict__ u, const float *__restrict__ bvec,
                        const float *__restrict__ tab, const int n, const int colour,
                        const float scale, const float inv_scale, const float relax,
                        const FastDiv div_x, const FastDiv div_xy)
    {
      using TB         = Tab<K>;
      constexpr int NP = TB::NP, NI = TB::NI;
      static_assert(NP <= 16 && NI <= NZ, "patch box must fit a 16-wide tile");
      static_assert(NW == 4 || NW == 8 || NW == 16, "4, 8 or 16 wavefronts per patch");
      constexpr int TT = 16 / NW;  // tiles per wavefront and stage
      __shared__ float sA[ARR], sB[ARR];

      // level sizes < 2^31 (host check): offsets from the patch's first
      // DoF are 32-bit
      const int N1   = K * n + 1;
      const int P    = N1 * N1;
      const int tid  = threadIdx.x;
      const int wave = tid >> 6, l = tid & 63, ln = l & 15, kq = l >> 4;
      const int px = colour & 1, py = (colour >> 1) & 1, pz = (colour >> 2) & 1;
      // patch coordinates: divisions by the patch counts as multiply-high
      // and shift (divisors set up on the host)
      const int q  = xcd_logical_id(blockIdx.x, gridDim.x);
      const int qz = (int)fastdiv((unsigned)q, div_xy), rq = q - qz * (int)div_xy.d;
      const int qy = (int)fastdiv((unsigned)rq, div_x), qx = rq - qy * (int)div_x.d;
      const long long first = (long long)K * (2 * qx + px) + (long long)K * (2 * qy + py) * N1 +
                              (long long)K * (2 * qz + pz) * P;
      float *__restrict__ up       = u + first;
      const float *__restrict__ bp = bvec + first;

      const float *pM = tab + TB::patchM, *pK = tab + TB::patchK, *S = tab + TB::S;
      const float *linv = tab + TB::lamInv3;
      float wM[4], wK[4], wST[4], wS[4];
      // interior rows of the patch matrices; S^T and S, zero-padded
      load_weights(wM, [&](int i, int j) { return i < NI && j < NP ? pM[(i + 1) * NP + j] : 0.f; });
      load_weights(wK, [&](int i, int j) { return i < NI && j < NP ? pK[(i + 1) * NP + j] : 0.f; });
      load_weights(wST, [&](int i, int j) { return i < NI && j < NI ? S[j * NI + i] : 0.f; });
      load_weights(wS, [&](int i, int j) { return i < NI && j < NI ? S[i * NI + j] : 0.f; });

      // the z pass's B operands: u on the box skeleton (interior zeroed) at
      // x = ln, y = t, z = 4(l/16) + s; this lane's interior b values (x
      // pass) and inverse eigenvalue sums (z scaling): every load issued
      // before the first wait
      float uz[TT][4], bv[TT][4];
#pragma unroll
      for (int tt = 0; tt < TT; ++tt)
#pragma unroll
        for (int s = 0; s < 4; ++s)
          {
            const int  t = wave + NW * tt, z = 4 * kq + s;
            const bool inner = ln > 0 && ln < NP - 1 && t > 0 && t < NP - 1 && z > 0 &&
                               z < NP - 1;
            uz[tt][s] = (ln < NP && t < NP && z < NP && !inner)
                          ? up[ln + t * N1 + z * P]
                          : 0.f;
          }
#pragma unroll
      for (int tt = 0; tt < TT; ++tt)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          {
            const int t = wave + NW * tt, xi = 4 * kq + r;
            bv[tt][r]   = (xi < NI && ln < NI && t < NI)
                            ? bp[(xi + 1) + (ln + 1) * N1 + (t + 1) * P]
                            : 0.f;
          }
      const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
      // z pass (lines x = ln, y = t): A = Mz u, B = Kz u
#pragma unroll
      for (int tt = 0; tt < TT; ++tt)
        {
          const int t = wave + NW * tt;
          floatx4   m = zero, k = zero;
          if (t < NP)
            {
              floatx4 v = {uz[tt][0], uz[tt][1], uz[tt][2], uz[tt][3]};
              m = contract_regs(v, wM, zero);
              k = contract_regs(v, wK, zero);
            }
          store_tile<2>(sA, t, m);
          store_tile<2>(sB, t, k);
        }
      __syncthreads();
      // The y pass, the x pass and S^T along y read and write the z-planes
      // t of their own tiles only (tile t: lines along y or x in plane
      // z = t), so they follow each other within the wavefront, in place,
      // with no work-group barrier: a wavefront's LDS operations complete
      // in order, and the fence keeps the compiler from moving them.
      auto wave_fence = []() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      };
      // y pass (lines x = ln, z = t < NZ): C = My A into B, A' = Ky A + My B
      // into A
#pragma unroll
      for (int tt = 0; tt < TT; ++tt)
        {
          const int t = wave + NW * tt;
          if (t < NZ)
            {
              const floatx4 c1 = contract<1>(sA, t, wM, zero);
              const floatx4 c2 = contract<1>(sB, t, wM, contract<1>(sA, t, wK, zero));
              store_tile<1>(sB, t, c1);
              store_tile<1>(sA, t, c2);
            }
        }
      wave_fence();
      // x pass (lines y = ln, z = t): r = b_I - h (Kx C + Mx A'), then S^T
      // along x in registers, into B
#pragma unroll
      for (int tt = 0; tt < TT; ++tt)
        {
          const int t = wave + NW * tt;
          if (t < NZ)
            {
              const floatx4 acc = contract<0>(sA, t, wM, contract<0>(sB, t, wK, zero));
              floatx4       res;
#pragma unroll
              for (int r = 0; r < 4; ++r)
                res[r] = (4 * kq + r < NI && ln < NI && t < NI) ? bv[tt][r] - scale * acc[r]
                                                                 : 0.f;
              store_tile<0>(sB, t, contract_regs(res, wST, zero));
            }
        }
      wave_fence();
      // the inverse eigenvalue sums of the scaling stage, loaded here (the b
      // values are no longer live): their latency passes behind S^T along y
      float lv[TT][4];
#pragma unroll
      for (int tt = 0; tt < TT; ++tt)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          {
            const int t = wave + NW * tt, zi = 4 * kq + r;
            lv[tt][r]   = (t < NI && zi < NI && ln < NI) ? linv[(zi * NI + t) * NI + ln] : 0.f;
          }
      // S^T along y (lines x = ln, z = t): B -> A
#pragma unroll
      for (int tt = 0; tt < TT; ++tt)
        {
          const int t = wave + NW * tt;
          if (t < NZ)
            store_tile<1>(sA, t, contract<1>(sB, t, wST, zero));
        }
      __syncthreads();
      // S^T along z (lines x = ln, y = t), eigenvalue scaling, S along z in
      // registers: A -> B (rows y >= NI stored as zeros)
#pragma unroll
      for (int tt = 0; tt < TT; ++tt)
        {
          const int t = wave + NW * tt;
          floatx4   out = zero;
          if (t < NI)
            {
              floatx4 acc = contract<2>(sA, t, wST, zero);
#pragma unroll
              for (int r = 0; r < 4; ++r)
                {
                  const int zi = 4 * kq + r;
                  acc[r] = (zi < NI && ln < NI) ? acc[r] * lv[tt][r] * inv_scale : 0.f;
                }
              out = contract_regs(acc, wS, zero);
            }
          store_tile<2>(sB, t, out);
        }
      __syncthreads();
      // S along y (lines x = ln, z = t): B -> A, and S along x on the same
      // planes of this wavefront
#pragma unroll
      for (int tt = 0; tt < TT; ++tt)
        {
          const int t = wave + NW * tt;
          if (t < NZ)
            store_tile<1>(sA, t, contract<1>(sB, t, wS, zero));
        }
      wave_fence();
      // S along x (lines y = ln, z = t), write the interior
#pragma unroll
      for (int tt = 0; tt < TT; ++tt)
        {
          const int t = wave + NW * tt;
          if (t < NI)
            {
              const floatx4 acc = contract<0>(sA, t, wS, zero);
#pragma unroll
              for (int r = 0; r < 4; ++r)
                {
                  const int xi = 4 * kq + r;
                  if (xi < NI && ln < NI)
                    up[(xi + 1) + (ln + 1) * N1 + (t + 1) * P] = relax * acc[r];
                }
            }
        }
    }
  } // namespace

  bool
  mfma_smoother_supported(int dim, int k, int prec)
  {
    return dim == 3 && prec == GTPS_FP32 && k >= 4 && k <= 7;
  }

  // one multiplicative colour pass (mode 0 form: b in src)
  void
  launch_patch_colour_mfma(const gtps_level *lvl, void *u, const void *src, int colour,
                           double relax)
  {
    GTPS_REQUIRE(mfma_smoother_supported(lvl->dim, lvl->k, lvl->prec),
                 "the matrix-core smoother is 3D FP32 with degree 4..7");
    const int n  = lvl->n;
    const int px = colour & 1, py = (colour >> 1) & 1, pz = (colour >> 2) & 1;
    const long long np = (long long)((n - px) / 2) * ((n - py) / 2) * ((n - pz) / 2);
    if (np <= 0)
      return;
    GTPS_REQUIRE(lvl->N < (1LL << 31), "level too large for the matrix-core smoother");
    const int     npx = (n - px) / 2, npy = (n - py) / 2;
    const FastDiv dx  = make_fastdiv((unsigned)npx);
    const FastDiv dxy = make_fastdiv((unsigned)(npx * npy));
    const float *tab   = reinterpret_cast<const float *>(lvl->dtab());
    const float  scale = (float)lvl->scale, inv = (float)(1.0 / lvl->scale);
    hipStream_t  s     = lvl->ctx->stream;
    const int nw = tuning().mfma_waves;
#define GTPS_MF(KK)                                                                      \
  do                                                                                     \
    {                                                                                    \
      if (nw == 16)                                                                      \
        hipLaunchKernelGGL((patch3d_mfma_kernel<KK, 16>), dim3((unsigned)np), dim3(1024), \
                           0, s, (float *)u, (const float *)src, tab, n, colour, scale,  \
                           inv, (float)relax, dx, dxy);                                  \
      else if (nw == 8)                                                                  \
        hipLaunchKernelGGL((patch3d_mfma_kernel<KK, 8>), dim3((unsigned)np), dim3(512),   \
                           0, s, (float *)u, (const float *)src, tab, n, colour, scale,  \
                           inv, (float)relax, dx, dxy);                                  \
      else                                                                               \
        hipLaunchKernelGGL((patch3d_mfma_kernel<KK, 4>), dim3((unsigned)np), dim3(256),   \
                           0, s, (float *)u, (const float *)src, tab, n, colour, scale,  \
                           inv, (float)relax, dx, dxy);                                  \
    }                                                                                    \
  while (0)
    switch (lvl->k)
      {
        case 4: GTPS_MF(4); break;
        case 5: GTPS_MF(5); break;
        case 6: GTPS_MF(6); break;
        case 7: GTPS_MF(7); break;
        default: break;
      }
#undef GTPS_MF
    GTPS_LAUNCH_CHECK();
  }
} // namespace gtps
