// smoother_mfma.hip - the multiplicative vertex-patch smoother for high
// degrees (3D, FP32, k = 4..7) with every 1D contraction on the matrix
// cores.
//
// A patch box has NP = 2k+1 <= 15 values per direction and NI = 2k-1
// interior ones.  Padded to 16, each step of the local solve is one 16x16
// matrix applied along one direction to the 256 lines of a 16^3 array:
// 16 GEMM tiles of v_mfma_f32_16x16x4f32 (4 k-steps each), 4 tiles per
// wavefront.  The steps are those of the LDS/FMA kernel
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
// LDS holds three 16^3 arrays, x fastest, z-plane stride 272 floats so that
// the strided reads of a half-wavefront fall in distinct banks.
#include "gtps_internal.h"

namespace gtps
{
  namespace
  {
    typedef float floatx4 __attribute__((ext_vector_type(4)));

    constexpr int MF_NT = 256;  // 4 wavefronts, one patch
    constexpr int ZS    = 272;  // z-plane stride of an LDS array
    constexpr int ARR   = 16 * ZS;

    __device__ __forceinline__ int
    lds_index(int x, int y, int z)
    {
      return x + 16 * y + ZS * z;
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
    // applied along D to the 16 lines of tile t of `in`
    template <int D>
    __device__ __forceinline__ floatx4
    contract(const float *in, int t, const float (&w)[4], floatx4 acc)
    {
      const int l = threadIdx.x & 63, n = l & 15, kq = l >> 4;
#pragma unroll
      for (int s = 0; s < 4; ++s)
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(w[s], in[line_index<D>(t, n, 4 * kq + s)],
                                                   acc, 0, 0, 0);
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

    template <int D>
    __device__ __forceinline__ void
    store_tile(float *out, int t, const floatx4 &v)
    {
      const int l = threadIdx.x & 63, n = l & 15, kq = l >> 4;
#pragma unroll
      for (int r = 0; r < 4; ++r)
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

    template <int K>
    __global__ void __launch_bounds__(MF_NT)
    patch3d_mfma_kernel(float *__restrict__ u, const float *__restrict__ bvec,
                        const float *__restrict__ tab, const int n, const int colour,
                        const float scale, const float inv_scale, const float relax)
    {
      using TB         = Tab<K>;
      constexpr int NP = TB::NP, NI = TB::NI;
      static_assert(NP <= 16, "patch box must fit a 16-wide tile");
      __shared__ float sP[ARR], sQ[ARR], sR[ARR];

      const int       N1   = K * n + 1;
      const long long P    = (long long)N1 * N1;
      const int       tid  = threadIdx.x;
      const int       wave = tid >> 6, l = tid & 63, ln = l & 15, kq = l >> 4;
      const int px = colour & 1, py = (colour >> 1) & 1, pz = (colour >> 2) & 1;
      const int npx = (n - px) / 2, npy = (n - py) / 2;
      const int q   = xcd_logical_id(blockIdx.x, gridDim.x);
      const int qx = q % npx, qy = (q / npx) % npy, qz = q / (npx * npy);
      const long long first = (long long)K * (2 * qx + px) + (long long)K * (2 * qy + py) * N1 +
                              (long long)K * (2 * qz + pz) * P;

      const float *pM = tab + TB::patchM, *pK = tab + TB::patchK, *S = tab + TB::S;
      const float *linv = tab + TB::lamInv3;
      float wM[4], wK[4], wST[4], wS[4];
      // interior rows of the patch matrices; S^T and S, zero-padded
      load_weights(wM, [&](int i, int j) { return i < NI && j < NP ? pM[(i + 1) * NP + j] : 0.f; });
      load_weights(wK, [&](int i, int j) { return i < NI && j < NP ? pK[(i + 1) * NP + j] : 0.f; });
      load_weights(wST, [&](int i, int j) { return i < NI && j < NI ? S[j * NI + i] : 0.f; });
      load_weights(wS, [&](int i, int j) { return i < NI && j < NI ? S[i * NI + j] : 0.f; });

      // u on the box skeleton (interior zeroed), zero padding, and this
      // lane's interior b values (consumed by the x pass): every load issued
      // before the first wait
      constexpr int IT = 16 * 16 * 16 / MF_NT;
      float         uv[IT], bv[4][4];
#pragma unroll
      for (int it = 0; it < IT; ++it)
        {
          const int e = tid + it * MF_NT;
          const int x = e & 15, y = (e >> 4) & 15, z = e >> 8;
          const bool inner = x > 0 && x < NP - 1 && y > 0 && y < NP - 1 && z > 0 && z < NP - 1;
          uv[it] = (x < NP && y < NP && z < NP && !inner)
                     ? u[first + x + (long long)y * N1 + (long long)z * P]
                     : 0.f;
        }
#pragma unroll
      for (int tt = 0; tt < 4; ++tt)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          {
            const int t = wave * 4 + tt, xi = 4 * kq + r;
            bv[tt][r]   = (xi < NI && ln < NI && t < NI)
                            ? bvec[first + (xi + 1) + (long long)(ln + 1) * N1 +
                                   (long long)(t + 1) * P]
                            : 0.f;
          }
#pragma unroll
      for (int it = 0; it < IT; ++it)
        {
          const int e = tid + it * MF_NT;
          sP[lds_index(e & 15, (e >> 4) & 15, e >> 8)] = uv[it];
        }
      __syncthreads();
      const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
      // z pass: Q = Mz u, R = Kz u
#pragma unroll
      for (int tt = 0; tt < 4; ++tt)
        {
          const int t = wave * 4 + tt;
          store_tile<2>(sQ, t, contract<2>(sP, t, wM, zero));
          store_tile<2>(sR, t, contract<2>(sP, t, wK, zero));
        }
      __syncthreads();
      // y pass: P = My Q, Q = Ky Q + My R (Q rewritten after every wave has
      // read it)
      floatx4 c2[4];
#pragma unroll
      for (int tt = 0; tt < 4; ++tt)
        {
          const int t = wave * 4 + tt;
          store_tile<1>(sP, t, contract<1>(sQ, t, wM, zero));
          c2[tt] = contract<1>(sR, t, wM, contract<1>(sQ, t, wK, zero));
        }
      __syncthreads();
#pragma unroll
      for (int tt = 0; tt < 4; ++tt)
        store_tile<1>(sQ, wave * 4 + tt, c2[tt]);
      __syncthreads();
      // x pass: r = b_I - h (Kx P + Mx Q), then S^T along x (registers)
#pragma unroll
      for (int tt = 0; tt < 4; ++tt)
        {
          const int t   = wave * 4 + tt;  // interior z
          floatx4   acc = contract<0>(sQ, t, wM, contract<0>(sP, t, wK, zero));
          floatx4   res;
#pragma unroll
          for (int r = 0; r < 4; ++r)
            {
              const int xi = 4 * kq + r;  // interior x; interior y = ln
              res[r] = (xi < NI && ln < NI && t < NI) ? bv[tt][r] - scale * acc[r] : 0.f;
            }
          store_tile<0>(sR, t, contract_regs(res, wST, zero));
        }
      __syncthreads();
      // S^T along y
#pragma unroll
      for (int tt = 0; tt < 4; ++tt)
        {
          const int t = wave * 4 + tt;
          store_tile<1>(sP, t, contract<1>(sR, t, wST, zero));
        }
      __syncthreads();
      // S^T along z, eigenvalue scaling, S along z (registers)
#pragma unroll
      for (int tt = 0; tt < 4; ++tt)
        {
          const int t   = wave * 4 + tt;  // y (eigen index); x = ln
          floatx4   acc = contract<2>(sP, t, wST, zero);
#pragma unroll
          for (int r = 0; r < 4; ++r)
            {
              const int zi = 4 * kq + r;
              acc[r] = (zi < NI && ln < NI && t < NI)
                         ? acc[r] * linv[(zi * NI + t) * NI + ln] * inv_scale
                         : 0.f;
            }
          store_tile<2>(sQ, t, contract_regs(acc, wS, zero));
        }
      __syncthreads();
      // S along y
#pragma unroll
      for (int tt = 0; tt < 4; ++tt)
        {
          const int t = wave * 4 + tt;
          store_tile<1>(sR, t, contract<1>(sQ, t, wS, zero));
        }
      __syncthreads();
      // S along x, write the interior
#pragma unroll
      for (int tt = 0; tt < 4; ++tt)
        {
          const int t   = wave * 4 + tt;  // interior z
          floatx4   acc = contract<0>(sR, t, wS, zero);
#pragma unroll
          for (int r = 0; r < 4; ++r)
            {
              const int xi = 4 * kq + r;
              if (xi < NI && ln < NI && t < NI)
                u[first + (xi + 1) + (long long)(ln + 1) * N1 + (long long)(t + 1) * P] =
                  relax * acc[r];
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
    const float *tab   = reinterpret_cast<const float *>(lvl->dtab());
    const float  scale = (float)lvl->scale, inv = (float)(1.0 / lvl->scale);
    hipStream_t  s     = lvl->ctx->stream;
#define GTPS_MF(KK)                                                                   \
  hipLaunchKernelGGL((patch3d_mfma_kernel<KK>), dim3((unsigned)np), dim3(MF_NT), 0, s, \
                     (float *)u, (const float *)src, tab, n, colour, scale, inv,       \
                     (float)relax)
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
