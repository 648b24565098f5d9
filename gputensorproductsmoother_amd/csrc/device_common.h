// device_common.h - shared device-side definitions for the gfx950 kernels.
//
// Table layout: every degree K owns one packed array of 1D data per
// precision (built once per context by setup.hip), passed to kernels as a
// `const T* __restrict__` so that uniform, compile-time-indexed reads become
// scalar loads.  Offsets are compile-time functions of K.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gtps
{
  constexpr int MAX_DEGREE = 7;
  constexpr int MAX_QUAD   = 16;  // quadrature points per direction (error norms)
  constexpr int MAX_DENSE_COARSE = 4096;  // interior DoFs of a tabulated level-2 V-cycle

  template <int K>
  struct Tab
  {
    static constexpr int N  = K + 1;      // cell DoFs per direction
    static constexpr int NP = 2 * K + 1;  // patch box per direction
    static constexpr int NI = 2 * K - 1;  // patch interior per direction
    static constexpr int W  = 2 * K + 1;  // width of a global 1D matrix row
    // reference-cell mass and stiffness [N*N]
    static constexpr int cellM = 0;
    static constexpr int cellK = cellM + N * N;
    // 2-cell patch mass and stiffness [NP*NP]
    static constexpr int patchM = cellK + N * N;
    static constexpr int patchK = patchM + NP * NP;
    // M-orthonormal interior eigenvectors S[i*NI+j] (vector j) and values
    static constexpr int S   = patchK + NP * NP;
    static constexpr int lam = S + NI * NI;
    // embedding (prolongation) matrix E[q*N+a], q in [0,2K], a in [0,K]
    static constexpr int E = lam + NI;
    // rows of the assembled global 1D matrices, by node position p in its
    // cell (p = 0: vertex row), band[p*W + off + K] for off in [-K, K]
    static constexpr int bandM = E + NP * N;
    static constexpr int bandK = bandM + K * W;
    // 1 / (l_i + l_j + l_k), [k][j][i], and 1 / (l_i + l_j), [j][i]
    static constexpr int lamInv3 = bandK + K * W;
    static constexpr int lamInv2 = lamInv3 + NI * NI * NI;
    // Even/odd split of the eigenbasis.  The interior pencil is invariant
    // under the reflection i -> NI-1-i, so every eigenvector is even or odd.
    // With MH = K-1 the middle index, the K even vectors are stored by their
    // rows 0..MH (SE[i*K+e]) and the K-1 odd ones by rows 0..MH-1
    // (SO[i*(K-1)+o]); eigen coordinates are renumbered even first, then
    // odd, and lamInv3p holds 1/(l_a+l_b+l_c) in that numbering [c][b][a].
    static constexpr int MH      = K - 1;
    static constexpr int NE = K, NO = K - 1;
    static constexpr int SE      = lamInv2 + NI * NI;
    static constexpr int SO      = SE + (MH + 1) * NE;
    static constexpr int lamInv3p = SO + (MH > 0 ? MH * NO : 0);
    static constexpr int size    = lamInv3p + NI * NI * NI;
  };

  // Column range of row r of a 2-cell patch matrix (block sparsity: rows
  // inside cell 0 couple to nodes 0..K, rows inside cell 1 to K..2K, the
  // shared vertex row to all).
  template <int K>
  __host__ __device__ constexpr int
  prow_lo(int r)
  {
    return r <= K ? 0 : K;
  }
  template <int K>
  __host__ __device__ constexpr int
  prow_hi(int r)
  {
    return r < K ? K : 2 * K;
  }
  // Column offsets of a global 1D row for node position p within its cell.
  template <int K>
  __host__ __device__ constexpr int
  band_lo(int p)
  {
    return p == 0 ? -K : -p;
  }
  template <int K>
  __host__ __device__ constexpr int
  band_hi(int p)
  {
    return p == 0 ? K : K - p;
  }

  template <typename T>
  __device__ __forceinline__ T
  wave_sum(T v)
  {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
      v += __shfl_xor(v, off, 64);
    return v;
  }

  // Sum over the block; result valid in thread 0.  `red` holds >= NT/64
  // doubles of LDS.
  template <int NT>
  __device__ __forceinline__ double
  block_sum(double v, double *red)
  {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
      red[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
      {
#pragma unroll
        for (int w = 0; w < NT / 64; ++w)
          s += red[w];
      }
    return s;
  }

  // division by an invariant divisor d >= 1 for 0 <= n < 2^31:
  // n / d = (umulhi(n, m) + n) >> s with s = ceil(log2 d),
  // m = floor(2^32 (2^s - d) / d) + 1 (set up on the host by make_fastdiv)
  struct FastDiv
  {
    unsigned d, m;
    int      s;
  };
  inline FastDiv
  make_fastdiv(unsigned d)
  {
    int s = 0;
    while ((1ull << s) < d)
      ++s;
    const unsigned long long m = ((1ull << 32) * ((1ull << s) - d)) / d + 1;
    return FastDiv{d, (unsigned)m, s};
  }
  __device__ __forceinline__ unsigned
  fastdiv(unsigned n, const FastDiv &f)
  {
    return (__umulhi(n, f.m) + n) >> f.s;
  }

  __device__ __forceinline__ int
  linear_block_id()
  {
    return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  }

  // MI355X dispatches consecutive work-group ids round-robin over its 8
  // XCDs, each with a private L2.  This bijection maps the hardware id to a
  // logical id such that every XCD works through one contiguous range of
  // logical ids, so neighbouring tiles (which share halo / face data) run
  // on the same XCD and hit the same L2.
  constexpr int N_XCD = 8;
  __host__ __device__ __forceinline__ int
  xcd_logical_id(int hw, int n)
  {
    // unsigned: the divisions by 8 are shifts and masks
    const unsigned un = (unsigned)n, uh = (unsigned)hw;
    const unsigned q = un / N_XCD, r = un % N_XCD;
    const unsigned xcd = uh % N_XCD, slot = uh / N_XCD;
    return (int)((xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + slot);
  }
} // namespace gtps
