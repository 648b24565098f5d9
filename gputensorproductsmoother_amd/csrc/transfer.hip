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
// Each direction is one pass (x, then y, then z) through two level-sized
// temporaries owned by the transfer object; the last pass accumulates into
// the destination, so no atomics are needed and the result is deterministic.
#include "gtps_internal.h"

namespace gtps
{
  namespace
  {
    constexpr int TR_NT = 256;

    // OUT: 0 store, 1 add, 2 add except on coarse Dirichlet nodes
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
                out[o] += acc;
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
        T *coarse, const T *coarse_in, bool prolong)
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
              pass<K, T, false, false, 2>(s, t1, coarse, tab, dim, 1, e1, e2, nc, Nc1);
            }
          else
            {
              pass<K, T, false, false, 0>(s, fine_in, t1, tab, dim, 0, e0, e1, nc, Nc1);
              pass<K, T, false, false, 0>(s, t1, t2, tab, dim, 1, e1, e2, nc, Nc1);
              pass<K, T, false, false, 2>(s, t2, coarse, tab, dim, 2, e2, e3, nc, Nc1);
            }
        }
    }

    template <typename T>
    void
    run_t(const gtps_transfer *tr, int fine_level, void *fine,
          const void *fine_in, void *coarse, const void *coarse_in,
          bool prolong)
    {
      T       *f  = reinterpret_cast<T *>(fine);
      const T *fi = reinterpret_cast<const T *>(fine_in);
      T       *c  = reinterpret_cast<T *>(coarse);
      const T *ci = reinterpret_cast<const T *>(coarse_in);
      switch (tr->k)
        {
          case 1: return run<1, T>(tr, fine_level, f, fi, c, ci, prolong);
          case 2: return run<2, T>(tr, fine_level, f, fi, c, ci, prolong);
          case 3: return run<3, T>(tr, fine_level, f, fi, c, ci, prolong);
          case 4: return run<4, T>(tr, fine_level, f, fi, c, ci, prolong);
          case 5: return run<5, T>(tr, fine_level, f, fi, c, ci, prolong);
          case 6: return run<6, T>(tr, fine_level, f, fi, c, ci, prolong);
          case 7: return run<7, T>(tr, fine_level, f, fi, c, ci, prolong);
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
} // namespace gtps
