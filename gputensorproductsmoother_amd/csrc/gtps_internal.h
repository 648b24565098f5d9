// gtps_internal.h - host-side objects and kernel launchers of libgtps.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gtps.h"
#include "device_common.h"

namespace gtps
{
  struct Error : std::runtime_error
  {
    int code;
    Error(int c, const std::string &m)
      : std::runtime_error(m), code(c)
    {}
  };

#define GTPS_HIP(call)                                                       \
  do                                                                         \
    {                                                                        \
      hipError_t e_ = (call);                                                \
      if (e_ != hipSuccess)                                                  \
        throw ::gtps::Error(GTPS_EHIP, std::string(#call) + ": " +           \
                                         hipGetErrorString(e_));             \
    }                                                                        \
  while (0)

#define GTPS_REQUIRE(cond, msg)                                              \
  do                                                                         \
    {                                                                        \
      if (!(cond))                                                           \
        throw ::gtps::Error(GTPS_EINVAL, msg);                               \
    }                                                                        \
  while (0)

#define GTPS_LAUNCH_CHECK() GTPS_HIP(hipGetLastError())

  // FP64 1D data of one degree (host copy); eigenpairs come from the GPU.
  struct Basis1D
  {
    int                 k = 0;
    std::vector<double> nodes, qp, qw;
    std::vector<double> cellM, cellK, patchM, patchK, E, lam, S;
  };

  // host-side 1D basis (GL nodes, Gauss rule, cell and patch matrices,
  // embedding); eigenpairs are filled by setup_tables on the device.
  Basis1D make_basis_host(int k);

  struct DegreeTables
  {
    Basis1D basis;
    double *d64 = nullptr;  // packed Tab<K> layout, FP64
    float  *d32 = nullptr;  // same, FP32
    // eigenpairs of the 2-cell interior pencils (K_cell0 + rho K_cell1,
    // M_cell0 + rho M_cell1) on the ratio grid of the variable-coefficient
    // smoother (bins -SEP_HALF..SEP_HALF, built on first use): per bin
    // S (ni x ni, S^T M S = I) then lambda (ni), SEP_W(k) values
    double *sep64 = nullptr;
    float  *sep32 = nullptr;
  };

  // Separable fit of a vertex patch's 2^3 cell coefficients (variable-
  // coefficient smoother): per-direction log ratios rounded to multiples of
  // SEP_STEP (|bin| <= SEP_HALF), and the scale
  //   sigma = max(max_K c_K / f_K / SEP_THETA, exp(mean_K log(c_K / f_K))),
  // which bounds the largest eigenvalue of the local correction by
  // SEP_THETA < 2 (see coeff.hip).
  constexpr int    SEP_HALF  = 64;
  constexpr int    SEP_BINS  = 2 * SEP_HALF + 1;
  constexpr double SEP_STEP  = 1.0 / 16.0;
  constexpr double SEP_THETA = 1.5;
  constexpr int
  SEP_W(int k)
  {
    return (2 * k - 1) * (2 * k - 1) + (2 * k - 1);
  }
  // FP64 or FP32 ratio-grid tables of degree k (built on first use)
  const void *get_sep_tables(gtps_ctx *ctx, int k, int precision);

  struct Workspace
  {
    double *partials = nullptr;  // per-block partial sums
    int     capacity = 0;
    double *scalars  = nullptr;  // device scalars (CG state)
    double *pinned   = nullptr;  // host-pinned mirror
    // coherent, device-mapped host memory that reduction kernels write
    // directly (residual norms read by the host without a copy);
    // readback_dev is its device address
    double *readback     = nullptr;
    double *readback_dev = nullptr;
  };
} // namespace gtps

struct gtps_ctx
{
  int                                                     device = 0;
  hipStream_t                                             stream = nullptr;
  bool                                                    own_stream = false;
  gtps::Workspace                                         ws;
  std::map<int, std::unique_ptr<gtps::DegreeTables>>      tables;
};

struct gtps_level
{
  gtps_ctx                *ctx = nullptr;
  int                      dim = 3, k = 1, level = 1, n = 2, N1 = 3;
  long long                N   = 0;
  int                      prec = GTPS_FP64;
  gtps::DegreeTables      *tab  = nullptr;
  double                   scale = 1.0;  // h^(dim-2)
  // variable coefficient (level precision, one value per cell, cells
  // cx + n cy + n^2 cz) and the element-vector buffer of the cell-wise
  // operator; both null for the constant-coefficient Laplacian
  void                    *coef  = nullptr;
  void                    *celem = nullptr;
  // per vertex patch (lexicographic over (n-1)^3): the separable fit of the
  // variable-coefficient smoother, ratio bins packed as (bx + SEP_HALF) |
  // (by + SEP_HALF) << 8 | (bz + SEP_HALF) << 16, and the scale sigma (level
  // precision); set with the coefficient
  int                     *sep_bins  = nullptr;
  void                    *sep_sigma = nullptr;
  ~gtps_level()
  {
    if (coef)
      (void)hipFree(coef);
    if (celem)
      (void)hipFree(celem);
    if (sep_bins)
      (void)hipFree(sep_bins);
    if (sep_sigma)
      (void)hipFree(sep_sigma);
  }
  const void *
  dtab() const
  {
    return prec == GTPS_FP32 ? (const void *)tab->d32 : (const void *)tab->d64;
  }
  size_t
  elem() const
  {
    return prec == GTPS_FP32 ? sizeof(float) : sizeof(double);
  }
};

struct gtps_smoother
{
  gtps_level *lvl      = nullptr;
  int         variant  = GTPS_MULT_FWD;
  double      relax    = 1.0;
  void       *resid    = nullptr;  // additive variant: global residual
};

// SIPG-DG level (3D): n cells per direction, DoFs on one tensor grid
// X = cx (k+1) + ix, index X + N1 Y + N1^2 Z.
struct gtps_dg_level
{
  gtps_ctx *ctx = nullptr;
  int       dim = 3, k = 1, n = 2, N1 = 4;
  long long N   = 0;
  int       prec = GTPS_FP64;
  double    h = 0.5, scale = 0.5, pf = 1.0;
  void     *tab = nullptr;                     // packed DG table (prec)
  void     *t[4] = {nullptr, nullptr, nullptr, nullptr};  // operator scratch
  int       tcap = 1;  // vectors each operator scratch buffer holds
  ~gtps_dg_level()
  {
    if (tab)
      (void)hipFree(tab);
    for (void *p : t)
      if (p)
        (void)hipFree(p);
  }
  size_t
  elem() const
  {
    return prec == GTPS_FP32 ? sizeof(float) : sizeof(double);
  }
};

struct gtps_transfer
{
  gtps_ctx *ctx = nullptr;
  int       dim = 3, k = 1, min_level = 1, max_level = 1, prec = GTPS_FP64;
  void     *tmp1 = nullptr, *tmp2 = nullptr;
  gtps::DegreeTables *tab = nullptr;
};

namespace gtps
{
  // process-wide tuning knobs (gtps_set_tuning)
  struct Tuning
  {
    // smoother colours with at most this many patches use the single-patch
    // kernel, larger ones the multi-patch kernel (k <= 3).  Measured on
    // MI355X at 64^3 Q3 the single-patch kernel is at least as fast on every
    // level (4.3-7.7 us per coarse-level launch instead of 11-13 us; level 6
    // FP32 30.7 vs 31.2 us, FP64 41 vs 104 us), so it is the default.
    long long small_colour_patches = 1LL << 62;
    // patch-group kernel: colours with more than group_min_patches patches
    // pack patches_per_wave (1, 2 or 4) patches into each wavefront
    long long group_min_patches = 2048;
    int       patches_per_wave  = 2;
    // patch-group kernel: wavefronts per work-group (1 or 4; 4 applies to 1
    // or 2 patches per wavefront, each wavefront with its own patches; 0:
    // 4 in FP32, 1 in FP64).  Measured on MI355X at 64^3 Q3, level-6 colour
    // (HIP events): FP32 21.0 us with 4 against 21.4-23.1 us with 1 (fewer,
    // larger work-groups to dispatch), FP64 34.7 against 32.8-34.2 us.
    int       group_waves = 0;
    // 3D operator: levels with at most this many cells per direction (half
    // as many for k >= 5) use the one-round box kernel instead of
    // z-marching.  Measured on MI355X (residual, HIP events): Q7 FP32 8^3
    // 17.7 us box against 29.5 us marching, 16^3 77.7 against 30.6 us
    long long box_vmult_max_n = 16;
    // patch-group smoother with row-staged global accesses (1) or the
    // lane-pencil form (0).  Measured on MI355X at 64^3 Q3 FP32: 29.7 us
    // (row-staged) against 24.8 us (lane form) per colour: the LDS staging
    // and its extra barrier cost more than the fewer cache-line accesses
    // save, so the lane form is the default.
    int       row_staged = 0;
    // tabulate the V-cycle from level 2 down as a dense matrix (hierarchies
    // of at least 3 levels whose level-2 interior has <= MAX_DENSE_COARSE
    // DoFs)
    int       dense_coarse = 1;
    // 3D FP32 degree 4..7 multiplicative smoother on the matrix cores
    int       mfma_smoother = 1;
    // matrix-core smoother: wavefronts per patch (4, 8 or 16).  Measured on
    // MI355X at Q7 32^3 (C3 FP32 solve): 12.44 ms with 8, 13.0 with 4,
    // 13.4 with 16; 8 is at least as fast on every level (level-5 colour
    // 64-76 us, a level-4 colour 13.4-13.8, a level 1-3 chain 8.2-8.7)
    int       mfma_waves = 8;
    // z-marching vmult: target work-group count divisor (cells per chunk =
    // ceil(n tx ty / vmult_zdiv)); 0: by precision
    long long vmult_zdiv = 0;
    // z-marching vmult tile for k = 3: 0 = 16x8 cells / 512 threads,
    // 1 = 8x8 / 256, 2 = 16x8 / 256; 3: by precision.  k = 7: 0 = 8x2 /
    // 512, 1 = 4x4 / 1024, 2 and 3 = 4x4 / 512
    int       vmult_tile = 3;
    // bumped by every successful gtps_set_tuning: HIP graphs captured under
    // an older generation are re-captured before their next replay
    unsigned long long generation = 0;
  };
  Tuning &tuning();

  // setup.hip
  DegreeTables *get_tables(gtps_ctx *ctx, int k);
  // values (and derivatives, der may be NULL) of the Lagrange basis on
  // `nodes` at x
  void lagrange_values(const std::vector<double> &nodes, double x,
                       double *val, double *der);
  // n-point Gauss-Legendre rule on [0, 1]
  void gauss_rule(int n, std::vector<double> &p, std::vector<double> &w);
  // FP64 generalized eigenpairs K s = lam M s (n <= 16) on the GPU;
  // S column j = eigenvector j, S^T M S = I, ascending eigenvalues
  void gpu_gen_eig(gtps_ctx *ctx, int n, const std::vector<double> &K,
                   const std::vector<double> &M, std::vector<double> &lam,
                   std::vector<double> &S);

  // laplace.hip: y = A x (resid: y = b - A x); optional fused partial dot
  // sum_i x_i y_i written per block to ctx->ws.partials; returns #blocks.
  int launch_laplace(const gtps_level *lvl, void *y, const void *x,
                     const void *b, bool fused_dot);

  // coeff.hip: variable-coefficient operator (3D). y = A(c) x (resid:
  // y = b - A(c) x; transpose: y = A(c)^T x); a null level coefficient
  // means c = 1. Needs lvl->celem. Returns #dot partial blocks.
  int  launch_coeff_apply(const gtps_level *lvl, void *y, const void *x,
                          const void *b, bool fused_dot, bool transpose);
  // grad[cell] = -lambda~_K . A^1_K u_K (lambda's Dirichlet entries masked)
  void launch_coeff_gradient(const gtps_level *lvl, void *grad,
                             const void *lambda, const void *u);
  // one colour of the inexact-FD multiplicative smoother (lvl->coef set)
  void launch_patch_coeff_colour(const gtps_level *lvl, void *u,
                                 const void *b, int colour, double relax);
  // the per-patch separable fit of the level's coefficient into
  // lvl->sep_bins / lvl->sep_sigma (allocated by the caller)
  void launch_separable_fit(const gtps_level *lvl);
  // coarse[c] = mean of the 8 children (FP64), nc coarse cells per direction
  void launch_coarsen_coefficient(hipStream_t s, int nc, const double *fine,
                                  double *coarse);
  bool coeff_smoother_supported(int k, int prec);

  // dg.hip: SIPG-DG level set-up and kernels
  void dg_level_setup(gtps_dg_level *lvl);  // tables + scratch
  // operator scratch for batches of nb vectors
  void dg_level_reserve(gtps_dg_level *lvl, int nb);
  bool dg_supported(int k, int prec);
  // Every DG launcher takes a batch of nb contiguous vectors (stride lvl->N
  // on the level, the coarse level's N for coarse operands); one launch.
  // y = A x (b != NULL: y = b - A x); fused partial dot x.y -> #blocks per
  // vector (vector j's partials at j * #blocks)
  int  launch_dg_apply(const gtps_dg_level *lvl, void *y, const void *x,
                       const void *b, bool fused_dot, int nb = 1);
  // one sweep over the 27 colours (variant GTPS_MULT_FWD / _REV)
  void launch_dg_sweep(const gtps_dg_level *lvl, void *u, const void *b,
                       int variant, double relax, int nb = 1);
  // n_sweeps sweeps alternating colour order from `first` (meshes of at
  // most 4 cells per direction run them in one launch)
  void launch_dg_sweeps(const gtps_dg_level *lvl, void *u, const void *b,
                        int first, int n_sweeps, double relax, int nb = 1);
  // the same on `columns` independent (u, b) column pairs of lvl->N values
  // each, one work-group per column (meshes of 2..4 cells per direction)
  void launch_dg_sweeps_batched(const gtps_dg_level *lvl, void *u, const void *b,
                                int first, int n_sweeps, double relax, int columns);
  // fine += P coarse / coarse (+)= P^T fine (store: coarse = P^T fine)
  void launch_dg_prolongate_add(const gtps_dg_level *fine_lvl, void *fine,
                                const void *coarse, int nb = 1);
  void launch_dg_restrict(const gtps_dg_level *fine_lvl, void *coarse,
                          const void *fine, bool store, int nb = 1);

  // smoother_mfma.hip: one multiplicative colour pass (b in src) with the
  // 1D contractions on v_mfma_f32_16x16x4f32 (3D, FP32, k = 4..7)
  bool mfma_smoother_supported(int dim, int k, int prec);
  void launch_patch_colour_mfma(const gtps_level *lvl, void *u, const void *src,
                                int colour, double relax);

  // coarse.hip: the level-2 V-cycle as a dense matrix over the interior
  // DoFs ((N1-2)^dim of them, N1 = DoFs per direction of level 2)
  int  dense_coarse_size(int dim, int N1);
  // row length of the stored matrices: n rounded up to a multiple of 4
  // (16-byte rows; the padding columns hold zeros)
  int  dense_ld(int n);
  void launch_unit_defects(hipStream_t s, int dim, int N1, int prec, void *D);
  void launch_gather_columns(hipStream_t s, int dim, int N1, int prec,
                             const void *U, void *B);
  void launch_dense_apply(hipStream_t s, int dim, int N1, int prec,
                          const void *B, const void *d, void *u);
  // dense forms without Dirichlet rows (DG): D[j][i] = delta_ij over n
  // DoFs, B = U^T, u = B d
  void launch_unit_defects_plain(hipStream_t s, int n, int prec, void *D);
  void launch_transpose_plain(hipStream_t s, int n, int prec, const void *U, void *B);
  void launch_dense_apply_plain(hipStream_t s, int n, int prec, const void *B,
                                const void *d, void *u, int nb = 1);

  // post.hip (FP64): error norms of u against prod sin(wave pi x_d) with
  // QGauss(nq) (nq <= MAX_QUAD), and the load vector from f at the
  // QGauss(k+1) points (mode 0: fq[cell][q]; 1: the sine problem's f;
  // 2: constant fc); n cells per direction, CG or DG layout
  void compute_error_norms(gtps_ctx *ctx, int dim, int k, int n, bool dg,
                           double wave, int nq, const double *u, double *l2,
                           double *h1);
  void assemble_rhs(gtps_ctx *ctx, int dim, int k, int n, bool dg, int mode,
                    const double *fq, double wave, double fc, double *b);

  // smoother.hip: one colour pass. mode 0 multiplicative (src = b),
  // mode 1 additive (src = global residual), mode 2 multiplicative on
  // u = 0 that also zeroes the patch skeleton (replaces "u = 0" + colour 0
  // because colour-0 patch boxes cover the mesh).
  void launch_patch_colour(const gtps_level *lvl, void *u, const void *src,
                           int colour, int mode, double relax);

  // transfer.hip
  void launch_prolongate_add(const gtps_transfer *tr, int to_level, void *fine,
                             const void *coarse);
  void launch_restrict_add(const gtps_transfer *tr, int from_level,
                           void *coarse, const void *fine);
  // coarse = P^T fine on non-Dirichlet coarse DoFs (Dirichlet entries are
  // left untouched)
  void launch_restrict_store(const gtps_transfer *tr, int from_level,
                             void *coarse, const void *fine);
  size_t transfer_tmp_elems(int dim, int k, int max_level);

  // vector.hip
  void launch_fill(hipStream_t s, long long n, int prec, void *x, double v);
  void launch_axpby(hipStream_t s, long long n, double a, const double *x,
                    double b, double *y);
  void launch_convert_f64_f32(hipStream_t s, long long n, const double *x,
                              float *y);
  void launch_convert_f32_f64(hipStream_t s, long long n, const float *x,
                              double *y);
  int  launch_dot_partials(hipStream_t s, long long n, const double *x,
                           const double *y, const float *y32, double *partials,
                           int capacity);
  // sum partials[0..nb) into scalars[slot] (deterministic, one block); a
  // non-NULL host_out (device address of mapped host memory) receives the
  // same value
  void launch_finalize_sum(hipStream_t s, const double *partials, int nb,
                           double *scalars, int slot, double *host_out = nullptr);
  // partial sums of b^2 over the Dirichlet DoFs of the level (the faces
  // enumerated one by one, edges counted more than once); returns #blocks
  int launch_boundary_sumsq(const gtps_level *lvl, const double *b, double *partials,
                            int capacity);
  // CG start from x = 0: x = 0, r = b, r32 = (float) b (r32 may be NULL),
  // p = 0, partial sums of b^2, and the CG scalars rz = 1, alpha = 0;
  // returns #blocks
  int launch_cg_init(hipStream_t s, long long n, const double *b, double *x, double *r,
                     float *r32, double *p, double *scalars, double *partials,
                     int capacity);
  // the same, followed by the scalar op (cg::OP_*) in the same kernel
  void launch_finalize_op(hipStream_t s, const double *partials, int nb,
                          double *scalars, int slot, int op);
  void launch_set_boundary(const gtps_level *lvl, void *v, double value);
  void launch_lift_boundary(const gtps_level *lvl, const double *b, double *x);
  // dst_B = src_B on the Dirichlet DoFs, interior untouched
  void launch_copy_boundary(const gtps_level *lvl, const void *src, void *dst);
  // y += (double) x
  void launch_add_f32_to_f64(hipStream_t s, long long n, const float *x,
                             double *y);

  // CG building blocks (vector.hip); scalar slots are defined there
  void launch_cg_scalar(hipStream_t s, double *scalars, int op);
  int  launch_cg_update_xr(hipStream_t s, long long n, double *x, double *r,
                           const double *p, const double *q, float *r32,
                           const double *scalars, double *partials,
                           int capacity);
  // x == nullptr in launch_cg_update_xr defers x += alpha p to the next
  // launch_cg_update_p with x (or to launch_cg_add_alpha_p)
  void launch_cg_update_p(hipStream_t s, long long n, double *p,
                          const double *z, const float *z32,
                          const double *scalars, double *x = nullptr);
  void launch_cg_add_alpha_p(hipStream_t s, long long n, double *x,
                             const double *p, const double *scalars);

  namespace cg
  {
    // device scalar slots
    constexpr int RZ = 0, RZ_OLD = 1, PQ = 2, ALPHA = 3, BETA = 4, RR = 5, BB = 6,
                  N_SLOTS = 8;
    // scalar ops
    constexpr int OP_ALPHA = 0;  // alpha = rz / pq
    constexpr int OP_BETA  = 1;  // beta = rz / rz_old
    constexpr int OP_SHIFT = 3;  // rz_old = rz
    constexpr int OP_INIT  = 4;  // rz = 1, alpha = 0 (first direction: p = 0 + z)
    // finalize_op only: rz_old = rz; rz = sum; beta = rz / rz_old
    constexpr int OP_SHIFT_BETA = 5;
  } // namespace cg
} // namespace gtps
