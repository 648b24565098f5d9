/*
 * gtps.h - C ABI of the MI355X-native geometric-multigrid hot path
 * (vertex-patch Schwarz smoother with fast-diagonalisation local solves,
 * matrix-free sum-factorised Laplace, level transfer, MG-preconditioned CG).
 *
 * Mesh and data layout (the contract every entry point shares):
 *   - unit square (dim 2) or unit cube (dim 3), uniform Cartesian mesh with
 *     2^level cells per direction, continuous Q_k Lagrange elements on
 *     Gauss-Lobatto nodes;
 *   - a level vector holds N = (k 2^level + 1)^dim values in lexicographic
 *     order (x fastest), boundary DoFs included; it is a plain device array
 *     owned by the caller (float for GTPS_FP32 objects, double for GTPS_FP64);
 *   - Dirichlet DoFs have identity rows: vmult copies them (y_B = x_B).
 *
 * Conventions: every call returns 0 on success or a GTPS_E* code, never
 * throws, and records a message readable with gtps_last_error() (per thread).
 * Work is enqueued asynchronously on the context's HIP stream; the caller
 * synchronises (gtps_ctx_synchronize) before reading results on the host.
 * Handles are not thread-safe.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository).
 */
#ifndef GTPS_H
#define GTPS_H

#ifdef __cplusplus
extern "C" {
#endif

#define GTPS_OK 0
#define GTPS_EINVAL 1   /* bad argument (dimension, degree, level, size) */
#define GTPS_EHIP 2     /* HIP runtime error */
#define GTPS_ENOMEM 3   /* device allocation failed */
#define GTPS_ENOTCONV 4 /* solver reached maxit without convergence */
#define GTPS_ENAN 5     /* non-finite residual detected */

#define GTPS_FP32 0
#define GTPS_FP64 1

/* smoother variants */
#define GTPS_MULT_FWD 0  /* multiplicative, colours 0..2^dim-1 */
#define GTPS_MULT_REV 1  /* multiplicative, colours reversed */
#define GTPS_ADDITIVE 2  /* additive (one residual, all patches, damped) */

/* multigrid smoothing schedules */
#define GTPS_MG_SYMMETRIC 0       /* pre MULT_FWD, post MULT_REV (CG-safe) */
#define GTPS_MG_REFERENCE_ORDER 1 /* pre and post MULT_FWD (solver.cuh:1969,1997) */
#define GTPS_MG_ADDITIVE 2        /* pre and post ADDITIVE */

typedef struct gtps_ctx gtps_ctx;
typedef struct gtps_level gtps_level;
typedef struct gtps_smoother gtps_smoother;
typedef struct gtps_transfer gtps_transfer;
typedef struct gtps_mg gtps_mg;
typedef struct gtps_dg_level gtps_dg_level;
typedef struct gtps_dg_mg gtps_dg_mg;

/* ---------------------------------------------------------------- context */
/* Device selection and stream; replaces the reference's implicit CUDA
 * default stream and findCudaDevice (apps/poisson.cu:488-489).
 * stream == NULL creates a stream owned by the context. */
int gtps_ctx_create(int device, void *stream, gtps_ctx **out);
int gtps_ctx_destroy(gtps_ctx *ctx);
int gtps_ctx_synchronize(gtps_ctx *ctx);
void *gtps_ctx_stream(gtps_ctx *ctx);
const char *gtps_last_error(void);
const char *gtps_version(void);
/* Process-wide tuning knobs; returns the previous value, or -1 (with
 * gtps_last_error set) for an unknown key or invalid value.  The knobs are
 * shared by every context and thread without locking: change them only
 * while no other thread is inside the library.  They act on the next
 * launch; a solver's replayed HIP graph is captured again after any change.
 * Keys:
 *   "small_colour_patches"  smoother colours with at most this many patches
 *                           run the single-patch (latency-bound, coarse
 *                           level) kernel, larger ones the multi-patch kernel
 *                           (default 2^62, i.e. always the
 *                           single-patch kernel where it applies: 3D,
 *                           k <= 3).
 *   "group_min_patches"     colours with more patches than this pack
 *                           several patches into each wavefront (default
 *                           2048: the finest two levels at 64^3 Q3);
 *   "patches_per_wave"      how many (1 to 5; default 2);
 *   "group_waves"           wavefronts per work-group of the patch-group
 *                           kernel, each with its own patches: 1 or 4 (4
 *                           applies to 1 or 2 patches per wavefront); 0
 *                           (default): 4 in FP32, 1 in FP64;
 *   "box_vmult_max_n"       3D operator on levels with at most this many
 *                           cells per direction (half as many for k >= 5):
 *                           one-round box kernel instead of the z-marching
 *                           one (default 16; k = 6, 7 in FP64 always
 *                           march);
 *   "row_staged"            patch-group smoother with its global accesses
 *                           staged through contiguous x-runs (1) or
 *                           per-lane pencils and rows (0, default).
 *   "dense_coarse"          1 (default): a multigrid hierarchy of at least 3
 *                           levels tabulates its V-cycle from level 2 down
 *                           (a fixed linear map of the level-2 defect) as a
 *                           dense matrix over the level-2 interior DoFs on
 *                           first use (at most 4096 of them: k <= 4 in 3D)
 *                           and applies it with one launch (constant
 *                           coefficient only); 0: the ordinary
 *                           level-2 smoothing and level-1 solve.
 *   "vmult_zdiv"            z-marching vmult work-group count target: z-chunks
 *                           of ceil(n tx ty / value) cells (default 0: 512
 *                           in FP32, 2048 in FP64, 256 in FP64 at k = 7);
 *   "vmult_tile"            z-marching vmult tile for k = 3: 0: 16x8 cells
 *                           and 512 threads, 1: 8x8 and 256, 2: 16x8 and
 *                           256, 3 (default): 0 in FP32, 1 in FP64; for
 *                           k = 7: 0: 8x2 cells and 512 threads, 1: 4x4
 *                           and 1024, 2 and 3 (default): 4x4 and 512;
 *   "mfma_smoother"         1 (default): 3D FP32 multiplicative smoothing of
 *                           degree 4..7 runs its 1D contractions on the
 *                           matrix cores (v_mfma_f32_16x16x4f32, patches
 *                           padded to 16^3); 0: the LDS/FMA kernel.
 *   "mfma_waves"            matrix-core smoother wavefronts per patch: 4,
 *                           8 (default) or 16.
 * Results do not depend on the knobs beyond round-off.  No reference
 * counterpart (the reference picks kernels at compile time,
 * scripts/ct_parameter.py.in:146-148). */
long long gtps_set_tuning(const char *key, long long value);

/* Current value of a tuning knob (-1 and gtps_last_error for an unknown key). */
long long gtps_get_tuning(const char *key);

/* ------------------------------------------------------ level / operator */
/* One MG level with its matrix-free Laplace operator.
 * Replaces PSMF::LaplaceOperator<dim,k,Number,DoFLayout::Q>::initialize
 * (include/laplace_operator.cuh:83-84) together with the level
 * PSMF::MatrixFree::reinit it wraps (include/dealii/cuda_mf.template.cuh:856). */
int gtps_level_create(gtps_ctx *ctx, int dim, int degree, int level,
                      int precision, gtps_level **out);
int gtps_level_destroy(gtps_level *lvl);
long long gtps_level_n_dofs(const gtps_level *lvl);

/* dst = A src (Dirichlet rows copied).
 * Replaces LaplaceOperator::vmult (include/laplace_operator.cuh:86-101). */
int gtps_laplace_vmult(gtps_level *lvl, void *dst, const void *src);

/* dst = b - A u (the V-cycle's defect update t = d - A u).
 * Replaces matrix[l].vmult + t.sadd(-1,1,d) (include/solver.cuh:1976-1977). */
int gtps_laplace_residual(gtps_level *lvl, void *dst, const void *b,
                          const void *u);

/* ------------------------------------------------- variable coefficient */
/* Variable-coefficient diffusion -div(c grad u) with c constant per cell
 * (3D only; cells numbered cx + n cy + n^2 cz, n = 2^level).  The reference
 * takes a coefficient in its solver constructor and ignores it
 * (include/solver.cuh:1548); its matrix-free hook for per-cell data is
 * MatrixFree::evaluate_coefficients (include/dealii/cuda_mf.template.cuh:
 * 815-824), never called on the hot path.  These entry points have no
 * further reference counterpart.
 *
 * Set (coef = device FP64 array of n^3 values, copied and rounded to the
 * level precision) or clear (coef = NULL) the level's coefficient.  While
 * set, gtps_laplace_vmult / gtps_laplace_residual apply A(c) =
 * sum_K c_K A^1_K (same Dirichlet identity rows), and smoothers on the
 * level use the inexact local solve cmax_P A^1_P (correction form
 * u_I <- u_I + relaxation A~_P^{-1} (b - A u)_I; multiplicative only). */
int gtps_level_set_coefficient(gtps_level *lvl, const double *coef);

/* dst = A^T src for the matrix of gtps_laplace_vmult (identity Dirichlet
 * rows, boundary columns kept): interior rows A_II src_I, boundary rows
 * A_IB^T src_I + src_B.  Needed for the adjoint of the constrained solve.
 * Uses the level's coefficient if set, else c = 1. */
int gtps_laplace_vmult_transpose(gtps_level *lvl, void *dst, const void *src);

/* grad[K] = -sum_{i,j in K} lambda~_i (A^1_K)_ij u_j = -int_K grad(lambda).
 * grad(u) per cell (lambda~ = lambda with its Dirichlet entries zeroed):
 * dJ/dc_K for J(u(c)) when lambda solves the adjoint system.  grad has n^3
 * entries in the level precision. */
int gtps_level_coefficient_gradient(gtps_level *lvl, void *grad,
                                    const void *lambda, const void *u);

/* Host copy (FP64) of the 1D data the level was set up with: reference-cell
 * mass/stiffness [(k+1)^2 each], patch mass/stiffness [(2k+1)^2 each],
 * interior eigenvalues [2k-1] and M-orthonormal eigenvectors [(2k-1)^2,
 * column j = vector j] computed on the GPU in FP64.  Any pointer may be NULL.
 * Replaces TensorProductData::get_{mass_matrix,derivative_matrix,
 * eigenvalues,eigenvectors} (include/tensor_product.h:205-260). */
int gtps_level_tensor_product_data(gtps_level *lvl, double *cell_mass,
                                   double *cell_laplace, double *patch_mass,
                                   double *patch_laplace, double *eigenvalues,
                                   double *eigenvectors);

/* --------------------------------------------------------------- smoother */
/* Vertex-patch Schwarz smoother with exact fast-diagonalisation local solves.
 * Replaces PSMF::PatchSmoother::initialize (include/patch_smoother.cuh:313-335)
 * with AdditionalData{relaxation} (patch_smoother.cuh:86-124). */
int gtps_smoother_create(gtps_level *lvl, int variant, double relaxation,
                         gtps_smoother **out);
int gtps_smoother_destroy(gtps_smoother *sm);
/* One sweep in place: u <- S(u, b).  Multiplicative: per colour, per patch,
 * u_I <- relaxation * (u_I + A_II^{-1} (b - A u)_I).  Additive:
 * u <- u + relaxation * sum_P R_P^T A_P^{-1} R_P (b - A u).
 * Replaces PatchSmoother::step (patch_smoother.cuh:92-124 -> step_impl
 * 187-292). */
int gtps_smoother_step(gtps_smoother *sm, void *u, const void *b);
/* u <- 0, then one sweep.  Replaces PatchSmoother::vmult
 * (patch_smoother.cuh:157-163). */
int gtps_smoother_vmult(gtps_smoother *sm, void *u, const void *b);

/* --------------------------------------------------------------- transfer */
/* Level transfer for levels min_level..max_level of one (dim, degree).
 * Replaces PSMF::MGTransferCUDA<dim,Number,Q> (include/cuda_mg_transfer.cuh:57)
 * with build() (cuda_mg_transfer.template.cuh:453-559). */
int gtps_transfer_create(gtps_ctx *ctx, int dim, int degree, int min_level,
                         int max_level, int precision, gtps_transfer **out);
int gtps_transfer_destroy(gtps_transfer *tr);
/* fine(to_level) += P coarse(to_level-1); coarse Dirichlet values ignored.
 * Replaces MGTransferCUDA::prolongate_and_add (.template.cuh:573-621). */
int gtps_transfer_prolongate_add(gtps_transfer *tr, int to_level, void *fine,
                                 const void *coarse);
/* fine = P coarse.  Replaces MGTransferCUDA::prolongate (.template.cuh:561). */
int gtps_transfer_prolongate(gtps_transfer *tr, int to_level, void *fine,
                             const void *coarse);
/* coarse(from_level-1) += P^T fine(from_level) on non-Dirichlet coarse DoFs.
 * Replaces MGTransferCUDA::restrict_and_add (.template.cuh:623-671). */
int gtps_transfer_restrict_add(gtps_transfer *tr, int from_level, void *coarse,
                               const void *fine);

/* ---------------------------------------------------------------- vectors */
/* FP64 vector kernels on the context stream (the deal.II
 * LinearAlgebra::distributed::Vector<double,CUDA> operations the solver uses,
 * include/solver.cuh:1712,1830,1853,1977, and the precision conversion
 * copy_vector, include/solver.cuh:64-78, 2070-2095).
 * Reductions are deterministic (fixed two-stage order) and return their value
 * to the host (they synchronise the stream). */
int gtps_vec_fill(gtps_ctx *ctx, long long n, int precision, void *x,
                  double value);
int gtps_vec_axpby(gtps_ctx *ctx, long long n, double a, const double *x,
                   double b, double *y);
int gtps_vec_dot(gtps_ctx *ctx, long long n, const double *x, const double *y,
                 double *result);
int gtps_vec_nrm2(gtps_ctx *ctx, long long n, const double *x, double *result);
int gtps_vec_convert_f64_f32(gtps_ctx *ctx, long long n, const double *x,
                             float *y);
int gtps_vec_convert_f32_f64(gtps_ctx *ctx, long long n, const float *x,
                             double *y);
/* Allocation-free variant of gtps_vec_dot: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): it calls
 * no hipMalloc/hipFree or allocating routine, does not synchronise, and
 * writes x.y to `result`, ONE double in DEVICE memory, on the context stream
 * (same kernels, grid and summation order as gtps_vec_dot, so the same
 * bits).  Its partial sums live in a static device array of
 * GTPS_NOALLOC_PARTIALS doubles (16 KiB per GPU, kept until the process
 * ends); the grid is capped at that many work-groups, so any
 * 0 <= n <= GTPS_NOALLOC_MAX_N is accepted and larger n is rejected with
 * GTPS_EINVAL.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams, other threads) race; calls ordered on
 * one stream do not. */
#define GTPS_NOALLOC_PARTIALS 2048
#define GTPS_NOALLOC_MAX_N (1LL << 62)
int gtps_vec_dot_noalloc(gtps_ctx *ctx, long long n, const double *x,
                         const double *y, double *result);
/* v_B = value on the level's Dirichlet DoFs.  Replaces
 * set_inhomogeneous_dofs / set_constrained_dofs (include/solver.cuh:45-62,
 * dealii/cuda_mf.template.cuh:514-528). */
int gtps_vec_set_boundary(gtps_level *lvl, void *v, double value);

/* ---------------------------------------------------------- multigrid+CG */
/* Multigrid hierarchy on levels 1..max_level (level 1: one vertex patch, the
 * exact coarse solve).  vcycle_precision: GTPS_FP32 (FP32 smoother, level
 * operators and transfers under FP64 CG) or GTPS_FP64 (FP64 everywhere).
 * Replaces the PSMF::MultigridSolvers constructor (include/solver.cuh:1545-1744). */
int gtps_mg_create(gtps_ctx *ctx, int dim, int degree, int max_level,
                   int vcycle_precision, int schedule, double relaxation,
                   gtps_mg **out);
int gtps_mg_destroy(gtps_mg *mg);
long long gtps_mg_n_dofs(const gtps_mg *mg);
/* z = V(r) on the finest level, FP64 in and out (conversion fused).
 * Replaces MultigridSolvers::vmult (include/solver.cuh:1915-1930). */
int gtps_mg_vcycle(gtps_mg *mg, double *z, const double *r);
/* MG-preconditioned FP64 CG for A x = b from x = lift(b) (x_B = b_B,
 * x_I = 0) until ||r_k|| <= rtol ||r_0|| or maxit iterations.  iters and
 * relres (optional) receive the iteration count and ||r||/||r_0||;
 * history (optional, maxit+1 entries) receives ||r_k||.
 * Replaces MultigridSolvers::solve_gmres (include/solver.cuh:1865-1910): the
 * outer Krylov method is CG with a symmetric V-cycle instead of GMRES. */
int gtps_mg_pcg_solve(gtps_mg *mg, double *x, const double *b, double rtol,
                      int maxit, int *iters, double *relres, double *history);
/* FP64 finest-level operator of the hierarchy: dst = A src.  Replaces
 * MultigridSolvers::do_matvec (include/solver.cuh:1934-1938). */
int gtps_mg_matvec(gtps_mg *mg, double *dst, const double *src);

/* Variable coefficient for the whole hierarchy: the finest levels take
 * coef (device FP64, n^3 cells of the finest level), each coarser level the
 * mean of its 8 children; NULL restores c = 1.  Multiplicative schedules
 * only.  No reference counterpart (see gtps_level_set_coefficient). */
int gtps_mg_set_coefficient(gtps_mg *mg, const double *coef);

/* Full multigrid, MultigridSolvers::solve (include/solver.cuh:1784-1860):
 * exact level-1 solve plus one more smoothing step, then for each finer
 * level: prolongate the coarser solution (FP64), impose the level's
 * Dirichlet values (the boundary entries of its right-hand side), form the
 * FP64 residual and add the correction of n_cycles V-cycles in the V-cycle
 * precision.  rhs is a host array of max_level device pointers, rhs[l-1]
 * the FP64 right-hand side of level l (the reference assembles the load
 * vector on every level).  x (FP64, finest level) receives the solution.
 * The finest level's residual goes through the hierarchy's CG scratch, so
 * no solve may run concurrently on the same hierarchy. */
int gtps_mg_fmg_solve(gtps_mg *mg, double *x, const double *const *rhs,
                      int n_cycles);

/* FP64 finest-level dst = A^T src (see gtps_laplace_vmult_transpose). */
int gtps_mg_matvec_transpose(gtps_mg *mg, double *dst, const double *src);

/* FP64 finest-level coefficient gradient (see
 * gtps_level_coefficient_gradient); grad has n^3 doubles. */
int gtps_mg_coefficient_gradient(gtps_mg *mg, double *grad,
                                 const double *lambda, const double *u);

/* Replay the per-iteration preconditioning sequence (V-cycle, r.z, beta,
 * p update; about 130 kernels for 6 levels) as one HIP graph captured on the
 * first solve (enable != 0, the default), or launch it kernel by kernel.
 * The graph uses no memset nodes.  Stage timers force kernel-by-kernel
 * launches.  No reference counterpart (the reference launches every kernel
 * on the default stream and synchronises after each stage,
 * include/solver.cuh:1959-1998). */
int gtps_mg_set_graph(gtps_mg *mg, int enable);

/* ---------------------------------------------------------- profiling */
/* Stage timers with HIP events on the context stream, the counterpart of the
 * reference's per-level timing buckets (MultigridSolvers::print_wall_times,
 * include/solver.cuh:1747-1771, filled at 1957-1999) without its per-stage
 * device synchronisations.  `mask` enables groups (bit g = group g):
 *   GTPS_PROF_SMOOTH_FINE   finest-level sweeps, one event pair per sweep,
 *                           counted as its colour launches (so total/count
 *                           is the mean colour-launch time)
 *   GTPS_PROF_SMOOTH_COARSE smoothing sweeps on the coarser levels
 *   GTPS_PROF_RESIDUAL_FINE t = d - A u on the finest level
 *   GTPS_PROF_TRANSFER      restriction and prolongation, all levels
 *   GTPS_PROF_CG_MATVEC     the FP64 q = A p of each CG iteration
 *   GTPS_PROF_VCYCLE        one whole V-cycle
 * gtps_mg_profile synchronises the stream and returns, for one group, the
 * number of timed intervals and their total in milliseconds since the last
 * reset. */
#define GTPS_PROF_SMOOTH_FINE 0
#define GTPS_PROF_SMOOTH_COARSE 1
#define GTPS_PROF_RESIDUAL_FINE 2
#define GTPS_PROF_TRANSFER 3
#define GTPS_PROF_CG_MATVEC 4
#define GTPS_PROF_VCYCLE 5
#define GTPS_PROF_N_GROUPS 6
int gtps_mg_set_profiling(gtps_mg *mg, int mask);
int gtps_mg_profile(gtps_mg *mg, int group, long long *count,
                    double *total_ms);
int gtps_mg_profile_reset(gtps_mg *mg);

/* ============================================================= SIPG-DG */
/* Symmetric interior penalty DG Laplacian, FE_DGQ(k) on Gauss-Lobatto
 * points, uniform mesh of n cells per direction on the unit cube (3D).  The
 * form is the reference's matrix-free DG test operator (tests/mf_dg.cu:
 * 63-274: cell term, interior faces (sigma [u] - {du/dn}) [v] - [u]{dv/dn},
 * boundary faces (2 sigma u - du/dn) v - u dv/dn) with sigma = penalty / h;
 * the reference uses penalty = k (k+1) there and has no DG smoother or
 * solver, so every entry point below is an addition.  Vectors hold
 * (n (k+1))^3 values on one tensor grid: X = cx (k+1) + ix, index
 * X + N1 Y + N1^2 Z with N1 = n (k+1) (x fastest); there are no Dirichlet
 * rows (boundary conditions are weak). */
int gtps_dg_level_create(gtps_ctx *ctx, int dim, int degree, int n_cells,
                         int precision, double penalty, gtps_dg_level **out);
int gtps_dg_level_destroy(gtps_dg_level *lvl);
long long gtps_dg_level_n_dofs(const gtps_dg_level *lvl);
/* dst = A src; dst = b - A u */
int gtps_dg_vmult(gtps_dg_level *lvl, void *dst, const void *src);
int gtps_dg_residual(gtps_dg_level *lvl, void *dst, const void *b,
                     const void *u);
/* One multiplicative sweep of the vertex-patch smoother (patch = the 2^3
 * cells around an interior vertex, all their DoFs; 27 colours = patch index
 * mod 3 per direction; exact local solves by fast diagonalisation).
 * variant GTPS_MULT_FWD or GTPS_MULT_REV; u <- (1-w) u + w u_P per patch.
 * Degrees: k <= 5 in FP64 and k <= 7 in FP32 (the patch of (2k+2)^3 DoFs
 * and its tables must fit one work-group's LDS); other levels are rejected
 * with GTPS_EINVAL.  The operator and transfer take every k <= 7. */
int gtps_dg_smoother_step(gtps_dg_level *lvl, void *u, const void *b,
                          int variant, double relaxation);
/* Cell-wise embedding from n/2 to n cells (fine_lvl is the fine level):
 * fine += P coarse; coarse += P^T fine. */
int gtps_dg_prolongate_add(gtps_dg_level *fine_lvl, void *fine,
                           const void *coarse);
int gtps_dg_restrict_add(gtps_dg_level *fine_lvl, void *coarse,
                         const void *fine);

/* DG multigrid on n_coarse 2^l cells per direction, l = 0..max_level:
 * V(1,1) (forward pre-, reverse post-smoothing) in vcycle_precision, the
 * coarsest level solved approximately by coarse_sweeps symmetric sweep pairs
 * from zero; FP64 CG on the finest level.  Same degree limits as
 * gtps_dg_smoother_step (FP64 k <= 5, FP32 k <= 7). */
int gtps_dg_mg_create(gtps_ctx *ctx, int dim, int degree, int n_coarse,
                      int max_level, int vcycle_precision, double penalty,
                      int coarse_sweeps, gtps_dg_mg **out);
int gtps_dg_mg_destroy(gtps_dg_mg *mg);
/* enable (default) / disable the HIP-graph replay of the per-iteration
 * preconditioning sequence in gtps_dg_mg_pcg_solve */
int gtps_dg_mg_set_graph(gtps_dg_mg *mg, int enable);
/* nrhs systems A x_j = b_j solved together (x, b: nrhs contiguous FP64
 * vectors of gtps_dg_mg_n_dofs values): every V-cycle kernel runs once on
 * the whole batch (a second grid dimension), the CG scalars are per vector,
 * and the loop runs until every vector reaches rtol (iterations[j]: the
 * first iteration at which vector j did; relres[j] its final relative
 * residual).  The first call with a larger batch reallocates the
 * hierarchy's vectors.  No reference counterpart (the reference has no DG
 * solver); this serves the batched adjoint of the DG workload. */
int gtps_dg_mg_pcg_solve_batch(gtps_dg_mg *mg, double *x, const double *b,
                               int nrhs, double rtol, int maxit,
                               int *iterations, double *relres);
long long gtps_dg_mg_n_dofs(const gtps_dg_mg *mg);
int gtps_dg_mg_matvec(gtps_dg_mg *mg, double *dst, const double *src);
int gtps_dg_mg_vcycle(gtps_dg_mg *mg, double *z, const double *r);
/* x = A^{-1} b by MG-preconditioned FP64 CG from x = 0 (same contract as
 * gtps_mg_pcg_solve). */
int gtps_dg_mg_pcg_solve(gtps_dg_mg *mg, double *x, const double *b,
                         double rtol, int maxit, int *iterations,
                         double *relres, double *history);

/* ====================================================== post-processing */
/* FP64 device vectors; n_cells per direction (2^level for the CG layout).
 *
 * Load vector b_i = sum_cells int f phi_i with QGauss(k+1) per cell:
 * replaces the host cell loop of LaplaceOperator::compute_residual
 * (include/laplace_operator.cuh:119-173) with f given at the quadrature
 * points, f_q[cell][q], cell = cx + n (cy + n cz), q = qx + (k+1) (qy +
 * (k+1) qz), point ((c + xi_q) / n) per direction with xi_q the Gauss points
 * on [0,1].  Dirichlet rows of the CG layout are set to zero (the reference
 * leaves them; a CG residual needs them zero).  gtps_rhs_sine assembles the
 * reference's RightHandSide f = dim (w pi)^2 prod sin(w pi x_d)
 * (apps/poisson.cu:76-87, w = 3 there) without an f_q array;
 * gtps_rhs_constant assembles f = value. */
int gtps_assemble_rhs(gtps_ctx *ctx, int dim, int degree, int level,
                      const double *f_q, double *b);
int gtps_rhs_sine(gtps_ctx *ctx, int dim, int degree, int level, double wave,
                  double *b);
int gtps_rhs_constant(gtps_ctx *ctx, int dim, int degree, int level,
                      double value, double *b);
int gtps_dg_assemble_rhs(gtps_ctx *ctx, int dim, int degree, int n_cells,
                         const double *f_q, double *b);
int gtps_dg_rhs_sine(gtps_ctx *ctx, int dim, int degree, int n_cells,
                     double wave, double *b);
/* L2 norm and H1 seminorm (cell-wise gradient for DG) of u_h - u with
 * u = prod sin(w pi x_d), QGauss(n_q) per cell, n_q in 1..16: replaces
 * LaplaceProblem::compute_error (apps/poisson.cu:353-382: QGauss(k+1),
 * VectorTools::integrate_difference L2_norm / H1_seminorm against Solution,
 * apps/poisson.cu:44-74).  Blocks until the result is on the host. */
int gtps_error_norms(gtps_ctx *ctx, int dim, int degree, int level,
                     double wave, int n_q, const double *u, double *l2,
                     double *h1);
int gtps_dg_error_norms(gtps_ctx *ctx, int dim, int degree, int n_cells,
                        double wave, int n_q, const double *u, double *l2,
                        double *h1);

/* Solution file: a 64-byte little-endian header (magic "GTPSVEC1", int32
 * precision, dim, degree, n_cells, layout, 0; int64 n; uint64 FNV-1a hash
 * of the data bytes; 16 zero bytes) followed by the n values.  The
 * reference writes no solution files (its apps print norms and timings);
 * these are additions for checkpointing a solve.  x is a device array of n
 * = (k n_cells + 1)^dim (GTPS_LAYOUT_CG) or (n_cells (k+1))^dim
 * (GTPS_LAYOUT_DG) values of the given precision.  gtps_vector_info reads
 * the header only (no GPU); gtps_vector_load checks the header against n
 * and precision and the hash against the data. */
#define GTPS_LAYOUT_CG 0
#define GTPS_LAYOUT_DG 1
int gtps_vector_save(gtps_ctx *ctx, const char *path, const void *x,
                     int precision, int dim, int degree, int n_cells,
                     int layout);
int gtps_vector_info(const char *path, int *precision, int *dim, int *degree,
                     int *n_cells, int *layout, long long *n);
int gtps_vector_load(gtps_ctx *ctx, const char *path, void *x, long long n,
                     int precision);

#ifdef __cplusplus
}
#endif

#endif /* GTPS_H */
