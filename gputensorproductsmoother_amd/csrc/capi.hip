// capi.hip - the extern "C" boundary (include/gtps.h) and the host-side
// multigrid / CG orchestration.
//
// Host code owns no numerics: it validates arguments, allocates level
// vectors once, and enqueues kernels on the context stream.  The V-cycle
// follows MultigridSolvers::v_cycle (include/solver.cuh:1952-2001) without
// the per-stage device synchronisations; CG replaces SolverGMRES
// (solver.cuh:1865-1910) and keeps its scalars on the device.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "gtps_internal.h"

using namespace gtps;

// Partial sums of gtps_vec_dot_noalloc: a file-scope device array of
// GTPS_NOALLOC_PARTIALS doubles (16 KiB), one copy on every GPU the library
// runs on, held until the process ends.  Host code never passes the array's
// name to a kernel (on the host it is only a stand-in for the symbol): it
// uses the per-device address that hipGetSymbolAddress returns once per
// device when the first context on that device is created.
static __device__ __align__(256) double scratch[GTPS_NOALLOC_PARTIALS];

gtps::Tuning &
gtps::tuning()
{
  static Tuning t;
  return t;
}

namespace
{
  thread_local std::string g_last_error;

  template <typename F>
  int
  guard(F &&f)
  {
    try
      {
        f();
        g_last_error.clear();
        return GTPS_OK;
      }
    catch (const Error &e)
      {
        g_last_error = e.what();
        return e.code;
      }
    catch (const std::bad_alloc &)
      {
        g_last_error = "host allocation failed";
        return GTPS_ENOMEM;
      }
    catch (const std::exception &e)
      {
        g_last_error = e.what();
        return GTPS_EHIP;
      }
    catch (...)
      {
        g_last_error = "unknown error";
        return GTPS_EHIP;
      }
  }

  void
  set_device(const gtps_ctx *ctx)
  {
    GTPS_HIP(hipSetDevice(ctx->device));
  }

  // device address of `scratch` per device (set up by gtps_ctx_create with
  // that device current)
  // (a failed lookup leaves the pointer NULL and keeps the HIP error for
  // gtps_vec_dot_noalloc to report; the context itself does not need it)
  constexpr int MAX_SCRATCH_DEVICES = 64;
  std::mutex    g_scratch_mutex;
  double       *g_scratch_ptr[MAX_SCRATCH_DEVICES] = {};
  hipError_t    g_scratch_err[MAX_SCRATCH_DEVICES]  = {};

  void
  scratch_setup(int device)
  {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (device < 0 || device >= MAX_SCRATCH_DEVICES || g_scratch_ptr[device])
      return;
    void            *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      {
        (void)hipGetLastError();
        g_scratch_err[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
        return;
      }
    g_scratch_ptr[device] = (double *)p;
  }

  double *
  scratch_of(int device)
  {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    GTPS_REQUIRE(device >= 0 && device < MAX_SCRATCH_DEVICES,
                 "no scratch array on devices with index >= 64");
    if (!g_scratch_ptr[device])
      throw Error(GTPS_EHIP, std::string("scratch array address unavailable: ") +
                               hipGetErrorString(g_scratch_err[device]));
    return g_scratch_ptr[device];
  }

  void *
  dmalloc(size_t bytes)
  {
    void *p = nullptr;
    if (bytes == 0)
      bytes = 16;
    if (hipMalloc(&p, bytes) != hipSuccess)
      {
        (void)hipGetLastError();
        throw Error(GTPS_ENOMEM, "hipMalloc of " + std::to_string(bytes) +
                                   " bytes failed");
      }
    return p;
  }

  void
  dfree(void *p)
  {
    if (p)
      (void)hipFree(p);
  }

  void
  check_level_args(int dim, int degree, int level, int precision)
  {
    GTPS_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
    GTPS_REQUIRE(degree >= 1 && degree <= MAX_DEGREE, "degree must be in 1..7");
    GTPS_REQUIRE(level >= 1, "level must be >= 1");
    GTPS_REQUIRE(precision == GTPS_FP32 || precision == GTPS_FP64,
                 "precision must be GTPS_FP32 or GTPS_FP64");
    // N1 must fit the kernels' 32-bit per-direction indices and the vector
    // length a 64-bit index
    GTPS_REQUIRE(level <= (dim == 3 ? 12 : 20), "level too large");
  }

  gtps_level *
  make_level(gtps_ctx *ctx, int dim, int degree, int level, int precision)
  {
    check_level_args(dim, degree, level, precision);
    auto lvl   = std::make_unique<gtps_level>();
    lvl->ctx   = ctx;
    lvl->dim   = dim;
    lvl->k     = degree;
    lvl->level = level;
    lvl->n     = 1 << level;
    lvl->N1    = degree * lvl->n + 1;
    lvl->N     = dim == 3 ? (long long)lvl->N1 * lvl->N1 * lvl->N1
                          : (long long)lvl->N1 * lvl->N1;
    lvl->prec  = precision;
    lvl->tab   = get_tables(ctx, degree);
    lvl->scale = std::pow(std::ldexp(1.0, -level), dim - 2);
    return lvl.release();
  }

  void
  ensure_celem(gtps_level *lvl)
  {
    if (!lvl->celem)
      {
        const long long ncells = (long long)lvl->n * lvl->n * lvl->n;
        const long long n3     = (long long)(lvl->k + 1) * (lvl->k + 1) * (lvl->k + 1);
        lvl->celem             = dmalloc(lvl->elem() * ncells * n3);
      }
  }

  // copy (and round) a device FP64 per-cell coefficient into the level
  void
  set_level_coefficient(gtps_level *lvl, const double *c)
  {
    GTPS_REQUIRE(lvl->dim == 3, "the variable-coefficient operator is 3D");
    const long long ncells = (long long)lvl->n * lvl->n * lvl->n;
    if (!lvl->coef)
      lvl->coef = dmalloc(lvl->elem() * ncells);
    ensure_celem(lvl);
    // the smoother's ratio-grid eigenpairs, built here (setup) and not on
    // the first smoothing step, which may run under HIP graph capture
    (void)get_sep_tables(lvl->ctx, lvl->k, lvl->prec);
    if (lvl->prec == GTPS_FP32)
      launch_convert_f64_f32(lvl->ctx->stream, ncells, c, (float *)lvl->coef);
    else
      GTPS_HIP(hipMemcpyAsync(lvl->coef, c, sizeof(double) * ncells,
                              hipMemcpyDeviceToDevice, lvl->ctx->stream));
    // the smoother's per-patch separable fit, from the level-precision
    // coefficient the operator uses
    const long long np = (long long)(lvl->n - 1) * (lvl->n - 1) * (lvl->n - 1);
    if (np > 0)
      {
        if (!lvl->sep_bins)
          {
            lvl->sep_bins  = (int *)dmalloc(sizeof(int) * np);
            lvl->sep_sigma = dmalloc(lvl->elem() * np);
          }
        launch_separable_fit(lvl);
      }
  }

  void
  clear_level_coefficient(gtps_level *lvl)
  {
    if (lvl->coef)
      {
        GTPS_HIP(hipStreamSynchronize(lvl->ctx->stream));
        dfree(lvl->coef);
        lvl->coef = nullptr;
        dfree(lvl->sep_bins);
        dfree(lvl->sep_sigma);
        lvl->sep_bins  = nullptr;
        lvl->sep_sigma = nullptr;
      }
  }

  // brackets the colour launches of a sweep when a timer is attached (one
  // event pair per sweep, credited to the number of launches it holds)
  struct LaunchTimer
  {
    virtual long tic_launch()                  = 0;
    virtual void toc_launch(long e, int count) = 0;
    virtual ~LaunchTimer()                     = default;
  };

  // one smoothing sweep; `resid` is level-sized scratch for the additive
  // form.  from_zero: u is taken as zero on entry (vmult semantics); the
  // forward multiplicative sweep then starts with the colour-0 pass that
  // writes the whole mesh (interiors solved, skeleton zeroed) instead of a
  // separate fill.
  void
  sweep(const gtps_level *lvl, void *u, const void *b, int variant,
        double relax, void *resid, LaunchTimer *timer = nullptr,
        bool from_zero = false)
  {
    const int ncol = 1 << lvl->dim;
    if (lvl->coef)
      {
        // variable coefficient: correction-form sweep with the inexact local
        // solve (no "from zero" colour pass, no additive form)
        GTPS_REQUIRE(variant != GTPS_ADDITIVE,
                     "the additive smoother has no variable-coefficient form");
        GTPS_REQUIRE(coeff_smoother_supported(lvl->k, lvl->prec),
                     "the variable-coefficient smoother supports degree <= 5 "
                     "in FP64 and <= 7 in FP32");
        if (from_zero)
          launch_fill(lvl->ctx->stream, lvl->N, lvl->prec, u, 0.0);
        const long e = timer ? timer->tic_launch() : -1;
        for (int cc = 0; cc < ncol; ++cc)
          launch_patch_coeff_colour(lvl, u, b,
                                    variant == GTPS_MULT_REV ? ncol - 1 - cc : cc,
                                    relax);
        if (timer)
          timer->toc_launch(e, ncol);
        return;
      }
    if (from_zero && variant != GTPS_MULT_FWD)
      launch_fill(lvl->ctx->stream, lvl->N, lvl->prec, u, 0.0);
    if (variant == GTPS_ADDITIVE)
      {
        GTPS_REQUIRE(resid != nullptr, "additive sweep needs a residual buffer");
        launch_laplace(lvl, resid, u, b, false);
        const long e = timer ? timer->tic_launch() : -1;
        for (int c = 0; c < ncol; ++c)
          launch_patch_colour(lvl, u, resid, c, 1, relax);
        if (timer)
          timer->toc_launch(e, ncol);
        return;
      }
    const long e = timer ? timer->tic_launch() : -1;
    for (int cc = 0; cc < ncol; ++cc)
      {
        const int c = variant == GTPS_MULT_REV ? ncol - 1 - cc : cc;
        launch_patch_colour(lvl, u, b, c,
                            (from_zero && variant == GTPS_MULT_FWD && cc == 0) ? 2 : 0,
                            relax);
      }
    if (timer)
      timer->toc_launch(e, ncol);
  }
} // namespace

struct gtps_mg : LaunchTimer
{
  gtps_ctx                  *ctx = nullptr;
  int                        dim = 3, k = 1, maxl = 1, prec = GTPS_FP32;
  int                        schedule = GTPS_MG_SYMMETRIC;
  double                     relax    = 1.0;
  std::vector<gtps_level *>  lv;       // [0..maxl], index 0 unused
  gtps_level                *fine64 = nullptr;
  bool                       own_fine64 = false;
  gtps_transfer             *tr = nullptr;
  std::vector<void *>        d, u, t;  // level vectors in prec
  double                    *r = nullptr, *p = nullptr, *q = nullptr;
  hipEvent_t                 ev = nullptr;

  // variable coefficient: FP64 per-level chain (index l; the finest entry
  // is a copy of the caller's array), empty for c = 1
  std::vector<double *>      coef64;
  // full multigrid: FP64 operators and solutions per level, FP64 transfer
  std::vector<gtps_level *>  lv64;
  std::vector<bool>          own_lv64;
  std::vector<double *>      sol;
  gtps_transfer             *tr64     = nullptr;
  bool                       own_tr64 = false;

  // the per-iteration preconditioning sequence (V-cycle, r.z, beta,
  // deferred x update, p update) replayed as one HIP graph; captured lazily
  // on a private stream, again when the solution vector changes
  bool                       use_graph  = true;
  hipStream_t                cap_stream = nullptr;
  hipGraphExec_t             dir_exec   = nullptr;
  unsigned long long         dir_gen    = 0;  // tuning generation at capture
  double                    *dir_x      = nullptr;  // solution at capture

  // the V-cycle from level 2 down as a dense matrix (coarse.hip), built
  // on first use and dropped when the operators change
  void                      *dense2      = nullptr;
  bool                       dense_valid = false;
  bool                       dense_off   = false;  // while tabulating

  // stage timers (HIP events on the context stream)
  int                        prof_mask = 0;
  std::vector<hipEvent_t>    ev_pool;
  size_t                     ev_used = 0;
  struct Interval
  {
    int    group;
    size_t start, stop;  // event indices
    int    count;        // launches (or intervals) it stands for
  };
  std::vector<Interval>      pending;
  long long                  prof_count[GTPS_PROF_N_GROUPS] = {};
  double                     prof_ms[GTPS_PROF_N_GROUPS]    = {};

  hipEvent_t
  next_event()
  {
    if (ev_used == ev_pool.size())
      {
        hipEvent_t e;
        GTPS_HIP(hipEventCreate(&e));
        ev_pool.push_back(e);
      }
    return ev_pool[ev_used++];
  }
  // returns the start-event index, or -1 when the group is disabled
  long
  tic(int group)
  {
    if (!(prof_mask & (1 << group)))
      return -1;
    const size_t i = ev_used;
    GTPS_HIP(hipEventRecord(next_event(), ctx->stream));
    return (long)i;
  }
  void
  toc(int group, long start, int count = 1)
  {
    if (start < 0)
      return;
    const size_t stop = ev_used;
    GTPS_HIP(hipEventRecord(next_event(), ctx->stream));
    pending.push_back({group, (size_t)start, stop, count});
  }
  // after the stream is synchronised: fold pending intervals into totals
  void
  collect()
  {
    for (const auto &iv : pending)
      {
        float ms = 0.f;
        GTPS_HIP(hipEventElapsedTime(&ms, ev_pool[iv.start], ev_pool[iv.stop]));
        prof_count[iv.group] += iv.count;
        prof_ms[iv.group] += ms;
      }
    pending.clear();
    ev_used = 0;
  }

  long
  tic_launch() override
  {
    return tic(GTPS_PROF_SMOOTH_FINE);
  }
  void
  toc_launch(long e, int count) override
  {
    toc(GTPS_PROF_SMOOTH_FINE, e, count);
  }

  int
  pre_variant() const
  {
    return schedule == GTPS_MG_ADDITIVE ? GTPS_ADDITIVE : GTPS_MULT_FWD;
  }
  int
  post_variant() const
  {
    if (schedule == GTPS_MG_ADDITIVE)
      return GTPS_ADDITIVE;
    return schedule == GTPS_MG_SYMMETRIC ? GTPS_MULT_REV : GTPS_MULT_FWD;
  }

  // V-cycle on level l with defect dl and correction ul (level precision);
  // n_cycles > 1 repeats smoothing + coarse correction on the same defect
  // (MultigridSolvers::v_cycle(level, my_n_cycles), solver.cuh:1950-2001)
  void
  vcycle(int l, const void *dl, void *ul, int n_cycles = 1)
  {
    const gtps_level *L = lv[l];
    if (l == 2 && n_cycles == 1 && dense_valid && !dense_off && tuning().dense_coarse != 0)
      {
        launch_dense_apply(ctx->stream, dim, L->N1, prec, dense2, dl, ul);
        return;
      }
    if (l == 1)
      {
        // one vertex patch covers the domain: exact solve
        sweep(L, ul, dl, GTPS_MULT_FWD, 1.0, nullptr, nullptr, true);
        return;
      }
    for (int c = 0; c < n_cycles; ++c)
      cycle(l, L, dl, ul, c == 0);
  }

  void
  cycle(int l, const gtps_level *L, const void *dl, void *ul, bool from_zero)
  {
    // finest level: every colour launch is timed on its own; coarser
    // levels: whole sweeps
    const bool   fine = l == maxl;
    LaunchTimer *lt   = fine ? this : nullptr;
    long         e    = fine ? -1 : tic(GTPS_PROF_SMOOTH_COARSE);
    sweep(L, ul, dl, pre_variant(), relax, t[l], lt, from_zero);
    toc(GTPS_PROF_SMOOTH_COARSE, e);
    e = fine ? tic(GTPS_PROF_RESIDUAL_FINE) : -1;
    launch_laplace(L, t[l], ul, dl, false);  // t = d - A u
    toc(GTPS_PROF_RESIDUAL_FINE, e);
    // d_{l-1} = R t: the store form writes every interior coarse DoF; its
    // Dirichlet entries were zeroed at setup and are never written
    e = tic(GTPS_PROF_TRANSFER);
    launch_restrict_store(tr, l, d[l - 1], t[l]);
    toc(GTPS_PROF_TRANSFER, e);
    vcycle(l - 1, d[l - 1], u[l - 1]);
    e = tic(GTPS_PROF_TRANSFER);
    launch_prolongate_add(tr, l, ul, u[l - 1]);
    toc(GTPS_PROF_TRANSFER, e);
    e = fine ? -1 : tic(GTPS_PROF_SMOOTH_COARSE);
    sweep(L, ul, dl, post_variant(), relax, t[l], lt);
    toc(GTPS_PROF_SMOOTH_COARSE, e);
  }

  // z = V(r), rz_old = rz, rz = r.z, beta = rz / rz_old, x += alpha p (the
  // update of the last step, deferred so that p is read once), p = z +
  // beta p; every other buffer is owned by the hierarchy, so the sequence
  // can be captured
  void
  enqueue_direction(double *x)
  {
    hipStream_t     s   = ctx->stream;
    double         *sc  = ctx->ws.scalars;
    double         *pt  = ctx->ws.partials;
    const long long N   = fine64->N;
    const bool      f32 = prec == GTPS_FP32;
    float          *z32 = f32 ? (float *)u[maxl] : nullptr;
    double         *z64 = f32 ? nullptr : (double *)u[maxl];
    const long      e   = tic(GTPS_PROF_VCYCLE);
    if (f32)
      vcycle(maxl, d[maxl], z32);
    else
      vcycle(maxl, r, z64);
    toc(GTPS_PROF_VCYCLE, e);
    const int nb = launch_dot_partials(s, N, r, z64, z32, pt, ctx->ws.capacity);
    launch_finalize_op(s, pt, nb, sc, cg::RZ, cg::OP_SHIFT_BETA);
    launch_cg_update_p(s, N, p, z64, z32, sc, x);
  }

  // a graph captured before the last gtps_set_tuning may replay other
  // kernels than an eager launch would, and one captured for another
  // solution vector writes there: capture it again
  void
  drop_stale_graph(const double *x)
  {
    if (dir_exec && (dir_gen != tuning().generation || dir_x != x))
      {
        GTPS_HIP(hipStreamSynchronize(ctx->stream));
        (void)hipGraphExecDestroy(dir_exec);
        dir_exec = nullptr;
      }
  }

  void
  capture_direction(double *x)
  {
    if (!cap_stream)
      GTPS_HIP(hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking));
    hipStream_t saved = ctx->stream;
    ctx->stream       = cap_stream;
    hipGraph_t graph  = nullptr;
    try
      {
        GTPS_HIP(hipStreamBeginCapture(cap_stream,
                                       hipStreamCaptureModeThreadLocal));
        enqueue_direction(x);
        GTPS_HIP(hipStreamEndCapture(cap_stream, &graph));
      }
    catch (...)
      {
        ctx->stream = saved;
        if (!graph)
          {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(cap_stream, &g);
            if (g)
              (void)hipGraphDestroy(g);
          }
        else
          (void)hipGraphDestroy(graph);
        throw;
      }
    ctx->stream = saved;
    hipError_t e = hipGraphInstantiate(&dir_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    GTPS_HIP(e);
    dir_gen = tuning().generation;
    dir_x   = x;
  }

  // x: the solution, which receives the deferred x += alpha p
  void
  direction(double *x)
  {
    ensure_dense();
    // stage timers need eager launches
    if (!use_graph || prof_mask != 0)
      {
        enqueue_direction(x);
        return;
      }
    drop_stale_graph(x);
    if (!dir_exec)
      capture_direction(x);
    GTPS_HIP(hipGraphLaunch(dir_exec, ctx->stream));
  }

  // z = V(r) in FP64, r and z finest-level vectors
  void
  precondition(const double *rin, double *zout, bool convert_in)
  {
    ensure_dense();
    const long long N = lv[maxl]->N;
    if (prec == GTPS_FP32)
      {
        if (convert_in)
          launch_convert_f64_f32(ctx->stream, N, rin, (float *)d[maxl]);
        vcycle(maxl, d[maxl], u[maxl]);
        if (zout)
          launch_convert_f32_f64(ctx->stream, N, (const float *)u[maxl], zout);
      }
    else
      vcycle(maxl, rin, zout ? (void *)zout : u[maxl]);
  }

  bool
  dense_eligible() const
  {
    // constant coefficient only: a coefficient typically changes between
    // solves (adjoint loops), and tabulating costs one sub-cycle per column
    return tuning().dense_coarse != 0 && maxl >= 3 && coef64.empty() &&
           dense_coarse_size(dim, lv[2]->N1) <= MAX_DENSE_COARSE;
  }

  void
  drop_dense()
  {
    dfree(dense2);
    dense2      = nullptr;
    dense_valid = false;
  }

  // tabulate V(level 2) column by column: unit defects in, the ordinary
  // sub-cycle applied to each, interior values gathered into rows
  void
  ensure_dense()
  {
    if (dense_valid || !dense_eligible())
      return;
    hipStream_t     s  = ctx->stream;
    const int       N1 = lv[2]->N1;
    const long long N  = lv[2]->N;
    const int       ni = dense_coarse_size(dim, N1);
    const size_t    es = prec == GTPS_FP32 ? sizeof(float) : sizeof(double);
    void *D = dmalloc(es * N * ni), *U = nullptr;
    const int saved_mask = prof_mask;
    try
      {
        U = dmalloc(es * N * ni);
        if (!dense2)
          dense2 = dmalloc(es * (size_t)ni * dense_ld(ni));
        launch_unit_defects(s, dim, N1, prec, D);
        dense_off = true;
        prof_mask = 0;
        for (int j = 0; j < ni; ++j)
          vcycle(2, (const char *)D + es * N * j, (char *)U + es * N * j);
        dense_off = false;
        prof_mask = saved_mask;
        launch_gather_columns(s, dim, N1, prec, U, dense2);
        GTPS_HIP(hipStreamSynchronize(s));
      }
    catch (...)
      {
        dense_off = false;
        prof_mask = saved_mask;
        (void)hipStreamSynchronize(s);
        dfree(D);
        dfree(U);
        drop_dense();
        throw;
      }
    dfree(D);
    dfree(U);
    dense_valid = true;
  }

  // FP64 levels, solutions and transfer for full multigrid (lazily)
  void
  ensure_fmg()
  {
    if (!sol.empty())
      return;
    lv64.assign(maxl + 1, nullptr);
    own_lv64.assign(maxl + 1, false);
    sol.assign(maxl + 1, nullptr);
    for (int l = 1; l <= maxl; ++l)
      {
        if (l == maxl)
          lv64[l] = fine64;
        else if (prec == GTPS_FP64)
          lv64[l] = lv[l];
        else
          {
            lv64[l]     = make_level(ctx, dim, k, l, GTPS_FP64);
            own_lv64[l] = true;
            if (!coef64.empty())
              set_level_coefficient(lv64[l], coef64[l]);
          }
        sol[l] = (double *)dmalloc(sizeof(double) * lv64[l]->N);
      }
    if (maxl >= 2)
      {
        if (prec == GTPS_FP64)
          tr64 = tr;
        else
          {
            tr64            = new gtps_transfer();
            own_tr64        = true;
            tr64->ctx       = ctx;
            tr64->dim       = dim;
            tr64->k         = k;
            tr64->min_level = 1;
            tr64->max_level = maxl;
            tr64->prec      = GTPS_FP64;
            tr64->tab       = get_tables(ctx, k);
            const size_t e  = transfer_tmp_elems(dim, k, maxl);
            tr64->tmp1      = dmalloc(e * sizeof(double));
            tr64->tmp2      = dmalloc(e * sizeof(double));
          }
      }
  }

  // full multigrid (MultigridSolvers::solve, solver.cuh:1784-1860)
  void
  fmg(double *x, const double *const *rhs, int n_cycles)
  {
    ensure_fmg();
    ensure_dense();
    hipStream_t s   = ctx->stream;
    const bool  f32 = prec == GTPS_FP32;
    // level 1: exact coarse solve, then one more smoothing step
    const long long N1v = lv[1]->N;
    if (f32)
      launch_convert_f64_f32(s, N1v, rhs[0], (float *)d[1]);
    else
      GTPS_HIP(hipMemcpyAsync(d[1], rhs[0], sizeof(double) * N1v,
                              hipMemcpyDeviceToDevice, s));
    vcycle(1, d[1], u[1]);
    sweep(lv[1], u[1], d[1], GTPS_MULT_FWD, 1.0, nullptr);
    if (f32)
      launch_convert_f32_f64(s, N1v, (const float *)u[1], sol[1]);
    else
      GTPS_HIP(hipMemcpyAsync(sol[1], u[1], sizeof(double) * N1v,
                              hipMemcpyDeviceToDevice, s));
    for (int l = 2; l <= maxl; ++l)
      {
        const gtps_level *L64 = lv64[l];
        const long long   N   = L64->N;
        // prolongate (coarse Dirichlet entries excluded), then impose this
        // level's Dirichlet values
        launch_fill(s, N, GTPS_FP64, sol[l], 0.0);
        launch_prolongate_add(tr64, l, sol[l], sol[l - 1]);
        launch_copy_boundary(L64, rhs[l - 1], sol[l]);
        // FP64 residual into q (a finest-level-sized scratch vector)
        launch_laplace(L64, q, sol[l], rhs[l - 1], false);
        if (f32)
          launch_convert_f64_f32(s, N, q, (float *)d[l]);
        else
          GTPS_HIP(hipMemcpyAsync(d[l], q, sizeof(double) * N,
                                  hipMemcpyDeviceToDevice, s));
        vcycle(l, d[l], u[l], n_cycles);
        if (f32)
          launch_add_f32_to_f64(s, N, (const float *)u[l], sol[l]);
        else
          launch_axpby(s, N, 1.0, (const double *)u[l], 1.0, sol[l]);
      }
    GTPS_HIP(hipMemcpyAsync(x, sol[maxl], sizeof(double) * lv64[maxl]->N,
                            hipMemcpyDeviceToDevice, s));
  }

  void
  free_coef64()
  {
    for (double *c : coef64)
      dfree(c);
    coef64.clear();
  }

  ~gtps_mg()
  {
    drop_dense();
    for (size_t l = 0; l < lv64.size(); ++l)
      {
        if (own_lv64[l])
          delete lv64[l];
        dfree(sol[l]);
      }
    if (own_tr64 && tr64)
      {
        dfree(tr64->tmp1);
        dfree(tr64->tmp2);
        delete tr64;
      }
    free_coef64();
    for (size_t l = 0; l < lv.size(); ++l)
      {
        dfree(d.size() > l ? d[l] : nullptr);
        dfree(u.size() > l ? u[l] : nullptr);
        dfree(t.size() > l ? t[l] : nullptr);
        delete lv[l];
      }
    if (own_fine64)
      delete fine64;
    if (tr)
      {
        dfree(tr->tmp1);
        dfree(tr->tmp2);
        delete tr;
      }
    dfree(r);
    dfree(p);
    dfree(q);
    if (ev)
      (void)hipEventDestroy(ev);
    for (hipEvent_t e : ev_pool)
      (void)hipEventDestroy(e);
    if (dir_exec)
      (void)hipGraphExecDestroy(dir_exec);
    if (cap_stream)
      (void)hipStreamDestroy(cap_stream);
  }
};

namespace
{
  gtps_dg_level *
  make_dg_level(gtps_ctx *ctx, int dim, int degree, int n, int precision,
                double pf)
  {
    GTPS_REQUIRE(dim == 3, "the DG path is 3D");
    GTPS_REQUIRE(degree >= 1 && degree <= MAX_DEGREE, "degree must be in 1..7");
    GTPS_REQUIRE(n >= 2 && n <= 4096, "cells per direction must be in 2..4096");
    GTPS_REQUIRE(precision == GTPS_FP32 || precision == GTPS_FP64,
                 "precision must be GTPS_FP32 or GTPS_FP64");
    GTPS_REQUIRE(std::isfinite(pf) && pf > 0.0, "penalty factor must be positive");
    auto lvl   = std::make_unique<gtps_dg_level>();
    lvl->ctx   = ctx;
    lvl->dim   = dim;
    lvl->k     = degree;
    lvl->n     = n;
    lvl->N1    = n * (degree + 1);
    lvl->N     = (long long)lvl->N1 * lvl->N1 * lvl->N1;
    lvl->prec  = precision;
    lvl->h     = 1.0 / n;
    lvl->scale = lvl->h;  // h^(dim-2)
    lvl->pf    = pf;
    dg_level_setup(lvl.get());
    return lvl.release();
  }

  // solution-file header (64 bytes, little-endian hosts only)
  struct VecHeader
  {
    char               magic[8];
    int                precision, dim, degree, n_cells, layout, reserved;
    long long          n;
    unsigned long long hash;
    unsigned char      pad[16];
  };
  static_assert(sizeof(VecHeader) == 64, "solution-file header must be 64 bytes");

  unsigned long long
  fnv1a(const unsigned char *p, size_t n)
  {
    unsigned long long h = 0xcbf29ce484222325ull;
    for (size_t i = 0; i < n; ++i)
      {
        h ^= p[i];
        h *= 0x100000001b3ull;
      }
    return h;
  }

  long long
  layout_dofs(int dim, int degree, int n_cells, int layout)
  {
    GTPS_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
    GTPS_REQUIRE(degree >= 1 && degree <= MAX_DEGREE, "degree must be in 1..7");
    GTPS_REQUIRE(n_cells >= 1 && n_cells <= 4096, "n_cells must be in 1..4096");
    GTPS_REQUIRE(layout == GTPS_LAYOUT_CG || layout == GTPS_LAYOUT_DG, "bad layout");
    const long long n1 = layout == GTPS_LAYOUT_CG ? (long long)degree * n_cells + 1
                                                  : (long long)n_cells * (degree + 1);
    return dim == 3 ? n1 * n1 * n1 : n1 * n1;
  }

  VecHeader
  make_vec_header(int precision, int dim, int degree, int n_cells, int layout)
  {
    GTPS_REQUIRE(precision == GTPS_FP32 || precision == GTPS_FP64, "bad precision");
    VecHeader hd;
    std::memset(&hd, 0, sizeof hd);
    std::memcpy(hd.magic, "GTPSVEC1", 8);
    hd.precision = precision;
    hd.dim       = dim;
    hd.degree    = degree;
    hd.n_cells   = n_cells;
    hd.layout    = layout;
    hd.n         = layout_dofs(dim, degree, n_cells, layout);
    return hd;
  }

  // reads and validates the header; with data != NULL also the values
  // (checked against the stored hash)
  VecHeader
  read_vec_header(const char *path, std::vector<unsigned char> *data)
  {
    FILE *f = std::fopen(path, "rb");
    GTPS_REQUIRE(f != nullptr, std::string("cannot open: ") + path);
    VecHeader hd;
    const bool got = std::fread(&hd, sizeof hd, 1, f) == 1;
    bool       ok  = got && std::memcmp(hd.magic, "GTPSVEC1", 8) == 0 &&
               (hd.precision == GTPS_FP32 || hd.precision == GTPS_FP64) &&
               hd.dim >= 2 && hd.dim <= 3 && hd.degree >= 1 &&
               hd.degree <= MAX_DEGREE && hd.n_cells >= 1 && hd.n_cells <= 4096 &&
               (hd.layout == GTPS_LAYOUT_CG || hd.layout == GTPS_LAYOUT_DG) &&
               hd.n == layout_dofs(hd.dim, hd.degree, hd.n_cells, hd.layout);
    if (ok && data)
      {
        const size_t bytes = (size_t)hd.n * (hd.precision == GTPS_FP32 ? 4 : 8);
        data->resize(bytes);
        ok = (bytes == 0 || std::fread(data->data(), bytes, 1, f) == 1) &&
             std::fgetc(f) == EOF;
        if (ok && fnv1a(data->data(), bytes) != hd.hash)
          {
            std::fclose(f);
            throw Error(GTPS_EINVAL, std::string("data hash mismatch: ") + path);
          }
      }
    std::fclose(f);
    GTPS_REQUIRE(ok, std::string("not a valid solution file: ") + path);
    return hd;
  }
} // namespace

// SIPG-DG multigrid: levels 0..maxl with n0 2^l cells per direction;
// V(1,1) with the 27-colour vertex-patch smoother (forward pre-, reverse
// post-smoothing), coarse level solved approximately by symmetric sweep
// pairs from zero; FP64 CG on the finest level.
struct gtps_dg_mg
{
  gtps_ctx                     *ctx = nullptr;
  int                           dim = 3, k = 1, n0 = 3, maxl = 0;
  int                           prec = GTPS_FP32, coarse_sweeps = 2;
  double                        pf = 1.0, relax = 1.0;
  std::vector<gtps_dg_level *>  lv;
  gtps_dg_level                *fine64     = nullptr;
  bool                          own_fine64 = false;
  std::vector<void *>           d, u, t;
  double                       *r = nullptr, *p = nullptr, *q = nullptr;
  hipEvent_t                    ev = nullptr;
  // the per-iteration preconditioning sequence replayed as one HIP graph
  // (about 180 launches per V-cycle), captured lazily on a private stream
  bool                          use_graph  = true;
  hipStream_t                   cap_stream = nullptr;
  hipGraphExec_t                dir_exec   = nullptr;
  unsigned long long            dir_gen    = 0;  // tuning generation at capture
  // the batched V-cycle as a graph, for batches of vb_nb rows
  hipGraphExec_t                vb_exec    = nullptr;
  int                           vb_nb      = 0;
  unsigned long long            vb_gen     = 0;
  // the coarsest solve (symmetric sweep pairs from zero, a fixed linear
  // map) tabulated as a dense matrix on first use (coarse.hip)
  void                         *dense0       = nullptr;
  bool                          dense0_valid = false;
  // batched solves: every level vector, r, p, q hold bcap contiguous
  // vectors; per-vector CG scalars and host copies of the residual norms
  int                           bcap    = 1;
  double                       *bsc     = nullptr;
  double                       *bpinned = nullptr;

  // room for batches of nb right-hand sides (drops the captured graph)
  void
  ensure_batch(int nb)
  {
    if (nb <= bcap && bsc)
      return;
    hipStream_t s = ctx->stream;
    GTPS_HIP(hipStreamSynchronize(s));
    if (dir_exec)
      {
        (void)hipGraphExecDestroy(dir_exec);
        dir_exec = nullptr;
      }
    if (vb_exec)
      {
        (void)hipGraphExecDestroy(vb_exec);
        vb_exec = nullptr;
      }
    const int    nc = std::max(nb, bcap);
    const size_t es = prec == GTPS_FP32 ? sizeof(float) : sizeof(double);
    if (nc > bcap)
      {
        for (int l = 0; l <= maxl; ++l)
          {
            for (auto *vec : {&d, &u, &t})
              {
                dfree((*vec)[l]);
                (*vec)[l] = nullptr;
              }
            d[l] = dmalloc(es * lv[l]->N * nc);
            u[l] = dmalloc(es * lv[l]->N * nc);
            t[l] = dmalloc(es * lv[l]->N * nc);
            dg_level_reserve(lv[l], nc);
          }
        if (own_fine64)
          dg_level_reserve(fine64, nc);
        const long long N = fine64->N;
        for (double **v : {&r, &p, &q})
          {
            dfree(*v);
            *v = nullptr;
            *v = (double *)dmalloc(sizeof(double) * N * nc);
          }
        bcap = nc;
      }
    dfree(bsc);
    bsc = (double *)dmalloc(sizeof(double) * cg::N_SLOTS * bcap);
    if (bpinned)
      (void)hipHostFree(bpinned);
    bpinned = nullptr;
    GTPS_HIP(hipHostMalloc((void **)&bpinned, sizeof(double) * bcap, hipHostMallocDefault));
  }

  // V-cycle on nb contiguous vectors (the level's N apart)
  void
  vcycle_batch(int l, const void *dl, void *ul, int nb)
  {
    const gtps_dg_level *L = lv[l];
    if (l == 0 && dense0_valid && tuning().dense_coarse != 0)
      {
        launch_dense_apply_plain(ctx->stream, (int)L->N, prec, dense0, dl, ul, nb);
        return;
      }
    launch_fill(ctx->stream, L->N * nb, prec, ul, 0.0);
    if (l == 0)
      {
        launch_dg_sweeps(L, ul, dl, GTPS_MULT_FWD, 2 * coarse_sweeps, 1.0, nb);
        return;
      }
    launch_dg_sweep(L, ul, dl, GTPS_MULT_FWD, relax, nb);
    launch_dg_apply(L, t[l], ul, dl, false, nb);  // t = d - A u
    launch_dg_restrict(L, d[l - 1], t[l], true, nb);
    vcycle_batch(l - 1, d[l - 1], u[l - 1], nb);
    launch_dg_prolongate_add(L, ul, u[l - 1], nb);
    launch_dg_sweep(L, ul, dl, GTPS_MULT_REV, relax, nb);
  }

  // the batched V-cycle z_j = V(r_j) of every row
  void
  enqueue_vcycle_batch(int nb)
  {
    if (prec == GTPS_FP32)
      vcycle_batch(maxl, d[maxl], u[maxl], nb);
    else
      vcycle_batch(maxl, r, u[maxl], nb);
  }

  // the batched V-cycle replayed as one HIP graph per batch size (it uses
  // only the hierarchy's own buffers); captured again when the batch size or
  // the tuning knobs change, dropped when the buffers are reallocated
  void
  run_vcycle_batch(int nb)
  {
    if (!use_graph)
      {
        enqueue_vcycle_batch(nb);
        return;
      }
    if (vb_exec && (vb_nb != nb || vb_gen != tuning().generation))
      {
        GTPS_HIP(hipStreamSynchronize(ctx->stream));
        (void)hipGraphExecDestroy(vb_exec);
        vb_exec = nullptr;
      }
    if (!vb_exec)
      {
        if (!cap_stream)
          GTPS_HIP(hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking));
        hipStream_t saved = ctx->stream;
        ctx->stream       = cap_stream;
        hipGraph_t graph  = nullptr;
        try
          {
            GTPS_HIP(hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal));
            enqueue_vcycle_batch(nb);
            GTPS_HIP(hipStreamEndCapture(cap_stream, &graph));
          }
        catch (...)
          {
            ctx->stream = saved;
            if (!graph)
              {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(cap_stream, &g);
                if (g)
                  (void)hipGraphDestroy(g);
              }
            else
              (void)hipGraphDestroy(graph);
            throw;
          }
        ctx->stream  = saved;
        hipError_t e = hipGraphInstantiate(&vb_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        GTPS_HIP(e);
        vb_nb  = nb;
        vb_gen = tuning().generation;
      }
    GTPS_HIP(hipGraphLaunch(vb_exec, ctx->stream));
  }

  // z = V(r) for the whole batch, then rz, beta and p = z + beta p for the
  // rows still iterating (itc[j] >= 0: converged, left frozen)
  void
  direction_batch(int nb, const std::vector<int> &itc)
  {
    hipStream_t     s   = ctx->stream;
    const long long N   = fine64->N;
    const bool      f32 = prec == GTPS_FP32;
    double         *pt  = ctx->ws.partials;
    run_vcycle_batch(nb);
    for (int j = 0; j < nb; ++j)
      {
        if (itc[j] >= 0)
          continue;
        const float  *z32 = f32 ? (const float *)u[maxl] + N * j : nullptr;
        const double *z64 = f32 ? nullptr : (const double *)u[maxl] + N * j;
        double       *sc  = bsc + cg::N_SLOTS * j;
        const int nbk = launch_dot_partials(s, N, r + N * j, z64, z32, pt, ctx->ws.capacity);
        launch_finalize_op(s, pt, nbk, sc, cg::RZ, cg::OP_SHIFT_BETA);
        launch_cg_update_p(s, N, p + N * j, z64, z32, sc);
      }
  }

  // residual norms of the batch to the host (one synchronisation)
  void
  read_norms(int nb, double *out)
  {
    hipStream_t s = ctx->stream;
    for (int j = 0; j < nb; ++j)
      GTPS_HIP(hipMemcpyAsync(bpinned + j, bsc + cg::N_SLOTS * j + cg::RR, sizeof(double),
                              hipMemcpyDeviceToHost, s));
    GTPS_HIP(hipStreamSynchronize(s));
    for (int j = 0; j < nb; ++j)
      out[j] = std::sqrt(bpinned[j]);
  }

  void
  ensure_dense0()
  {
    const gtps_dg_level *L = lv[0];
    if (dense0_valid || tuning().dense_coarse == 0 || L->n > 4 || L->N > MAX_DENSE_COARSE)
      return;
    hipStream_t  s  = ctx->stream;
    const int    N  = (int)L->N;
    const size_t es = prec == GTPS_FP32 ? sizeof(float) : sizeof(double);
    void *D = dmalloc(es * N * N), *U = nullptr;
    try
      {
        U = dmalloc(es * N * N);
        if (!dense0)
          dense0 = dmalloc(es * (size_t)N * dense_ld(N));
        launch_unit_defects_plain(s, N, prec, D);
        launch_fill(s, (long long)N * N, prec, U, 0.0);
        launch_dg_sweeps_batched(L, U, D, GTPS_MULT_FWD, 2 * coarse_sweeps, 1.0, N);
        launch_transpose_plain(s, N, prec, U, dense0);
        GTPS_HIP(hipStreamSynchronize(s));
      }
    catch (...)
      {
        (void)hipStreamSynchronize(s);
        dfree(D);
        dfree(U);
        dfree(dense0);
        dense0 = nullptr;
        throw;
      }
    dfree(D);
    dfree(U);
    dense0_valid = true;
  }

  void
  vcycle(int l, const void *dl, void *ul)
  {
    const gtps_dg_level *L = lv[l];
    if (l == 0 && dense0_valid && tuning().dense_coarse != 0)
      {
        launch_dense_apply_plain(ctx->stream, (int)L->N, prec, dense0, dl, ul);
        return;
      }
    launch_fill(ctx->stream, L->N, prec, ul, 0.0);
    if (l == 0)
      {
        launch_dg_sweeps(L, ul, dl, GTPS_MULT_FWD, 2 * coarse_sweeps, 1.0);
        return;
      }
    launch_dg_sweep(L, ul, dl, GTPS_MULT_FWD, relax);
    launch_dg_apply(L, t[l], ul, dl, false);  // t = d - A u
    launch_dg_restrict(L, d[l - 1], t[l], true);
    vcycle(l - 1, d[l - 1], u[l - 1]);
    launch_dg_prolongate_add(L, ul, u[l - 1]);
    launch_dg_sweep(L, ul, dl, GTPS_MULT_REV, relax);
  }

  // a graph captured before the last gtps_set_tuning may replay other
  // kernels than an eager launch would: capture it again
  void
  drop_stale_graph()
  {
    if (dir_exec && dir_gen != tuning().generation)
      {
        GTPS_HIP(hipStreamSynchronize(ctx->stream));
        (void)hipGraphExecDestroy(dir_exec);
        dir_exec = nullptr;
      }
  }

  void
  capture_direction()
  {
    if (!cap_stream)
      GTPS_HIP(hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking));
    hipStream_t saved = ctx->stream;
    ctx->stream       = cap_stream;
    hipGraph_t graph  = nullptr;
    try
      {
        GTPS_HIP(hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal));
        enqueue_direction();
        GTPS_HIP(hipStreamEndCapture(cap_stream, &graph));
      }
    catch (...)
      {
        ctx->stream = saved;
        if (!graph)
          {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(cap_stream, &g);
            if (g)
              (void)hipGraphDestroy(g);
          }
        else
          (void)hipGraphDestroy(graph);
        throw;
      }
    ctx->stream  = saved;
    hipError_t e = hipGraphInstantiate(&dir_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    GTPS_HIP(e);
    dir_gen = tuning().generation;
  }

  void
  direction()
  {
    ensure_dense0();
    if (!use_graph)
      {
        enqueue_direction();
        return;
      }
    drop_stale_graph();
    if (!dir_exec)
      capture_direction();
    GTPS_HIP(hipGraphLaunch(dir_exec, ctx->stream));
  }

  // z = V(r); rz_old = rz; rz = r.z; beta; p = z + beta p
  void
  enqueue_direction()
  {
    hipStream_t     s   = ctx->stream;
    const long long N   = fine64->N;
    const bool      f32 = prec == GTPS_FP32;
    double         *sc  = ctx->ws.scalars;
    double         *pt  = ctx->ws.partials;
    float          *z32 = f32 ? (float *)u[maxl] : nullptr;
    double         *z64 = f32 ? nullptr : (double *)u[maxl];
    if (f32)
      vcycle(maxl, d[maxl], z32);
    else
      vcycle(maxl, r, z64);
    const int nb = launch_dot_partials(s, N, r, z64, z32, pt, ctx->ws.capacity);
    launch_finalize_op(s, pt, nb, sc, cg::RZ, cg::OP_SHIFT_BETA);
    launch_cg_update_p(s, N, p, z64, z32, sc);
  }

  ~gtps_dg_mg()
  {
    dfree(dense0);
    dfree(bsc);
    if (bpinned)
      (void)hipHostFree(bpinned);
    if (dir_exec)
      (void)hipGraphExecDestroy(dir_exec);
    if (vb_exec)
      (void)hipGraphExecDestroy(vb_exec);
    if (cap_stream)
      (void)hipStreamDestroy(cap_stream);
    for (size_t l = 0; l < lv.size(); ++l)
      {
        dfree(d.size() > l ? d[l] : nullptr);
        dfree(u.size() > l ? u[l] : nullptr);
        dfree(t.size() > l ? t[l] : nullptr);
        delete lv[l];
      }
    if (own_fine64)
      delete fine64;
    dfree(r);
    dfree(p);
    dfree(q);
    if (ev)
      (void)hipEventDestroy(ev);
  }
};

extern "C"
{
  const char *
  gtps_last_error(void)
  {
    return g_last_error.c_str();
  }

  const char *
  gtps_version(void)
  {
    return "gtps 0.1.0 gfx950";
  }

  int
  gtps_ctx_create(int device, void *stream, gtps_ctx **out)
  {
    return guard([&] {
      GTPS_REQUIRE(out != nullptr, "out is NULL");
      int ndev = 0;
      GTPS_HIP(hipGetDeviceCount(&ndev));
      GTPS_REQUIRE(device >= 0 && device < ndev, "no such device");
      GTPS_HIP(hipSetDevice(device));
      scratch_setup(device);
      auto ctx    = std::make_unique<gtps_ctx>();
      ctx->device = device;
      if (stream)
        ctx->stream = (hipStream_t)stream;
      else
        {
          GTPS_HIP(hipStreamCreate(&ctx->stream));
          ctx->own_stream = true;
        }
      ctx->ws.capacity = 65536;
      ctx->ws.partials =
        (double *)dmalloc(sizeof(double) * ctx->ws.capacity);
      ctx->ws.scalars = (double *)dmalloc(sizeof(double) * cg::N_SLOTS);
      GTPS_HIP(hipHostMalloc((void **)&ctx->ws.pinned,
                             sizeof(double) * cg::N_SLOTS,
                             hipHostMallocDefault));
      GTPS_HIP(hipHostMalloc((void **)&ctx->ws.readback, sizeof(double) * cg::N_SLOTS,
                             hipHostMallocMapped | hipHostMallocCoherent));
      GTPS_HIP(hipHostGetDevicePointer((void **)&ctx->ws.readback_dev,
                                       ctx->ws.readback, 0));
      *out = ctx.release();
    });
  }

  int
  gtps_ctx_destroy(gtps_ctx *ctx)
  {
    return guard([&] {
      if (!ctx)
        return;
      (void)hipSetDevice(ctx->device);
      (void)hipStreamSynchronize(ctx->stream);
      for (auto &kv : ctx->tables)
        {
          dfree(kv.second->d64);
          dfree(kv.second->d32);
          dfree(kv.second->sep64);
          dfree(kv.second->sep32);
        }
      dfree(ctx->ws.partials);
      dfree(ctx->ws.scalars);
      if (ctx->ws.pinned)
        (void)hipHostFree(ctx->ws.pinned);
      if (ctx->ws.readback)
        (void)hipHostFree(ctx->ws.readback);
      if (ctx->own_stream)
        (void)hipStreamDestroy(ctx->stream);
      delete ctx;
    });
  }

  int
  gtps_ctx_synchronize(gtps_ctx *ctx)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx, "ctx is NULL");
      set_device(ctx);
      GTPS_HIP(hipStreamSynchronize(ctx->stream));
    });
  }

  void *
  gtps_ctx_stream(gtps_ctx *ctx)
  {
    return ctx ? (void *)ctx->stream : nullptr;
  }

  // ------------------------------------------------------------- levels
  int
  gtps_level_create(gtps_ctx *ctx, int dim, int degree, int level,
                    int precision, gtps_level **out)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && out, "NULL argument");
      set_device(ctx);
      *out = make_level(ctx, dim, degree, level, precision);
    });
  }

  int
  gtps_level_destroy(gtps_level *lvl)
  {
    return guard([&] { delete lvl; });
  }

  long long
  gtps_level_n_dofs(const gtps_level *lvl)
  {
    return lvl ? lvl->N : -1;
  }

  int
  gtps_laplace_vmult(gtps_level *lvl, void *dst, const void *src)
  {
    return guard([&] {
      GTPS_REQUIRE(lvl && dst && src, "NULL argument");
      GTPS_REQUIRE(dst != src, "vmult cannot run in place");
      set_device(lvl->ctx);
      launch_laplace(lvl, dst, src, nullptr, false);
    });
  }

  int
  gtps_laplace_residual(gtps_level *lvl, void *dst, const void *b,
                        const void *u)
  {
    return guard([&] {
      GTPS_REQUIRE(lvl && dst && b && u, "NULL argument");
      GTPS_REQUIRE(dst != u, "residual output cannot alias u");
      set_device(lvl->ctx);
      launch_laplace(lvl, dst, u, b, false);
    });
  }

  int
  gtps_laplace_vmult_transpose(gtps_level *lvl, void *dst, const void *src)
  {
    return guard([&] {
      GTPS_REQUIRE(lvl && dst && src, "NULL argument");
      GTPS_REQUIRE(dst != src, "vmult cannot run in place");
      GTPS_REQUIRE(lvl->dim == 3, "the transpose operator is 3D");
      set_device(lvl->ctx);
      ensure_celem(lvl);
      launch_coeff_apply(lvl, dst, src, nullptr, false, true);
    });
  }

  int
  gtps_level_set_coefficient(gtps_level *lvl, const double *coef)
  {
    return guard([&] {
      GTPS_REQUIRE(lvl, "NULL level");
      set_device(lvl->ctx);
      if (coef)
        set_level_coefficient(lvl, coef);
      else
        clear_level_coefficient(lvl);
    });
  }

  int
  gtps_level_coefficient_gradient(gtps_level *lvl, void *grad,
                                  const void *lambda, const void *u)
  {
    return guard([&] {
      GTPS_REQUIRE(lvl && grad && lambda && u, "NULL argument");
      GTPS_REQUIRE(lvl->dim == 3, "the coefficient gradient is 3D");
      set_device(lvl->ctx);
      launch_coeff_gradient(lvl, grad, lambda, u);
    });
  }

  int
  gtps_level_tensor_product_data(gtps_level *lvl, double *cell_mass,
                                 double *cell_laplace, double *patch_mass,
                                 double *patch_laplace, double *eigenvalues,
                                 double *eigenvectors)
  {
    return guard([&] {
      GTPS_REQUIRE(lvl, "NULL level");
      const Basis1D &b    = lvl->tab->basis;
      auto           copy = [](const std::vector<double> &v, double *o) {
        if (o)
          std::memcpy(o, v.data(), sizeof(double) * v.size());
      };
      copy(b.cellM, cell_mass);
      copy(b.cellK, cell_laplace);
      copy(b.patchM, patch_mass);
      copy(b.patchK, patch_laplace);
      copy(b.lam, eigenvalues);
      copy(b.S, eigenvectors);
    });
  }

  // ------------------------------------------------------------ smoother
  int
  gtps_smoother_create(gtps_level *lvl, int variant, double relaxation,
                       gtps_smoother **out)
  {
    return guard([&] {
      GTPS_REQUIRE(lvl && out, "NULL argument");
      GTPS_REQUIRE(variant == GTPS_MULT_FWD || variant == GTPS_MULT_REV ||
                     variant == GTPS_ADDITIVE,
                   "unknown smoother variant");
      GTPS_REQUIRE(std::isfinite(relaxation) && relaxation > 0.0,
                   "relaxation must be positive");
      set_device(lvl->ctx);
      auto sm     = std::make_unique<gtps_smoother>();
      sm->lvl     = lvl;
      sm->variant = variant;
      sm->relax   = relaxation;
      if (variant == GTPS_ADDITIVE)
        sm->resid = dmalloc(lvl->elem() * lvl->N);
      *out = sm.release();
    });
  }

  int
  gtps_smoother_destroy(gtps_smoother *sm)
  {
    return guard([&] {
      if (!sm)
        return;
      dfree(sm->resid);
      delete sm;
    });
  }

  int
  gtps_smoother_step(gtps_smoother *sm, void *u, const void *b)
  {
    return guard([&] {
      GTPS_REQUIRE(sm && u && b, "NULL argument");
      GTPS_REQUIRE(u != b, "u and b must differ");
      set_device(sm->lvl->ctx);
      sweep(sm->lvl, u, b, sm->variant, sm->relax, sm->resid);
    });
  }

  int
  gtps_smoother_vmult(gtps_smoother *sm, void *u, const void *b)
  {
    return guard([&] {
      GTPS_REQUIRE(sm && u && b, "NULL argument");
      GTPS_REQUIRE(u != b, "u and b must differ");
      set_device(sm->lvl->ctx);
      sweep(sm->lvl, u, b, sm->variant, sm->relax, sm->resid, nullptr, true);
    });
  }

  // ------------------------------------------------------------ transfer
  int
  gtps_transfer_create(gtps_ctx *ctx, int dim, int degree, int min_level,
                       int max_level, int precision, gtps_transfer **out)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && out, "NULL argument");
      check_level_args(dim, degree, max_level, precision);
      GTPS_REQUIRE(min_level >= 1 && min_level < max_level,
                   "need 1 <= min_level < max_level");
      set_device(ctx);
      auto tr       = std::make_unique<gtps_transfer>();
      tr->ctx       = ctx;
      tr->dim       = dim;
      tr->k         = degree;
      tr->min_level = min_level;
      tr->max_level = max_level;
      tr->prec      = precision;
      tr->tab       = get_tables(ctx, degree);
      const size_t e = transfer_tmp_elems(dim, degree, max_level);
      const size_t s = precision == GTPS_FP32 ? sizeof(float) : sizeof(double);
      tr->tmp1       = dmalloc(e * s);
      tr->tmp2       = dmalloc(e * s);
      *out           = tr.release();
    });
  }

  int
  gtps_transfer_destroy(gtps_transfer *tr)
  {
    return guard([&] {
      if (!tr)
        return;
      dfree(tr->tmp1);
      dfree(tr->tmp2);
      delete tr;
    });
  }

  int
  gtps_transfer_prolongate_add(gtps_transfer *tr, int to_level, void *fine,
                               const void *coarse)
  {
    return guard([&] {
      GTPS_REQUIRE(tr && fine && coarse, "NULL argument");
      GTPS_REQUIRE(to_level > tr->min_level && to_level <= tr->max_level,
                   "to_level out of range");
      set_device(tr->ctx);
      launch_prolongate_add(tr, to_level, fine, coarse);
    });
  }

  int
  gtps_transfer_prolongate(gtps_transfer *tr, int to_level, void *fine,
                           const void *coarse)
  {
    return guard([&] {
      GTPS_REQUIRE(tr && fine && coarse, "NULL argument");
      GTPS_REQUIRE(to_level > tr->min_level && to_level <= tr->max_level,
                   "to_level out of range");
      set_device(tr->ctx);
      const long long nf = 1LL * tr->k * (1LL << to_level) + 1;
      launch_fill(tr->ctx->stream, tr->dim == 3 ? nf * nf * nf : nf * nf,
                  tr->prec, fine, 0.0);
      launch_prolongate_add(tr, to_level, fine, coarse);
    });
  }

  int
  gtps_transfer_restrict_add(gtps_transfer *tr, int from_level, void *coarse,
                             const void *fine)
  {
    return guard([&] {
      GTPS_REQUIRE(tr && fine && coarse, "NULL argument");
      GTPS_REQUIRE(from_level > tr->min_level && from_level <= tr->max_level,
                   "from_level out of range");
      set_device(tr->ctx);
      launch_restrict_add(tr, from_level, coarse, fine);
    });
  }

  // ------------------------------------------------------------- vectors
  int
  gtps_vec_fill(gtps_ctx *ctx, long long n, int precision, void *x,
                double value)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && (x || n == 0) && n >= 0, "bad argument");
      GTPS_REQUIRE(precision == GTPS_FP32 || precision == GTPS_FP64,
                   "precision must be GTPS_FP32 or GTPS_FP64");
      set_device(ctx);
      launch_fill(ctx->stream, n, precision, x, value);
    });
  }

  int
  gtps_vec_axpby(gtps_ctx *ctx, long long n, double a, const double *x,
                 double b, double *y)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && n >= 0 && (n == 0 || (x && y)), "bad argument");
      set_device(ctx);
      launch_axpby(ctx->stream, n, a, x, b, y);
    });
  }

  int
  gtps_vec_dot(gtps_ctx *ctx, long long n, const double *x, const double *y,
               double *result)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && result && n >= 0 && (n == 0 || (x && y)),
                   "bad argument");
      set_device(ctx);
      if (n == 0)
        {
          *result = 0.0;
          return;
        }
      const int nb = launch_dot_partials(ctx->stream, n, x, y, nullptr,
                                         ctx->ws.partials, ctx->ws.capacity);
      launch_finalize_sum(ctx->stream, ctx->ws.partials, nb, ctx->ws.scalars,
                          cg::N_SLOTS - 1);
      GTPS_HIP(hipMemcpyAsync(ctx->ws.pinned, ctx->ws.scalars + cg::N_SLOTS - 1,
                              sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
      GTPS_HIP(hipStreamSynchronize(ctx->stream));
      *result = ctx->ws.pinned[0];
    });
  }

  int
  gtps_vec_dot_noalloc(gtps_ctx *ctx, long long n, const double *x, const double *y,
                       double *result)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && result && n >= 0 && (n == 0 || (x && y)), "bad argument");
      GTPS_REQUIRE(n <= GTPS_NOALLOC_MAX_N,
                   "n exceeds GTPS_NOALLOC_MAX_N (2^62) for gtps_vec_dot_noalloc");
      double *pt = scratch_of(ctx->device);
      set_device(ctx);
      if (n == 0)
        {
          launch_fill(ctx->stream, 1, GTPS_FP64, result, 0.0);
          return;
        }
      // the same kernels and grid as gtps_vec_dot: the partial count is
      // capped at GTPS_NOALLOC_PARTIALS work-groups (grid-stride loops), so
      // every partial slot a launch uses is written before it is read
      const int nb = launch_dot_partials(ctx->stream, n, x, y, nullptr, pt,
                                         GTPS_NOALLOC_PARTIALS);
      GTPS_REQUIRE(nb >= 1 && nb <= GTPS_NOALLOC_PARTIALS, "partial count out of range");
      launch_finalize_sum(ctx->stream, pt, nb, result, 0);
    });
  }

  int
  gtps_vec_nrm2(gtps_ctx *ctx, long long n, const double *x, double *result)
  {
    double d  = 0.0;
    int    rc = gtps_vec_dot(ctx, n, x, x, &d);
    if (rc == GTPS_OK && result)
      *result = std::sqrt(d);
    return rc;
  }

  int
  gtps_vec_convert_f64_f32(gtps_ctx *ctx, long long n, const double *x,
                           float *y)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && n >= 0 && (n == 0 || (x && y)), "bad argument");
      set_device(ctx);
      launch_convert_f64_f32(ctx->stream, n, x, y);
    });
  }

  int
  gtps_vec_convert_f32_f64(gtps_ctx *ctx, long long n, const float *x,
                           double *y)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && n >= 0 && (n == 0 || (x && y)), "bad argument");
      set_device(ctx);
      launch_convert_f32_f64(ctx->stream, n, x, y);
    });
  }

  int
  gtps_vec_set_boundary(gtps_level *lvl, void *v, double value)
  {
    return guard([&] {
      GTPS_REQUIRE(lvl && v, "NULL argument");
      set_device(lvl->ctx);
      launch_set_boundary(lvl, v, value);
    });
  }

  // ------------------------------------------------------------ multigrid
  int
  gtps_mg_create(gtps_ctx *ctx, int dim, int degree, int max_level,
                 int vcycle_precision, int schedule, double relaxation,
                 gtps_mg **out)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && out, "NULL argument");
      check_level_args(dim, degree, max_level, vcycle_precision);
      GTPS_REQUIRE(max_level >= 1, "max_level must be >= 1");
      GTPS_REQUIRE(schedule == GTPS_MG_SYMMETRIC ||
                     schedule == GTPS_MG_REFERENCE_ORDER ||
                     schedule == GTPS_MG_ADDITIVE,
                   "unknown schedule");
      GTPS_REQUIRE(std::isfinite(relaxation) && relaxation > 0.0,
                   "relaxation must be positive");
      set_device(ctx);
      auto mg      = std::make_unique<gtps_mg>();
      mg->ctx      = ctx;
      mg->dim      = dim;
      mg->k        = degree;
      mg->maxl     = max_level;
      mg->prec     = vcycle_precision;
      mg->schedule = schedule;
      mg->relax    = relaxation;
      mg->lv.assign(max_level + 1, nullptr);
      mg->d.assign(max_level + 1, nullptr);
      mg->u.assign(max_level + 1, nullptr);
      mg->t.assign(max_level + 1, nullptr);
      const size_t es =
        vcycle_precision == GTPS_FP32 ? sizeof(float) : sizeof(double);
      for (int l = 1; l <= max_level; ++l)
        {
          mg->lv[l] = make_level(ctx, dim, degree, l, vcycle_precision);
          const size_t bytes = es * mg->lv[l]->N;
          mg->d[l]           = dmalloc(bytes);
          mg->u[l]           = dmalloc(bytes);
          mg->t[l]           = dmalloc(bytes);
        }
      if (vcycle_precision == GTPS_FP64)
        mg->fine64 = mg->lv[max_level];
      else
        {
          mg->fine64     = make_level(ctx, dim, degree, max_level, GTPS_FP64);
          mg->own_fine64 = true;
        }
      if (max_level >= 2)
        {
          mg->tr            = new gtps_transfer();
          mg->tr->ctx       = ctx;
          mg->tr->dim       = dim;
          mg->tr->k         = degree;
          mg->tr->min_level = 1;
          mg->tr->max_level = max_level;
          mg->tr->prec      = vcycle_precision;
          mg->tr->tab       = get_tables(ctx, degree);
          const size_t e    = transfer_tmp_elems(dim, degree, max_level);
          mg->tr->tmp1      = dmalloc(e * es);
          mg->tr->tmp2      = dmalloc(e * es);
        }
      const long long N = mg->lv[max_level]->N;
      mg->r             = (double *)dmalloc(sizeof(double) * N);
      mg->p             = (double *)dmalloc(sizeof(double) * N);
      mg->q             = (double *)dmalloc(sizeof(double) * N);
      GTPS_HIP(hipEventCreateWithFlags(&mg->ev, hipEventDisableTiming));
      // zero boundary entries of the level vectors once: kernels never write
      // them except through vmult / residual
      for (int l = 1; l <= max_level; ++l)
        {
          launch_fill(ctx->stream, mg->lv[l]->N, vcycle_precision, mg->d[l], 0.0);
          launch_fill(ctx->stream, mg->lv[l]->N, vcycle_precision, mg->u[l], 0.0);
          launch_fill(ctx->stream, mg->lv[l]->N, vcycle_precision, mg->t[l], 0.0);
        }
      GTPS_HIP(hipStreamSynchronize(ctx->stream));
      *out = mg.release();
    });
  }

  int
  gtps_mg_destroy(gtps_mg *mg)
  {
    return guard([&] {
      if (!mg)
        return;
      (void)hipSetDevice(mg->ctx->device);
      (void)hipStreamSynchronize(mg->ctx->stream);
      delete mg;
    });
  }

  long long
  gtps_mg_n_dofs(const gtps_mg *mg)
  {
    return mg ? mg->lv[mg->maxl]->N : -1;
  }

  int
  gtps_mg_vcycle(gtps_mg *mg, double *z, const double *r)
  {
    return guard([&] {
      GTPS_REQUIRE(mg && z && r, "NULL argument");
      GTPS_REQUIRE((const void *)z != (const void *)r, "z and r must differ");
      set_device(mg->ctx);
      const long e = mg->tic(GTPS_PROF_VCYCLE);
      mg->precondition(r, z, true);
      mg->toc(GTPS_PROF_VCYCLE, e);
    });
  }

  long long
  gtps_set_tuning(const char *key, long long value)
  {
    long long prev = -1;
    int rc = guard([&] {
      GTPS_REQUIRE(key != nullptr, "NULL key");
      const std::string k(key);
      if (k == "small_colour_patches")
        {
          GTPS_REQUIRE(value >= 0, "value must be >= 0");
          prev                              = tuning().small_colour_patches;
          tuning().small_colour_patches     = value;
        }
      else if (k == "group_min_patches")
        {
          GTPS_REQUIRE(value >= 0, "value must be >= 0");
          prev                          = tuning().group_min_patches;
          tuning().group_min_patches    = value;
        }
      else if (k == "patches_per_wave")
        {
          GTPS_REQUIRE(value >= 1 && value <= 5, "patches_per_wave must be in 1..5");
          prev                         = tuning().patches_per_wave;
          tuning().patches_per_wave    = (int)value;
        }
      else if (k == "group_waves")
        {
          GTPS_REQUIRE(value == 0 || value == 1 || value == 4, "group_waves must be 0, 1 or 4");
          prev                  = tuning().group_waves;
          tuning().group_waves  = (int)value;
        }
      else if (k == "box_vmult_max_n")
        {
          GTPS_REQUIRE(value >= 0, "value must be >= 0");
          prev                       = tuning().box_vmult_max_n;
          tuning().box_vmult_max_n   = value;
        }
      else if (k == "row_staged")
        {
          GTPS_REQUIRE(value == 0 || value == 1, "row_staged must be 0 or 1");
          prev                  = tuning().row_staged;
          tuning().row_staged   = (int)value;
        }
      else if (k == "dense_coarse")
        {
          GTPS_REQUIRE(value == 0 || value == 1, "dense_coarse must be 0 or 1");
          prev                  = tuning().dense_coarse;
          tuning().dense_coarse = (int)value;
        }
      else if (k == "vmult_tile")
        {
          GTPS_REQUIRE(value >= 0 && value <= 3, "vmult_tile must be 0, 1, 2 or 3");
          prev                = tuning().vmult_tile;
          tuning().vmult_tile = (int)value;
        }
      else if (k == "vmult_zdiv")
        {
          GTPS_REQUIRE(value >= 0 && value <= (1 << 20), "vmult_zdiv must be in 0..2^20");
          prev                = tuning().vmult_zdiv;
          tuning().vmult_zdiv = value;
        }
      else if (k == "mfma_smoother")
        {
          GTPS_REQUIRE(value == 0 || value == 1, "mfma_smoother must be 0 or 1");
          prev                   = tuning().mfma_smoother;
          tuning().mfma_smoother = (int)value;
        }
      else if (k == "mfma_waves")
        {
          GTPS_REQUIRE(value == 4 || value == 8 || value == 16,
                       "mfma_waves must be 4, 8 or 16");
          prev                = tuning().mfma_waves;
          tuning().mfma_waves = (int)value;
        }
      else
        throw Error(GTPS_EINVAL, "unknown tuning key " + k);
      ++tuning().generation;
    });
    return rc == GTPS_OK ? prev : -1;
  }

  long long
  gtps_get_tuning(const char *key)
  {
    long long v = -1;
    int rc = guard([&] {
      GTPS_REQUIRE(key != nullptr, "NULL key");
      const std::string k(key);
      if (k == "small_colour_patches")
        v = tuning().small_colour_patches;
      else if (k == "group_min_patches")
        v = tuning().group_min_patches;
      else if (k == "patches_per_wave")
        v = tuning().patches_per_wave;
      else if (k == "group_waves")
        v = tuning().group_waves;
      else if (k == "box_vmult_max_n")
        v = tuning().box_vmult_max_n;
      else if (k == "row_staged")
        v = tuning().row_staged;
      else if (k == "dense_coarse")
        v = tuning().dense_coarse;
      else if (k == "mfma_smoother")
        v = tuning().mfma_smoother;
      else if (k == "mfma_waves")
        v = tuning().mfma_waves;
      else if (k == "vmult_zdiv")
        v = tuning().vmult_zdiv;
      else if (k == "vmult_tile")
        v = tuning().vmult_tile;
      else
        throw Error(GTPS_EINVAL, "unknown tuning key " + k);
    });
    return rc == GTPS_OK ? v : -1;
  }

  int
  gtps_mg_set_graph(gtps_mg *mg, int enable)
  {
    return guard([&] {
      GTPS_REQUIRE(mg, "NULL argument");
      mg->use_graph = enable != 0;
    });
  }

  int
  gtps_mg_set_profiling(gtps_mg *mg, int mask)
  {
    return guard([&] {
      GTPS_REQUIRE(mg, "NULL argument");
      GTPS_REQUIRE(mask >= 0 && mask < (1 << GTPS_PROF_N_GROUPS), "bad mask");
      mg->prof_mask = mask;
    });
  }

  int
  gtps_mg_profile(gtps_mg *mg, int group, long long *count, double *total_ms)
  {
    return guard([&] {
      GTPS_REQUIRE(mg, "NULL argument");
      GTPS_REQUIRE(group >= 0 && group < GTPS_PROF_N_GROUPS, "bad group");
      set_device(mg->ctx);
      GTPS_HIP(hipStreamSynchronize(mg->ctx->stream));
      mg->collect();
      if (count)
        *count = mg->prof_count[group];
      if (total_ms)
        *total_ms = mg->prof_ms[group];
    });
  }

  int
  gtps_mg_profile_reset(gtps_mg *mg)
  {
    return guard([&] {
      GTPS_REQUIRE(mg, "NULL argument");
      set_device(mg->ctx);
      GTPS_HIP(hipStreamSynchronize(mg->ctx->stream));
      mg->pending.clear();
      mg->ev_used = 0;
      for (int g = 0; g < GTPS_PROF_N_GROUPS; ++g)
        {
          mg->prof_count[g] = 0;
          mg->prof_ms[g]    = 0.0;
        }
    });
  }

  int
  gtps_mg_matvec(gtps_mg *mg, double *dst, const double *src)
  {
    return guard([&] {
      GTPS_REQUIRE(mg && dst && src && (const void *)dst != (const void *)src,
                   "bad argument");
      set_device(mg->ctx);
      launch_laplace(mg->fine64, dst, src, nullptr, false);
    });
  }

  int
  gtps_mg_matvec_transpose(gtps_mg *mg, double *dst, const double *src)
  {
    return guard([&] {
      GTPS_REQUIRE(mg && dst && src && (const void *)dst != (const void *)src,
                   "bad argument");
      GTPS_REQUIRE(mg->dim == 3, "the transpose operator is 3D");
      set_device(mg->ctx);
      ensure_celem(mg->fine64);
      launch_coeff_apply(mg->fine64, dst, src, nullptr, false, true);
    });
  }

  int
  gtps_mg_set_coefficient(gtps_mg *mg, const double *coef)
  {
    return guard([&] {
      GTPS_REQUIRE(mg, "NULL argument");
      set_device(mg->ctx);
      hipStream_t s = mg->ctx->stream;
      // kernels change: the captured preconditioning graph is stale
      GTPS_HIP(hipStreamSynchronize(s));
      if (mg->dir_exec)
        {
          (void)hipGraphExecDestroy(mg->dir_exec);
          mg->dir_exec = nullptr;
        }
      mg->drop_dense();
      auto each_level = [&](auto &&f) {
        for (int l = 1; l <= mg->maxl; ++l)
          f(mg->lv[l], l);
        if (mg->own_fine64)
          f(mg->fine64, mg->maxl);
        for (size_t l = 0; l < mg->lv64.size(); ++l)
          if (mg->own_lv64[l])
            f(mg->lv64[l], (int)l);
      };
      if (!coef)
        {
          each_level([](gtps_level *L, int) { clear_level_coefficient(L); });
          mg->free_coef64();
          return;
        }
      GTPS_REQUIRE(mg->dim == 3, "the variable-coefficient operator is 3D");
      GTPS_REQUIRE(mg->schedule != GTPS_MG_ADDITIVE,
                   "the additive smoother has no variable-coefficient form");
      GTPS_REQUIRE(coeff_smoother_supported(mg->k, mg->prec),
                   "the variable-coefficient smoother supports degree <= 5 in "
                   "FP64 and <= 7 in FP32");
      // FP64 chain of averaged coefficients, owned by the hierarchy
      if (mg->coef64.empty())
        {
          mg->coef64.assign(mg->maxl + 1, nullptr);
          for (int l = 1; l <= mg->maxl; ++l)
            mg->coef64[l] = (double *)dmalloc(sizeof(double) * (1LL << (3 * l)));
        }
      GTPS_HIP(hipMemcpyAsync(mg->coef64[mg->maxl], coef,
                              sizeof(double) * (1LL << (3 * mg->maxl)),
                              hipMemcpyDeviceToDevice, s));
      for (int l = mg->maxl - 1; l >= 1; --l)
        launch_coarsen_coefficient(s, 1 << l, mg->coef64[l + 1], mg->coef64[l]);
      each_level([&](gtps_level *L, int l) { set_level_coefficient(L, mg->coef64[l]); });
      GTPS_HIP(hipStreamSynchronize(s));
    });
  }

  int
  gtps_mg_fmg_solve(gtps_mg *mg, double *x, const double *const *rhs,
                    int n_cycles)
  {
    return guard([&] {
      GTPS_REQUIRE(mg && x && rhs, "NULL argument");
      GTPS_REQUIRE(n_cycles >= 1, "n_cycles must be >= 1");
      for (int l = 1; l <= mg->maxl; ++l)
        GTPS_REQUIRE(rhs[l - 1] != nullptr, "missing right-hand side of a level");
      GTPS_REQUIRE((const void *)x != (const void *)rhs[mg->maxl - 1],
                   "x and the finest right-hand side must differ");
      set_device(mg->ctx);
      mg->fmg(x, rhs, n_cycles);
    });
  }

  int
  gtps_mg_coefficient_gradient(gtps_mg *mg, double *grad, const double *lambda,
                               const double *u)
  {
    return guard([&] {
      GTPS_REQUIRE(mg && grad && lambda && u, "NULL argument");
      GTPS_REQUIRE(mg->dim == 3, "the coefficient gradient is 3D");
      set_device(mg->ctx);
      launch_coeff_gradient(mg->fine64, grad, lambda, u);
    });
  }

  int
  gtps_mg_pcg_solve(gtps_mg *mg, double *x, const double *b, double rtol,
                    int maxit, int *iters, double *relres, double *history)
  {
    int  it_done = 0;
    double rel   = 0.0;
    bool converged = false, nonfinite = false;
    int rc = guard([&] {
      GTPS_REQUIRE(mg && x && b && (const void *)x != (const void *)b,
                   "bad argument");
      GTPS_REQUIRE(maxit >= 0 && rtol >= 0.0, "bad rtol / maxit");
      set_device(mg->ctx);
      gtps_ctx      *ctx = mg->ctx;
      hipStream_t    s   = ctx->stream;
      const long long N  = mg->fine64->N;
      double        *sc  = ctx->ws.scalars;
      double        *pt  = ctx->ws.partials;
      const int      cap = ctx->ws.capacity;
      const bool     f32 = mg->prec == GTPS_FP32;
      float         *r32 = f32 ? (float *)mg->d[mg->maxl] : nullptr;

      // the finalize kernel of every residual norm also writes it to mapped
      // host memory; the host waits on an event recorded right after it
      double *rb      = ctx->ws.readback_dev;
      auto    read_rr = [&](int nb) {
        launch_finalize_sum(s, pt, nb, sc, cg::RR, rb);
        GTPS_HIP(hipEventRecord(mg->ev, s));
      };
      auto wait_rr = [&]() {
        GTPS_HIP(hipEventSynchronize(mg->ev));
        return std::sqrt(ctx->ws.readback[0]);
      };

      // x0 = lift(b), r0 = b - A x0.  When b vanishes on the Dirichlet DoFs
      // (the usual right-hand side) this is x0 = 0, r0 = b, one fused pass;
      // a sum of b^2 over the boundary faces decides.
      {
        const int nbb = launch_boundary_sumsq(mg->fine64, b, pt, cap);
        launch_finalize_sum(s, pt, nbb, sc, cg::BB, rb + 1);
        GTPS_HIP(hipEventRecord(mg->ev, s));
        GTPS_HIP(hipEventSynchronize(mg->ev));
      }
      if (ctx->ws.readback[1] == 0.0)
        read_rr(launch_cg_init(s, N, b, x, mg->r, r32, mg->p, sc, pt, cap));
      else
        {
          launch_lift_boundary(mg->fine64, b, x);
          launch_laplace(mg->fine64, mg->r, x, b, false);
          if (f32)
            launch_convert_f64_f32(s, N, mg->r, r32);
          // p = 0 and rz = 1 make the first direction update p = z + beta 0
          launch_fill(s, N, GTPS_FP64, mg->p, 0.0);
          launch_cg_scalar(s, sc, cg::OP_INIT);
          read_rr(launch_dot_partials(s, N, mg->r, mg->r, nullptr, pt, cap));
        }
      // the first preconditioning step is enqueued before the host learns
      // r0 (a zero or non-finite r0 discards it; alpha = 0 here, so its x
      // update leaves x as it is)
      if (maxit > 0)
        mg->direction(x);
      const double r0 = wait_rr();
      if (history)
        history[0] = r0;
      if (!std::isfinite(r0))
        {
          nonfinite = true;
          GTPS_HIP(hipStreamSynchronize(s));
          return;
        }
      if (r0 == 0.0)
        {
          converged = true;
          GTPS_HIP(hipStreamSynchronize(s));
          return;
        }
      double rn = r0, rn_prev = -1.0;
      // x += alpha p of the last step not yet enqueued (the next direction
      // step applies it)
      bool x_pending = false;
      while (it_done < maxit)
        {
          // q = A p with fused p.q, alpha = rz / pq
          const long e  = mg->tic(GTPS_PROF_CG_MATVEC);
          const int nbq = launch_laplace(mg->fine64, mg->q, mg->p, nullptr, true);
          mg->toc(GTPS_PROF_CG_MATVEC, e);
          launch_finalize_op(s, pt, nbq, sc, cg::PQ, cg::OP_ALPHA);
          read_rr(launch_cg_update_xr(s, N, nullptr, mg->r, mg->p, mg->q, r32, sc, pt, cap));
          x_pending = true;
          ++it_done;
          // Enqueue the next preconditioning step while the host checks,
          // unless the residual history predicts that this iteration
          // converges (then the step would only be waited for and dropped).
          // The prediction assumes the last reduction factor again, with a
          // margin of 4; a wrong guess costs one host round trip.
          bool speculate = it_done < maxit;
          if (speculate && rn_prev > 0.0 && rn < rn_prev)
            speculate = rn * (rn / rn_prev) > 0.25 * rtol * r0;
          if (speculate)
            {
              mg->direction(x);
              x_pending = false;
            }
          rn_prev = rn;
          rn      = wait_rr();
          if (history)
            history[it_done] = rn;
          if (!std::isfinite(rn))
            {
              nonfinite = true;
              break;
            }
          if (rn <= rtol * r0)
            {
              converged = true;
              break;
            }
          if (!speculate && it_done < maxit)
            {
              mg->direction(x);
              x_pending = false;
            }
        }
      if (x_pending)
        launch_cg_add_alpha_p(s, N, x, mg->p, sc);
      rel = rn / r0;
      GTPS_HIP(hipStreamSynchronize(s));
      mg->collect();
    });
    if (iters)
      *iters = it_done;
    if (relres)
      *relres = rel;
    if (rc != GTPS_OK)
      return rc;
    if (nonfinite)
      {
        g_last_error = "non-finite residual in CG";
        return GTPS_ENAN;
      }
    if (!converged)
      {
        g_last_error = "CG reached maxit without convergence";
        return GTPS_ENOTCONV;
      }
    return GTPS_OK;
  }

  // ------------------------------------------------------------- SIPG-DG
  int
  gtps_dg_level_create(gtps_ctx *ctx, int dim, int degree, int n_cells,
                       int precision, double penalty, gtps_dg_level **out)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && out, "NULL argument");
      set_device(ctx);
      *out = make_dg_level(ctx, dim, degree, n_cells, precision, penalty);
    });
  }

  int
  gtps_dg_level_destroy(gtps_dg_level *lvl)
  {
    return guard([&] { delete lvl; });
  }

  long long
  gtps_dg_level_n_dofs(const gtps_dg_level *lvl)
  {
    return lvl ? lvl->N : -1;
  }

  int
  gtps_dg_vmult(gtps_dg_level *lvl, void *dst, const void *src)
  {
    return guard([&] {
      GTPS_REQUIRE(lvl && dst && src && dst != src, "bad argument");
      set_device(lvl->ctx);
      launch_dg_apply(lvl, dst, src, nullptr, false);
    });
  }

  int
  gtps_dg_residual(gtps_dg_level *lvl, void *dst, const void *b, const void *u)
  {
    return guard([&] {
      GTPS_REQUIRE(lvl && dst && b && u && dst != u, "bad argument");
      set_device(lvl->ctx);
      launch_dg_apply(lvl, dst, u, b, false);
    });
  }

  int
  gtps_dg_smoother_step(gtps_dg_level *lvl, void *u, const void *b, int variant,
                        double relaxation)
  {
    return guard([&] {
      GTPS_REQUIRE(lvl && u && b && u != b, "bad argument");
      GTPS_REQUIRE(variant == GTPS_MULT_FWD || variant == GTPS_MULT_REV,
                   "the DG smoother is multiplicative (forward or reverse)");
      GTPS_REQUIRE(dg_supported(lvl->k, lvl->prec),
                   "the DG smoother supports degree <= 5 in FP64 and <= 7 in FP32");
      set_device(lvl->ctx);
      launch_dg_sweep(lvl, u, b, variant, relaxation);
    });
  }

  int
  gtps_dg_prolongate_add(gtps_dg_level *fine_lvl, void *fine, const void *coarse)
  {
    return guard([&] {
      GTPS_REQUIRE(fine_lvl && fine && coarse, "NULL argument");
      set_device(fine_lvl->ctx);
      launch_dg_prolongate_add(fine_lvl, fine, coarse);
    });
  }

  int
  gtps_dg_restrict_add(gtps_dg_level *fine_lvl, void *coarse, const void *fine)
  {
    return guard([&] {
      GTPS_REQUIRE(fine_lvl && fine && coarse, "NULL argument");
      set_device(fine_lvl->ctx);
      launch_dg_restrict(fine_lvl, coarse, fine, false);
    });
  }

  int
  gtps_dg_mg_create(gtps_ctx *ctx, int dim, int degree, int n_coarse,
                    int max_level, int vcycle_precision, double penalty,
                    int coarse_sweeps, gtps_dg_mg **out)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && out, "NULL argument");
      GTPS_REQUIRE(max_level >= 0 && max_level <= 10, "max_level must be in 0..10");
      GTPS_REQUIRE(coarse_sweeps >= 1 && coarse_sweeps <= 64,
                   "coarse_sweeps must be in 1..64");
      GTPS_REQUIRE(dg_supported(degree, vcycle_precision),
                   "the DG smoother supports degree <= 5 in FP64 and <= 7 in FP32");
      GTPS_REQUIRE((long long)n_coarse << max_level <= 4096, "mesh too large");
      set_device(ctx);
      auto mg           = std::make_unique<gtps_dg_mg>();
      mg->ctx           = ctx;
      mg->dim           = dim;
      mg->k             = degree;
      mg->n0            = n_coarse;
      mg->maxl          = max_level;
      mg->prec          = vcycle_precision;
      mg->pf            = penalty;
      mg->coarse_sweeps = coarse_sweeps;
      mg->lv.assign(max_level + 1, nullptr);
      mg->d.assign(max_level + 1, nullptr);
      mg->u.assign(max_level + 1, nullptr);
      mg->t.assign(max_level + 1, nullptr);
      const size_t es = vcycle_precision == GTPS_FP32 ? sizeof(float) : sizeof(double);
      for (int l = 0; l <= max_level; ++l)
        {
          mg->lv[l] = make_dg_level(ctx, dim, degree, n_coarse << l,
                                    vcycle_precision, penalty);
          mg->d[l]  = dmalloc(es * mg->lv[l]->N);
          mg->u[l]  = dmalloc(es * mg->lv[l]->N);
          mg->t[l]  = dmalloc(es * mg->lv[l]->N);
        }
      if (vcycle_precision == GTPS_FP64)
        mg->fine64 = mg->lv[max_level];
      else
        {
          mg->fine64 = make_dg_level(ctx, dim, degree, n_coarse << max_level,
                                     GTPS_FP64, penalty);
          mg->own_fine64 = true;
        }
      const long long N = mg->fine64->N;
      mg->r = (double *)dmalloc(sizeof(double) * N);
      mg->p = (double *)dmalloc(sizeof(double) * N);
      mg->q = (double *)dmalloc(sizeof(double) * N);
      GTPS_HIP(hipEventCreateWithFlags(&mg->ev, hipEventDisableTiming));
      GTPS_HIP(hipStreamSynchronize(ctx->stream));
      *out = mg.release();
    });
  }

  int
  gtps_dg_mg_destroy(gtps_dg_mg *mg)
  {
    return guard([&] {
      if (!mg)
        return;
      (void)hipSetDevice(mg->ctx->device);
      (void)hipStreamSynchronize(mg->ctx->stream);
      delete mg;
    });
  }

  int
  gtps_dg_mg_pcg_solve_batch(gtps_dg_mg *mg, double *x, const double *b, int nrhs,
                             double rtol, int maxit, int *iterations, double *relres)
  {
    return guard([&] {
      GTPS_REQUIRE(mg && x && b && (const void *)x != (const void *)b, "bad argument");
      GTPS_REQUIRE(nrhs >= 1 && nrhs <= 64, "nrhs must be in 1..64");
      GTPS_REQUIRE(maxit >= 0 && rtol >= 0.0, "bad rtol / maxit");
      set_device(mg->ctx);
      mg->ensure_batch(nrhs);
      mg->ensure_dense0();
      hipStream_t     s   = mg->ctx->stream;
      const long long N   = mg->fine64->N;
      const bool      f32 = mg->prec == GTPS_FP32;
      double         *pt  = mg->ctx->ws.partials;
      const int       cap = mg->ctx->ws.capacity;
      launch_fill(s, N * nrhs, GTPS_FP64, x, 0.0);
      GTPS_HIP(hipMemcpyAsync(mg->r, b, sizeof(double) * N * nrhs, hipMemcpyDeviceToDevice, s));
      for (int j = 0; j < nrhs; ++j)
        {
          double *rj  = mg->r + N * j;
          const int nbk = launch_dot_partials(s, N, rj, rj, nullptr, pt, cap);
          launch_finalize_sum(s, pt, nbk, mg->bsc + cg::N_SLOTS * j, cg::RR);
          launch_cg_scalar(s, mg->bsc + cg::N_SLOTS * j, cg::OP_INIT);
        }
      if (f32)
        launch_convert_f64_f32(s, N * nrhs, mg->r, (float *)mg->d[mg->maxl]);
      launch_fill(s, N * nrhs, GTPS_FP64, mg->p, 0.0);
      std::vector<double> r0(nrhs), rn(nrhs);
      std::vector<int>    itc(nrhs, -1);
      mg->read_norms(nrhs, r0.data());
      for (int j = 0; j < nrhs; ++j)
        {
          if (!std::isfinite(r0[j]))
            throw Error(GTPS_ENAN, "non-finite right-hand side in batched CG");
          if (r0[j] == 0.0)
            itc[j] = 0;
        }
      auto all_done = [&]() {
        for (int j = 0; j < nrhs; ++j)
          if (itc[j] < 0)
            return false;
        return true;
      };
      // a row stops changing once it has converged (or had a zero right-hand
      // side): its x, r, p and CG scalars are left as they are, so each row
      // is the single-vector solve of that row; the batched kernels still run
      // over every row
      int it = 0;
      if (!all_done())
        mg->direction_batch(nrhs, itc);
      while (!all_done() && it < maxit)
        {
          const int nbq = launch_dg_apply(mg->fine64, mg->q, mg->p, nullptr, true, nrhs);
          for (int j = 0; j < nrhs; ++j)
            if (itc[j] < 0)
              launch_finalize_op(s, pt + (long long)nbq * j, nbq, mg->bsc + cg::N_SLOTS * j,
                                 cg::PQ, cg::OP_ALPHA);
          for (int j = 0; j < nrhs; ++j)
            {
              if (itc[j] >= 0)
                continue;
              float *r32 = f32 ? (float *)mg->d[mg->maxl] + N * j : nullptr;
              const int nbr =
                launch_cg_update_xr(s, N, x + N * j, mg->r + N * j, mg->p + N * j,
                                    mg->q + N * j, r32, mg->bsc + cg::N_SLOTS * j, pt, cap);
              launch_finalize_sum(s, pt, nbr, mg->bsc + cg::N_SLOTS * j, cg::RR);
            }
          ++it;
          mg->read_norms(nrhs, rn.data());
          for (int j = 0; j < nrhs; ++j)
            {
              if (!std::isfinite(rn[j]))
                throw Error(GTPS_ENAN, "non-finite residual in batched CG");
              if (itc[j] < 0 && rn[j] <= rtol * r0[j])
                itc[j] = it;
            }
          if (!all_done() && it < maxit)
            mg->direction_batch(nrhs, itc);
        }
      for (int j = 0; j < nrhs; ++j)
        {
          if (iterations)
            iterations[j] = itc[j] < 0 ? it : itc[j];
          if (relres)
            relres[j] = r0[j] > 0.0 ? (it > 0 ? rn[j] : r0[j]) / r0[j] : 0.0;
        }
      GTPS_HIP(hipStreamSynchronize(s));
      if (!all_done())
        throw Error(GTPS_ENOTCONV, "batched CG reached maxit without convergence");
    });
  }

  int
  gtps_dg_mg_set_graph(gtps_dg_mg *mg, int enable)
  {
    return guard([&] {
      GTPS_REQUIRE(mg, "NULL argument");
      mg->use_graph = enable != 0;
    });
  }

  long long
  gtps_dg_mg_n_dofs(const gtps_dg_mg *mg)
  {
    return mg ? mg->fine64->N : -1;
  }

  int
  gtps_dg_mg_matvec(gtps_dg_mg *mg, double *dst, const double *src)
  {
    return guard([&] {
      GTPS_REQUIRE(mg && dst && src && (const void *)dst != (const void *)src,
                   "bad argument");
      set_device(mg->ctx);
      launch_dg_apply(mg->fine64, dst, src, nullptr, false);
    });
  }

  int
  gtps_dg_mg_vcycle(gtps_dg_mg *mg, double *z, const double *r)
  {
    return guard([&] {
      GTPS_REQUIRE(mg && z && r && (const void *)z != (const void *)r, "bad argument");
      set_device(mg->ctx);
      mg->ensure_dense0();
      const long long N = mg->fine64->N;
      if (mg->prec == GTPS_FP32)
        {
          launch_convert_f64_f32(mg->ctx->stream, N, r, (float *)mg->d[mg->maxl]);
          mg->vcycle(mg->maxl, mg->d[mg->maxl], mg->u[mg->maxl]);
          launch_convert_f32_f64(mg->ctx->stream, N, (const float *)mg->u[mg->maxl], z);
        }
      else
        mg->vcycle(mg->maxl, r, z);
    });
  }

  int
  gtps_dg_mg_pcg_solve(gtps_dg_mg *mg, double *x, const double *b, double rtol,
                       int maxit, int *iters, double *relres, double *history)
  {
    int    it_done = 0;
    double rel     = 0.0;
    bool   converged = false, nonfinite = false;
    int rc = guard([&] {
      GTPS_REQUIRE(mg && x && b && (const void *)x != (const void *)b, "bad argument");
      GTPS_REQUIRE(maxit >= 0 && rtol >= 0.0, "bad rtol / maxit");
      set_device(mg->ctx);
      gtps_ctx       *ctx = mg->ctx;
      hipStream_t     s   = ctx->stream;
      const long long N   = mg->fine64->N;
      double         *sc  = ctx->ws.scalars;
      double         *pt  = ctx->ws.partials;
      const int       cap = ctx->ws.capacity;
      const bool      f32 = mg->prec == GTPS_FP32;
      float          *r32 = f32 ? (float *)mg->d[mg->maxl] : nullptr;
      // residual norms reach the host through mapped memory written by
      // their finalize kernel (see gtps_mg_pcg_solve)
      double *rb      = ctx->ws.readback_dev;
      auto    read_rr = [&](int nb) {
        launch_finalize_sum(s, pt, nb, sc, cg::RR, rb);
        GTPS_HIP(hipEventRecord(mg->ev, s));
      };
      auto wait_rr = [&]() {
        GTPS_HIP(hipEventSynchronize(mg->ev));
        return std::sqrt(ctx->ws.readback[0]);
      };
      // x0 = 0, r0 = b (no Dirichlet rows: boundary conditions are weak),
      // p = 0, rz = 1 in one pass; the first preconditioning step is
      // enqueued before the host learns r0
      read_rr(launch_cg_init(s, N, b, x, mg->r, r32, mg->p, sc, pt, cap));
      if (maxit > 0)
        mg->direction();
      const double r0 = wait_rr();
      if (history)
        history[0] = r0;
      if (!std::isfinite(r0))
        {
          nonfinite = true;
          GTPS_HIP(hipStreamSynchronize(s));
          return;
        }
      if (r0 == 0.0)
        {
          converged = true;
          GTPS_HIP(hipStreamSynchronize(s));
          return;
        }
      double rn = r0, rn_prev = -1.0;
      while (it_done < maxit)
        {
          const int nbq = launch_dg_apply(mg->fine64, mg->q, mg->p, nullptr, true);
          launch_finalize_op(s, pt, nbq, sc, cg::PQ, cg::OP_ALPHA);
          read_rr(launch_cg_update_xr(s, N, x, mg->r, mg->p, mg->q, r32, sc, pt, cap));
          ++it_done;
          bool speculate = it_done < maxit;
          if (speculate && rn_prev > 0.0 && rn < rn_prev)
            speculate = rn * (rn / rn_prev) > 0.25 * rtol * r0;
          if (speculate)
            mg->direction();
          rn_prev = rn;
          rn      = wait_rr();
          if (history)
            history[it_done] = rn;
          if (!std::isfinite(rn))
            {
              nonfinite = true;
              break;
            }
          if (rn <= rtol * r0)
            {
              converged = true;
              break;
            }
          if (!speculate && it_done < maxit)
            mg->direction();
        }
      rel = rn / r0;
      GTPS_HIP(hipStreamSynchronize(s));
    });
    if (iters)
      *iters = it_done;
    if (relres)
      *relres = rel;
    if (rc != GTPS_OK)
      return rc;
    if (nonfinite)
      {
        g_last_error = "non-finite residual in CG";
        return GTPS_ENAN;
      }
    if (!converged)
      {
        g_last_error = "CG reached maxit without convergence";
        return GTPS_ENOTCONV;
      }
    return GTPS_OK;
  }

  // ------------------------------------------------------- post-processing
  int
  gtps_assemble_rhs(gtps_ctx *ctx, int dim, int degree, int level,
                    const double *f_q, double *b)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && f_q && b, "NULL argument");
      GTPS_REQUIRE(level >= 0 && level <= 12, "level must be in 0..12");
      set_device(ctx);
      assemble_rhs(ctx, dim, degree, 1 << level, false, 0, f_q, 0.0, 0.0, b);
    });
  }

  int
  gtps_rhs_sine(gtps_ctx *ctx, int dim, int degree, int level, double wave, double *b)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && b, "NULL argument");
      GTPS_REQUIRE(level >= 0 && level <= 12, "level must be in 0..12");
      GTPS_REQUIRE(std::isfinite(wave), "wave number must be finite");
      set_device(ctx);
      assemble_rhs(ctx, dim, degree, 1 << level, false, 1, nullptr, wave, 0.0, b);
    });
  }

  int
  gtps_rhs_constant(gtps_ctx *ctx, int dim, int degree, int level, double value,
                    double *b)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && b, "NULL argument");
      GTPS_REQUIRE(level >= 0 && level <= 12, "level must be in 0..12");
      GTPS_REQUIRE(std::isfinite(value), "value must be finite");
      set_device(ctx);
      assemble_rhs(ctx, dim, degree, 1 << level, false, 2, nullptr, 0.0, value, b);
    });
  }

  int
  gtps_dg_assemble_rhs(gtps_ctx *ctx, int dim, int degree, int n_cells,
                       const double *f_q, double *b)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && f_q && b, "NULL argument");
      set_device(ctx);
      assemble_rhs(ctx, dim, degree, n_cells, true, 0, f_q, 0.0, 0.0, b);
    });
  }

  int
  gtps_dg_rhs_sine(gtps_ctx *ctx, int dim, int degree, int n_cells, double wave,
                   double *b)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && b, "NULL argument");
      GTPS_REQUIRE(std::isfinite(wave), "wave number must be finite");
      set_device(ctx);
      assemble_rhs(ctx, dim, degree, n_cells, true, 1, nullptr, wave, 0.0, b);
    });
  }

  int
  gtps_error_norms(gtps_ctx *ctx, int dim, int degree, int level, double wave,
                   int n_q, const double *u, double *l2, double *h1)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && u && l2 && h1, "NULL argument");
      GTPS_REQUIRE(level >= 0 && level <= 12, "level must be in 0..12");
      set_device(ctx);
      compute_error_norms(ctx, dim, degree, 1 << level, false, wave, n_q, u, l2, h1);
    });
  }

  int
  gtps_dg_error_norms(gtps_ctx *ctx, int dim, int degree, int n_cells, double wave,
                      int n_q, const double *u, double *l2, double *h1)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && u && l2 && h1, "NULL argument");
      set_device(ctx);
      compute_error_norms(ctx, dim, degree, n_cells, true, wave, n_q, u, l2, h1);
    });
  }

  int
  gtps_vector_save(gtps_ctx *ctx, const char *path, const void *x, int precision,
                   int dim, int degree, int n_cells, int layout)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && path && x, "NULL argument");
      VecHeader hd = make_vec_header(precision, dim, degree, n_cells, layout);
      const size_t bytes = (size_t)hd.n * (precision == GTPS_FP32 ? 4 : 8);
      std::vector<unsigned char> data(bytes);
      set_device(ctx);
      GTPS_HIP(hipMemcpyAsync(data.data(), x, bytes, hipMemcpyDeviceToHost, ctx->stream));
      GTPS_HIP(hipStreamSynchronize(ctx->stream));
      hd.hash = fnv1a(data.data(), bytes);
      FILE *f = std::fopen(path, "wb");
      GTPS_REQUIRE(f != nullptr, std::string("cannot open for writing: ") + path);
      const bool ok = std::fwrite(&hd, sizeof hd, 1, f) == 1 &&
                      (bytes == 0 || std::fwrite(data.data(), bytes, 1, f) == 1);
      const bool closed = std::fclose(f) == 0;
      GTPS_REQUIRE(ok && closed, std::string("write failed: ") + path);
    });
  }

  int
  gtps_vector_info(const char *path, int *precision, int *dim, int *degree,
                   int *n_cells, int *layout, long long *n)
  {
    return guard([&] {
      GTPS_REQUIRE(path != nullptr, "NULL path");
      const VecHeader hd = read_vec_header(path, nullptr);
      if (precision) *precision = hd.precision;
      if (dim) *dim = hd.dim;
      if (degree) *degree = hd.degree;
      if (n_cells) *n_cells = hd.n_cells;
      if (layout) *layout = hd.layout;
      if (n) *n = hd.n;
    });
  }

  int
  gtps_vector_load(gtps_ctx *ctx, const char *path, void *x, long long n,
                   int precision)
  {
    return guard([&] {
      GTPS_REQUIRE(ctx && path && x, "NULL argument");
      std::vector<unsigned char> data;
      const VecHeader hd = read_vec_header(path, &data);
      GTPS_REQUIRE(hd.n == n, "file holds a vector of another size");
      GTPS_REQUIRE(hd.precision == precision, "file holds another precision");
      set_device(ctx);
      GTPS_HIP(hipMemcpyAsync(x, data.data(), data.size(), hipMemcpyHostToDevice,
                              ctx->stream));
      GTPS_HIP(hipStreamSynchronize(ctx->stream));
    });
  }
}
