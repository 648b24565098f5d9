"""Host-side mirror of the reference's hot-path API over the C ABI.

Class and method names follow the reference (PSMF namespace):

* :class:`LaplaceOperator`  - include/laplace_operator.cuh:71-205
* :class:`PatchSmoother`    - include/patch_smoother.cuh:86-124, 313-335
* :class:`MGTransferCUDA`   - include/cuda_mg_transfer.cuh:57-358
* :class:`MultigridSolvers` - include/solver.cuh:1541-2210
* :func:`rhs_sine`, :func:`assemble_rhs` - LaplaceOperator::compute_residual
  (include/laplace_operator.cuh:119-173) with apps/poisson.cu:76-87
* :func:`error_norms` - LaplaceProblem::compute_error (apps/poisson.cu:353-382)
* :class:`DGOperator`, :class:`DGMultigrid` - SIPG-DG (tests/mf_dg.cu:63-274)

Vectors are 1D ``torch`` tensors on the context's device (float32 for FP32
objects, float64 for FP64), lexicographic with boundary DoFs included.  All
work is enqueued on the context stream, which defaults to torch's current
stream so that torch and gtps operations are ordered.  Errors raise
:class:`GtpsError` (the reference raises deal.II ExcMessage / AssertCuda).
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib as L

FP32, FP64 = L.GTPS_FP32, L.GTPS_FP64

_DTYPES = {FP32: torch.float32, FP64: torch.float64}


def _prec(dtype) -> int:
    if dtype in (FP32, torch.float32, "float32", "fp32", "float"):
        return FP32
    if dtype in (FP64, torch.float64, "float64", "fp64", "double"):
        return FP64
    raise ValueError(f"unsupported precision {dtype!r}")


def default_relaxation(dim: int, schedule: int) -> float:
    """Damping used when none is given: 1 for the multiplicative schedules
    (relaxation = 1 at every call site of the reference,
    include/solver.cuh:1719-1737); for the additive schedule 0.4 in 2D and
    0.2 in 3D, since every cell-interior DoF lies in 2^dim overlapping
    vertex-patch interiors (the CG iteration counts behind this choice are
    recorded with the additive tests)."""
    if schedule == L.GTPS_MG_ADDITIVE:
        return 0.4 if dim == 2 else 0.2
    return 1.0


def get_tuning(key: str) -> int:
    """Current value of a tuning knob (see gtps_set_tuning)."""
    v = L.load().gtps_get_tuning(key.encode())
    if v < 0:
        L.check(L.GTPS_EINVAL, "get_tuning")
    return int(v)


def set_tuning(key: str, value: int) -> int:
    """Set a process-wide tuning knob (see gtps_set_tuning in
    include/gtps.h); returns the previous value."""
    prev = L.load().gtps_set_tuning(key.encode(), int(value))
    if prev < 0:
        L.check(L.GTPS_EINVAL, "set_tuning")
    return int(prev)


class Context:
    """Device + HIP stream shared by all objects built on it."""

    def __init__(self, device: int = 0, stream=None):
        lib = L.load()
        self.device = device
        if stream is None and torch.cuda.is_available():
            stream = torch.cuda.current_stream(device).cuda_stream
        h = ctypes.c_void_p()
        L.check(lib.gtps_ctx_create(device, ctypes.c_void_p(stream or 0), ctypes.byref(h)),
                "ctx_create")
        self._h = h

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        L.check(L.load().gtps_ctx_synchronize(self._h), "synchronize")

    def close(self):
        if getattr(self, "_h", None):
            L.load().gtps_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # FP64 vector operations (deal.II Vector<double,CUDA> counterparts)
    def dot(self, x: torch.Tensor, y: torch.Tensor) -> float:
        _chk_vec(x, torch.float64)
        _chk_vec(y, torch.float64, x.numel())
        out = ctypes.c_double()
        L.check(L.load().gtps_vec_dot(self._h, x.numel(), _ptr(x), _ptr(y), ctypes.byref(out)),
                "dot")
        return out.value

    def l2_norm(self, x: torch.Tensor) -> float:
        _chk_vec(x, torch.float64)
        out = ctypes.c_double()
        L.check(L.load().gtps_vec_nrm2(self._h, x.numel(), _ptr(x), ctypes.byref(out)), "nrm2")
        return out.value

    def sadd(self, s: float, a: float, x: torch.Tensor, y: torch.Tensor):
        """y = s*y + a*x (deal.II Vector::sadd(s, a, x))."""
        _chk_vec(x, torch.float64)
        _chk_vec(y, torch.float64, x.numel())
        L.check(L.load().gtps_vec_axpby(self._h, x.numel(), a, _ptr(x), s, _ptr(y)), "axpby")

    def convert(self, src: torch.Tensor, dst: torch.Tensor):
        """copy_vector: FP64 -> FP32 (round to nearest) or FP32 -> FP64."""
        lib = L.load()
        if src.dtype == torch.float64 and dst.dtype == torch.float32:
            rc = lib.gtps_vec_convert_f64_f32(self._h, src.numel(), _ptr(src), _ptr(dst))
        elif src.dtype == torch.float32 and dst.dtype == torch.float64:
            rc = lib.gtps_vec_convert_f32_f64(self._h, src.numel(), _ptr(src), _ptr(dst))
        else:
            raise ValueError("convert needs one float32 and one float64 tensor")
        _chk_vec(dst, dst.dtype, src.numel())
        L.check(rc, "convert")


def _ptr(t: torch.Tensor):
    return ctypes.c_void_p(t.data_ptr())


def _chk_vec(t: torch.Tensor, dtype, n: int | None = None):
    if not isinstance(t, torch.Tensor):
        raise TypeError("expected a torch.Tensor")
    if t.dtype != dtype:
        raise TypeError(f"expected {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError("vector must be contiguous")
    if t.device.type != "cuda":
        raise ValueError("vector must live on the GPU")
    if n is not None and t.numel() != n:
        raise ValueError(f"vector has {t.numel()} entries, expected {n}")


class LaplaceOperator:
    """Matrix-free Laplace operator of one level (LaplaceOperator<dim,k,Number>)."""

    def __init__(self, ctx: Context, dim: int, fe_degree: int, level: int, dtype=FP64):
        self.ctx, self.dim, self.fe_degree, self.level = ctx, dim, fe_degree, level
        self.prec = _prec(dtype)
        self.dtype = _DTYPES[self.prec]
        h = ctypes.c_void_p()
        L.check(L.load().gtps_level_create(ctx.handle, dim, fe_degree, level, self.prec,
                                           ctypes.byref(h)), "level_create")
        self._h = h
        self.n_dofs = int(L.load().gtps_level_n_dofs(h))
        self.n_dofs_1d = fe_degree * (1 << level) + 1

    @property
    def handle(self):
        return self._h

    def initialize_dof_vector(self) -> torch.Tensor:
        return torch.zeros(self.n_dofs, dtype=self.dtype, device=f"cuda:{self.ctx.device}")

    def vmult(self, dst: torch.Tensor, src: torch.Tensor):
        _chk_vec(dst, self.dtype, self.n_dofs)
        _chk_vec(src, self.dtype, self.n_dofs)
        L.check(L.load().gtps_laplace_vmult(self._h, _ptr(dst), _ptr(src)), "vmult")

    Tvmult = vmult

    def residual(self, dst: torch.Tensor, b: torch.Tensor, u: torch.Tensor):
        """dst = b - A u."""
        for t in (dst, b, u):
            _chk_vec(t, self.dtype, self.n_dofs)
        L.check(L.load().gtps_laplace_residual(self._h, _ptr(dst), _ptr(b), _ptr(u)),
                "residual")

    @property
    def n_cells(self) -> int:
        return (1 << self.level) ** self.dim

    def set_coefficient(self, coef: torch.Tensor | None):
        """Per-cell coefficient c (FP64 on the GPU, cells cx + n cy + n^2 cz)
        for -div(c grad u); None restores c = 1 (3D only)."""
        if coef is None:
            L.check(L.load().gtps_level_set_coefficient(self._h, None), "set_coefficient")
            return
        _chk_vec(coef, torch.float64, self.n_cells)
        L.check(L.load().gtps_level_set_coefficient(self._h, _ptr(coef)), "set_coefficient")

    def vmult_transpose(self, dst: torch.Tensor, src: torch.Tensor):
        """dst = A^T src (identity Dirichlet rows of A become A_IB^T columns)."""
        _chk_vec(dst, self.dtype, self.n_dofs)
        _chk_vec(src, self.dtype, self.n_dofs)
        L.check(L.load().gtps_laplace_vmult_transpose(self._h, _ptr(dst), _ptr(src)),
                "vmult_transpose")

    def coefficient_gradient(self, grad: torch.Tensor, lam: torch.Tensor, u: torch.Tensor):
        """grad_K = -int_K grad(lambda~) . grad(u) (lambda's Dirichlet entries
        masked)."""
        _chk_vec(grad, self.dtype, self.n_cells)
        _chk_vec(lam, self.dtype, self.n_dofs)
        _chk_vec(u, self.dtype, self.n_dofs)
        L.check(L.load().gtps_level_coefficient_gradient(self._h, _ptr(grad), _ptr(lam),
                                                         _ptr(u)), "coefficient_gradient")

    def set_constrained_values(self, v: torch.Tensor, value: float = 0.0):
        _chk_vec(v, self.dtype, self.n_dofs)
        L.check(L.load().gtps_vec_set_boundary(self._h, _ptr(v), value), "set_boundary")

    def tensor_product_data(self) -> dict:
        """FP64 1D data: cell/patch mass and Laplace, GPU eigenpairs."""
        import numpy as np
        k = self.fe_degree
        n, np_, ni = k + 1, 2 * k + 1, 2 * k - 1
        out = dict(cell_mass=np.zeros((n, n)), cell_laplace=np.zeros((n, n)),
                   patch_mass=np.zeros((np_, np_)), patch_laplace=np.zeros((np_, np_)),
                   eigenvalues=np.zeros(ni), eigenvectors=np.zeros((ni, ni)))
        p = [out[key].ctypes.data_as(L._pd) for key in
             ("cell_mass", "cell_laplace", "patch_mass", "patch_laplace",
              "eigenvalues", "eigenvectors")]
        L.check(L.load().gtps_level_tensor_product_data(self._h, *p), "tensor_product_data")
        return out

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                L.load().gtps_level_destroy(self._h)
                self._h = None
        except Exception:
            pass


class PatchSmoother:
    """Vertex-patch Schwarz smoother with FD local solves (PatchSmoother)."""

    MULTIPLICATIVE = L.GTPS_MULT_FWD
    MULTIPLICATIVE_REVERSE = L.GTPS_MULT_REV
    ADDITIVE = L.GTPS_ADDITIVE

    def __init__(self, op: LaplaceOperator, variant: int = L.GTPS_MULT_FWD,
                 relaxation: float = 1.0):
        self.op, self.variant, self.relaxation = op, variant, relaxation
        h = ctypes.c_void_p()
        L.check(L.load().gtps_smoother_create(op.handle, variant, relaxation, ctypes.byref(h)),
                "smoother_create")
        self._h = h

    def vmult(self, dst: torch.Tensor, src: torch.Tensor):
        """dst = 0, then one sweep."""
        _chk_vec(dst, self.op.dtype, self.op.n_dofs)
        _chk_vec(src, self.op.dtype, self.op.n_dofs)
        L.check(L.load().gtps_smoother_vmult(self._h, _ptr(dst), _ptr(src)), "smoother_vmult")

    def step(self, dst: torch.Tensor, src: torch.Tensor):
        """One sweep in place on dst."""
        _chk_vec(dst, self.op.dtype, self.op.n_dofs)
        _chk_vec(src, self.op.dtype, self.op.n_dofs)
        L.check(L.load().gtps_smoother_step(self._h, _ptr(dst), _ptr(src)), "smoother_step")

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                L.load().gtps_smoother_destroy(self._h)
                self._h = None
        except Exception:
            pass


class MGTransferCUDA:
    """Level transfer (MGTransferCUDA<dim,Number,Q>)."""

    def __init__(self, ctx: Context, dim: int, fe_degree: int, min_level: int,
                 max_level: int, dtype=FP64):
        self.ctx, self.dim, self.fe_degree = ctx, dim, fe_degree
        self.min_level, self.max_level = min_level, max_level
        self.prec = _prec(dtype)
        self.dtype = _DTYPES[self.prec]
        h = ctypes.c_void_p()
        L.check(L.load().gtps_transfer_create(ctx.handle, dim, fe_degree, min_level, max_level,
                                              self.prec, ctypes.byref(h)), "transfer_create")
        self._h = h

    def _n(self, level):
        n1 = self.fe_degree * (1 << level) + 1
        return n1 ** self.dim

    def prolongate(self, to_level: int, dst: torch.Tensor, src: torch.Tensor):
        _chk_vec(dst, self.dtype, self._n(to_level))
        _chk_vec(src, self.dtype, self._n(to_level - 1))
        L.check(L.load().gtps_transfer_prolongate(self._h, to_level, _ptr(dst), _ptr(src)),
                "prolongate")

    def prolongate_and_add(self, to_level: int, dst: torch.Tensor, src: torch.Tensor):
        _chk_vec(dst, self.dtype, self._n(to_level))
        _chk_vec(src, self.dtype, self._n(to_level - 1))
        L.check(L.load().gtps_transfer_prolongate_add(self._h, to_level, _ptr(dst), _ptr(src)),
                "prolongate_add")

    def restrict_and_add(self, from_level: int, dst: torch.Tensor, src: torch.Tensor):
        _chk_vec(dst, self.dtype, self._n(from_level - 1))
        _chk_vec(src, self.dtype, self._n(from_level))
        L.check(L.load().gtps_transfer_restrict_add(self._h, from_level, _ptr(dst), _ptr(src)),
                "restrict_add")

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                L.load().gtps_transfer_destroy(self._h)
                self._h = None
        except Exception:
            pass


class MultigridSolvers:
    """MG hierarchy (levels 1..max_level) with the FP64 MG-preconditioned CG.

    ``vcycle_dtype`` float32 gives the mixed-precision form (FP32 smoother,
    level operators and transfers under FP64 CG); float64 gives FP64
    everywhere.  ``schedule`` selects the smoothing order (symmetric by
    default, required for CG).
    """

    SYMMETRIC = L.GTPS_MG_SYMMETRIC
    REFERENCE_ORDER = L.GTPS_MG_REFERENCE_ORDER
    ADDITIVE = L.GTPS_MG_ADDITIVE

    def __init__(self, ctx: Context, dim: int, fe_degree: int, max_level: int,
                 vcycle_dtype=FP32, schedule: int = L.GTPS_MG_SYMMETRIC,
                 relaxation: float | None = None):
        self.ctx, self.dim, self.fe_degree, self.max_level = ctx, dim, fe_degree, max_level
        self.prec = _prec(vcycle_dtype)
        if relaxation is None:
            relaxation = default_relaxation(dim, schedule)
        self.relaxation = relaxation
        h = ctypes.c_void_p()
        L.check(L.load().gtps_mg_create(ctx.handle, dim, fe_degree, max_level, self.prec,
                                        schedule, relaxation, ctypes.byref(h)), "mg_create")
        self._h = h
        self.n_dofs = int(L.load().gtps_mg_n_dofs(h))
        self._coef = None
        self._bmask = None

    def vmult(self, dst: torch.Tensor, src: torch.Tensor):
        """Preconditioner z = V(r) (MultigridSolvers::vmult)."""
        _chk_vec(dst, torch.float64, self.n_dofs)
        _chk_vec(src, torch.float64, self.n_dofs)
        L.check(L.load().gtps_mg_vcycle(self._h, _ptr(dst), _ptr(src)), "vcycle")

    def do_matvec(self, dst: torch.Tensor, src: torch.Tensor):
        _chk_vec(dst, torch.float64, self.n_dofs)
        _chk_vec(src, torch.float64, self.n_dofs)
        L.check(L.load().gtps_mg_matvec(self._h, _ptr(dst), _ptr(src)), "matvec")

    def solve_cg(self, b: torch.Tensor, x: torch.Tensor | None = None, rtol: float = 1e-8,
                 maxit: int = 100, raise_on_failure: bool = True):
        """Solve A x = b; returns (x, iterations, relative residual, history)."""
        _chk_vec(b, torch.float64, self.n_dofs)
        if x is None:
            x = torch.empty_like(b)
        _chk_vec(x, torch.float64, self.n_dofs)
        it = ctypes.c_int()
        rel = ctypes.c_double()
        hist = (ctypes.c_double * (maxit + 1))()
        rc = L.load().gtps_mg_pcg_solve(self._h, _ptr(x), _ptr(b), rtol, maxit,
                                        ctypes.byref(it), ctypes.byref(rel), hist)
        if rc != L.GTPS_OK and (raise_on_failure or rc not in (L.GTPS_ENOTCONV,)):
            L.check(rc, "pcg_solve")
        return x, it.value, rel.value, list(hist[: it.value + 1])

    @property
    def n_cells(self) -> int:
        return (1 << self.max_level) ** self.dim

    def set_coefficient(self, coef: torch.Tensor | None):
        """Per-cell coefficient of the finest level (FP64 on the GPU); coarse
        levels get the mean of their children.  None restores c = 1."""
        if coef is None:
            L.check(L.load().gtps_mg_set_coefficient(self._h, None), "set_coefficient")
            self._coef = None
            return
        _chk_vec(coef, torch.float64, self.n_cells)
        L.check(L.load().gtps_mg_set_coefficient(self._h, _ptr(coef)), "set_coefficient")
        self._coef = coef

    def matvec_transpose(self, dst: torch.Tensor, src: torch.Tensor):
        _chk_vec(dst, torch.float64, self.n_dofs)
        _chk_vec(src, torch.float64, self.n_dofs)
        L.check(L.load().gtps_mg_matvec_transpose(self._h, _ptr(dst), _ptr(src)),
                "matvec_transpose")

    def coefficient_gradient(self, grad: torch.Tensor, lam: torch.Tensor, u: torch.Tensor):
        _chk_vec(grad, torch.float64, self.n_cells)
        _chk_vec(lam, torch.float64, self.n_dofs)
        _chk_vec(u, torch.float64, self.n_dofs)
        L.check(L.load().gtps_mg_coefficient_gradient(self._h, _ptr(grad), _ptr(lam), _ptr(u)),
                "coefficient_gradient")

    def boundary_mask(self) -> torch.Tensor:
        """Boolean mask of the Dirichlet DoFs of the finest level."""
        if getattr(self, "_bmask", None) is None:
            n1 = self.fe_degree * (1 << self.max_level) + 1
            i = torch.arange(self.n_dofs, device=f"cuda:{self.ctx.device}")
            x, y = i % n1, (i // n1) % n1
            m = (x == 0) | (x == n1 - 1) | (y == 0) | (y == n1 - 1)
            if self.dim == 3:
                z = i // (n1 * n1)
                m |= (z == 0) | (z == n1 - 1)
            self._bmask = m
        return self._bmask

    def solve(self, b: torch.Tensor, coef: torch.Tensor | None = None, rtol: float = 1e-10,
              maxit: int = 200) -> torch.Tensor:
        """Differentiable solve u = A(c)^{-1} b (torch.autograd): gradients
        flow to b and, when given, to the per-cell coefficient c."""
        return MGSolve.apply(self, b, coef, rtol, maxit)

    def solve_fmg(self, rhs_levels, x: torch.Tensor | None = None,
                  n_cycles: int = 1) -> torch.Tensor:
        """Full multigrid (MultigridSolvers::solve): rhs_levels[l-1] is the
        FP64 right-hand side of level l = 1..max_level."""
        if len(rhs_levels) != self.max_level:
            raise ValueError(f"need {self.max_level} right-hand sides (levels 1..max)")
        for l, r in enumerate(rhs_levels, start=1):
            _chk_vec(r, torch.float64, (self.fe_degree * (1 << l) + 1) ** self.dim)
        if x is None:
            x = torch.empty_like(rhs_levels[-1])
        _chk_vec(x, torch.float64, self.n_dofs)
        arr = (ctypes.c_void_p * self.max_level)(*[r.data_ptr() for r in rhs_levels])
        L.check(L.load().gtps_mg_fmg_solve(self._h, _ptr(x), arr, n_cycles), "fmg_solve")
        return x

    def set_graph(self, enable: bool = True):
        """Replay the per-iteration V-cycle sequence as one HIP graph
        (default) or launch its kernels one by one."""
        L.check(L.load().gtps_mg_set_graph(self._h, 1 if enable else 0), "set_graph")

    def set_profiling(self, groups=("smooth_fine",)):
        """Enable HIP-event stage timers for the named groups (see
        include/gtps.h GTPS_PROF_*); an empty sequence disables them."""
        mask = 0
        for g in groups:
            mask |= 1 << L.PROF_GROUPS[g]
        L.check(L.load().gtps_mg_set_profiling(self._h, mask), "set_profiling")

    def profile(self) -> dict:
        """{group: (intervals, total_ms)} since the last reset."""
        out = {}
        for name, g in L.PROF_GROUPS.items():
            c = ctypes.c_longlong()
            ms = ctypes.c_double()
            L.check(L.load().gtps_mg_profile(self._h, g, ctypes.byref(c), ctypes.byref(ms)),
                    "profile")
            out[name] = (c.value, ms.value)
        return out

    def profile_reset(self):
        L.check(L.load().gtps_mg_profile_reset(self._h), "profile_reset")

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                L.load().gtps_mg_destroy(self._h)
                self._h = None
        except Exception:
            pass


class MGSolve(torch.autograd.Function):
    """u = A(c)^{-1} b through the MG-preconditioned CG, differentiable in b
    and in the per-cell coefficient c (c = None: the solver's current
    operator, which must not change before the backward pass).

    A has identity Dirichlet rows and keeps its boundary columns, so with
    I/B the interior/Dirichlet DoFs, A^T lambda = g splits into
    A_II lambda_I = g_I (solved by the same symmetric MG-CG on a right-hand
    side with zero Dirichlet entries) and lambda_B = g_B - A_IB^T lambda_I
    (one transpose product).  Then dL/db = lambda and
    dL/dc_K = -lambda~_K^T A^1_K u_K with lambda~ = lambda_I (zero on B),
    computed by the cell-wise gradient kernel.
    """

    @staticmethod
    def forward(ctx, solver: "MultigridSolvers", b: torch.Tensor, coef, rtol: float,
                maxit: int):
        c = None
        if coef is not None:
            c = coef.detach().contiguous()
            solver.set_coefficient(c)
        u, it, rel, _ = solver.solve_cg(b.detach().contiguous(), rtol=rtol, maxit=maxit)
        ctx.solver, ctx.rtol, ctx.maxit = solver, rtol, maxit
        ctx.has_coef = c is not None
        ctx.forward_info = (it, rel)
        ctx.save_for_backward(u, c)
        return u

    @staticmethod
    def backward(ctx, grad_u: torch.Tensor):
        solver = ctx.solver
        u, coef = ctx.saved_tensors
        if ctx.has_coef and (solver._coef is None
                             or solver._coef.data_ptr() != coef.data_ptr()):
            solver.set_coefficient(coef.contiguous())
        g = grad_u.detach().to(torch.float64).contiguous()
        mask = solver.boundary_mask()
        rhs = torch.where(mask, torch.zeros_like(g), g)
        lam, _, _, _ = solver.solve_cg(rhs, rtol=ctx.rtol, maxit=ctx.maxit)
        grad_b = None
        if ctx.needs_input_grad[1]:
            t = torch.empty_like(lam)
            solver.matvec_transpose(t, lam)
            grad_b = torch.where(mask, g - t, lam)
        grad_c = None
        if ctx.has_coef and ctx.needs_input_grad[2]:
            grad_c = torch.empty(solver.n_cells, dtype=torch.float64, device=g.device)
            solver.coefficient_gradient(grad_c, lam, u)
        return None, grad_b, grad_c, None, None


# ------------------------------------------------------------------ SIPG-DG
def dg_penalty(fe_degree: int) -> float:
    """Default interior-penalty factor 2.5 (k+1)^2 (sigma = factor / h)."""
    return 2.5 * (fe_degree + 1) ** 2


class DGOperator:
    """SIPG-DG Laplacian of one level: n_cells per direction, FE_DGQ(k), DoFs
    on one tensor grid X = cx (k+1) + ix (x fastest), weak boundary
    conditions (no Dirichlet rows)."""

    def __init__(self, ctx: Context, dim: int, fe_degree: int, n_cells: int, dtype=FP64,
                 penalty: float | None = None):
        self.ctx, self.dim, self.fe_degree, self.n_cells = ctx, dim, fe_degree, n_cells
        self.prec = _prec(dtype)
        self.dtype = _DTYPES[self.prec]
        self.penalty = dg_penalty(fe_degree) if penalty is None else penalty
        h = ctypes.c_void_p()
        L.check(L.load().gtps_dg_level_create(ctx.handle, dim, fe_degree, n_cells, self.prec,
                                              self.penalty, ctypes.byref(h)), "dg_level_create")
        self._h = h
        self.n_dofs = int(L.load().gtps_dg_level_n_dofs(h))

    def vmult(self, dst: torch.Tensor, src: torch.Tensor):
        _chk_vec(dst, self.dtype, self.n_dofs)
        _chk_vec(src, self.dtype, self.n_dofs)
        L.check(L.load().gtps_dg_vmult(self._h, _ptr(dst), _ptr(src)), "dg_vmult")

    def residual(self, dst: torch.Tensor, b: torch.Tensor, u: torch.Tensor):
        for t in (dst, b, u):
            _chk_vec(t, self.dtype, self.n_dofs)
        L.check(L.load().gtps_dg_residual(self._h, _ptr(dst), _ptr(b), _ptr(u)), "dg_residual")

    def smoother_step(self, u: torch.Tensor, b: torch.Tensor, reverse: bool = False,
                      relaxation: float = 1.0):
        _chk_vec(u, self.dtype, self.n_dofs)
        _chk_vec(b, self.dtype, self.n_dofs)
        L.check(L.load().gtps_dg_smoother_step(
            self._h, _ptr(u), _ptr(b), L.GTPS_MULT_REV if reverse else L.GTPS_MULT_FWD,
            relaxation), "dg_smoother_step")

    def _coarse_n(self):
        n1 = (self.n_cells // 2) * (self.fe_degree + 1)
        return n1 ** self.dim

    def prolongate_and_add(self, fine: torch.Tensor, coarse: torch.Tensor):
        """fine += P coarse (this operator is the fine level)."""
        _chk_vec(fine, self.dtype, self.n_dofs)
        _chk_vec(coarse, self.dtype, self._coarse_n())
        L.check(L.load().gtps_dg_prolongate_add(self._h, _ptr(fine), _ptr(coarse)),
                "dg_prolongate_add")

    def restrict_and_add(self, coarse: torch.Tensor, fine: torch.Tensor):
        _chk_vec(fine, self.dtype, self.n_dofs)
        _chk_vec(coarse, self.dtype, self._coarse_n())
        L.check(L.load().gtps_dg_restrict_add(self._h, _ptr(coarse), _ptr(fine)),
                "dg_restrict_add")

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                L.load().gtps_dg_level_destroy(self._h)
                self._h = None
        except Exception:
            pass


class DGMultigrid:
    """SIPG-DG multigrid on n_coarse 2^l cells per direction (l = 0..max_level)
    with the 27-colour vertex-patch smoother and FP64 MG-CG."""

    def __init__(self, ctx: Context, dim: int, fe_degree: int, n_coarse: int, max_level: int,
                 vcycle_dtype=FP32, penalty: float | None = None, coarse_sweeps: int = 2):
        self.ctx, self.dim, self.fe_degree = ctx, dim, fe_degree
        self.n_coarse, self.max_level = n_coarse, max_level
        self.prec = _prec(vcycle_dtype)
        self.penalty = dg_penalty(fe_degree) if penalty is None else penalty
        h = ctypes.c_void_p()
        L.check(L.load().gtps_dg_mg_create(ctx.handle, dim, fe_degree, n_coarse, max_level,
                                           self.prec, self.penalty, coarse_sweeps,
                                           ctypes.byref(h)), "dg_mg_create")
        self._h = h
        self.n_dofs = int(L.load().gtps_dg_mg_n_dofs(h))

    def set_graph(self, enable: bool):
        """HIP-graph replay of the per-iteration preconditioning sequence."""
        L.check(L.load().gtps_dg_mg_set_graph(self._h, int(enable)), "dg_set_graph")

    def vmult(self, dst: torch.Tensor, src: torch.Tensor):
        """Preconditioner z = V(r) (FP64 in and out)."""
        _chk_vec(dst, torch.float64, self.n_dofs)
        _chk_vec(src, torch.float64, self.n_dofs)
        L.check(L.load().gtps_dg_mg_vcycle(self._h, _ptr(dst), _ptr(src)), "dg_vcycle")

    def do_matvec(self, dst: torch.Tensor, src: torch.Tensor):
        _chk_vec(dst, torch.float64, self.n_dofs)
        _chk_vec(src, torch.float64, self.n_dofs)
        L.check(L.load().gtps_dg_mg_matvec(self._h, _ptr(dst), _ptr(src)), "dg_matvec")

    def solve_cg(self, b: torch.Tensor, x: torch.Tensor | None = None, rtol: float = 1e-8,
                 maxit: int = 200, raise_on_failure: bool = True):
        _chk_vec(b, torch.float64, self.n_dofs)
        if x is None:
            x = torch.empty_like(b)
        _chk_vec(x, torch.float64, self.n_dofs)
        it = ctypes.c_int()
        rel = ctypes.c_double()
        hist = (ctypes.c_double * (maxit + 1))()
        rc = L.load().gtps_dg_mg_pcg_solve(self._h, _ptr(x), _ptr(b), rtol, maxit,
                                           ctypes.byref(it), ctypes.byref(rel), hist)
        if rc != L.GTPS_OK and (raise_on_failure or rc != L.GTPS_ENOTCONV):
            L.check(rc, "dg_pcg_solve")
        return x, it.value, rel.value, list(hist[: it.value + 1])

    def solve_cg_batch(self, B: torch.Tensor, X: torch.Tensor | None = None,
                       rtol: float = 1e-8, maxit: int = 200):
        """Solve A X_j = B_j for the rows of B (shape (m, N), FP64, contiguous)
        together; returns (X, iterations per row, relative residual per row)."""
        if B.dim() != 2 or B.shape[1] != self.n_dofs:
            raise ValueError(f"expected shape (m, {self.n_dofs}), got {tuple(B.shape)}")
        _chk_vec(B.reshape(-1), torch.float64)
        m = B.shape[0]
        if X is None:
            X = torch.empty_like(B)
        _chk_vec(X.reshape(-1), torch.float64, B.numel())
        its = (ctypes.c_int * m)()
        rel = (ctypes.c_double * m)()
        L.check(L.load().gtps_dg_mg_pcg_solve_batch(self._h, _ptr(X), _ptr(B), m, rtol, maxit,
                                                    its, rel), "dg_pcg_solve_batch")
        return X, list(its), list(rel)

    def solve(self, B: torch.Tensor, rtol: float = 1e-10, maxit: int = 300) -> torch.Tensor:
        """Differentiable solve of A U = B for one right-hand side (shape
        (N,)) or several (shape (m, N), solved together); the backward pass
        solves A G_B = G_U with the same MG-CG (A is symmetric and has no
        constrained rows)."""
        return DGSolve.apply(self, B, rtol, maxit)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                L.load().gtps_dg_mg_destroy(self._h)
                self._h = None
        except Exception:
            pass


class DGSolve(torch.autograd.Function):
    """U = A^{-1} B; the rows of a batch are solved together
    (DGMultigrid.solve_cg_batch), a single vector by DGMultigrid.solve_cg."""

    @staticmethod
    def _solve_rows(mg: "DGMultigrid", B: torch.Tensor, rtol, maxit):
        rows = B.reshape(-1, B.shape[-1]).contiguous()
        if rows.shape[0] == 1:
            out = torch.empty_like(rows)
            mg.solve_cg(rows[0], out[0], rtol=rtol, maxit=maxit)
        else:
            out, _, _ = mg.solve_cg_batch(rows, rtol=rtol, maxit=maxit)
        return out.reshape(B.shape)

    @staticmethod
    def forward(ctx, mg: "DGMultigrid", B: torch.Tensor, rtol: float, maxit: int):
        ctx.mg, ctx.rtol, ctx.maxit = mg, rtol, maxit
        return DGSolve._solve_rows(mg, B.detach().to(torch.float64).contiguous(), rtol, maxit)

    @staticmethod
    def backward(ctx, G: torch.Tensor):
        gB = DGSolve._solve_rows(ctx.mg, G.detach().to(torch.float64).contiguous(), ctx.rtol,
                                 ctx.maxit)
        return None, gB, None, None



# ------------------------------------------------------------ post-processing
def gauss_points(n_q: int):
    """(points, weights) of the n_q-point Gauss-Legendre rule on [0, 1]."""
    import numpy as np
    x, w = np.polynomial.legendre.leggauss(n_q)
    return 0.5 * (x + 1.0), 0.5 * w


def quadrature_points_1d(fe_degree: int, n_cells: int):
    """Coordinates of the QGauss(k+1) points along one direction, shape
    (n_cells, k+1): f_q of :func:`assemble_rhs` is f at the tensor products
    of these, ordered [cz][cy][cx][qz][qy][qx] (x fastest)."""
    import numpy as np
    q, _ = gauss_points(fe_degree + 1)
    return (np.arange(n_cells)[:, None] + q[None, :]) / n_cells


def _n_dofs(dim, k, n_cells, dg):
    n1 = n_cells * (k + 1) if dg else k * n_cells + 1
    return n1 ** dim


def _rhs_call(fn, ctx, dim, k, mesh, dg, *args, out=None):
    """mesh: 2^level cells (CG entry points take the level) or n_cells (DG)."""
    N = _n_dofs(dim, k, mesh if dg else 1 << mesh, dg)
    b = torch.empty(N, dtype=torch.float64, device=ctx.device) if out is None else out
    _chk_vec(b, torch.float64, N)
    L.check(fn(ctx.handle, dim, k, mesh, *args, _ptr(b)), fn.__name__)
    return b


def assemble_rhs(ctx: Context, dim: int, fe_degree: int, level: int, f_q: torch.Tensor,
                 out: torch.Tensor | None = None) -> torch.Tensor:
    """b_i = int f phi_i (QGauss(k+1)) on the Q_k space with 2^level cells per
    direction from f at the quadrature points (float64 device tensor of
    (2^level)^dim (k+1)^dim values); Dirichlet rows zero."""
    n = 1 << level
    _chk_vec(f_q, torch.float64, n ** dim * (fe_degree + 1) ** dim)
    return _rhs_call(L.load().gtps_assemble_rhs, ctx, dim, fe_degree, level, False,
                     _ptr(f_q), out=out)


def rhs_sine(ctx: Context, dim: int, fe_degree: int, level: int, wave: float = 3.0,
             out: torch.Tensor | None = None) -> torch.Tensor:
    """The reference Poisson problem's load vector, f = dim (w pi)^2 prod
    sin(w pi x_d) (apps/poisson.cu:76-87), solution prod sin(w pi x_d)."""
    return _rhs_call(L.load().gtps_rhs_sine, ctx, dim, fe_degree, level, False,
                     ctypes.c_double(wave), out=out)


def rhs_constant(ctx: Context, dim: int, fe_degree: int, level: int, value: float = 1.0,
                 out: torch.Tensor | None = None) -> torch.Tensor:
    return _rhs_call(L.load().gtps_rhs_constant, ctx, dim, fe_degree, level, False,
                     ctypes.c_double(value), out=out)


def dg_assemble_rhs(ctx: Context, dim: int, fe_degree: int, n_cells: int, f_q: torch.Tensor,
                    out: torch.Tensor | None = None) -> torch.Tensor:
    _chk_vec(f_q, torch.float64, n_cells ** dim * (fe_degree + 1) ** dim)
    return _rhs_call(L.load().gtps_dg_assemble_rhs, ctx, dim, fe_degree, n_cells, True,
                     _ptr(f_q), out=out)


def dg_rhs_sine(ctx: Context, dim: int, fe_degree: int, n_cells: int, wave: float = 1.0,
                out: torch.Tensor | None = None) -> torch.Tensor:
    return _rhs_call(L.load().gtps_dg_rhs_sine, ctx, dim, fe_degree, n_cells, True,
                     ctypes.c_double(wave), out=out)


def _norms(fn, ctx, dim, k, mesh, dg, u, wave, n_q):
    _chk_vec(u, torch.float64, _n_dofs(dim, k, mesh if dg else 1 << mesh, dg))
    l2, h1 = ctypes.c_double(), ctypes.c_double()
    L.check(fn(ctx.handle, dim, k, mesh, wave, k + 1 if n_q is None else n_q, _ptr(u),
               ctypes.byref(l2), ctypes.byref(h1)), fn.__name__)
    return l2.value, h1.value


def error_norms(ctx: Context, dim: int, fe_degree: int, level: int, u: torch.Tensor,
                wave: float = 3.0, n_q: int | None = None):
    """(L2 error, H1-seminorm error) of u against prod sin(w pi x_d) with
    QGauss(n_q), n_q = k+1 by default as in the reference."""
    return _norms(L.load().gtps_error_norms, ctx, dim, fe_degree, level, False, u, wave, n_q)


def dg_error_norms(ctx: Context, dim: int, fe_degree: int, n_cells: int, u: torch.Tensor,
                   wave: float = 1.0, n_q: int | None = None):
    return _norms(L.load().gtps_dg_error_norms, ctx, dim, fe_degree, n_cells, True, u, wave,
                  n_q)


def save_vector(ctx: Context, path: str, x: torch.Tensor, dim: int, fe_degree: int,
                n_cells: int, layout: int = L.GTPS_LAYOUT_CG):
    """Write a device vector with its mesh description (see gtps_vector_save)."""
    prec = _prec(x.dtype)
    _chk_vec(x, x.dtype, _n_dofs(dim, fe_degree, n_cells, layout == L.GTPS_LAYOUT_DG))
    L.check(L.load().gtps_vector_save(ctx.handle, str(path).encode(), _ptr(x), prec, dim,
                                      fe_degree, n_cells, layout), "vector_save")


def vector_info(path: str) -> dict:
    """Header of a solution file (no GPU needed)."""
    vals = [ctypes.c_int() for _ in range(5)]
    n = ctypes.c_longlong()
    L.check(L.load().gtps_vector_info(str(path).encode(), *[ctypes.byref(v) for v in vals],
                                      ctypes.byref(n)), "vector_info")
    prec, dim, k, nc, layout = (v.value for v in vals)
    return {"precision": prec, "dim": dim, "fe_degree": k, "n_cells": nc, "layout": layout,
            "n": n.value, "dtype": _DTYPES[prec]}


def load_vector(ctx: Context, path: str, out: torch.Tensor | None = None) -> torch.Tensor:
    info = vector_info(path)
    if out is None:
        out = torch.empty(info["n"], dtype=info["dtype"], device=ctx.device)
    _chk_vec(out, info["dtype"], info["n"])
    L.check(L.load().gtps_vector_load(ctx.handle, str(path).encode(), _ptr(out), info["n"],
                                      info["precision"]), "vector_load")
    return out
