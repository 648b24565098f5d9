"""The exported vector entry points, the transfer's plain prolongation and the
solvers' failure reports, called through the C ABI on the GPU.

* ``gtps_vec_fill / axpby / dot / nrm2 / convert_* / set_boundary`` replace
  the deal.II vector operations and helpers the reference's solver uses
  (``include/cuda_vector.cuh:79-118``, ``include/solver.cuh:45-78`` and
  ``2070-2095``).  Expected values come from numpy: fill, conversions and
  set_boundary bit-exact (numpy's float32 cast rounds to nearest even, like
  the kernel), axpby to the rounding of its two products and one sum, dot
  and nrm2 to 1e-14 relative against an exactly rounded sum of
  extended-precision products (``math.fsum``).
* ``gtps_transfer_prolongate`` (``MGTransferCUDA::prolongate``,
  ``cuda_mg_transfer.template.cuh:561``) against the oracle's
  prolongate-add into a zero vector.
* ``GTPS_ENOTCONV`` (maxit too small) and ``GTPS_ENAN`` (a NaN in the
  right-hand side) from the CG and DG solvers: both are host-side checks of
  the residual norm, so no kernel misbehaves.  The reference swallows the
  GMRES failure instead (``include/solver.cuh:1896-1904``).
* the batched DG CG with a zero right-hand side among its rows, and the DG
  adjoint whose loss uses one row only.

Sizes cover the empty vector, one element, lengths that are not multiples of
a work-group, and the metric configuration's 193^3 DoFs.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle as O
from helpers import boundary_mask, n_dofs, rel_err, rng

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SIZES = [0, 1, 255, 257, 4097, 1000003, 193 ** 3]


@pytest.fixture(scope="module")
def core():
    from gputensorproductsmoother_amd import core as C
    return C


@pytest.fixture(scope="module")
def lib():
    from gputensorproductsmoother_amd import _lib as L
    return L


@pytest.fixture(scope="module")
def ctx(core):
    return core.Context(0)


def _announce(*args):
    print("CASE", *args, flush=True)


def _vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() else None


def _dev(a, dtype=torch.float64):
    return torch.tensor(np.asarray(a), dtype=dtype, device="cuda:0")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("n", SIZES)
def test_vec_fill(lib, ctx, n, prec):
    _announce("vec-fill", n, prec)
    dt = torch.float64 if prec == "fp64" else torch.float32
    code = lib.GTPS_FP64 if prec == "fp64" else lib.GTPS_FP32
    # one guard element on each side must stay untouched
    buf = torch.full((n + 2,), 7.0, dtype=dt, device="cuda:0")
    x = buf[1:n + 1]
    v = 1.0 / 3.0
    ptr = ctypes.c_void_p(x.data_ptr()) if n else None
    lib.check(lib.load().gtps_vec_fill(ctx.handle, n, code, ptr, v), "fill")
    torch.cuda.synchronize()
    h = buf.cpu().numpy()
    want = np.float64(v) if prec == "fp64" else np.float32(v)
    assert h[0] == 7.0 and h[-1] == 7.0
    assert np.all(h[1:n + 1] == want)


def test_vec_fill_rejects_bad_precision(lib, ctx):
    _announce("vec-fill-bad-precision")
    x = torch.zeros(16, dtype=torch.float32, device="cuda:0")
    rc = lib.load().gtps_vec_fill(ctx.handle, 16, 7, ctypes.c_void_p(x.data_ptr()), 1.0)
    assert rc == lib.GTPS_EINVAL
    assert b"precision" in lib.load().gtps_last_error()
    torch.cuda.synchronize()
    assert torch.count_nonzero(x).item() == 0


@pytest.mark.parametrize("n", SIZES)
def test_vec_axpby(lib, ctx, n):
    _announce("vec-axpby", n)
    g = rng(n + 1)
    x, y = g.uniform(-1, 1, (2, n))
    a, b = 0.75, -1.25
    xd, yd = _dev(x), _dev(y)
    lib.check(lib.load().gtps_vec_axpby(ctx.handle, n, a, _vp(xd), b, _vp(yd)), "axpby")
    got = yd.cpu().numpy()
    want = a * x + b * y
    # both sides round each product and the sum at most once (a fused
    # multiply-add skips one rounding): within 2 eps (|a x| + |b y|) each
    bound = 4.0 * 2.0 ** -53 * (np.abs(a * x) + np.abs(b * y))
    assert np.all(np.abs(got - want) <= bound)


def _fsum_dot(x, y):
    # products in extended precision (64-bit significand), summed exactly
    # rounded: within ~1e-19 relative of the exact dot product
    p = x.astype(np.longdouble) * y.astype(np.longdouble)
    hi = p.astype(np.float64)
    lo = (p - hi).astype(np.float64)
    return math.fsum(np.concatenate([hi, lo]))


@pytest.mark.parametrize("n", SIZES)
def test_vec_dot_nrm2(lib, ctx, n):
    _announce("vec-dot-nrm2", n)
    g = rng(2 * n + 3)
    x, y = g.uniform(-1, 1, (2, n))
    xd, yd = _dev(x), _dev(y)
    d, r = ctypes.c_double(), ctypes.c_double()
    L = lib.load()
    lib.check(L.gtps_vec_dot(ctx.handle, n, _vp(xd), _vp(yd), ctypes.byref(d)), "dot")
    lib.check(L.gtps_vec_nrm2(ctx.handle, n, _vp(xd), ctypes.byref(r)), "nrm2")
    xx = _fsum_dot(x, x)
    want_d = _fsum_dot(x, y)
    want_r = math.sqrt(xx)
    if n == 0:
        assert d.value == 0.0 and r.value == 0.0
        return
    scale = math.sqrt(xx * _fsum_dot(y, y))
    assert abs(d.value - want_d) <= 1e-14 * scale
    assert abs(r.value - want_r) <= 1e-14 * want_r
    # the reduction order is fixed: the same call gives the same bits
    d2 = ctypes.c_double()
    lib.check(L.gtps_vec_dot(ctx.handle, n, _vp(xd), _vp(yd), ctypes.byref(d2)), "dot")
    assert d2.value == d.value


@pytest.mark.parametrize("n", SIZES)
def test_vec_convert(lib, ctx, n):
    _announce("vec-convert", n)
    g = rng(3 * n + 5)
    # values with more than 24 significant bits, ties of the float32
    # rounding (to even: 1 + 2^-24 -> 1, 1 + 3 2^-24 -> 1 + 2^-22) and
    # magnitudes across the normal float32 range
    x = g.uniform(-1, 1, n) * np.exp2(g.integers(-30, 30, n))
    if n >= 4:
        x[:4] = [1.0 + 2.0 ** -24, 1.0 + 3 * 2.0 ** -24, -(1.0 + 2.0 ** -24), 2.0 ** -100]
    xd = _dev(x)
    f = torch.empty(n, dtype=torch.float32, device="cuda:0")
    L = lib.load()
    lib.check(L.gtps_vec_convert_f64_f32(ctx.handle, n, _vp(xd), _vp(f)), "f64->f32")
    want = x.astype(np.float32)
    got = f.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    back = torch.empty(n, dtype=torch.float64, device="cuda:0")
    lib.check(L.gtps_vec_convert_f32_f64(ctx.handle, n, _vp(f), _vp(back)), "f32->f64")
    assert np.array_equal(back.cpu().numpy(), want.astype(np.float64))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("dim,k,level", [(2, 2, 3), (2, 7, 2), (3, 1, 2), (3, 3, 3), (3, 4, 2)])
def test_vec_set_boundary(core, lib, ctx, dim, k, level, prec):
    _announce("vec-set-boundary", dim, k, level, prec)
    dt = torch.float64 if prec == "fp64" else torch.float32
    op = core.LaplaceOperator(ctx, dim, k, level, dt)
    x = rng(level + k).uniform(-1, 1, op.n_dofs)
    if prec == "fp32":
        x = x.astype(np.float32).astype(np.float64)
    xd = _dev(x, dt)
    lib.check(lib.load().gtps_vec_set_boundary(op.handle, _vp(xd), 2.5), "set_boundary")
    got = xd.double().cpu().numpy()
    bm = boundary_mask(dim, k, level)
    assert np.all(got[bm] == 2.5)
    assert np.array_equal(got[~bm], x[~bm])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("dim,k,level", [(2, 2, 4), (2, 7, 2), (3, 1, 3), (3, 3, 4), (3, 4, 3),
                                         (3, 7, 2)])
def test_transfer_prolongate(core, ctx, dim, k, level, prec):
    """fine = P coarse (overwrites fine, including its Dirichlet DoFs)."""
    _announce("transfer-prolongate", dim, k, level, prec)
    dt = torch.float64 if prec == "fp64" else torch.float32
    coarse = rng(level * 7 + k).uniform(-1, 1, n_dofs(dim, k, level - 1))
    if prec == "fp32":
        coarse = coarse.astype(np.float32).astype(np.float64)
    ref = O.prolongate_add(dim, k, level, coarse, np.zeros(n_dofs(dim, k, level)))
    tr = core.MGTransferCUDA(ctx, dim, k, 1, level, dt)
    fine = torch.full((n_dofs(dim, k, level),), float("nan"), dtype=dt, device="cuda:0")
    tr.prolongate(level, fine, _dev(coarse, dt))
    torch.cuda.synchronize()
    got = fine.double().cpu().numpy()
    assert np.all(np.isfinite(got))
    assert rel_err(got, ref) < (1e-12 if prec == "fp64" else 2e-5)


# ------------------------------------------------------------ failure reports
def test_cg_reports_non_convergence(core, lib, ctx):
    _announce("cg-enotconv")
    dim, k, level = 3, 3, 3
    mg = core.MultigridSolvers(ctx, dim, k, level, torch.float32)
    b = _dev(O.rhs_random(dim, k, level))
    with pytest.raises(lib.GtpsError) as e:
        mg.solve_cg(b, rtol=1e-12, maxit=1)
    assert e.value.code == lib.GTPS_ENOTCONV
    assert b"maxit" in lib.load().gtps_last_error()
    # the iterate of the one step is still returned when not raising
    x, it, rel, hist = mg.solve_cg(b, rtol=1e-12, maxit=1, raise_on_failure=False)
    assert it == 1 and 1e-12 < rel < 1.0 and len(hist) == 2
    assert torch.isfinite(x).all()
    # the solver is usable afterwards
    _, it, rel, _ = mg.solve_cg(b, rtol=1e-8)
    assert rel <= 1e-8


@pytest.mark.parametrize("where", ["interior", "boundary"])
def test_cg_reports_nan(core, lib, ctx, where):
    _announce("cg-enan", where)
    dim, k, level = 3, 2, 3
    mg = core.MultigridSolvers(ctx, dim, k, level, torch.float64)
    bh = O.rhs_random(dim, k, level)
    idx = np.flatnonzero(~boundary_mask(dim, k, level) if where == "interior"
                         else boundary_mask(dim, k, level))[5]
    bh[idx] = np.nan
    with pytest.raises(lib.GtpsError) as e:
        mg.solve_cg(_dev(bh), rtol=1e-8)
    assert e.value.code == lib.GTPS_ENAN
    _, _, rel, _ = mg.solve_cg(_dev(O.rhs_random(dim, k, level)), rtol=1e-8)
    assert rel <= 1e-8


def test_dg_cg_failure_reports(core, lib, ctx):
    _announce("dg-cg-failures")
    mg = core.DGMultigrid(ctx, 3, 2, 3, 1, core.FP32)
    b = rng(3).uniform(-1, 1, mg.n_dofs)
    with pytest.raises(lib.GtpsError) as e:
        mg.solve_cg(_dev(b), rtol=1e-12, maxit=1)
    assert e.value.code == lib.GTPS_ENOTCONV
    bn = b.copy()
    bn[17] = np.inf
    with pytest.raises(lib.GtpsError) as e:
        mg.solve_cg(_dev(bn), rtol=1e-8)
    assert e.value.code == lib.GTPS_ENAN
    B = _dev(np.stack([b, b]))
    with pytest.raises(lib.GtpsError) as e:
        mg.solve_cg_batch(B, rtol=1e-12, maxit=1)
    assert e.value.code == lib.GTPS_ENOTCONV
    Bn = _dev(np.stack([b, bn]))
    with pytest.raises(lib.GtpsError) as e:
        mg.solve_cg_batch(Bn, rtol=1e-8)
    assert e.value.code == lib.GTPS_ENAN


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_dg_batch_with_zero_rows(core, ctx, prec):
    """A zero right-hand side in a batch gives a zero solution in zero
    iterations, and a row that converges before the others keeps the
    single-vector solution while the rest go on (FP64 oracle counts at
    1e-11 on this mesh: 5 iterations for the random row, 6 for the sine
    load vector)."""
    _announce("dg-batch-zero-rows", prec)
    dt = core.FP32 if prec == "fp32" else core.FP64
    mg = core.DGMultigrid(ctx, 3, 2, 3, 1, dt)
    N = mg.n_dofs
    b1 = rng(71).uniform(-1, 1, N)
    x1, it1, _, _ = mg.solve_cg(_dev(b1), rtol=1e-6)
    x1 = x1.clone()
    B = _dev(np.stack([np.zeros(N), b1, np.zeros(N), b1]))
    X, its, rel = mg.solve_cg_batch(B, rtol=1e-6)
    assert its == [0, it1, 0, it1]
    assert rel[0] == 0.0 and rel[2] == 0.0
    assert torch.count_nonzero(X[0]).item() == 0 and torch.count_nonzero(X[2]).item() == 0
    assert torch.equal(X[1], x1) and torch.equal(X[3], x1)
    smooth = _dev(O.dg_rhs(3, 2, 6, O.DG_RHS_SINE))
    xs, its_s, _, _ = mg.solve_cg(smooth, rtol=1e-11)
    xs = xs.clone()
    xb, its_b, _, _ = mg.solve_cg(_dev(b1), rtol=1e-11)
    xb = xb.clone()
    print("iterations sine", its_s, "random", its_b, flush=True)
    if prec == "fp64":
        assert (its_s, its_b) == (6, 5)
    X, its, rel = mg.solve_cg_batch(torch.stack([smooth, _dev(b1)]).contiguous(), rtol=1e-11)
    assert its == [its_s, its_b]
    assert torch.equal(X[0], xs) and torch.equal(X[1], xb)


def test_dg_adjoint_loss_on_one_row(core, ctx):
    """J = sum(U[0] * w): the cotangent of rows 1..3 is zero, so their
    gradients are exactly zero and row 0's is A^{-1} w."""
    _announce("dg-adjoint-one-row")
    k, n0, maxl = 2, 3, 1
    mg = core.DGMultigrid(ctx, 3, k, n0, maxl, core.FP32)
    N = mg.n_dofs
    g = rng(72)
    B = _dev(g.uniform(-1, 1, (4, N))).requires_grad_(True)
    w = g.uniform(-1, 1, N)
    U = mg.solve(B, rtol=1e-11)
    (U[0] * _dev(w)).sum().backward()
    gB = B.grad
    assert torch.isfinite(gB).all()
    for j in (1, 2, 3):
        assert torch.count_nonzero(gB[j]).item() == 0, j
    ref, _, _, _ = O.dg_pcg(3, k, n0, maxl, w, rtol=1e-12)
    assert rel_err(gB[0].cpu().numpy(), ref) < 1e-8
