"""GPU parity of the variable-coefficient path (operator, transpose, residual,
gradient kernel, inexact-FD smoother, MG-PCG) against the FP64 CPU oracle,
and the torch.autograd adjoint against central finite differences.

Tolerances as for the constant-coefficient kernels: FP64 1e-12 relative
(max-norm), FP32 2e-5 (x20 for a smoother sweep).  Every case prints what it
runs before touching the GPU.
"""
import numpy as np
import pytest

import oracle as O
from helpers import boundary_mask, n_dofs, rel_err, rng

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 2e-5}


@pytest.fixture(scope="module")
def core():
    from gputensorproductsmoother_amd import core as C
    return C


@pytest.fixture(scope="module")
def ctx(core):
    return core.Context(0)


def _dt(prec):
    return torch.float64 if prec == "fp64" else torch.float32


def _announce(*args):
    print("CASE", *args, flush=True)


def _dev(a, dtype=torch.float64):
    return torch.tensor(a, dtype=dtype, device="cuda:0")


def _coef(level, seed=5):
    return np.exp(0.5 * rng(seed).standard_normal((1 << level) ** 3))


OP_CASES = [(1, 3), (2, 2), (2, 3), (3, 2), (4, 2), (5, 1), (7, 1)]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("k,level", OP_CASES)
def test_coeff_vmult_residual_transpose(core, ctx, k, level, prec):
    _announce("coeff-vmult", k, level, prec)
    N = n_dofs(3, k, level)
    g = rng(level * 10 + k)
    x, b = g.uniform(-1, 1, (2, N))
    c = _coef(level)
    if prec == "fp32":
        x = x.astype(np.float32).astype(np.float64)
        b = b.astype(np.float32).astype(np.float64)
        c = c.astype(np.float32).astype(np.float64)
    op = core.LaplaceOperator(ctx, 3, k, level, _dt(prec))
    op.set_coefficient(_dev(c))
    xd, bd = _dev(x, _dt(prec)), _dev(b, _dt(prec))
    y = torch.full_like(xd, float("nan"))
    op.vmult(y, xd)
    r = torch.full_like(xd, float("nan"))
    op.residual(r, bd, xd)
    t = torch.full_like(xd, float("nan"))
    op.vmult_transpose(t, xd)
    torch.cuda.synchronize()
    ref = O.coeff_vmult(3, k, level, c, x)
    assert rel_err(y.cpu().numpy(), ref) < TOL[prec]
    assert rel_err(r.cpu().numpy(), b - ref) < TOL[prec]
    assert rel_err(t.cpu().numpy(), O.coeff_vmult(3, k, level, c, x, transpose=True)) < TOL[prec]
    # clearing restores the constant-coefficient operator
    op.set_coefficient(None)
    op.vmult(y, xd)
    torch.cuda.synchronize()
    assert rel_err(y.cpu().numpy(), O.laplace_vmult(3, k, level, x)) < TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("k,level", [(1, 3), (2, 2), (4, 2), (5, 1)])
def test_coeff_gradient_kernel(core, ctx, k, level, prec):
    _announce("coeff-gradient", k, level, prec)
    N = n_dofs(3, k, level)
    g = rng(level + 7 * k)
    lam, u = g.uniform(-1, 1, (2, N))
    if prec == "fp32":
        lam = lam.astype(np.float32).astype(np.float64)
        u = u.astype(np.float32).astype(np.float64)
    op = core.LaplaceOperator(ctx, 3, k, level, _dt(prec))
    grad = torch.full(((1 << level) ** 3,), float("nan"), dtype=_dt(prec), device="cuda:0")
    op.coefficient_gradient(grad, _dev(lam, _dt(prec)), _dev(u, _dt(prec)))
    torch.cuda.synchronize()
    assert rel_err(grad.cpu().numpy(), O.coeff_gradient(3, k, level, lam, u)) < TOL[prec] * 5


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("variant", ["fwd", "rev"])
@pytest.mark.parametrize("k,level", [(1, 3), (2, 2), (3, 2), (4, 2), (5, 1), (6, 1), (7, 1)])
def test_coeff_smoother_step(core, ctx, k, level, variant, prec):
    if prec == "fp64" and k > 5:
        pytest.skip("FP64 coefficient smoother supports k <= 5")
    _announce("coeff-smoother", k, level, variant, prec)
    N = n_dofs(3, k, level)
    g = rng(level * 3 + k)
    b, u = g.uniform(-1, 1, (2, N))
    c = _coef(level, 21)
    if prec == "fp32":
        b = b.astype(np.float32).astype(np.float64)
        u = u.astype(np.float32).astype(np.float64)
        c = c.astype(np.float32).astype(np.float64)
    ov, cv = {"fwd": (O.MULT_FWD, core.PatchSmoother.MULTIPLICATIVE),
              "rev": (O.MULT_REV, core.PatchSmoother.MULTIPLICATIVE_REVERSE)}[variant]
    ref = O.coeff_smoother_step(3, k, level, c, b, u, variant=ov, relax=0.9)
    op = core.LaplaceOperator(ctx, 3, k, level, _dt(prec))
    op.set_coefficient(_dev(c))
    sm = core.PatchSmoother(op, cv, 0.9)
    ud = _dev(u, _dt(prec))
    sm.step(ud, _dev(b, _dt(prec)))
    torch.cuda.synchronize()
    tol = TOL[prec] * (20 if prec == "fp32" else 1)
    assert rel_err(ud.cpu().numpy(), ref) < tol


def test_coeff_additive_rejected(core, ctx):
    op = core.LaplaceOperator(ctx, 3, 2, 2, torch.float64)
    op.set_coefficient(_dev(_coef(2)))
    sm = core.PatchSmoother(op, core.PatchSmoother.ADDITIVE, 0.2)
    u = op.initialize_dof_vector()
    with pytest.raises(Exception, match="additive"):
        sm.step(u, u.clone())


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("k,level", [(2, 3), (4, 3)])
def test_coeff_pcg_matches_oracle(core, ctx, k, level, prec):
    _announce("coeff-pcg", k, level, prec)
    c = _coef(level, 31)
    b = O.rhs_assembled(3, k, level)
    xo, ito, relo, _ = O.coeff_pcg(3, k, level, c, b, rtol=1e-10, maxit=200)
    mg = core.MultigridSolvers(ctx, 3, k, level, _dt(prec))
    mg.set_coefficient(_dev(c))
    x, it, rel, hist = mg.solve_cg(_dev(b), rtol=1e-10, maxit=200)
    torch.cuda.synchronize()
    print(f"iterations gpu {it} oracle {ito}", flush=True)
    assert rel <= 1e-10
    assert abs(it - ito) <= (0 if prec == "fp64" else 2)
    assert rel_err(x.cpu().numpy(), xo) < 1e-8
    # the graph replay gives the eager result bit for bit
    mg.set_graph(False)
    x2, it2, _, _ = mg.solve_cg(_dev(b), rtol=1e-10, maxit=200)
    torch.cuda.synchronize()
    assert it2 == it and torch.equal(x2, x)
    # clearing the coefficient restores the Poisson solver
    mg.set_coefficient(None)
    xp, itp, relp, _ = mg.solve_cg(_dev(b), rtol=1e-10, maxit=200)
    torch.cuda.synchronize()
    xpo, _, _, _ = O.pcg(3, k, level, b, rtol=1e-10)
    assert rel_err(xp.cpu().numpy(), xpo) < 1e-8


def _J(u, ud, mask):
    d = torch.where(mask, torch.zeros_like(u), u - ud)
    return 0.5 * (d * d).sum()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("k,level", [(2, 3), (4, 3)])
def test_autograd_adjoint_vs_finite_differences(core, ctx, k, level, prec):
    """dJ/dc and dJ/db from MGSolve.backward (V-cycle in `prec`) against
    central differences (h = 1e-6) of J evaluated with FP64-everywhere solves
    to 1e-14, and against the oracle's adjoint (3D Q2 and Q4 on 8^3 cells;
    the Q4 32^3 workload is checked against central differences in
    test_gpu_configs.py)."""
    _announce("autograd", k, level, prec)
    rtol = 1e-13 if prec == "fp64" else 1e-12
    mg = core.MultigridSolvers(ctx, 3, k, level, _dt(prec))
    mg64 = core.MultigridSolvers(ctx, 3, k, level, torch.float64)
    c0 = _dev(_coef(level, 41))
    g = rng(42)
    N = n_dofs(3, k, level)
    bm = boundary_mask(3, k, level)
    b0 = g.uniform(-1, 1, N)
    b0[bm] = 0.0
    ud = _dev(0.01 * g.uniform(0, 1, N))
    mask = mg.boundary_mask()
    c = c0.clone().requires_grad_(True)
    b = _dev(b0).requires_grad_(True)
    u = mg.solve(b, c, rtol=rtol, maxit=400)
    J = _J(u, ud, mask)
    J.backward()
    gc, gb = c.grad.cpu().numpy(), b.grad.cpu().numpy()
    # the oracle's adjoint gradient at the same state
    ref_u = O.coeff_pcg(3, k, level, c0.cpu().numpy(), b0, rtol=1e-14, maxit=400)[0]
    lam = O.coeff_pcg(3, k, level, c0.cpu().numpy(),
                      np.where(bm, 0.0, ref_u - ud.cpu().numpy()), rtol=1e-14, maxit=400)[0]
    ref_gc = O.coeff_gradient(3, k, level, lam, ref_u)
    assert rel_err(gc, ref_gc) < 1e-8
    assert rel_err(gb[~bm], lam[~bm]) < 1e-8

    mg64.set_coefficient(c0)
    u0, _, _, _ = mg64.solve_cg(_dev(b0), rtol=1e-14, maxit=400, raise_on_failure=False)
    u0 = u0.clone()

    def u64(bv, cv):
        # u(c', b') = u0 + A(c')^{-1} (b' - A(c') u0): the correction is O(h),
        # so the difference of two perturbed solutions keeps ~12 digits
        mg64.set_coefficient(cv)
        r = torch.empty_like(u0)
        mg64.do_matvec(r, u0)
        d, _, _, _ = mg64.solve_cg(bv - r, rtol=1e-12, maxit=400, raise_on_failure=False)
        return u0 + d

    def central(up, um):
        # J(+) - J(-) = sum (u+ - u-) ((u+ + u-)/2 - u_d) over the interior,
        # free of the cancellation of two large sums
        d = torch.where(mask, torch.zeros_like(up), (up - um) * (0.5 * (up + um) - ud))
        return d.sum().item() / (2 * h)

    h = 1e-6
    worst = 0.0
    for K in rng(43).choice(c0.numel(), 6, replace=False):
        cp, cm = c0.clone(), c0.clone()
        cp[K] += h
        cm[K] -= h
        fd = central(u64(_dev(b0), cp), u64(_dev(b0), cm))
        worst = max(worst, abs(gc[K] - fd) / abs(fd))
        assert abs(gc[K] - fd) <= 1e-6 * abs(fd), (K, gc[K], fd)
    for i in rng(44).choice(np.flatnonzero(~bm), 4, replace=False):
        bp, bmn = b0.copy(), b0.copy()
        bp[i] += h
        bmn[i] -= h
        fd = central(u64(_dev(bp), c0), u64(_dev(bmn), c0))
        assert abs(gb[i] - fd) <= 1e-6 * abs(fd), (i, gb[i], fd)
    print(f"worst relative FD deviation of dJ/dc: {worst:.2e}", flush=True)


def test_autograd_boundary_gradient_of_b(core, ctx):
    """dL/db on Dirichlet DoFs is g_B - A_IB^T lambda_I (identity rows)."""
    k, level = 2, 2
    mg = core.MultigridSolvers(ctx, 3, k, level, torch.float64)
    N = n_dofs(3, k, level)
    bm = boundary_mask(3, k, level)
    g = rng(45)
    b0 = g.uniform(-1, 1, N)
    w = _dev(g.uniform(-1, 1, N))
    b = _dev(b0).requires_grad_(True)
    u = mg.solve(b, None, rtol=1e-14, maxit=300)
    (u * w).sum().backward()
    # dense check: A^T lambda = w
    A = np.zeros((N, N))
    for j in range(N):
        e = np.zeros(N)
        e[j] = 1.0
        A[:, j] = O.laplace_vmult(3, k, level, e)
    lam = np.linalg.solve(A.T, w.cpu().numpy())
    assert rel_err(b.grad.cpu().numpy(), lam) < 1e-9
    assert np.abs(b.grad.cpu().numpy()[bm]).max() > 0
