"""Properties of the discrete problem that pin the oracle independently of
its implementation: known stencils, exactness on polynomials, adjointness,
the exact coarse solve, V-cycle symmetry, and the O(h^(k+1)) L2 error of
the reference's Poisson test problem u = prod_d sin(3 pi x_d)
(apps/poisson.cu:44-87)."""
import numpy as np
import pytest

import oracle as O
from helpers import boundary_mask, n_dofs, rel_err, rng


def coords(dim, k, level):
    n = 1 << level
    b = O.basis_1d(k)
    x1 = np.concatenate([(c + b["nodes"][:-1]) / n for c in range(n)] + [[1.0]])
    if dim == 2:
        Y, X = np.meshgrid(x1, x1, indexing="ij")
        return [X.ravel(), Y.ravel()]
    Z, Y, X = np.meshgrid(x1, x1, x1, indexing="ij")
    return [X.ravel(), Y.ravel(), Z.ravel()]


def test_q1_2d_stencil():
    """Bilinear elements on a uniform grid: 8/3 on the diagonal, -1/3 for
    the eight neighbours, independent of h."""
    for level in (2, 3):
        N1 = (1 << level) + 1
        x = np.zeros(N1 * N1)
        c = N1 // 2
        x[c + c * N1] = 1.0
        y = O.laplace_vmult(2, 1, level, x).reshape(N1, N1)
        assert abs(y[c, c] - 8 / 3) < 1e-14
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dx or dy:
                    assert abs(y[c + dy, c + dx] + 1 / 3) < 1e-14
        y[c - 1:c + 2, c - 1:c + 2] = 0
        assert np.abs(y).max() == 0


@pytest.mark.parametrize("dim,k,level", [(2, 1, 3), (2, 3, 3), (3, 2, 2), (3, 3, 2), (3, 5, 1)])
def test_vmult_kills_linear_fields(dim, k, level):
    """With boundary columns kept (reference convention), interior rows of
    A u vanish for any affine u: int grad(phi_i) . const = 0."""
    X = coords(dim, k, level)
    u = 0.3 + sum((d + 1) * X[d] for d in range(dim))
    y = O.laplace_vmult(dim, k, level, u)
    bm = boundary_mask(dim, k, level)
    assert np.abs(y[~bm]).max() < 1e-11
    assert np.array_equal(y[bm], u[bm])


@pytest.mark.parametrize("dim,k,level", [(2, 2, 3), (3, 3, 2)])
def test_vmult_symmetric_on_zero_boundary(dim, k, level):
    g = rng(2)
    bm = boundary_mask(dim, k, level)
    x, y = g.uniform(-1, 1, (2, n_dofs(dim, k, level)))
    x[bm] = 0
    y[bm] = 0
    a = O.laplace_vmult(dim, k, level, x) @ y
    b = x @ O.laplace_vmult(dim, k, level, y)
    assert abs(a - b) < 1e-12 * abs(a)
    assert x @ O.laplace_vmult(dim, k, level, x) > 0


@pytest.mark.parametrize("dim,k", [(2, 1), (2, 4), (3, 1), (3, 3), (3, 7)])
def test_level1_smoother_is_exact_solve(dim, k):
    N = n_dofs(dim, k, 1)
    bm = boundary_mask(dim, k, 1)
    b = rng(4).uniform(-1, 1, N)
    b[bm] = 0
    u = O.smoother_step(dim, k, 1, b, np.zeros(N))
    r = b - O.laplace_vmult(dim, k, 1, u)
    assert np.abs(r).max() < 1e-11 * np.abs(b).max()


@pytest.mark.parametrize("dim,k,level", [(2, 2, 3), (2, 5, 2), (3, 1, 3), (3, 3, 2), (3, 4, 2)])
def test_prolongation_interpolates_polynomials(dim, k, level):
    """Prolongation is interpolation: a degree-k polynomial vanishing on the
    boundary is reproduced exactly on the fine level."""
    Xc, Xf = coords(dim, k, level - 1), coords(dim, k, level)

    def f(X):
        v = 1.0
        for d in range(dim):
            v = v * X[d] * (1 - X[d]) * (0.5 + X[d] ** (k - 2) if k >= 2 else 1.0)
        return v if k >= 2 else v * 0  # Q1: only the zero polynomial vanishes

    if k >= 2:
        fine = O.prolongate_add(dim, k, level, f(Xc), np.zeros(len(Xf[0])))
        assert np.abs(fine - f(Xf)).max() < 1e-13
    # coarse Dirichlet values are ignored
    c = np.zeros(len(Xc[0]))
    c[boundary_mask(dim, k, level - 1)] = 1.0
    assert np.abs(O.prolongate_add(dim, k, level, c, np.zeros(len(Xf[0])))).max() == 0


@pytest.mark.parametrize("dim,k,level", [(2, 2, 4), (3, 2, 3), (3, 3, 2)])
def test_restriction_is_adjoint(dim, k, level):
    g = rng(9)
    nf, nc = n_dofs(dim, k, level), n_dofs(dim, k, level - 1)
    x = g.uniform(-1, 1, nf)
    y = g.uniform(-1, 1, nc)
    y[boundary_mask(dim, k, level - 1)] = 0
    Py = O.prolongate_add(dim, k, level, y, np.zeros(nf))
    Rx = O.restrict_add(dim, k, level, x, np.zeros(nc))
    assert abs(Rx @ y - x @ Py) < 1e-12 * abs(x @ Py)
    assert np.all(Rx[boundary_mask(dim, k, level - 1)] == 0)


def _vcycle_bilinear(dim, k, level, variant):
    g = rng(17)
    bm = boundary_mask(dim, k, level)
    r, s = g.uniform(-1, 1, (2, n_dofs(dim, k, level)))
    r[bm] = 0
    s[bm] = 0
    a = O.vcycle(dim, k, level, r, mg_variant=variant) @ s
    b = r @ O.vcycle(dim, k, level, s, mg_variant=variant)
    return a, b


def test_vcycle_symmetry():
    """Reversing the colours in post-smoothing makes V symmetric (CG needs
    it); the reference's order (solver.cuh:1969,1997) does not."""
    a, b = _vcycle_bilinear(3, 2, 3, O.MG_SYMMETRIC)
    assert abs(a - b) < 1e-12 * abs(a)
    a, b = _vcycle_bilinear(3, 2, 3, O.MG_REFERENCE_ORDER)
    assert abs(a - b) > 1e-6 * abs(a)
    a, b = _vcycle_bilinear(2, 3, 4, O.MG_ADDITIVE)
    assert abs(a - b) < 1e-12 * abs(a)


@pytest.mark.parametrize("dim,k,levels", [(2, 2, (3, 4, 5)), (2, 3, (2, 3, 4)), (3, 2, (2, 3, 4))])
def test_poisson_l2_convergence(dim, k, levels):
    """Discretisation error of the reference's test problem decreases like
    h^(k+1) (rate within 0.35 of k+1 over two refinements)."""
    errs = []
    for level in levels:
        b = O.rhs_assembled(dim, k, level, O.RHS_SINE)
        x, it, rel, _ = O.pcg(dim, k, level, b, rtol=1e-13, maxit=100)
        errs.append(O.l2_error_sine(dim, k, level, x))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(np.abs(rates[-1:] - (k + 1)) < 0.35), (errs, rates)


def test_planted_mistake_is_caught():
    """A plausible bug (reference colour order instead of reversed, i.e. a
    wrong post-smoothing index) changes the V-cycle by far more than the
    tolerances the GPU parity tests use."""
    r = O.rhs_random(3, 3, 3)
    z_sym = O.vcycle(3, 3, 3, r, mg_variant=O.MG_SYMMETRIC)
    z_ref = O.vcycle(3, 3, 3, r, mg_variant=O.MG_REFERENCE_ORDER)
    assert np.abs(z_sym - z_ref).max() > 1e-3 * np.abs(z_sym).max()


# ------------------------------------------------------------ full multigrid
@pytest.mark.parametrize("dim,k,level", [(2, 2, 4), (3, 2, 3), (3, 3, 3)])
def test_fmg_reaches_discretisation_error(dim, k, level):
    """One FMG pass (n_cycles = 1) on the sine problem is as accurate as the
    exact discrete solution (within 5%), the classical FMG property."""
    rhs = [O.rhs_assembled(dim, k, l, O.RHS_SINE) for l in range(1, level + 1)]
    x = O.fmg(dim, k, level, rhs)
    xs, _, _, _ = O.pcg(dim, k, level, rhs[-1], rtol=1e-13)
    e_fmg, e_ex = O.l2_error_sine(dim, k, level, x), O.l2_error_sine(dim, k, level, xs)
    assert e_fmg <= 1.05 * e_ex


@pytest.mark.parametrize("dim,k,level", [(2, 3, 3), (3, 2, 2)])
def test_fmg_cycles_converge_to_discrete_solution(dim, k, level):
    """With many V-cycles per level the FMG correction loop converges to the
    discrete solution (it is an iteration on the right system)."""
    rhs = [O.rhs_assembled(dim, k, l, O.RHS_CONST) for l in range(1, level + 1)]
    x = O.fmg(dim, k, level, rhs, n_cycles=25)
    xs, _, _, _ = O.pcg(dim, k, level, rhs[-1], rtol=1e-14)
    assert rel_err(x, xs) < 1e-11


def test_fmg_level_one_is_exact_solve():
    rhs = [O.rhs_assembled(3, 3, 1, O.RHS_CONST)]
    x = O.fmg(3, 3, 1, rhs)
    r = rhs[0] - O.laplace_vmult(3, 3, 1, x)
    assert np.abs(r).max() < 1e-13 * np.abs(rhs[0]).max()
