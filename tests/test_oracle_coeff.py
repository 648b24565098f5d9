"""Pin the oracle's variable-coefficient pieces (operator, transpose,
inexact-FD smoother, coefficient coarsening, MG-PCG, dJ/dc and the problem
generators) without trusting its own formulation:

* dense numpy assembly of sum_K c_K P_K^T (K (x) M (x) M + ...) P_K from the
  50-digit golden 1D cell matrices (tests/golden/basis_1d.json), with the
  identity Dirichlet rows;
* block Gauss-Seidel over the patch colours with the inexact local solve
  (sigma_P sum_K f_K A_K)^{-1} of the separable fit of the patch's cell
  coefficients, assembled densely from the golden cell matrices;
* exact identities: c = 1 gives the Poisson operator and the exact smoother,
  sum_K c_K dJ/dc_K = -lambda~ . A(c) u, <A x, y> = <x, A^T y>;
* central finite differences of J(c) = 1/2 ||u(c) - u_d||^2 with dense
  direct solves.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from helpers import boundary_mask, n_dofs, rel_err, rng

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "basis_1d.json")))


def _cell_1d(k):
    g = GOLD[str(k)]
    return np.array(g["M"], dtype=np.float64), np.array(g["K"], dtype=np.float64)


def _cell_dofs(dim, k, level, cx, cy, cz):
    n1 = k * (1 << level) + 1
    i = np.arange(k + 1)
    if dim == 2:
        return ((k * cy + i)[:, None] * n1 + (k * cx + i)[None, :]).ravel()
    return ((k * cz + i)[:, None, None] * n1 * n1 + (k * cy + i)[None, :, None] * n1
            + (k * cx + i)[None, None, :]).ravel()


def _cells(dim, level):
    n = 1 << level
    for cz in range(n if dim == 3 else 1):
        for cy in range(n):
            for cx in range(n):
                yield cx + n * cy + n * n * cz, cx, cy, cz


def _cell_matrix(dim, k, level):
    """h^(d-2) Kronecker-sum cell matrix, x fastest (kron order z, y, x)."""
    M, K = _cell_1d(k)
    h = 1.0 / (1 << level)
    if dim == 2:
        A = np.kron(M, K) + np.kron(K, M)
    else:
        A = np.kron(np.kron(M, M), K) + np.kron(np.kron(M, K), M) + np.kron(np.kron(K, M), M)
    return h ** (dim - 2) * A


def dense_full(dim, k, level, coef):
    """Unconstrained assembled sum_K c_K A_K."""
    N = n_dofs(dim, k, level)
    A = np.zeros((N, N))
    Ac = _cell_matrix(dim, k, level)
    for K, cx, cy, cz in _cells(dim, level):
        d = _cell_dofs(dim, k, level, cx, cy, cz)
        A[np.ix_(d, d)] += (1.0 if coef is None else coef[K]) * Ac
    return A


def dense_constrained(dim, k, level, coef):
    A = dense_full(dim, k, level, coef)
    bm = boundary_mask(dim, k, level)
    A[bm, :] = 0.0
    A[bm, bm] = 1.0
    return A


def _patches(dim, level, colour):
    n = 1 << level
    px, py, pz = colour & 1, (colour >> 1) & 1, (colour >> 2) & 1
    for iz in (range(pz, n - 1, 2) if dim == 3 else [0]):
        for iy in range(py, n - 1, 2):
            for ix in range(px, n - 1, 2):
                yield ix, iy, iz


SEP_STEP, SEP_HALF, SEP_THETA = 1.0 / 16.0, 64, 1.5


def separable_fit(dim, cells, coef):
    """Per-cell weights f_K of the rank-1 fit of the patch's cell
    coefficients (cells listed with the x offset fastest) and the scale
    sigma: log-ratios per direction rounded to the grid, then
    sigma = max(max c/f / theta, exp(mean log(c/f)))."""
    lc = np.log(np.array([coef[c] for c in cells]))
    nc = len(cells)
    bits = np.array([[(c >> a) & 1 for a in range(dim)] for c in range(nc)])
    lr = np.array([(lc[bits[:, a] == 1].mean() - lc[bits[:, a] == 0].mean())
                   for a in range(dim)])
    q = np.clip(np.rint(lr / SEP_STEP), -SEP_HALF, SEP_HALF) * SEP_STEP
    lf = bits @ q
    sigma = max(np.exp((lc - lf).max()) / SEP_THETA, np.exp((lc - lf).mean()))
    return np.exp(lf), sigma


def dense_coeff_sweep(dim, k, level, coef, b, u, reverse=False, relax=1.0):
    """Block Gauss-Seidel over the colours; local solve with the dense
    sigma sum_K f_K A_K over the patch's cells (interior rows and columns),
    the separable fit of the patch's coefficients, assembled from the golden
    cell matrices (no eigenvectors)."""
    A = dense_constrained(dim, k, level, coef)
    Ac = _cell_matrix(dim, k, level)
    n1 = k * (1 << level) + 1
    n = 1 << level
    u = u.copy()
    cols = list(range(1 << dim))
    if dim == 2:
        cols = [c for c in cols if c < 4]
    if reverse:
        cols = cols[::-1]
    for col in cols:
        for ix, iy, iz in _patches(dim, level, col):
            i = np.arange(1, 2 * k)
            if dim == 2:
                I = ((k * iy + i)[:, None] * n1 + (k * ix + i)[None, :]).ravel()
                cells = [(ix + a) + n * (iy + b_) for b_ in (0, 1) for a in (0, 1)]
            else:
                I = ((k * iz + i)[:, None, None] * n1 * n1 + (k * iy + i)[None, :, None] * n1
                     + (k * ix + i)[None, None, :]).ravel()
                cells = [(ix + a) + n * (iy + b_) + n * n * (iz + c)
                         for c in (0, 1) for b_ in (0, 1) for a in (0, 1)]
            f, sigma = separable_fit(dim, cells, coef)
            At = np.zeros((A.shape[0], A.shape[0]))
            for w, K in zip(f, cells):
                cx, cy = K % n, (K // n) % n
                cz = K // (n * n) if dim == 3 else 0
                d = _cell_dofs(dim, k, level, cx, cy, cz)
                At[np.ix_(d, d)] += sigma * w * Ac
            r = b - A @ u
            u[I] += relax * np.linalg.solve(At[np.ix_(I, I)], r[I])
    return u


def _coef(dim, level, seed=5):
    return np.exp(0.5 * rng(seed).standard_normal((1 << level) ** dim))


CASES = [(3, 1, 2), (3, 2, 1), (3, 2, 2), (2, 2, 2), (2, 3, 2)]


@pytest.mark.parametrize("dim,k,level", CASES)
def test_coeff_vmult_dense(dim, k, level):
    c = _coef(dim, level)
    x = rng(1).uniform(-1, 1, n_dofs(dim, k, level))
    A = dense_constrained(dim, k, level, c)
    assert rel_err(O.coeff_vmult(dim, k, level, c, x), A @ x) < 1e-13
    assert rel_err(O.coeff_vmult(dim, k, level, c, x, transpose=True), A.T @ x) < 1e-13


@pytest.mark.parametrize("dim,k,level", [(3, 2, 2), (2, 3, 3)])
def test_coeff_vmult_unit_and_scaled(dim, k, level):
    x = rng(2).uniform(-1, 1, n_dofs(dim, k, level))
    ref = O.laplace_vmult(dim, k, level, x)
    ones = np.ones((1 << level) ** dim)
    assert np.array_equal(O.coeff_vmult(dim, k, level, None, x), ref)
    assert rel_err(O.coeff_vmult(dim, k, level, ones, x), ref) < 1e-15
    bm = boundary_mask(dim, k, level)
    y = O.coeff_vmult(dim, k, level, 3.0 * ones, x)
    assert rel_err(y[~bm], 3.0 * ref[~bm]) < 1e-14
    assert np.array_equal(y[bm], x[bm])


def test_coeff_transpose_adjoint_identity():
    dim, k, level = 3, 3, 2
    c = _coef(dim, level, 9)
    g = rng(3)
    x, y = g.uniform(-1, 1, (2, n_dofs(dim, k, level)))
    lhs = O.coeff_vmult(dim, k, level, c, x) @ y
    rhs = x @ O.coeff_vmult(dim, k, level, c, y, transpose=True)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


@pytest.mark.parametrize("dim,k,level", [(3, 1, 2), (3, 2, 2), (2, 2, 3), (2, 3, 2)])
@pytest.mark.parametrize("reverse", [False, True])
def test_coeff_smoother_dense(dim, k, level, reverse):
    c = _coef(dim, level, 11)
    g = rng(4)
    b, u = g.uniform(-1, 1, (2, n_dofs(dim, k, level)))
    ref = dense_coeff_sweep(dim, k, level, c, b, u, reverse=reverse, relax=0.8)
    got = O.coeff_smoother_step(dim, k, level, c, b, u,
                                variant=O.MULT_REV if reverse else O.MULT_FWD, relax=0.8)
    assert rel_err(got, ref) < 1e-12


def _exact_coeff_sweep(dim, k, level, coef, b, u):
    """Block Gauss-Seidel with exact local solves A_P,II^{-1}."""
    A = dense_constrained(dim, k, level, coef)
    n1, n = k * (1 << level) + 1, 1 << level
    u = u.copy()
    for col in range(1 << dim):
        if dim == 2 and col >= 4:
            continue
        for ix, iy, iz in _patches(dim, level, col):
            i = np.arange(1, 2 * k)
            if dim == 2:
                I = ((k * iy + i)[:, None] * n1 + (k * ix + i)[None, :]).ravel()
            else:
                I = ((k * iz + i)[:, None, None] * n1 * n1 + (k * iy + i)[None, :, None] * n1
                     + (k * ix + i)[None, None, :]).ravel()
            r = b - A @ u
            u[I] += np.linalg.solve(A[np.ix_(I, I)], r[I])
    return u


@pytest.mark.parametrize("dim,k,level", [(3, 2, 2), (3, 1, 3), (2, 3, 3)])
def test_coeff_smoother_separable_coefficient_is_exact(dim, k, level):
    """A separable coefficient c = p(cx) q(cy) r(cz) whose neighbour ratios lie
    on the fit's grid makes every patch's fit exact: the sweep is block
    Gauss-Seidel with exact local solves."""
    n = 1 << level
    g = rng(13)
    logs = [SEP_STEP * g.integers(-12, 13, n) for _ in range(dim)]
    cz_ = np.exp(logs[2]) if dim == 3 else np.ones(1)
    c = (cz_[:, None, None] * np.exp(logs[1])[None, :, None]
         * np.exp(logs[0])[None, None, :]).ravel()
    b, u = g.uniform(-1, 1, (2, n_dofs(dim, k, level)))
    got = O.coeff_smoother_step(dim, k, level, c, b, u)
    assert rel_err(got, _exact_coeff_sweep(dim, k, level, c, b, u)) < 1e-11


@pytest.mark.parametrize("k", [1, 2, 3])
def test_separable_local_solve_bounds(k):
    """For random patch coefficients (lognormal, sigma up to 1.5) the local
    operator sigma A~ of the fit satisfies A_P,II <= 1.5 sigma A~ (largest
    generalized eigenvalue at most the fit's theta), the bound that keeps the
    symmetric V-cycle positive definite."""
    dim, level = 3, 1
    n1 = 2 * k + 1
    Ac = _cell_matrix(dim, k, level)
    i = np.arange(1, 2 * k)
    I = ((i)[:, None, None] * n1 * n1 + i[None, :, None] * n1 + i[None, None, :]).ravel()
    g = rng(17)
    worst = 0.0
    for trial in range(40):
        sd = [0.25, 0.5, 1.0, 1.5][trial % 4]
        c = np.exp(sd * g.standard_normal(8))
        f, sigma = separable_fit(dim, list(range(8)), c)
        A = np.zeros((n1 ** 3, n1 ** 3))
        At = np.zeros_like(A)
        for K in range(8):
            d = _cell_dofs(dim, k, level, K & 1, (K >> 1) & 1, K >> 2)
            A[np.ix_(d, d)] += c[K] * Ac
            At[np.ix_(d, d)] += sigma * f[K] * Ac
        Ai, Ati = A[np.ix_(I, I)], At[np.ix_(I, I)]
        L = np.linalg.cholesky(Ati)
        Li = np.linalg.inv(L)
        lam = np.linalg.eigvalsh(Li @ Ai @ Li.T)
        worst = max(worst, lam.max())
        assert lam.max() <= SEP_THETA * (1 + 1e-12)
        assert lam.min() > 0
    assert worst > 1.0


@pytest.mark.parametrize("dim,k,level", [(3, 2, 2), (2, 3, 3)])
def test_coeff_smoother_unit_coefficient_is_exact_smoother(dim, k, level):
    g = rng(6)
    b, u = g.uniform(-1, 1, (2, n_dofs(dim, k, level)))
    ones = np.ones((1 << level) ** dim)
    got = O.coeff_smoother_step(dim, k, level, ones, b, u)
    ref = O.smoother_step(dim, k, level, b, u)
    assert rel_err(got, ref) < 1e-12


def test_coeff_coarsen_is_child_mean():
    for dim in (2, 3):
        f = _coef(dim, 3, 12)
        n = 4
        shp = (n, 2, n, 2, n, 2) if dim == 3 else (n, 2, n, 2)
        ref = f.reshape(shp).mean(axis=(1, 3, 5) if dim == 3 else (1, 3)).ravel()
        assert rel_err(O.coeff_coarsen(dim, 2, f), ref) < 1e-15


@pytest.mark.parametrize("dim,k,level", [(3, 2, 2), (3, 3, 2), (2, 3, 3)])
def test_coeff_pcg_matches_direct_solve(dim, k, level):
    c = _coef(dim, level, 13)
    b = O.rhs_assembled(dim, k, level)
    x, it, rel, hist = O.coeff_pcg(dim, k, level, c, b, rtol=1e-13, maxit=200)
    xs = np.linalg.solve(dense_constrained(dim, k, level, c), b)
    assert rel <= 1e-13
    assert rel_err(x, xs) < 1e-11
    assert it <= 60
    # the V-cycle is a symmetric operator (reversed post-smoothing)
    g = rng(14)
    r, s = g.uniform(-1, 1, (2, b.size))
    bm = boundary_mask(dim, k, level)
    r[bm] = s[bm] = 0
    v1 = O.coeff_vcycle(dim, k, level, c, r) @ s
    v2 = r @ O.coeff_vcycle(dim, k, level, c, s)
    assert abs(v1 - v2) < 1e-11 * abs(v1)


def _dense_J(A, b, ud, bm):
    u = np.linalg.solve(A, b)
    d = (u - ud)[~bm]
    return 0.5 * d @ d, u


@pytest.mark.parametrize("dim,k,level", [(3, 2, 1), (3, 1, 2), (2, 3, 2)])
def test_coeff_gradient_matches_finite_differences(dim, k, level):
    c = _coef(dim, level, 15)
    g = rng(16)
    N = n_dofs(dim, k, level)
    bm = boundary_mask(dim, k, level)
    b = g.uniform(-1, 1, N)
    b[bm] = 0.0
    ud = 0.01 * g.uniform(0, 1, N)
    J0, u = _dense_J(dense_constrained(dim, k, level, c), b, ud, bm)
    A = dense_constrained(dim, k, level, c)
    lam = np.linalg.solve(A.T, np.where(bm, 0.0, u - ud))
    grad = O.coeff_gradient(dim, k, level, lam, u)
    h = 1e-6
    for K in range(c.size):
        cp, cm = c.copy(), c.copy()
        cp[K] += h
        cm[K] -= h
        fd = (_dense_J(dense_constrained(dim, k, level, cp), b, ud, bm)[0]
              - _dense_J(dense_constrained(dim, k, level, cm), b, ud, bm)[0]) / (2 * h)
        assert abs(grad[K] - fd) <= 1e-7 * max(abs(fd), 1e-3 * np.abs(grad).max())


def test_coeff_gradient_sum_identity():
    """sum_K c_K dJ/dc_K = -lambda~ . (A(c) u) restricted to the interior rows."""
    dim, k, level = 3, 3, 2
    c = _coef(dim, level, 17)
    g = rng(18)
    lam, u = g.uniform(-1, 1, (2, n_dofs(dim, k, level)))
    bm = boundary_mask(dim, k, level)
    grad = O.coeff_gradient(dim, k, level, lam, u)
    Au = O.coeff_vmult(dim, k, level, c, u)
    ref = -(lam[~bm] @ Au[~bm])
    assert abs(c @ grad - ref) < 1e-11 * abs(ref)


def test_coeff_problem_generators_agree():
    from gputensorproductsmoother_amd import problems
    dim, k, level = 3, 4, 3
    c1, u1, f1 = O.coeff_problem(dim, k, level)
    c2, u2, f2 = problems.coeff_problem(dim, k, level)
    assert rel_err(c2, c1) < 1e-15
    assert np.array_equal(u1, u2)
    assert np.array_equal(f1, f2)
    assert len(set(f1.tolist())) == 16
    # c = exp(0.5 g), g ~ N(0, 1): log c has mean ~0 and std ~0.5
    big, _, _ = problems.coeff_problem(3, 1, 5)
    lg = np.log(big)
    assert abs(lg.mean()) < 0.02 and abs(lg.std() - 0.5) < 0.01
    assert u1.min() >= 0.0 and u1.max() <= 0.01


def test_problem_rhs_generators_agree():
    from gputensorproductsmoother_amd import problems
    for dim, k, level in [(2, 2, 3), (3, 3, 2), (3, 4, 2)]:
        assert np.array_equal(problems.rhs_random(dim, k, level), O.rhs_random(dim, k, level))
        M = np.array(GOLD[str(k)]["M"], dtype=np.float64)
        assert rel_err(problems.rhs_constant(dim, k, level, M),
                       O.rhs_assembled(dim, k, level)) < 1e-14
