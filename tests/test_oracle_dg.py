"""Pin the oracle's SIPG-DG restatement (cell + face quadrature loops,
vertex-patch smoother, DG transfer, multigrid, PCG) without trusting its own
formulation:

* the operator equals h^(d-2) sum_d L_d (x) M (x) M assembled in numpy from
  the 50-digit golden cell matrices and the face formulas of the symmetric
  interior penalty method evaluated with an independent numpy Lagrange
  derivative (barycentric form), and it is symmetric positive definite;
* consistency: for k >= 2 the exact solution prod x_d (1 - x_d) lies in the
  DG space, so the discrete solution reproduces it to round-off;
* convergence: L2 error O(h^(k+1)) on the sine problem;
* the smoother sweep equals dense block Gauss-Seidel over the 3^d colours
  of vertex patches with direct local solves;
* prolongation is exact on coarse polynomials and restriction is its
  transpose; the V-cycle is symmetric; PCG matches a dense solve.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from helpers import rel_err, rng

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "basis_1d.json")))


def _lagrange_deriv_at(nodes, i, x):
    """d/dx of the i-th Lagrange polynomial on `nodes` at x (product rule)."""
    n = len(nodes)
    s = 0.0
    for m in range(n):
        if m == i:
            continue
        p = 1.0 / (nodes[i] - nodes[m])
        for l in range(n):
            if l != i and l != m:
                p *= (x - nodes[l]) / (nodes[i] - nodes[l])
        s += p
    return s


def numpy_L1d(k, n, pf):
    """Global 1D SIPG matrix (unit-scaled) from golden K and face formulas."""
    g = GOLD[str(k)]
    K = np.array(g["K"], dtype=np.float64)
    nodes = np.array(g["nodes"], dtype=np.float64)
    n1 = k + 1
    v0 = np.array([1.0 if i == 0 else 0.0 for i in range(n1)])
    v1 = np.array([1.0 if i == k else 0.0 for i in range(n1)])
    d0 = np.array([_lagrange_deriv_at(nodes, i, 0.0) for i in range(n1)])
    d1 = np.array([_lagrange_deriv_at(nodes, i, 1.0) for i in range(n1)])
    N = n * n1
    L = np.zeros((N, N))
    for c in range(n):
        L[c * n1:(c + 1) * n1, c * n1:(c + 1) * n1] += K
    for c in range(n - 1):
        # jump and average of (minus = cell c at x=1, plus = cell c+1 at x=0)
        jump = np.concatenate([v1, -v0])
        avg = 0.5 * np.concatenate([d1, d0])
        blk = pf * np.outer(jump, jump) - np.outer(jump, avg) - np.outer(avg, jump)
        L[c * n1:(c + 2) * n1, c * n1:(c + 2) * n1] += blk
    L[:n1, :n1] += 2 * pf * np.outer(v0, v0) + np.outer(v0, d0) + np.outer(d0, v0)
    o = (n - 1) * n1
    L[o:, o:] += 2 * pf * np.outer(v1, v1) - np.outer(v1, d1) - np.outer(d1, v1)
    return L


def numpy_dg_matrix(dim, k, n, pf):
    g = GOLD[str(k)]
    M = np.kron(np.eye(n), np.array(g["M"], dtype=np.float64))
    L = numpy_L1d(k, n, pf)
    h = 1.0 / n
    if dim == 2:
        A = np.kron(M, L) + np.kron(L, M)
    else:
        A = np.kron(np.kron(M, M), L) + np.kron(np.kron(M, L), M) + np.kron(np.kron(L, M), M)
    return h ** (dim - 2) * A


def oracle_matrix(dim, k, n, pf):
    N = O.dg_n_dofs(dim, k, n)
    return np.stack([O.dg_vmult(dim, k, n, e, pf) for e in np.eye(N)], axis=1)


@pytest.mark.parametrize("dim,k,n", [(2, 1, 3), (2, 3, 3), (3, 1, 3), (3, 2, 2)])
def test_dg_operator_is_kronecker_sum(dim, k, n):
    pf = O.dg_penalty(k)
    A = oracle_matrix(dim, k, n, pf)
    Aref = numpy_dg_matrix(dim, k, n, pf)
    assert np.abs(A - Aref).max() < 1e-12 * np.abs(Aref).max()
    assert np.abs(A - A.T).max() < 1e-12 * np.abs(A).max()
    assert np.linalg.eigvalsh(A).min() > 0
    assert np.abs(O.dg_global_1d(k, n, pf) - numpy_L1d(k, n, pf)).max() < 1e-12


def test_dg_reference_penalty_form():
    """The penalty factor is a plain parameter: the reference's k(k+1)
    (tests/mf_dg.cu:78-81) gives the same Kronecker structure."""
    k, n = 3, 3
    pf = k * (k + 1.0)
    assert np.abs(oracle_matrix(2, k, n, pf) - numpy_dg_matrix(2, k, n, pf)).max() < 1e-11


def _poly_rhs(dim, k, n):
    """b_i = int f phi_i for f = -lap prod x(1-x) (numpy, Gauss k+3)."""
    g = GOLD[str(k)]
    nodes = np.array(g["nodes"], dtype=np.float64)
    qp, qw = np.polynomial.legendre.leggauss(k + 3)
    qp, qw = 0.5 * (qp + 1), 0.5 * qw
    n1 = k + 1
    phi = np.array([[np.prod([(x - nodes[m]) / (nodes[i] - nodes[m]) for m in range(n1) if m != i])
                     for i in range(n1)] for x in qp])          # [q, i]
    h = 1.0 / n
    # per direction: m0_i = int phi_i * x(1-x), m1_i = int phi_i * 1 (cell-wise)
    m_p = np.zeros(n * n1)
    m_1 = np.zeros(n * n1)
    for c in range(n):
        x = (c + qp) * h
        m_p[c * n1:(c + 1) * n1] = h * (qw * x * (1 - x)) @ phi
        m_1[c * n1:(c + 1) * n1] = h * qw @ phi
    if dim == 2:
        b = 2 * (np.multiply.outer(m_p, m_1) + np.multiply.outer(m_1, m_p))
    else:
        b = 2 * (np.multiply.outer(m_p, np.multiply.outer(m_p, m_1))
                 + np.multiply.outer(m_p, np.multiply.outer(m_1, m_p))
                 + np.multiply.outer(m_1, np.multiply.outer(m_p, m_p)))
    u = np.zeros(n * n1)
    for c in range(n):
        x = (c + nodes) * h
        u[c * n1:(c + 1) * n1] = x * (1 - x)
    if dim == 2:
        ue = np.multiply.outer(u, u)
    else:
        ue = np.multiply.outer(u, np.multiply.outer(u, u))
    return b.ravel(), ue.ravel()


@pytest.mark.parametrize("dim,k,n", [(2, 2, 3), (2, 3, 6), (3, 2, 3)])
def test_dg_consistency_exact_polynomial(dim, k, n):
    b, ue = _poly_rhs(dim, k, n)
    lvl = {3: 0, 6: 1}[n]
    x, it, rel, _ = O.dg_pcg(dim, k, 3, lvl, b, rtol=1e-14, maxit=300)
    assert rel_err(x, ue) < 1e-11


def test_dg_convergence_rate():
    dim, k = 2, 2
    errs = []
    for lvl, n in enumerate((3, 6, 12, 24)):
        b = O.dg_rhs(dim, k, n)
        x, _, _, _ = O.dg_pcg(dim, k, 3, lvl, b, rtol=1e-13, maxit=300)
        errs.append(O.dg_l2_error_sine(dim, k, n, x))
    rates = [np.log2(errs[i] / errs[i + 1]) for i in range(3)]
    assert min(rates) > k + 1 - 0.1


def dense_dg_sweep(dim, k, n, pf, b, u, reverse=False):
    A = numpy_dg_matrix(dim, k, n, pf)
    n1 = k + 1
    N1 = n * n1
    u = u.copy()
    cols = [(cx, cy, cz) for cz in (range(3) if dim == 3 else [0]) for cy in range(3)
            for cx in range(3)]
    if reverse:
        cols = cols[::-1]
    for (cx, cy, cz) in cols:
        for pz in (range(cz, n - 1, 3) if dim == 3 else [0]):
            for py in range(cy, n - 1, 3):
                for px in range(cx, n - 1, 3):
                    r = np.arange(2 * n1)
                    X = px * n1 + r
                    Y = py * n1 + r
                    if dim == 2:
                        I = (Y[:, None] * N1 + X[None, :]).ravel()
                    else:
                        Z = pz * n1 + r
                        I = (Z[:, None, None] * N1 * N1 + Y[None, :, None] * N1
                             + X[None, None, :]).ravel()
                    res = b - A @ u
                    u[I] += np.linalg.solve(A[np.ix_(I, I)], res[I])
    return u


@pytest.mark.parametrize("dim,k,n", [(2, 2, 5), (2, 1, 7), (3, 1, 4)])
@pytest.mark.parametrize("reverse", [False, True])
def test_dg_smoother_is_block_gauss_seidel(dim, k, n, reverse):
    pf = O.dg_penalty(k)
    g = rng(5)
    N = O.dg_n_dofs(dim, k, n)
    b, u = g.uniform(-1, 1, (2, N))
    ref = dense_dg_sweep(dim, k, n, pf, b, u, reverse)
    got = O.dg_smoother_step(dim, k, n, b, u, variant=O.MULT_REV if reverse else O.MULT_FWD)
    assert rel_err(got, ref) < 1e-11


@pytest.mark.parametrize("dim,k", [(2, 2), (3, 1), (3, 3)])
def test_dg_transfer_exact_and_adjoint(dim, k):
    nc, nf = 3, 6
    g = GOLD[str(k)]
    nodes = np.array(g["nodes"], dtype=np.float64)
    n1 = k + 1

    def grid_coords(n):
        return np.concatenate([(c + nodes) / n for c in range(n)])

    rr = rng(7)
    coeffs = rr.uniform(-1, 1, (k + 1, k + 1)) if dim == 2 else rr.uniform(-1, 1, (k + 1,) * 3)

    def poly(*xs):
        # sum c_ab.. x^a y^b (z^c), degree <= k per variable
        if dim == 2:
            X, Y = xs
            return np.polynomial.polynomial.polyval2d(X, Y, coeffs)
        X, Y, Z = xs
        return np.polynomial.polynomial.polyval3d(X, Y, Z, coeffs)

    def sample(n):
        x = grid_coords(n)
        if dim == 2:
            Y, X = np.meshgrid(x, x, indexing="ij")
            return poly(X, Y).ravel()
        Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
        return poly(X, Y, Z).ravel()

    pc = sample(nc)
    pfine = O.dg_prolongate_add(dim, k, nf, pc, np.zeros(O.dg_n_dofs(dim, k, nf)))
    assert rel_err(pfine, sample(nf)) < 1e-12
    a = rr.uniform(-1, 1, O.dg_n_dofs(dim, k, nc))
    bf = rr.uniform(-1, 1, O.dg_n_dofs(dim, k, nf))
    lhs = O.dg_prolongate_add(dim, k, nf, a, np.zeros_like(bf)) @ bf
    rhs = a @ O.dg_restrict_add(dim, k, nf, bf, np.zeros_like(a))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    del n1


@pytest.mark.parametrize("dim,k,lvl", [(2, 2, 2), (3, 1, 1)])
def test_dg_vcycle_symmetric_and_pcg_direct(dim, k, lvl):
    n = 3 << lvl
    N = O.dg_n_dofs(dim, k, n)
    g = rng(9)
    r, s = g.uniform(-1, 1, (2, N))
    v1 = O.dg_vcycle(dim, k, 3, lvl, r) @ s
    v2 = r @ O.dg_vcycle(dim, k, 3, lvl, s)
    assert abs(v1 - v2) < 1e-11 * abs(v1)
    b = g.uniform(-1, 1, N)
    x, it, rel, _ = O.dg_pcg(dim, k, 3, lvl, b, rtol=1e-13, maxit=300)
    xs = np.linalg.solve(numpy_dg_matrix(dim, k, n, O.dg_penalty(k)), b)
    assert rel_err(x, xs) < 1e-10
    assert it <= 20
