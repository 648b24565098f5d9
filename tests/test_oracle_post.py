"""Pins of the oracle's post-processing (load-vector assembly from f at the
Gauss points, L2 / H1-seminorm errors against prod sin(w pi x_d)).

Each check is chosen so that a plausible slip fails it: closed-form norms
of the exact solution (u_h = 0), the separable 1D-product assembly already
pinned elsewhere, partition of unity, interpolation convergence orders
k+1 (L2) and k (H1), and CG/DG agreement on the same piecewise polynomial.
"""
import numpy as np
import pytest

import oracle as O


def gauss01(nq):
    x, w = np.polynomial.legendre.leggauss(nq)
    return 0.5 * (x + 1.0), 0.5 * w


def f_at_points(dim, k, n, f):
    """f at the QGauss(k+1) points of every cell, shape (n^dim, (k+1)^dim)."""
    q, _ = gauss01(k + 1)
    x1 = (np.arange(n)[:, None] + q[None, :]) / n           # [cell][q]
    if dim == 2:
        X = x1[None, :, None, :]
        Y = x1[:, None, :, None]
        v = f(X, Y, None)                                  # [cy][cx][qy][qx]
        return np.broadcast_to(v, (n, n, k + 1, k + 1)).reshape(n * n, -1)
    X = x1[None, None, :, None, None, :]
    Y = x1[None, :, None, None, :, None]
    Z = x1[:, None, None, :, None, None]
    v = np.broadcast_to(f(X, Y, Z), (n, n, n, k + 1, k + 1, k + 1))
    return v.reshape(n ** 3, -1)


def nodal_coords(k, n, dg):
    nodes = O.basis_1d(k)["nodes"]
    if dg:
        return ((np.arange(n)[:, None] + nodes[None, :]) / n).ravel()
    X = np.arange(k * n + 1)
    c = np.minimum(X // k, n - 1)
    return (c + nodes[X - k * c]) / n


def interpolant(dim, k, n, wave, dg=False):
    x1 = np.sin(wave * np.pi * nodal_coords(k, n, dg))
    out = np.multiply.outer(x1, x1)
    if dim == 3:
        out = np.multiply.outer(x1, out)
    return out.ravel()


@pytest.mark.parametrize("dim,k,n", [(2, 1, 4), (2, 3, 4), (3, 2, 2), (3, 4, 2)])
def test_rhs_constant_matches_separable_assembly(dim, k, n):
    level = int(np.log2(n))
    fq = f_at_points(dim, k, n, lambda x, y, z: np.ones_like(x))
    b = O.assemble_rhs(dim, k, n, fq)
    ref = O.rhs_assembled(dim, k, level, O.RHS_CONST)
    assert np.max(np.abs(b - ref)) < 1e-15 * 10


@pytest.mark.parametrize("dim,k,n", [(2, 2, 4), (3, 3, 2), (3, 1, 4)])
def test_rhs_sine_matches_separable_assembly(dim, k, n):
    level = int(np.log2(n))
    a = 3 * np.pi

    def f(x, y, z):
        v = dim * a * a * np.sin(a * x) * np.sin(a * y)
        return v if z is None else v * np.sin(a * z)
    b = O.assemble_rhs(dim, k, n, f_at_points(dim, k, n, f))
    ref = O.rhs_assembled(dim, k, level, O.RHS_SINE)
    assert np.max(np.abs(b - ref)) < 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("dim,k,n", [(2, 3, 3), (3, 2, 3)])
def test_dg_rhs_partition_of_unity(dim, k, n):
    """f = 1 on a DG space: sum_i b_i = int 1 = 1, and every cell's block
    sums to its volume h^dim."""
    b = O.assemble_rhs(dim, k, n, f_at_points(dim, k, n, lambda x, y, z: np.ones_like(x)),
                       dg=True)
    assert abs(b.sum() - 1.0) < 1e-14
    if dim == 3:
        ref = O.dg_rhs(3, k, n, O.DG_RHS_CONST)
        assert np.max(np.abs(b - ref)) < 1e-15 * 10


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("wave", [1.0, 3.0])
def test_error_norms_of_zero_are_closed_forms(dim, wave):
    """u_h = 0: ||u||_L2^2 = 2^-dim, |u|_H1^2 = dim (w pi)^2 2^-dim
    (each factor integrates sin^2 or cos^2 over [0,1] to 1/2)."""
    k, n = 2, 4
    u = np.zeros((k * n + 1) ** dim)
    l2, h1 = O.error_norms(dim, k, n, u, wave=wave, n_q=12)
    assert abs(l2 - 2.0 ** (-dim / 2)) < 1e-13
    assert abs(h1 - np.sqrt(dim * (wave * np.pi) ** 2 * 2.0 ** -dim)) < 1e-12


def test_error_norms_l2_agrees_with_independent_loop():
    k, n = 3, 4
    rng = np.random.default_rng(3)
    u = rng.uniform(-1, 1, (k * n + 1) ** 3)
    l2, _ = O.error_norms(3, k, n, u, wave=3.0, n_q=k + 3)
    assert abs(l2 - O.l2_error_sine(3, k, 2, u)) < 1e-13 * l2


@pytest.mark.parametrize("dim,k", [(2, 1), (2, 2), (3, 2), (3, 3)])
def test_interpolation_error_orders(dim, k):
    """Nodal interpolant of sin(pi x) sin(pi y) (sin(pi z)): L2 error falls
    like h^(k+1), H1 seminorm like h^k."""
    ns = (4, 8) if dim == 3 else (8, 16)
    errs = [O.error_norms(dim, k, n, interpolant(dim, k, n, 1.0), wave=1.0, n_q=k + 4)
            for n in ns]
    r0 = np.log2(errs[0][0] / errs[1][0])
    r1 = np.log2(errs[0][1] / errs[1][1])
    assert abs(r0 - (k + 1)) < 0.35
    assert abs(r1 - k) < 0.35


@pytest.mark.parametrize("dim,k,n", [(2, 2, 3), (3, 2, 2), (3, 3, 3)])
def test_dg_and_cg_layouts_agree_on_continuous_functions(dim, k, n):
    for wave in (1.0, 3.0):
        cg = O.error_norms(dim, k, n, interpolant(dim, k, n, wave), wave=wave)
        dg = O.error_norms(dim, k, n, interpolant(dim, k, n, wave, dg=True), wave=wave, dg=True)
        assert np.allclose(cg, dg, rtol=1e-13, atol=0)
