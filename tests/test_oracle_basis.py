"""Pin the oracle's 1D data (GL nodes, cell mass/stiffness, embedding,
patch eigenpairs) against closed forms, invariants and the 50-digit mpmath
golden fixture tests/golden/basis_1d.json (adaptive quadrature, not a Gauss
rule, so the Gauss(k+1)-exactness the oracle relies on is checked too)."""
import json
import os

import numpy as np
import pytest
import scipy.linalg

import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "basis_1d.json")))
KS = range(1, 8)


@pytest.mark.parametrize("k", KS)
def test_against_golden(k):
    b = O.basis_1d(k)
    g = GOLD[str(k)]
    assert np.abs(b["nodes"] - np.array(g["nodes"])).max() < 2e-16
    for key in ("M", "K", "E"):
        ref = np.array(g[key])
        assert np.abs(b[key] - ref).max() <= 1e-13 * np.abs(ref).max(), key
    lam = np.array(g["lam"])
    assert np.abs(b["lam"] - lam).max() <= 1e-12 * lam.max()


def test_closed_forms_q1_q2():
    b1 = O.basis_1d(1)
    assert np.allclose(b1["M"], np.array([[2, 1], [1, 2]]) / 6, atol=1e-16, rtol=0)
    assert np.allclose(b1["K"], np.array([[1, -1], [-1, 1]]), atol=1e-15, rtol=0)
    # Q1 two-cell patch interior: one DoF, lambda = K_I / M_I = 2 / (2/3)
    assert abs(b1["lam"][0] - 3.0) < 1e-14
    b2 = O.basis_1d(2)
    assert np.allclose(b2["nodes"], [0, 0.5, 1], atol=1e-16)
    assert np.allclose(b2["M"], np.array([[4, 2, -1], [2, 16, 2], [-1, 2, 4]]) / 30,
                       atol=1e-16, rtol=0)
    assert np.allclose(b2["K"], np.array([[7, -8, 1], [-8, 16, -8], [1, -8, 7]]) / 3,
                       atol=1e-14, rtol=0)
    # Q2 embedding: the midpoint node of a child is interpolated by the
    # coarse quadratic: values at x = 1/4 are (3/8, 3/4, -1/8)
    assert np.allclose(b2["E"][1], [3 / 8, 3 / 4, -1 / 8], atol=1e-15)


@pytest.mark.parametrize("k", KS)
def test_invariants(k):
    b = O.basis_1d(k)
    M, K, E = b["M"], b["K"], b["E"]
    assert np.allclose(M, M.T, atol=1e-16) and np.allclose(K, K.T, atol=1e-14)
    assert abs(M.sum() - 1.0) < 1e-14          # int_0^1 1 = 1
    assert np.abs(K @ np.ones(k + 1)).max() < 1e-12  # constants in the kernel
    assert np.all(np.linalg.eigvalsh(M) > 0)
    # K x = boundary flux of the linear function: (K nodes)_i = phi_i(1) - phi_i(0)
    flux = np.zeros(k + 1)
    flux[0], flux[-1] = -1.0, 1.0
    assert np.abs(K @ b["nodes"] - flux).max() < 1e-12
    # M x = first moments int phi_i x, and partition of unity in E
    assert abs(np.ones(k + 1) @ M @ b["nodes"] - 0.5) < 1e-14
    assert np.abs(E.sum(axis=1) - 1.0).max() < 1e-13
    # E interpolates degree-k polynomials exactly on the two children
    nodes = b["nodes"]
    fine = np.concatenate([nodes / 2, 0.5 + nodes[1:] / 2])
    for p in range(k + 1):
        assert np.abs(E @ nodes ** p - fine ** p).max() < 1e-12


@pytest.mark.parametrize("k", KS)
def test_gl_nodes_are_roots_of_dPk(k):
    nodes = O.basis_1d(k)["nodes"]
    if k > 1:
        dP = np.polynomial.legendre.Legendre.basis(k).deriv()
        roots = np.sort((dP.roots().real + 1) / 2)
        assert np.abs(nodes[1:-1] - roots).max() < 1e-14


@pytest.mark.parametrize("k", KS)
def test_patch_eigenpairs(k):
    b = O.basis_1d(k)
    MI, KI = b["MP"][1:-1, 1:-1], b["KP"][1:-1, 1:-1]
    S, lam = b["S"], b["lam"]
    n = 2 * k - 1
    assert np.abs(S.T @ MI @ S - np.eye(n)).max() < 1e-13
    assert np.abs(S.T @ KI @ S - np.diag(lam)).max() < 1e-12 * lam.max()
    ref = scipy.linalg.eigh(KI, MI, eigvals_only=True)
    assert np.abs(np.sort(lam) - ref).max() < 1e-11 * ref.max()
    # FD inverse equals the dense inverse of the interior Kronecker sum (2D, 3D)
    for dim in (2, 3):
        if dim == 2:
            A = np.kron(MI, KI) + np.kron(KI, MI)
        else:
            A = np.kron(MI, np.kron(MI, KI)) + np.kron(MI, np.kron(KI, MI)) + \
                np.kron(KI, np.kron(MI, MI))
        Sd = np.kron(S, S) if dim == 2 else np.kron(S, np.kron(S, S))
        L = (lam[:, None] + lam[None, :]).ravel() if dim == 2 else \
            (lam[:, None, None] + lam[None, :, None] + lam[None, None, :]).ravel()
        Ainv = Sd @ np.diag(1.0 / L) @ Sd.T
        assert np.abs(Ainv @ A - np.eye(n ** dim)).max() < 1e-10


def test_gen_eig_random_pencils():
    g = np.random.default_rng(5)
    for n in (1, 2, 5, 9, 13):
        X = g.standard_normal((n, n))
        M = X @ X.T + n * np.eye(n)
        Y = g.standard_normal((n, n))
        K = Y @ Y.T + 0.1 * np.eye(n)
        lam, S = O.gen_eig(K, M)
        assert np.abs(S.T @ M @ S - np.eye(n)).max() < 1e-12
        assert np.abs(S.T @ K @ S - np.diag(lam)).max() < 1e-11 * lam.max()
        assert np.all(np.diff(lam) >= 0)
        assert np.abs(lam - scipy.linalg.eigh(K, M, eigvals_only=True)).max() < 1e-11 * lam.max()
