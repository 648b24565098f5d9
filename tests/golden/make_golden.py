#!/usr/bin/env python3
"""Generate the golden fixtures under tests/golden/.

Independent of the C++ oracle and of the HIP library:

* basis_1d.json - for k = 1..7, with 50-digit mpmath: Gauss-Lobatto nodes on
  [0,1] (roots of P_k'), cell mass and stiffness integrals of the Lagrange
  basis by adaptive tanh-sinh quadrature (not a Gauss rule), the
  (2k+1) x (k+1) embedding matrix (coarse basis at the fine nodes of the two
  children) and the eigenvalues of the 2-cell patch interior pencil
  K_I s = lambda M_I s (Cholesky + mpmath eigsy).
* dense_<dim>d_q<k>_l<level>.npz - small systems restated with dense numpy
  float64 algebra from the basis_1d data: the global matrix as a Kronecker
  sum with identity Dirichlet rows (the reference keeps boundary columns,
  dealii/cuda_fee.cuh:273-290, and copies Dirichlet rows,
  dealii/cuda_mf.template.cuh:494-508), one multiplicative sweep as block
  Gauss-Seidel over the 2^dim patch colours with direct dense local solves
  (the reference's "Exact" local solver idea, loop_kernel.cuh:248-327), the
  additive sweep, prolongation as a Kronecker product of 1D interpolation
  matrices and restriction as its transpose, and the direct solution of
  A x = b.

Run: python tests/golden/make_golden.py  (a few seconds)
"""
from __future__ import annotations

import json
import os

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
mp.mp.dps = 50
KMAX = 7


def legendre_coeffs(n):
    """Exact coefficients (ascending powers) of P_n by Bonnet's recurrence."""
    from fractions import Fraction
    p0, p1 = [Fraction(1)], [Fraction(0), Fraction(1)]
    if n == 0:
        return p0
    for m in range(1, n):
        nxt = [Fraction(0)] * (m + 2)
        for i, c in enumerate(p1):
            nxt[i + 1] += Fraction(2 * m + 1, m + 1) * c
        for i, c in enumerate(p0):
            nxt[i] -= Fraction(m, m + 1) * c
        p0, p1 = p1, nxt
    return p1


def gl_nodes(k):
    """0, 1 and the roots of P_k' mapped from [-1,1] to [0,1]."""
    if k == 1:
        return [mp.mpf(0), mp.mpf(1)]
    c = legendre_coeffs(k)
    d = [i * c[i] for i in range(1, len(c))]  # P_k', ascending
    roots = mp.polyroots([mp.mpf(x.numerator) / x.denominator for x in reversed(d)],
                         maxsteps=200, extraprec=200)
    xs = sorted((mp.re(r) + 1) / 2 for r in roots)
    return [mp.mpf(0)] + xs + [mp.mpf(1)]


def lagrange(nodes, i, x):
    v = mp.mpf(1)
    for j, xj in enumerate(nodes):
        if j != i:
            v *= (x - xj) / (nodes[i] - xj)
    return v


def lagrange_d(nodes, i, x):
    s = mp.mpf(0)
    for m, xm in enumerate(nodes):
        if m == i:
            continue
        v = 1 / (nodes[i] - xm)
        for j, xj in enumerate(nodes):
            if j != i and j != m:
                v *= (x - xj) / (nodes[i] - xj)
        s += v
    return s


def basis(k):
    nodes = gl_nodes(k)
    n = k + 1
    M = mp.matrix(n, n)
    K = mp.matrix(n, n)
    for i in range(n):
        for j in range(i, n):
            M[i, j] = M[j, i] = mp.quad(lambda x: lagrange(nodes, i, x) * lagrange(nodes, j, x),
                                        [0, 1])
            K[i, j] = K[j, i] = mp.quad(lambda x: lagrange_d(nodes, i, x) * lagrange_d(nodes, j, x),
                                        [0, 1])
    E = mp.matrix(2 * k + 1, n)
    for q in range(2 * k + 1):
        x = nodes[q] / 2 if q <= k else (1 + nodes[q - k]) / 2
        for a in range(n):
            E[q, a] = lagrange(nodes, a, x)
    npch = 2 * k + 1
    MP = mp.matrix(npch, npch)
    KP = mp.matrix(npch, npch)
    for c in range(2):
        for i in range(n):
            for j in range(n):
                MP[c * k + i, c * k + j] += M[i, j]
                KP[c * k + i, c * k + j] += K[i, j]
    ni = 2 * k - 1
    MI = mp.matrix(ni, ni)
    KI = mp.matrix(ni, ni)
    for i in range(ni):
        for j in range(ni):
            MI[i, j] = MP[i + 1, j + 1]
            KI[i, j] = KP[i + 1, j + 1]
    L = mp.cholesky(MI)
    Li = L ** -1
    C = Li * KI * Li.T
    C = (C + C.T) / 2
    lam = sorted(mp.eigsy(C, eigvals_only=True))
    f = lambda A: [[float(A[i, j]) for j in range(A.cols)] for i in range(A.rows)]
    return dict(nodes=[float(x) for x in nodes], M=f(M), K=f(K), E=f(E),
                lam=[float(x) for x in lam])


# ------------------------------------------------------------ dense systems
def glob_1d(C, k, n):
    """Assemble a reference-cell matrix over n cells (no h scaling)."""
    N1 = k * n + 1
    G = np.zeros((N1, N1))
    for c in range(n):
        G[c * k:c * k + k + 1, c * k:c * k + k + 1] += C
    return G


def operator(dim, k, level, B):
    n = 1 << level
    h = 1.0 / n
    M = glob_1d(np.array(B["M"]), k, n)
    K = glob_1d(np.array(B["K"]), k, n)
    if dim == 2:
        A = np.kron(M, K) + np.kron(K, M)
    else:
        A = (np.kron(M, np.kron(M, K)) + np.kron(M, np.kron(K, M)) + np.kron(K, np.kron(M, M)))
        A *= h
    bd = boundary(dim, k, level)
    A[bd, :] = 0.0
    A[bd, bd] = 1.0
    return A


def boundary(dim, k, level):
    N1 = k * (1 << level) + 1
    e = np.zeros(N1, bool)
    e[[0, -1]] = True
    if dim == 2:
        return (e[:, None] | e[None, :]).ravel()
    return (e[:, None, None] | e[None, :, None] | e[None, None, :]).ravel()


def patches(dim, k, level, colour):
    """Interior global indices of every patch of a colour (x parity fastest)."""
    n = 1 << level
    N1 = k * n + 1
    par = [(colour >> j) & 1 for j in range(dim)]
    rng = [range(par[j], n - 1, 2) for j in range(dim)]
    out = []
    ii = np.arange(1, 2 * k)
    if dim == 2:
        for cy in rng[1]:
            for cx in rng[0]:
                X = k * cx + ii
                Y = k * cy + ii
                out.append((X[None, :] + N1 * Y[:, None]).ravel())
    else:
        for cz in rng[2]:
            for cy in rng[1]:
                for cx in rng[0]:
                    X, Y, Z = k * cx + ii, k * cy + ii, k * cz + ii
                    out.append((X[None, None, :] + N1 * Y[None, :, None]
                                + N1 * N1 * Z[:, None, None]).ravel())
    return out


def sweep(A, b, u, dim, k, level, mode, relax):
    u = u.copy()
    cols = list(range(1 << dim))
    if mode == "rev":
        cols = cols[::-1]
    r_fixed = b - A @ u if mode == "add" else None
    for c in cols:
        for I in patches(dim, k, level, c):
            AII = A[np.ix_(I, I)]
            if mode == "add":
                u[I] += relax * np.linalg.solve(AII, r_fixed[I])
            else:
                r = b[I] - A[I, :] @ u
                u[I] = relax * (u[I] + np.linalg.solve(AII, r))
    return u


def prolongation(dim, k, fine_level, B):
    nc = 1 << (fine_level - 1)
    Nc1, Nf1 = k * nc + 1, 2 * k * nc + 1
    E = np.array(B["E"])
    P1 = np.zeros((Nf1, Nc1))
    for j in range(Nf1):
        cc = min(j // (2 * k), nc - 1)
        q = j - 2 * k * cc
        P1[j, k * cc:k * cc + k + 1] = E[q]
    P = np.kron(P1, P1) if dim == 2 else np.kron(P1, np.kron(P1, P1))
    mask = ~boundary(dim, k, fine_level - 1)
    return P * mask[None, :]


CASES = [(2, 1, 3), (2, 2, 3), (2, 3, 2), (3, 1, 3), (3, 2, 2), (3, 3, 2), (3, 4, 1)]


def main():
    data = {str(k): basis(k) for k in range(1, KMAX + 1)}
    with open(os.path.join(HERE, "basis_1d.json"), "w") as f:
        json.dump(data, f, indent=1)
    for dim, k, level in CASES:
        B = data[str(k)]
        g = np.random.default_rng(1000 * dim + 10 * k + level)
        A = operator(dim, k, level, B)
        N = A.shape[0]
        bd = boundary(dim, k, level)
        x = g.uniform(-1, 1, N)
        b = g.uniform(-1, 1, N)
        u0 = g.uniform(-1, 1, N)
        out = dict(x=x, Ax=A @ x, b=b, u0=u0,
                   u_fwd=sweep(A, b, u0, dim, k, level, "fwd", 1.0),
                   u_rev=sweep(A, b, u0, dim, k, level, "rev", 1.0),
                   u_add=sweep(A, b, u0, dim, k, level, "add", 0.5))
        b0 = b.copy()
        b0[bd] = 0.0
        out["b0"] = b0
        out["xsol"] = np.linalg.solve(A, b0)
        if level >= 2:
            P = prolongation(dim, k, level, B)
            nc = P.shape[1]
            coarse = g.uniform(-1, 1, nc)
            fine0 = g.uniform(-1, 1, N)
            fine = g.uniform(-1, 1, N)
            coarse0 = g.uniform(-1, 1, nc)
            cmask = ~boundary(dim, k, level - 1)
            out.update(coarse=coarse, fine0=fine0, prolong=fine0 + P @ coarse, fine=fine,
                       coarse0=coarse0, restrict=coarse0 + cmask * (P.T @ fine))
        np.savez_compressed(os.path.join(HERE, f"dense_{dim}d_q{k}_l{level}.npz"), **out)
        print("wrote", dim, k, level, N)


if __name__ == "__main__":
    main()
