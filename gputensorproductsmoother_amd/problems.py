"""Synthetic test-problem data (numpy, deterministic, no CPU solver).

Every generator derives from SplitMix64, u01(j) = (splitmix64(seed ^ j) >>
11) 2^-53, so the same data can be produced on any host:

* ``rhs_random``: b_i = 2 u01(i) - 1 per DoF, zero on Dirichlet DoFs (the
  3D Q3 Poisson workload);
* ``rhs_constant``: the assembled load vector of f = 1, b_i = int phi_i,
  zero on Dirichlet DoFs (the reference's Poisson right-hand side with
  f = 1, built like its RightHandSide, apps/poisson.cu:44-87);
* ``coeff_problem``: per-cell coefficient c_K = exp(sigma g_K) with g_K
  standard normal by Box-Muller on stream 2^32, target u_d,i = 0.01 u01(2^33
  + i) on interior DoFs, and distinct finite-difference cells from stream
  2^34 (the variable-coefficient adjoint problem);
* ``rhs_dg_random``: b_i = 2 u01(i + m 2^36) - 1 on every DoF of a DG
  space (weak boundary conditions, no constrained rows), m = 0, 1, ... for
  the right-hand sides of a batch.
"""
from __future__ import annotations

import numpy as np

SEED = 20261202


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser of uint64 inputs (wrapping arithmetic)."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def u01(j: np.ndarray, seed: int = SEED) -> np.ndarray:
    j = np.asarray(j, dtype=np.uint64)
    return (splitmix64(j ^ np.uint64(seed)) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def boundary_mask(dim: int, k: int, level: int) -> np.ndarray:
    n1 = k * (1 << level) + 1
    e = np.zeros(n1, dtype=bool)
    e[0] = e[-1] = True
    if dim == 3:
        bd = e[:, None, None] | e[None, :, None] | e[None, None, :]
    else:
        bd = e[:, None] | e[None, :]
    return bd.ravel()


def rhs_random(dim: int, k: int, level: int, seed: int = SEED) -> np.ndarray:
    n = (k * (1 << level) + 1) ** dim
    b = 2.0 * u01(np.arange(n, dtype=np.uint64), seed) - 1.0
    b[boundary_mask(dim, k, level)] = 0.0
    return b


def rhs_constant(dim: int, k: int, level: int, cell_mass: np.ndarray) -> np.ndarray:
    """b = m (x) m (x) m with m_i = int phi_i^{1D}: the reference-cell mass
    matrix (on [0,1], e.g. from LaplaceOperator.tensor_product_data) times
    the ones vector, scaled by h and assembled over the cells."""
    n = 1 << level
    h = 1.0 / n
    mloc = np.asarray(cell_mass, dtype=np.float64).sum(axis=1) * h
    m = np.zeros(k * n + 1)
    for c in range(n):
        m[k * c: k * c + k + 1] += mloc
    b = np.multiply.outer(m, m)
    if dim == 3:
        b = np.multiply.outer(m, b)
    b = b.ravel().copy()
    b[boundary_mask(dim, k, level)] = 0.0
    return b


def coeff_problem(dim: int, k: int, level: int, seed: int = SEED, sigma: float = 0.5,
                  n_fd_cells: int = 16):
    """(coef, u_d, fd_cells) of the variable-coefficient problem."""
    n = 1 << level
    ncell = n ** dim
    K = np.arange(ncell, dtype=np.uint64)
    a = u01(np.uint64(1 << 32) + np.uint64(2) * K, seed)
    c = u01(np.uint64(1 << 32) + np.uint64(2) * K + np.uint64(1), seed)
    g = np.sqrt(-2.0 * np.log(1.0 - a)) * np.cos(2.0 * np.pi * c)
    coef = np.exp(sigma * g)
    nd = (k * n + 1) ** dim
    ud = 0.01 * u01(np.uint64(1 << 33) + np.arange(nd, dtype=np.uint64), seed)
    ud[boundary_mask(dim, k, level)] = 0.0
    cells = []
    j = 0
    while len(cells) < n_fd_cells:
        cell = int(u01(np.uint64(1 << 34) + np.uint64(j), seed) * ncell)
        if cell not in cells:
            cells.append(cell)
        j += 1
    return coef, ud, np.array(cells, dtype=np.int64)


def rhs_dg_random(dim: int, k: int, n_cells: int, batch: int = 1,
                  seed: int = SEED) -> np.ndarray:
    """Random DG right-hand sides, shape (batch, (n_cells (k+1))^dim)."""
    nd = (n_cells * (k + 1)) ** dim
    i = np.arange(nd, dtype=np.uint64)
    return np.stack([2.0 * u01(i + np.uint64(m << 36), seed) - 1.0 for m in range(batch)])
