"""ctypes front-end of the FP64 CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, always as the checker.  The product
library never routes through it.  The C++ restatement lives in
``oracle/gtps_oracle.cpp``; ``build()`` compiles it with the committed
Makefile.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libgtps_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def build() -> str:
    """Compile the oracle shared library in-tree (make -C oracle)."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "libgtps_oracle.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        i, d, ll, ull = ctypes.c_int, ctypes.c_double, ctypes.c_longlong, ctypes.c_ulonglong
        L.oracle_basis_1d.argtypes = [i] + [_dp] * 8
        L.oracle_gen_eig.argtypes = [i, _dp, _dp, _dp, _dp]
        L.oracle_n_dofs.argtypes = [i, i, i]
        L.oracle_n_dofs.restype = ll
        L.oracle_laplace_vmult.argtypes = [i, i, i, _dp, _dp]
        L.oracle_smoother_step.argtypes = [i, i, i, i, d, _dp, _dp]
        L.oracle_prolongate_add.argtypes = [i, i, i, _dp, _dp]
        L.oracle_restrict_add.argtypes = [i, i, i, _dp, _dp]
        L.oracle_vcycle.argtypes = [i, i, i, i, d, _dp, _dp]
        L.oracle_pcg.argtypes = [i, i, i, i, d, _dp, _dp, d, i,
                                 ctypes.POINTER(i), _dp, _dp]
        L.oracle_rhs_random.argtypes = [i, i, i, ull, _dp]
        L.oracle_rhs_assembled.argtypes = [i, i, i, i, _dp]
        L.oracle_l2_error_sine.argtypes = [i, i, i, _dp]
        L.oracle_l2_error_sine.restype = d
        L.oracle_set_threads.argtypes = [i]
        L.oracle_coeff_pcg.argtypes = [i, i, i, _dp, i, d, _dp, _dp, d, i,
                                       ctypes.POINTER(i), _dp, _dp]
        L.oracle_coeff_vmult.argtypes = [i, i, i, _dp, _dp, _dp, i]
        L.oracle_coeff_smoother_step.argtypes = [i, i, i, _dp, i, d, _dp, _dp]
        L.oracle_coeff_coarsen.argtypes = [i, i, _dp, _dp]
        L.oracle_coeff_vcycle.argtypes = [i, i, i, _dp, i, d, _dp, _dp]
        L.oracle_coeff_gradient.argtypes = [i, i, i, _dp, _dp, _dp]
        L.oracle_fmg.argtypes = [i, i, i, _dp, i, d, _dp, i, _dp]
        L.oracle_dg_n_dofs.argtypes = [i, i, i]
        L.oracle_dg_n_dofs.restype = ll
        L.oracle_dg_vmult.argtypes = [i, i, i, d, _dp, _dp]
        L.oracle_dg_global_1d.argtypes = [i, i, d, _dp]
        L.oracle_dg_smoother_step.argtypes = [i, i, i, d, i, d, _dp, _dp]
        L.oracle_dg_prolongate_add.argtypes = [i, i, i, _dp, _dp]
        L.oracle_dg_restrict_add.argtypes = [i, i, i, _dp, _dp]
        L.oracle_dg_vcycle.argtypes = [i, i, i, i, d, i, d, i, _dp, _dp]
        L.oracle_dg_pcg.argtypes = [i, i, i, i, d, i, d, i, _dp, _dp, d, i,
                                    ctypes.POINTER(i), _dp, _dp]
        L.oracle_dg_rhs.argtypes = [i, i, i, i, _dp]
        L.oracle_dg_l2_error_sine.argtypes = [i, i, i, _dp]
        L.oracle_dg_l2_error_sine.restype = d
        L.oracle_assemble_rhs.argtypes = [i, i, i, i, _dp, _dp]
        L.oracle_error_norms.argtypes = [i, i, i, i, d, i, _dp, ctypes.POINTER(d),
                                         ctypes.POINTER(d)]
        L.oracle_coeff_problem.argtypes = [i, i, i, ull, d, _dp, _dp, i,
                                           ctypes.POINTER(ll)]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _chk(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed (rc={rc})")


def n_dofs(dim: int, k: int, level: int) -> int:
    return int(lib().oracle_n_dofs(dim, k, level))


def basis_1d(k: int) -> dict:
    n, np_, ni = k + 1, 2 * k + 1, 2 * k - 1
    out = dict(nodes=np.zeros(n), M=np.zeros((n, n)), K=np.zeros((n, n)),
               E=np.zeros((np_, n)), MP=np.zeros((np_, np_)), KP=np.zeros((np_, np_)),
               lam=np.zeros(ni), S=np.zeros((ni, ni)))
    _chk(lib().oracle_basis_1d(k, *[_p(out[key]) for key in
                                    ("nodes", "M", "K", "E", "MP", "KP", "lam", "S")]),
         "basis_1d")
    return out


def gen_eig(K: np.ndarray, M: np.ndarray):
    n = K.shape[0]
    K = np.ascontiguousarray(K, dtype=np.float64)
    M = np.ascontiguousarray(M, dtype=np.float64)
    lam, S = np.zeros(n), np.zeros((n, n))
    _chk(lib().oracle_gen_eig(n, _p(K), _p(M), _p(lam), _p(S)), "gen_eig")
    return lam, S


def laplace_vmult(dim, k, level, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    _chk(lib().oracle_laplace_vmult(dim, k, level, _p(x), _p(y)), "vmult")
    return y


MULT_FWD, MULT_REV, ADDITIVE = 0, 1, 2


def smoother_step(dim, k, level, b, u, variant=MULT_FWD, relax=1.0):
    b = np.ascontiguousarray(b, dtype=np.float64)
    u = np.array(u, dtype=np.float64, copy=True)
    _chk(lib().oracle_smoother_step(dim, k, level, variant, relax, _p(b), _p(u)), "smoother")
    return u


def prolongate_add(dim, k, fine_level, coarse, fine):
    coarse = np.ascontiguousarray(coarse, dtype=np.float64)
    fine = np.array(fine, dtype=np.float64, copy=True)
    _chk(lib().oracle_prolongate_add(dim, k, fine_level, _p(coarse), _p(fine)), "prolongate")
    return fine


def restrict_add(dim, k, fine_level, fine, coarse):
    fine = np.ascontiguousarray(fine, dtype=np.float64)
    coarse = np.array(coarse, dtype=np.float64, copy=True)
    _chk(lib().oracle_restrict_add(dim, k, fine_level, _p(fine), _p(coarse)), "restrict")
    return coarse


MG_SYMMETRIC, MG_REFERENCE_ORDER, MG_ADDITIVE = 0, 1, 2


def vcycle(dim, k, level, r, mg_variant=MG_SYMMETRIC, relax=1.0):
    r = np.ascontiguousarray(r, dtype=np.float64)
    z = np.zeros_like(r)
    _chk(lib().oracle_vcycle(dim, k, level, mg_variant, relax, _p(r), _p(z)), "vcycle")
    return z


def pcg(dim, k, level, b, rtol=1e-8, maxit=100, mg_variant=MG_SYMMETRIC, relax=1.0):
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    it = ctypes.c_int(0)
    rel = ctypes.c_double(0.0)
    hist = np.zeros(maxit + 1)
    rc = lib().oracle_pcg(dim, k, level, mg_variant, relax, _p(b), _p(x), rtol, maxit,
                          ctypes.byref(it), ctypes.byref(rel), _p(hist))
    _chk(rc, "pcg")
    return x, it.value, rel.value, hist[: it.value + 1]


def rhs_random(dim, k, level, seed=20261202):
    b = np.zeros(n_dofs(dim, k, level))
    _chk(lib().oracle_rhs_random(dim, k, level, seed, _p(b)), "rhs_random")
    return b


RHS_CONST, RHS_SINE = 0, 1


def rhs_assembled(dim, k, level, kind=RHS_CONST):
    b = np.zeros(n_dofs(dim, k, level))
    _chk(lib().oracle_rhs_assembled(dim, k, level, kind, _p(b)), "rhs_assembled")
    return b


def l2_error_sine(dim, k, level, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    return float(lib().oracle_l2_error_sine(dim, k, level, _p(u)))


def set_threads(n: int) -> int:
    return int(lib().oracle_set_threads(n))


# ---------------------------------------------------------------- variable
# coefficient (cells numbered cx + n cy + n^2 cz, n = 2^level)

def n_cells(dim: int, level: int) -> int:
    return (1 << level) ** dim


def _coef(c):
    return None if c is None else _p(np.ascontiguousarray(c, dtype=np.float64))


def coeff_vmult(dim, k, level, coef, x, transpose=False):
    """y = A(c) x with identity Dirichlet rows (or its transpose)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    c = None if coef is None else np.ascontiguousarray(coef, dtype=np.float64)
    y = np.zeros_like(x)
    _chk(lib().oracle_coeff_vmult(dim, k, level, None if c is None else _p(c), _p(x), _p(y),
                                  1 if transpose else 0), "coeff_vmult")
    return y


def coeff_smoother_step(dim, k, level, coef, b, u, variant=MULT_FWD, relax=1.0):
    b = np.ascontiguousarray(b, dtype=np.float64)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    u = np.array(u, dtype=np.float64, copy=True)
    _chk(lib().oracle_coeff_smoother_step(dim, k, level, _p(c), variant, relax, _p(b), _p(u)),
         "coeff_smoother")
    return u


def coeff_coarsen(dim, coarse_level, fine):
    fine = np.ascontiguousarray(fine, dtype=np.float64)
    out = np.zeros(n_cells(dim, coarse_level))
    _chk(lib().oracle_coeff_coarsen(dim, coarse_level, _p(fine), _p(out)), "coeff_coarsen")
    return out


def coeff_vcycle(dim, k, level, coef, r, mg_variant=MG_SYMMETRIC, relax=1.0):
    r = np.ascontiguousarray(r, dtype=np.float64)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    z = np.zeros_like(r)
    _chk(lib().oracle_coeff_vcycle(dim, k, level, _p(c), mg_variant, relax, _p(r), _p(z)),
         "coeff_vcycle")
    return z


def coeff_pcg(dim, k, level, coef, b, rtol=1e-8, maxit=100, mg_variant=MG_SYMMETRIC,
              relax=1.0):
    b = np.ascontiguousarray(b, dtype=np.float64)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    x = np.zeros_like(b)
    it = ctypes.c_int(0)
    rel = ctypes.c_double(0.0)
    hist = np.zeros(maxit + 1)
    rc = lib().oracle_coeff_pcg(dim, k, level, _p(c), mg_variant, relax, _p(b), _p(x), rtol,
                                maxit, ctypes.byref(it), ctypes.byref(rel), _p(hist))
    _chk(rc, "coeff_pcg")
    return x, it.value, rel.value, hist[: it.value + 1]


def coeff_gradient(dim, k, level, lam, u):
    """dJ/dc_K = -lambda~_K^T A^1_K u_K (lambda's Dirichlet entries masked)."""
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    g = np.zeros(n_cells(dim, level))
    _chk(lib().oracle_coeff_gradient(dim, k, level, _p(lam), _p(u), _p(g)), "coeff_gradient")
    return g


def coeff_problem(dim, k, level, seed=20261202, sigma=0.5, n_fd_cells=16):
    """Coefficient exp(sigma g), target u_d and the finite-difference cells of
    the variable-coefficient problem (SplitMix64 streams 2^32, 2^33, 2^34)."""
    coef = np.zeros(n_cells(dim, level))
    ud = np.zeros(n_dofs(dim, k, level))
    cells = np.zeros(n_fd_cells, dtype=np.int64)
    _chk(lib().oracle_coeff_problem(dim, k, level, seed, sigma, _p(coef), _p(ud), n_fd_cells,
                                    cells.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))),
         "coeff_problem")
    return coef, ud, cells


def fmg(dim, k, level, rhs_levels, n_cycles=1, mg_variant=MG_SYMMETRIC, relax=1.0, coef=None):
    """Full multigrid from level 1 to `level`; rhs_levels[l-1] is the
    right-hand side of level l (its Dirichlet entries are the boundary
    values)."""
    assert len(rhs_levels) == level
    allr = np.ascontiguousarray(np.concatenate([np.asarray(r, dtype=np.float64)
                                                for r in rhs_levels]))
    x = np.zeros(n_dofs(dim, k, level))
    c = None if coef is None else np.ascontiguousarray(coef, dtype=np.float64)
    _chk(lib().oracle_fmg(dim, k, level, None if c is None else _p(c), mg_variant, relax,
                          _p(allr), n_cycles, _p(x)), "fmg")
    return x


# ---------------------------------------------------------------- SIPG-DG
# DoFs on one tensor grid X = cx (k+1) + ix (n cells per direction), index
# X + N1 Y + N1^2 Z; penalty sigma = pf / h.

def dg_penalty(k):
    """The penalty factor of the DG workload, 2.5 (k+1)^2."""
    return 2.5 * (k + 1) ** 2


def dg_n_dofs(dim, k, n):
    return int(lib().oracle_dg_n_dofs(dim, k, n))


def dg_vmult(dim, k, n, x, pf=None):
    pf = dg_penalty(k) if pf is None else pf
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    _chk(lib().oracle_dg_vmult(dim, k, n, pf, _p(x), _p(y)), "dg_vmult")
    return y


def dg_global_1d(k, n, pf=None):
    pf = dg_penalty(k) if pf is None else pf
    N = n * (k + 1)
    L = np.zeros((N, N))
    _chk(lib().oracle_dg_global_1d(k, n, pf, _p(L)), "dg_global_1d")
    return L


def dg_smoother_step(dim, k, n, b, u, variant=MULT_FWD, relax=1.0, pf=None):
    pf = dg_penalty(k) if pf is None else pf
    b = np.ascontiguousarray(b, dtype=np.float64)
    u = np.array(u, dtype=np.float64, copy=True)
    _chk(lib().oracle_dg_smoother_step(dim, k, n, pf, variant, relax, _p(b), _p(u)),
         "dg_smoother")
    return u


def dg_prolongate_add(dim, k, n_fine, coarse, fine):
    coarse = np.ascontiguousarray(coarse, dtype=np.float64)
    fine = np.array(fine, dtype=np.float64, copy=True)
    _chk(lib().oracle_dg_prolongate_add(dim, k, n_fine, _p(coarse), _p(fine)), "dg_prolong")
    return fine


def dg_restrict_add(dim, k, n_fine, fine, coarse):
    fine = np.ascontiguousarray(fine, dtype=np.float64)
    coarse = np.array(coarse, dtype=np.float64, copy=True)
    _chk(lib().oracle_dg_restrict_add(dim, k, n_fine, _p(fine), _p(coarse)), "dg_restrict")
    return coarse


def dg_vcycle(dim, k, n0, max_level, r, pf=None, mg_variant=MG_SYMMETRIC, relax=1.0,
              coarse_sweeps=2):
    pf = dg_penalty(k) if pf is None else pf
    r = np.ascontiguousarray(r, dtype=np.float64)
    z = np.zeros_like(r)
    _chk(lib().oracle_dg_vcycle(dim, k, n0, max_level, pf, mg_variant, relax, coarse_sweeps,
                                _p(r), _p(z)), "dg_vcycle")
    return z


def dg_pcg(dim, k, n0, max_level, b, rtol=1e-8, maxit=200, pf=None, mg_variant=MG_SYMMETRIC,
           relax=1.0, coarse_sweeps=2):
    pf = dg_penalty(k) if pf is None else pf
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    it = ctypes.c_int(0)
    rel = ctypes.c_double(0.0)
    hist = np.zeros(maxit + 1)
    _chk(lib().oracle_dg_pcg(dim, k, n0, max_level, pf, mg_variant, relax, coarse_sweeps,
                             _p(b), _p(x), rtol, maxit, ctypes.byref(it), ctypes.byref(rel),
                             _p(hist)), "dg_pcg")
    return x, it.value, rel.value, hist[: it.value + 1]


DG_RHS_CONST, DG_RHS_SINE = 0, 1


def dg_rhs(dim, k, n, kind=DG_RHS_SINE):
    b = np.zeros(dg_n_dofs(dim, k, n))
    _chk(lib().oracle_dg_rhs(dim, k, n, kind, _p(b)), "dg_rhs")
    return b


def dg_l2_error_sine(dim, k, n, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    return float(lib().oracle_dg_l2_error_sine(dim, k, n, _p(u)))


# ---------------------------------------------------------------- post-processing
def assemble_rhs(dim, k, n_cells, fq, dg=False):
    """b_i = int f phi_i from f at the QGauss(k+1) points, fq shaped
    (n_cells^dim, (k+1)^dim) (cell and point lexicographic, x fastest)."""
    fq = np.ascontiguousarray(fq, dtype=np.float64).ravel()
    n1 = n_cells * (k + 1) if dg else k * n_cells + 1
    b = np.zeros(n1 ** dim)
    _chk(lib().oracle_assemble_rhs(dim, k, n_cells, int(dg), _p(fq), _p(b)), "assemble_rhs")
    return b


def error_norms(dim, k, n_cells, u, wave=3.0, n_q=None, dg=False):
    """(L2, H1 seminorm) of u_h - prod sin(wave pi x_d), QGauss(n_q) per cell."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    l2, h1 = ctypes.c_double(), ctypes.c_double()
    _chk(lib().oracle_error_norms(dim, k, n_cells, int(dg), wave, k + 1 if n_q is None else n_q,
                                  _p(u), ctypes.byref(l2), ctypes.byref(h1)), "error_norms")
    return l2.value, h1.value
