"""Pin the oracle's operator, smoother sweeps, transfer and solve against the
dense numpy restatements in tests/golden/dense_*.npz (Kronecker-sum matrix,
block Gauss-Seidel with direct local solves, Kronecker interpolation)."""
import glob
import os
import re

import numpy as np
import pytest

import oracle as O
from helpers import rel_err

FILES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "dense_*.npz")))


def _case(path):
    m = re.search(r"dense_(\d)d_q(\d)_l(\d)\.npz", path)
    return tuple(int(v) for v in m.groups())


def _load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(params=FILES, ids=[os.path.basename(f) for f in FILES])
def case(request):
    return _case(request.param), _load(request.param)


def test_fixtures_present():
    assert len(FILES) >= 7


def test_vmult(case):
    (dim, k, level), d = case
    assert rel_err(O.laplace_vmult(dim, k, level, d["x"]), d["Ax"]) < 1e-13


@pytest.mark.parametrize("mode", ["fwd", "rev", "add"])
def test_smoother_sweep(case, mode):
    (dim, k, level), d = case
    var, relax = {"fwd": (O.MULT_FWD, 1.0), "rev": (O.MULT_REV, 1.0),
                  "add": (O.ADDITIVE, 0.5)}[mode]
    u = O.smoother_step(dim, k, level, d["b"], d["u0"], variant=var, relax=relax)
    assert rel_err(u, d["u_" + mode]) < 1e-11


def test_transfer(case):
    (dim, k, level), d = case
    if level < 2:
        pytest.skip("no coarser level")
    p = O.prolongate_add(dim, k, level, d["coarse"], d["fine0"])
    assert rel_err(p, d["prolong"]) < 1e-13
    r = O.restrict_add(dim, k, level, d["fine"], d["coarse0"])
    assert rel_err(r, d["restrict"]) < 1e-13


def test_pcg_matches_direct_solve(case):
    (dim, k, level), d = case
    x, it, rel, hist = O.pcg(dim, k, level, d["b0"], rtol=1e-13, maxit=100)
    assert rel <= 1e-13
    assert rel_err(x, d["xsol"]) < 1e-11
    assert it <= 15
