"""GPU parity of the post-processing entry points (load-vector assembly,
L2 / H1 error norms, solution files) against the FP64 CPU oracle.

Assembly and norms are FP64 sums in a different order than the oracle's:
relative 1e-13 (max-norm for vectors).  Every case prints what it runs
before touching the GPU.
"""
import numpy as np
import pytest

import oracle as O
from helpers import rng

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def core():
    from gputensorproductsmoother_amd import core as C
    return C


@pytest.fixture(scope="module")
def ctx(core):
    return core.Context(0)


def _announce(*args):
    print("CASE", *args, flush=True)


def _dev(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")


def _host(t):
    return t.double().cpu().numpy()


def _maxrel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = [(2, 1, 3), (2, 2, 4), (2, 7, 2), (3, 1, 3), (3, 2, 2), (3, 3, 3), (3, 5, 1),
         (3, 7, 1)]


@pytest.mark.parametrize("dim,k,level", CASES)
def test_rhs_sine_and_constant(core, ctx, dim, k, level):
    _announce("rhs-sine", dim, k, level)
    b = core.rhs_sine(ctx, dim, k, level)
    assert _maxrel(_host(b), O.rhs_assembled(dim, k, level, O.RHS_SINE)) < 1e-13
    c = core.rhs_constant(ctx, dim, k, level, 2.5)
    assert _maxrel(_host(c), 2.5 * O.rhs_assembled(dim, k, level, O.RHS_CONST)) < 1e-13


@pytest.mark.parametrize("dim,k,level", CASES)
def test_assemble_rhs_from_points(core, ctx, dim, k, level):
    _announce("assemble-rhs", dim, k, level)
    n = 1 << level
    fq = rng(40 + 10 * k + level).uniform(-1, 1, n ** dim * (k + 1) ** dim)
    b = core.assemble_rhs(ctx, dim, k, level, _dev(fq))
    assert _maxrel(_host(b), O.assemble_rhs(dim, k, n, fq)) < 1e-13


@pytest.mark.parametrize("dim,k,n", [(2, 2, 3), (3, 1, 2), (3, 4, 3), (3, 7, 2)])
def test_dg_assemble_rhs(core, ctx, dim, k, n):
    _announce("dg-assemble-rhs", dim, k, n)
    fq = rng(60 + k + n).uniform(-1, 1, n ** dim * (k + 1) ** dim)
    b = core.dg_assemble_rhs(ctx, dim, k, n, _dev(fq))
    assert _maxrel(_host(b), O.assemble_rhs(dim, k, n, fq, dg=True)) < 1e-13
    a = np.pi
    q = core.quadrature_points_1d(k, n).ravel()
    s = np.sin(a * q)
    if dim == 3:
        f = dim * a * a * np.einsum("i,j,k->ijk", s, s, s)
        f = f.reshape(n, k + 1, n, k + 1, n, k + 1).transpose(0, 2, 4, 1, 3, 5)
    else:
        f = dim * a * a * np.outer(s, s)
        f = f.reshape(n, k + 1, n, k + 1).transpose(0, 2, 1, 3)
    bs = core.dg_rhs_sine(ctx, dim, k, n, wave=1.0)
    assert _maxrel(_host(bs), O.assemble_rhs(dim, k, n, f.ravel(), dg=True)) < 1e-13


@pytest.mark.parametrize("dim,k,level,nq,wave", [(2, 2, 3, 3, 3.0), (2, 4, 2, 8, 1.0),
                                                 (3, 1, 3, 2, 3.0), (3, 3, 2, 4, 3.0),
                                                 (3, 4, 2, 12, 2.0), (3, 7, 1, 8, 3.0)])
def test_error_norms(core, ctx, dim, k, level, nq, wave):
    _announce("error-norms", dim, k, level, nq, wave)
    n = 1 << level
    u = rng(80 + k).uniform(-1, 1, (k * n + 1) ** dim)
    l2, h1 = core.error_norms(ctx, dim, k, level, _dev(u), wave=wave, n_q=nq)
    rl2, rh1 = O.error_norms(dim, k, n, u, wave=wave, n_q=nq)
    assert abs(l2 - rl2) < 1e-13 * rl2 and abs(h1 - rh1) < 1e-13 * rh1


@pytest.mark.parametrize("dim,k,n", [(2, 3, 3), (3, 2, 3), (3, 4, 2)])
def test_dg_error_norms(core, ctx, dim, k, n):
    _announce("dg-error-norms", dim, k, n)
    u = rng(90 + k).uniform(-1, 1, (n * (k + 1)) ** dim)
    l2, h1 = core.dg_error_norms(ctx, dim, k, n, _dev(u), wave=1.0, n_q=k + 2)
    rl2, rh1 = O.error_norms(dim, k, n, u, wave=1.0, n_q=k + 2, dg=True)
    assert abs(l2 - rl2) < 1e-13 * rl2 and abs(h1 - rh1) < 1e-13 * rh1


def test_poisson_sine_convergence_orders(core, ctx):
    """The reference's Poisson problem (u = prod sin(3 pi x_d)) solved with
    the GPU load vector and MG-CG: L2 error falls like h^(k+1) and H1 like
    h^k under refinement, with the errors from the GPU norm kernel."""
    dim, k = 2, 2
    errs = []
    for level in (4, 5):
        _announce("poisson-sine", dim, k, level)
        mg = core.MultigridSolvers(ctx, dim, k, level, torch.float64)
        b = core.rhs_sine(ctx, dim, k, level)
        x, it, rel, _ = mg.solve_cg(b, rtol=1e-12)
        errs.append(core.error_norms(ctx, dim, k, level, x))
        assert rel < 1e-12
    assert abs(np.log2(errs[0][0] / errs[1][0]) - (k + 1)) < 0.3
    assert abs(np.log2(errs[0][1] / errs[1][1]) - k) < 0.3


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_vector_file_round_trip(core, ctx, tmp_path, dtype):
    _announce("vector-io", dtype)
    dim, k, n = 3, 2, 4
    N = (k * n + 1) ** dim
    x = torch.tensor(rng(5).uniform(-1, 1, N), dtype=dtype, device="cuda:0")
    path = tmp_path / "u.gtps"
    core.save_vector(ctx, path, x, dim, k, n)
    info = core.vector_info(path)
    assert (info["dim"], info["fe_degree"], info["n_cells"], info["n"]) == (dim, k, n, N)
    y = core.load_vector(ctx, path)
    assert y.dtype == dtype and torch.equal(x, y)
    raw = bytearray(path.read_bytes())
    raw[64 + 17] ^= 0x01
    bad = tmp_path / "bad.gtps"
    bad.write_bytes(bytes(raw))
    with pytest.raises(RuntimeError, match="hash"):
        core.load_vector(ctx, bad)
    wrong = torch.empty(N - 1, dtype=dtype, device="cuda:0")
    with pytest.raises((RuntimeError, ValueError)):
        core.load_vector(ctx, path, out=wrong)


def test_dg_solution_file_and_error(core, ctx, tmp_path):
    """DG layout: solve the sine problem on 6^3 Q2 cells, save, reload, and
    the reloaded solution has the same (small) error."""
    _announce("dg-io-error")
    mg = core.DGMultigrid(ctx, 3, 2, 3, 1, torch.float32)
    b = core.dg_rhs_sine(ctx, 3, 2, 6, wave=1.0)
    x, _, rel, _ = mg.solve_cg(b, rtol=1e-10)
    e = core.dg_error_norms(ctx, 3, 2, 6, x, wave=1.0)
    path = tmp_path / "dg.gtps"
    core.save_vector(ctx, path, x, 3, 2, 6, layout=core.L.GTPS_LAYOUT_DG)
    y = core.load_vector(ctx, path)
    assert torch.equal(x, y)
    assert core.dg_error_norms(ctx, 3, 2, 6, y, wave=1.0) == e
    assert e[0] < 2e-3
