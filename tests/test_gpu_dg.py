"""GPU parity of the SIPG-DG path (face-term operator, 27-colour vertex-patch
smoother, DG transfer, V-cycle, FP64 MG-CG and the batched adjoint with
respect to the right-hand side) against the FP64 CPU oracle.

Tolerances (relative, max-norm): operator and transfer FP64 1e-12, FP32
2e-5; one smoother sweep and one V-cycle FP64 1e-10 (27 dependent colour
stages, each an exact local solve of condition up to ~1e5), FP32 4e-4.
Every case prints what it runs before touching the GPU.
"""
import numpy as np
import pytest

import oracle as O
from helpers import rel_err, rng

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL_OP = {"fp64": 1e-12, "fp32": 2e-5}
TOL_SMOOTH = {"fp64": 1e-10, "fp32": 4e-4}


@pytest.fixture(scope="module")
def core():
    from gputensorproductsmoother_amd import core as C
    return C


@pytest.fixture(scope="module")
def ctx(core):
    return core.Context(0)


def _dt(prec):
    return torch.float64 if prec == "fp64" else torch.float32


def _prec(core, prec):
    return core.FP64 if prec == "fp64" else core.FP32


def _announce(*args):
    print("CASE", *args, flush=True)


def _dev(a, dtype=torch.float64):
    return torch.tensor(a, dtype=dtype, device="cuda:0")


def _host(t):
    return t.double().cpu().numpy()


def _round(a, prec):
    return a.astype(np.float32).astype(np.float64) if prec == "fp32" else a


OP_CASES = [(1, 2), (1, 3), (2, 3), (2, 5), (3, 4), (4, 3), (4, 6), (5, 3), (6, 2), (7, 2)]


def _skip_unsupported(k, prec):
    if prec == "fp64" and k > 5:
        pytest.skip("the FP64 DG smoother supports degree <= 5")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("k,n", OP_CASES)
def test_dg_vmult_and_residual(core, ctx, k, n, prec):
    _announce("dg-vmult", k, n, prec)
    op = core.DGOperator(ctx, 3, k, n, _prec(core, prec))
    N = O.dg_n_dofs(3, k, n)
    assert op.n_dofs == N
    g = rng(100 * k + n)
    x = _round(g.uniform(-1, 1, N), prec)
    b = _round(g.uniform(-1, 1, N), prec)
    ref = O.dg_vmult(3, k, n, x)
    y = torch.empty(N, dtype=_dt(prec), device="cuda:0")
    op.vmult(y, _dev(x, _dt(prec)))
    assert rel_err(_host(y), ref) < TOL_OP[prec]
    op.residual(y, _dev(b, _dt(prec)), _dev(x, _dt(prec)))
    ref_r = b - ref
    assert np.max(np.abs(_host(y) - ref_r)) < TOL_OP[prec] * 10 * np.max(np.abs(ref))


def test_dg_vmult_symmetric(core, ctx):
    """The SIPG operator is symmetric: u^T A v = v^T A u on random vectors."""
    _announce("dg-symmetry", 3, 4)
    op = core.DGOperator(ctx, 3, 3, 4, core.FP64)
    g = rng(7)
    u, v = g.uniform(-1, 1, (2, op.n_dofs))
    Au = torch.empty(op.n_dofs, dtype=torch.float64, device="cuda:0")
    Av = torch.empty_like(Au)
    op.vmult(Au, _dev(u))
    op.vmult(Av, _dev(v))
    a = float(np.dot(v, _host(Au)))
    b = float(np.dot(u, _host(Av)))
    assert abs(a - b) < 1e-12 * max(abs(a), 1.0) * 10


SMOOTH_CASES = [(1, 2), (1, 4), (2, 3), (2, 5), (3, 4), (4, 3), (4, 5), (5, 4), (7, 3)]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("k,n", SMOOTH_CASES)
def test_dg_smoother_step(core, ctx, k, n, reverse, prec):
    _skip_unsupported(k, prec)
    _announce("dg-smoother", k, n, "rev" if reverse else "fwd", prec)
    op = core.DGOperator(ctx, 3, k, n, _prec(core, prec))
    N = op.n_dofs
    g = rng(300 + 10 * k + n)
    b = _round(g.uniform(-1, 1, N), prec)
    u0 = _round(g.uniform(-1, 1, N), prec)
    relax = 0.8
    ref = O.dg_smoother_step(3, k, n, b, u0, O.MULT_REV if reverse else O.MULT_FWD, relax)
    u = _dev(u0, _dt(prec))
    op.smoother_step(u, _dev(b, _dt(prec)), reverse=reverse, relaxation=relax)
    assert rel_err(_host(u), ref) < TOL_SMOOTH[prec]


def test_dg_smoother_reduces_error(core, ctx):
    """One symmetric sweep pair contracts the A-norm error of A u = b."""
    _announce("dg-smoother-contract", 3, 6)
    op = core.DGOperator(ctx, 3, 3, 6, core.FP64)
    N = op.n_dofs
    g = rng(17)
    xs = g.uniform(-1, 1, N)
    b = O.dg_vmult(3, 3, 6, xs)
    u = torch.zeros(N, dtype=torch.float64, device="cuda:0")
    bd = _dev(b)

    def anorm(e):
        return float(np.sqrt(np.dot(e, O.dg_vmult(3, 3, 6, e))))

    e0 = anorm(xs)
    op.smoother_step(u, bd)
    op.smoother_step(u, bd, reverse=True)
    e1 = anorm(xs - _host(u))
    assert e1 < 0.9 * e0


TRANSFER_CASES = [(1, 2), (2, 4), (3, 6), (4, 4), (5, 6), (7, 4)]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("k,n", TRANSFER_CASES)
def test_dg_transfer(core, ctx, k, n, prec):
    _announce("dg-transfer", k, n, prec)
    op = core.DGOperator(ctx, 3, k, n, _prec(core, prec))
    Nf = op.n_dofs
    Nc = O.dg_n_dofs(3, k, n // 2)
    g = rng(500 + 10 * k + n)
    c = _round(g.uniform(-1, 1, Nc), prec)
    f0 = _round(g.uniform(-1, 1, Nf), prec)
    ref_f = O.dg_prolongate_add(3, k, n, c, f0)
    f = _dev(f0, _dt(prec))
    op.prolongate_and_add(f, _dev(c, _dt(prec)))
    assert rel_err(_host(f), ref_f) < TOL_OP[prec]
    ref_c = O.dg_restrict_add(3, k, n, f0, c)
    cc = _dev(c, _dt(prec))
    op.restrict_and_add(cc, _dev(f0, _dt(prec)))
    assert rel_err(_host(cc), ref_c) < TOL_OP[prec]


def test_dg_transfer_reproduces_coarse_polynomials(core, ctx):
    """Prolongation is the embedding of nested DG spaces: a global quadratic
    given by its coarse nodal values comes out as its fine nodal values."""
    k, n = 2, 4
    _announce("dg-transfer-poly", k, n)
    op = core.DGOperator(ctx, 3, k, n, core.FP64)
    nodes = O.basis_1d(k)["nodes"]

    def field(nc):
        x1 = ((np.arange(nc)[:, None] + nodes[None, :]) / nc).ravel()
        z, y, x = np.meshgrid(x1, x1, x1, indexing="ij")
        return (x * x - 0.5 * y + z * x).ravel()

    c = field(n // 2)
    f = torch.zeros(op.n_dofs, dtype=torch.float64, device="cuda:0")
    op.prolongate_and_add(f, _dev(c))
    assert rel_err(_host(f), field(n)) < 1e-12


MG_CASES = [(1, 3, 1), (2, 3, 1), (3, 3, 1), (4, 3, 1), (2, 2, 2)]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("k,n0,maxl", MG_CASES)
def test_dg_vcycle(core, ctx, k, n0, maxl, prec):
    _announce("dg-vcycle", k, n0, maxl, prec)
    mg = core.DGMultigrid(ctx, 3, k, n0, maxl, _prec(core, prec))
    g = rng(700 + 10 * k + maxl)
    r = g.uniform(-1, 1, mg.n_dofs)
    ref = O.dg_vcycle(3, k, n0, maxl, r)
    z = torch.empty(mg.n_dofs, dtype=torch.float64, device="cuda:0")
    mg.vmult(z, _dev(r))
    assert rel_err(_host(z), ref) < TOL_SMOOTH[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_dg_pcg_matches_oracle(core, ctx, prec):
    k, n0, maxl = 3, 3, 1
    _announce("dg-pcg", k, n0, maxl, prec)
    mg = core.DGMultigrid(ctx, 3, k, n0, maxl, _prec(core, prec))
    b = O.dg_rhs(3, k, n0 << maxl, O.DG_RHS_SINE)
    xr, itr, relr, _ = O.dg_pcg(3, k, n0, maxl, b, rtol=1e-8)
    x, it, rel, hist = mg.solve_cg(_dev(b), rtol=1e-8)
    assert rel < 1e-8 and abs(it - itr) <= (0 if prec == "fp64" else 1)
    assert len(hist) == it + 1
    xh = _host(x)
    assert rel_err(xh, xr) < 1e-6
    Ax = O.dg_vmult(3, k, n0 << maxl, xh)
    assert np.linalg.norm(b - Ax) <= 1.0001e-8 * np.linalg.norm(b)
    # discretisation check: the sine solution converges at O(h^(k+1))
    assert O.dg_l2_error_sine(3, k, n0 << maxl, xh) < 1e-3


def test_dg_pcg_iterations_bounded_across_levels(core, ctx):
    """h-independent iteration counts: 1, 2 and 3 refinements of a 3^3 mesh."""
    k = 2
    its = []
    for maxl in (1, 2, 3):
        _announce("dg-pcg-levels", k, 3, maxl)
        mg = core.DGMultigrid(ctx, 3, k, 3, maxl, core.FP32)
        b = torch.ones(mg.n_dofs, dtype=torch.float64, device="cuda:0")
        _, it, rel, _ = mg.solve_cg(b, rtol=1e-8)
        assert rel < 1e-8
        its.append(it)
    assert max(its) - min(its) <= 3 and max(its) <= 25


def test_dg_batched_adjoint_wrt_rhs(core, ctx):
    """U = A^{-1} B for 4 right-hand sides; d<G, U>/dB = A^{-1} G (A is
    symmetric), compared with the oracle's FP64 PCG and with a directional
    finite difference (the map is linear, so the difference is exact up to
    the solver tolerance)."""
    k, n0, maxl = 2, 3, 1
    _announce("dg-adjoint", k, n0, maxl)
    mg = core.DGMultigrid(ctx, 3, k, n0, maxl, core.FP32)
    N = mg.n_dofs
    g = rng(901)
    Bh = g.uniform(-1, 1, (4, N))
    Gh = g.uniform(-1, 1, (4, N))
    B = _dev(Bh).requires_grad_(True)
    U = mg.solve(B, rtol=1e-11)
    J = (U * _dev(Gh)).sum()
    J.backward()
    gB = _host(B.grad)
    for i in range(4):
        ref_u, _, _, _ = O.dg_pcg(3, k, n0, maxl, Bh[i], rtol=1e-12)
        ref_g, _, _, _ = O.dg_pcg(3, k, n0, maxl, Gh[i], rtol=1e-12)
        assert rel_err(_host(U[i].detach()), ref_u) < 1e-8
        assert rel_err(gB[i], ref_g) < 1e-8
    dB = g.uniform(-1, 1, (4, N))
    h = 1e-6
    with torch.no_grad():
        Jp = float((mg.solve(_dev(Bh + h * dB), rtol=1e-11) * _dev(Gh)).sum())
        Jm = float((mg.solve(_dev(Bh - h * dB), rtol=1e-11) * _dev(Gh)).sum())
    fd = (Jp - Jm) / (2 * h)
    an = float(np.sum(gB * dB))
    assert abs(fd - an) < 1e-5 * max(abs(an), 1.0)


def test_dg_rejects_bad_arguments(core, ctx):
    _announce("dg-errors")
    big = core.DGOperator(ctx, 3, 6, 4, core.FP64)  # operator yes, smoother no
    v = torch.zeros(big.n_dofs, dtype=torch.float64, device="cuda:0")
    with pytest.raises(RuntimeError):
        big.smoother_step(v, v.clone())
    with pytest.raises(RuntimeError):
        core.DGOperator(ctx, 3, 2, 1, core.FP64)
    op = core.DGOperator(ctx, 3, 2, 3, core.FP64)
    c = torch.zeros(27, dtype=torch.float64, device="cuda:0")
    f = torch.zeros(op.n_dofs, dtype=torch.float64, device="cuda:0")
    with pytest.raises((RuntimeError, ValueError)):
        op.prolongate_and_add(f, c)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_dg_graph_replays_match_eager(core, ctx, prec):
    """The DG preconditioning sequence replayed as a HIP graph gives bitwise
    the same CG iterates as launch-by-launch execution, over several replays
    with eager torch work in between."""
    _announce("dg-graph-vs-eager", prec)
    dt = core.FP32 if prec == "fp32" else core.FP64
    b = _dev(O.dg_rhs(3, 2, 6, O.DG_RHS_SINE))
    mg_e = core.DGMultigrid(ctx, 3, 2, 3, 1, dt)
    mg_e.set_graph(False)
    x_e, it_e, _, h_e = mg_e.solve_cg(b, rtol=1e-10)
    x_e = x_e.clone()
    mg_g = core.DGMultigrid(ctx, 3, 2, 3, 1, dt)
    for rep in range(3):
        x_g, it_g, _, h_g = mg_g.solve_cg(b, rtol=1e-10)
        torch.cuda.synchronize()
        assert np.isfinite(float((b * 2.0).sum().item()))
        assert it_g == it_e and h_g == h_e, rep
        assert torch.equal(x_g, x_e), rep


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("k,n0,maxl", [(2, 3, 1), (4, 3, 1), (1, 2, 2)])
def test_dg_tabulated_coarse_solve(core, ctx, k, n0, maxl, prec):
    """The coarsest DG solve tabulated as a dense matrix gives the same
    V-cycle as the sweep kernels (FP64 1e-12, FP32 2e-5 relative) and as
    the oracle."""
    _announce("dg-dense-coarse", k, n0, maxl, prec)
    dt = core.FP32 if prec == "fp32" else core.FP64
    r = rng(40 + k).uniform(-1, 1, O.dg_n_dofs(3, k, n0 << maxl))
    out = {}
    for dense in (1, 0):
        prev = core.set_tuning("dense_coarse", dense)
        try:
            mg = core.DGMultigrid(ctx, 3, k, n0, maxl, dt)
            z = torch.empty(mg.n_dofs, dtype=torch.float64, device="cuda:0")
            mg.vmult(z, _dev(r))
            torch.cuda.synchronize()
            out[dense] = _host(z)
        finally:
            core.set_tuning("dense_coarse", prev)
    assert rel_err(out[1], out[0]) < (1e-12 if prec == "fp64" else 2e-5)
    assert rel_err(out[1], O.dg_vcycle(3, k, n0, maxl, r)) < TOL_SMOOTH[prec]


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_dg_batched_solve(core, ctx, prec):
    """Rows solved together: identical rows reproduce the single-vector
    MG-CG bitwise (same kernels per vector, same iteration count); distinct
    rows each reach the tolerance and match the oracle's solution."""
    _announce("dg-batch", prec)
    dt = core.FP32 if prec == "fp32" else core.FP64
    k, n0, maxl = 2, 3, 1
    mg = core.DGMultigrid(ctx, 3, k, n0, maxl, dt)
    N = mg.n_dofs
    b = rng(61).uniform(-1, 1, N)
    x1, it1, _, _ = mg.solve_cg(_dev(b), rtol=1e-8)
    x1 = x1.clone()
    B = _dev(np.stack([b] * 3))
    X, its, rel = mg.solve_cg_batch(B, rtol=1e-8)
    assert its == [it1] * 3
    for j in range(3):
        assert torch.equal(X[j], x1), j
    Bh = rng(62).uniform(-1, 1, (4, N))
    X, its, rel = mg.solve_cg_batch(_dev(Bh), rtol=1e-8)
    Xh = _host(X)
    for j in range(4):
        assert rel[j] <= 1e-8
        r = Bh[j] - O.dg_vmult(3, k, n0 << maxl, Xh[j])
        assert np.linalg.norm(r) <= 1.0001e-8 * np.linalg.norm(Bh[j])
        xr, _, _, _ = O.dg_pcg(3, k, n0, maxl, Bh[j], rtol=1e-12)
        assert rel_err(Xh[j], xr) < 1e-6


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_dg_batch_graph_replays_match_eager(core, ctx, prec):
    """The batched V-cycle replayed as a HIP graph gives bitwise the batched
    solve launched kernel by kernel, for batch sizes 3, 2 and 3 in a row (a
    new size is captured again), over repeated replays with eager torch work
    and synchronisations in between."""
    _announce("dg-batch-graph-vs-eager", prec)
    dt = core.FP32 if prec == "fp32" else core.FP64
    mg_e = core.DGMultigrid(ctx, 3, 2, 3, 1, dt)
    mg_e.set_graph(False)
    mg_g = core.DGMultigrid(ctx, 3, 2, 3, 1, dt)
    N = mg_e.n_dofs
    g = rng(63)
    for m in (3, 2, 3):
        B = _dev(g.uniform(-1, 1, (m, N)))
        Xe, its_e, rel_e = mg_e.solve_cg_batch(B, rtol=1e-9)
        Xe = Xe.clone()
        for rep in range(2):
            Xg, its_g, rel_g = mg_g.solve_cg_batch(B, rtol=1e-9)
            torch.cuda.synchronize()
            assert np.isfinite(float((B * 2.0).sum().item()))
            assert its_g == its_e and rel_g == rel_e, (m, rep)
            assert torch.equal(Xg, Xe), (m, rep)
