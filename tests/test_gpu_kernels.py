"""GPU parity of the level kernels against the FP64 CPU oracle.

Each case prints (and flushes) what it runs before touching the GPU.
Tolerances: FP64 results agree with the oracle to 1e-12 relative (max-norm,
relative to the oracle's max entry); FP32 results to 2e-5 (inputs are
rounded to FP32 once, operators accumulate O(10-100) terms).
"""
import numpy as np
import pytest

import oracle as O
from helpers import boundary_mask, n_dofs, rel_err, rng

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 2e-5}


@pytest.fixture(scope="module")
def core():
    from gputensorproductsmoother_amd import core as C
    return C


@pytest.fixture(scope="module")
def ctx(core):
    return core.Context(0)


def _dt(prec):
    return torch.float64 if prec == "fp64" else torch.float32


def _announce(*args):
    print("CASE", *args, flush=True)


def _to_dev(a, prec):
    return torch.tensor(a, dtype=_dt(prec), device="cuda:0")


VMULT_CASES = [
    (2, 1, 3), (2, 2, 4), (2, 3, 3), (2, 5, 2), (2, 7, 2),
    (3, 1, 3), (3, 2, 3), (3, 3, 2), (3, 3, 3), (3, 3, 5), (3, 4, 2), (3, 4, 3),
    (3, 5, 2), (3, 6, 2), (3, 7, 2),
]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("dim,k,level", VMULT_CASES)
def test_laplace_vmult(core, ctx, dim, k, level, prec):
    _announce("vmult", dim, k, level, prec)
    x = rng(level * 100 + k).uniform(-1, 1, n_dofs(dim, k, level))
    if prec == "fp32":
        x = x.astype(np.float32).astype(np.float64)
    ref = O.laplace_vmult(dim, k, level, x)
    op = core.LaplaceOperator(ctx, dim, k, level, _dt(prec))
    xd = _to_dev(x, prec)
    yd = torch.full_like(xd, float("nan"))
    op.vmult(yd, xd)
    torch.cuda.synchronize()
    assert rel_err(yd.cpu().numpy(), ref) < TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("dim,k,level", [(2, 2, 3), (3, 3, 3), (3, 2, 4)])
def test_laplace_residual(core, ctx, dim, k, level, prec):
    _announce("residual", dim, k, level, prec)
    g = rng(7)
    u = g.uniform(-1, 1, n_dofs(dim, k, level))
    b = g.uniform(-1, 1, n_dofs(dim, k, level))
    if prec == "fp32":
        u = u.astype(np.float32).astype(np.float64)
        b = b.astype(np.float32).astype(np.float64)
    ref = b - O.laplace_vmult(dim, k, level, u)
    op = core.LaplaceOperator(ctx, dim, k, level, _dt(prec))
    out = torch.full((len(u),), float("nan"), dtype=_dt(prec), device="cuda:0")
    op.residual(out, _to_dev(b, prec), _to_dev(u, prec))
    torch.cuda.synchronize()
    assert rel_err(out.cpu().numpy(), ref) < TOL[prec]


SMOOTH_CASES = [(2, 1, 3), (2, 2, 3), (2, 3, 2), (2, 6, 2), (3, 1, 3), (3, 2, 2),
                (3, 3, 2), (3, 3, 3), (3, 4, 2), (3, 5, 2), (3, 7, 2)]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("variant", ["fwd", "rev", "add"])
@pytest.mark.parametrize("dim,k,level", SMOOTH_CASES)
def test_smoother_step(core, ctx, dim, k, level, variant, prec):
    _announce("smoother", dim, k, level, variant, prec)
    g = rng(level * 10 + k)
    N = n_dofs(dim, k, level)
    b = g.uniform(-1, 1, N)
    u = g.uniform(-1, 1, N)
    if prec == "fp32":
        u = u.astype(np.float32).astype(np.float64)
        b = b.astype(np.float32).astype(np.float64)
    var = {"fwd": (O.MULT_FWD, core.PatchSmoother.MULTIPLICATIVE, 1.0),
           "rev": (O.MULT_REV, core.PatchSmoother.MULTIPLICATIVE_REVERSE, 1.0),
           "add": (O.ADDITIVE, core.PatchSmoother.ADDITIVE, 0.5)}[variant]
    ref = O.smoother_step(dim, k, level, b, u, variant=var[0], relax=var[2])
    op = core.LaplaceOperator(ctx, dim, k, level, _dt(prec))
    sm = core.PatchSmoother(op, var[1], var[2])
    ud = _to_dev(u, prec)
    sm.step(ud, _to_dev(b, prec))
    torch.cuda.synchronize()
    tol = TOL[prec] * (20 if prec == "fp32" else 1)
    assert rel_err(ud.cpu().numpy(), ref) < tol


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("dim,k", [(2, 2), (2, 5), (3, 1), (3, 3), (3, 7)])
def test_smoother_level1_exact_solve(core, ctx, dim, k, prec):
    """On level 1 one vertex patch covers the domain: vmult is A^-1 on the
    interior (the reference's coarse solve, solver.cuh:80-102, 1955-1963)."""
    _announce("coarse-exact", dim, k, prec)
    N = n_dofs(dim, k, 1)
    bm = boundary_mask(dim, k, 1)
    b = rng(3).uniform(-1, 1, N)
    b[bm] = 0
    op = core.LaplaceOperator(ctx, dim, k, 1, _dt(prec))
    sm = core.PatchSmoother(op)
    ud = torch.full((N,), float("nan"), dtype=_dt(prec), device="cuda:0")
    sm.vmult(ud, _to_dev(b, prec))
    torch.cuda.synchronize()
    u = ud.cpu().numpy().astype(np.float64)
    assert np.all(u[bm] == 0)
    res = b - O.laplace_vmult(dim, k, 1, u)
    tol = 1e-11 if prec == "fp64" else 1e-4
    assert np.abs(res[~bm]).max() / np.abs(b).max() < tol


TRANSFER_CASES = [(2, 1, 3), (2, 2, 4), (2, 4, 3), (2, 7, 2), (3, 1, 3), (3, 2, 3),
                  (3, 3, 2), (3, 3, 4), (3, 4, 3), (3, 5, 2), (3, 7, 2)]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("dim,k,level", TRANSFER_CASES)
def test_transfer(core, ctx, dim, k, level, prec):
    _announce("transfer", dim, k, level, prec)
    g = rng(level + 31 * k)
    nf, nc = n_dofs(dim, k, level), n_dofs(dim, k, level - 1)
    coarse = g.uniform(-1, 1, nc)
    fine = g.uniform(-1, 1, nf)
    fine0 = g.uniform(-1, 1, nf)
    coarse0 = g.uniform(-1, 1, nc)
    if prec == "fp32":
        coarse, fine, fine0, coarse0 = (a.astype(np.float32).astype(np.float64)
                                        for a in (coarse, fine, fine0, coarse0))
    ref_p = O.prolongate_add(dim, k, level, coarse, fine0)
    ref_r = O.restrict_add(dim, k, level, fine, coarse0)
    tr = core.MGTransferCUDA(ctx, dim, k, 1, level, _dt(prec))
    fd = _to_dev(fine0, prec)
    tr.prolongate_and_add(level, fd, _to_dev(coarse, prec))
    cd = _to_dev(coarse0, prec)
    tr.restrict_and_add(level, cd, _to_dev(fine, prec))
    torch.cuda.synchronize()
    assert rel_err(fd.cpu().numpy(), ref_p) < TOL[prec]
    assert rel_err(cd.cpu().numpy(), ref_r) < TOL[prec] * (5 if prec == "fp32" else 1)


def test_tensor_product_data(core, ctx):
    """GPU FP64 eigenpairs: S^T M S = I, S^T K S = diag(lambda), and the
    eigenvalues agree with the oracle's CPU Jacobi."""
    for k in range(1, 8):
        _announce("eig", k)
        op = core.LaplaceOperator(ctx, 3, k, 2, torch.float64)
        d = op.tensor_product_data()
        MI = d["patch_mass"][1:-1, 1:-1]
        KI = d["patch_laplace"][1:-1, 1:-1]
        S, lam = d["eigenvectors"], d["eigenvalues"]
        n = 2 * k - 1
        assert np.abs(S.T @ MI @ S - np.eye(n)).max() < 1e-12
        assert np.abs(S.T @ KI @ S - np.diag(lam)).max() < 1e-11 * lam.max()
        ref = O.basis_1d(k)
        assert np.abs(lam - ref["lam"]).max() < 1e-11 * lam.max()
        assert np.abs(d["cell_mass"] - ref["M"]).max() < 1e-14
        assert np.abs(d["cell_laplace"] - ref["K"]).max() < 1e-12


SMOOTHER_PATHS = {
    # small_colour_patches, group_min_patches, patches_per_wave, row_staged,
    # group_waves
    "single": (1 << 40, 1 << 40, 2, 1, 1),
    "group2": (1 << 40, 0, 2, 1, 1),
    "group4": (1 << 40, 0, 4, 1, 1),
    "single_lane": (1 << 40, 1 << 40, 2, 0, 1),
    "group2_lane": (1 << 40, 0, 2, 0, 1),
    "group4_lane": (1 << 40, 0, 4, 0, 1),
    "single_lane_4waves": (1 << 40, 1 << 40, 2, 0, 4),
    "group2_lane_4waves": (1 << 40, 0, 2, 0, 4),
    "multi": (0, 4096, 2, 1, 1),
}


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("variant", ["fwd", "add"])
@pytest.mark.parametrize("path", list(SMOOTHER_PATHS))
@pytest.mark.parametrize("dim,k,level", [(3, 1, 4), (3, 2, 3), (3, 3, 3), (3, 3, 4)])
def test_smoother_both_kernel_paths(core, ctx, dim, k, level, variant, prec, path):
    """The patch-group kernel at 1, 2 and 4 patches per wavefront (coarse and
    fine levels), with row-staged or per-lane global accesses, with one or
    four wavefronts per work-group, and the multi-wave kernel agree with the
    oracle on the same level; odd patch counts per row leave a partly empty
    last group."""
    _announce("smoother-path", dim, k, level, variant, prec, path)
    g = rng(level * 7 + k)
    N = n_dofs(dim, k, level)
    b, u = g.uniform(-1, 1, (2, N))
    if prec == "fp32":
        u = u.astype(np.float32).astype(np.float64)
        b = b.astype(np.float32).astype(np.float64)
    ov, cv, relax = {"fwd": (O.MULT_FWD, core.PatchSmoother.MULTIPLICATIVE, 1.0),
                     "add": (O.ADDITIVE, core.PatchSmoother.ADDITIVE, 0.5)}[variant]
    ref = O.smoother_step(dim, k, level, b, u, variant=ov, relax=relax)
    keys = ("small_colour_patches", "group_min_patches", "patches_per_wave", "row_staged",
            "group_waves")
    prev = [core.set_tuning(kk, v) for kk, v in zip(keys, SMOOTHER_PATHS[path])]
    try:
        op = core.LaplaceOperator(ctx, dim, k, level, _dt(prec))
        sm = core.PatchSmoother(op, cv, relax)
        ud = _to_dev(u, prec)
        sm.step(ud, _to_dev(b, prec))
        torch.cuda.synchronize()
    finally:
        for kk, v in zip(keys, prev):
            core.set_tuning(kk, v)
    tol = TOL[prec] * (20 if prec == "fp32" else 1)
    assert rel_err(ud.cpu().numpy(), ref) < tol


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("path", ["box", "march"])
@pytest.mark.parametrize("k,level", [(1, 3), (2, 3), (3, 2), (3, 4), (4, 3), (3, 5), (5, 2), (6, 2), (7, 2)])
def test_vmult_box_and_march_paths(core, ctx, k, level, path, prec):
    """The one-round box kernel (small levels; k >= 6 in FP32 only) and the z-marching
    kernel agree with the oracle on the same level (vmult and residual)."""
    _announce("vmult-path", k, level, path, prec)
    g = rng(level * 5 + k)
    N = n_dofs(3, k, level)
    x, b = g.uniform(-1, 1, (2, N))
    if prec == "fp32":
        x = x.astype(np.float32).astype(np.float64)
        b = b.astype(np.float32).astype(np.float64)
    prev = core.set_tuning("box_vmult_max_n", 1 << 20 if path == "box" else 0)
    try:
        op = core.LaplaceOperator(ctx, 3, k, level, _dt(prec))
        xd, bd = _to_dev(x, prec), _to_dev(b, prec)
        y = torch.full_like(xd, float("nan"))
        r = torch.full_like(xd, float("nan"))
        op.vmult(y, xd)
        op.residual(r, bd, xd)
        torch.cuda.synchronize()
    finally:
        core.set_tuning("box_vmult_max_n", prev)
    ref = O.laplace_vmult(3, k, level, x)
    assert rel_err(y.cpu().numpy(), ref) < TOL[prec]
    assert rel_err(r.cpu().numpy(), b - ref) < TOL[prec]


@pytest.mark.parametrize("mfma", [4, 0, 8, 16])
@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("k,level", [(4, 2), (4, 3), (5, 2), (6, 2), (7, 1), (7, 2), (7, 3)])
def test_smoother_matrix_core_path(core, ctx, k, level, reverse, mfma):
    """3D FP32 multiplicative smoothing of degree 4..7 on the matrix cores
    (v_mfma_f32_16x16x4f32, patches padded to 16^3) with mfma = 4, 8 or 16
    wavefronts per patch, and on the LDS/FMA kernel (mfma = 0), agree with
    the oracle (FP32 tolerance x20, relaxation 0.8)."""
    _announce("smoother-mfma", k, level, reverse, mfma)
    g = rng(level * 11 + k)
    N = n_dofs(3, k, level)
    b, u = g.uniform(-1, 1, (2, N))
    u = u.astype(np.float32).astype(np.float64)
    b = b.astype(np.float32).astype(np.float64)
    relax = 0.8
    ref = O.smoother_step(3, k, level, b, u, variant=O.MULT_REV if reverse else O.MULT_FWD,
                          relax=relax)
    prev = core.set_tuning("mfma_smoother", 1 if mfma else 0)
    prev_waves = core.set_tuning("mfma_waves", mfma or 8)
    try:
        op = core.LaplaceOperator(ctx, 3, k, level, torch.float32)
        sm = core.PatchSmoother(op, core.PatchSmoother.MULTIPLICATIVE_REVERSE if reverse
                                else core.PatchSmoother.MULTIPLICATIVE, relax)
        ud = _to_dev(u, "fp32")
        sm.step(ud, _to_dev(b, "fp32"))
        torch.cuda.synchronize()
    finally:
        core.set_tuning("mfma_smoother", prev)
        core.set_tuning("mfma_waves", prev_waves)
    assert rel_err(ud.cpu().numpy(), ref) < TOL["fp32"] * 20


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("tile", [0, 1, 2])
@pytest.mark.parametrize("level", [3, 4])
def test_vmult_tiles_q7(core, ctx, level, tile, prec):
    """The three z-marching tiles of the k = 7 vmult (8x2 cells / 512
    threads, 4x4 / 1024, 4x4 / 512) agree with the oracle, vmult and
    residual (level 3: a partial 8x2 tile row; the box kernel is off)."""
    _announce("vmult-tile-q7", level, tile, prec)
    k = 7
    g = rng(level * 13 + tile)
    N = n_dofs(3, k, level)
    x, b = g.uniform(-1, 1, (2, N))
    if prec == "fp32":
        x = x.astype(np.float32).astype(np.float64)
        b = b.astype(np.float32).astype(np.float64)
    prev = core.set_tuning("vmult_tile", tile)
    prev_box = core.set_tuning("box_vmult_max_n", 0)
    try:
        op = core.LaplaceOperator(ctx, 3, k, level, _dt(prec))
        xd, bd = _to_dev(x, prec), _to_dev(b, prec)
        y = torch.full_like(xd, float("nan"))
        r = torch.full_like(xd, float("nan"))
        op.vmult(y, xd)
        op.residual(r, bd, xd)
        torch.cuda.synchronize()
    finally:
        core.set_tuning("vmult_tile", prev)
        core.set_tuning("box_vmult_max_n", prev_box)
    ref = O.laplace_vmult(3, k, level, x)
    assert rel_err(y.cpu().numpy(), ref) < TOL[prec]
    assert rel_err(r.cpu().numpy(), b - ref) < TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("tile", [0, 1, 2])
@pytest.mark.parametrize("level", [3, 5])
def test_vmult_tiles(core, ctx, level, tile, prec):
    """The three z-marching tiles of the k = 3 vmult (16x8 cells / 512
    threads, 8x8 / 256, 16x8 / 256) agree with the oracle, vmult and
    residual, with a partial last tile (level 3: 8 cells per direction)."""
    _announce("vmult-tile", level, tile, prec)
    k = 3
    g = rng(level * 13 + tile)
    N = n_dofs(3, k, level)
    x, b = g.uniform(-1, 1, (2, N))
    if prec == "fp32":
        x = x.astype(np.float32).astype(np.float64)
        b = b.astype(np.float32).astype(np.float64)
    prev = [core.set_tuning("vmult_tile", tile), core.set_tuning("box_vmult_max_n", 0)]
    try:
        op = core.LaplaceOperator(ctx, 3, k, level, _dt(prec))
        y = torch.empty(N, dtype=_dt(prec), device="cuda:0")
        op.vmult(y, _to_dev(x, prec))
        r = torch.empty_like(y)
        op.residual(r, _to_dev(b, prec), _to_dev(x, prec))
        torch.cuda.synchronize()
    finally:
        core.set_tuning("vmult_tile", prev[0])
        core.set_tuning("box_vmult_max_n", prev[1])
    ref = O.laplace_vmult(3, k, level, x)
    assert rel_err(y.cpu().numpy(), ref) < TOL[prec]
    assert rel_err(r.cpu().numpy(), b - ref) < TOL[prec] * 10


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("tile", [0, 2])
@pytest.mark.parametrize("level", [2, 3])
def test_vmult_tiles_k7(core, ctx, level, tile, prec):
    """The two z-marching tiles of the k = 7 vmult (8x2 cells and 4x4 cells,
    512 threads) agree with the oracle, vmult and residual; level 2 (4 cells
    per direction) leaves the 8x2 tile partial in x."""
    _announce("vmult-tile-k7", level, tile, prec)
    k = 7
    g = rng(level * 17 + tile)
    N = n_dofs(3, k, level)
    x, b = g.uniform(-1, 1, (2, N))
    if prec == "fp32":
        x = x.astype(np.float32).astype(np.float64)
        b = b.astype(np.float32).astype(np.float64)
    prev = core.set_tuning("vmult_tile", tile)
    try:
        op = core.LaplaceOperator(ctx, 3, k, level, _dt(prec))
        y = torch.empty(N, dtype=_dt(prec), device="cuda:0")
        op.vmult(y, _to_dev(x, prec))
        r = torch.empty_like(y)
        op.residual(r, _to_dev(b, prec), _to_dev(x, prec))
        torch.cuda.synchronize()
    finally:
        core.set_tuning("vmult_tile", prev)
    ref = O.laplace_vmult(3, k, level, x)
    assert rel_err(y.cpu().numpy(), ref) < TOL[prec]
    assert rel_err(r.cpu().numpy(), b - ref) < TOL[prec] * 10
