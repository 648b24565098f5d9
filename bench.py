#!/usr/bin/env python3
"""Benchmark of the geometric-multigrid hot path on one MI355X.

Workload (one "step"): one MG-preconditioned FP64 CG solve of 3D Q3 Poisson
on the unit cube with 64^3 cells (N = 193^3 = 7,189,057 DoFs, zero
Dirichlet), from x = 0 to ||r|| <= 1e-8 ||r_0||, right-hand side
b_i = 2 u01(i) - 1 per DoF with u01 from SplitMix64 (seed 20261202) and
b = 0 on Dirichlet DoFs.  The V-cycle V(1,1) uses the multiplicative
(8-colour) vertex-patch smoother with fast-diagonalisation local solves,
reversed colour order in post-smoothing, exact solve on level 1.

metric: MDoF/s = N / t_solve / 1e6 with the FP32 V-cycle (smoother, level
operators, transfers in FP32) under FP64 CG.  The FP64-everywhere form is
timed beside it.  Inputs are resident in HBM before the timed region; setup
(hierarchy, GPU FP64 eigen-solve) is excluded and reported separately.

roofline: the dominant kernel is the finest-level patch-smoother colour
launch; its average duration is measured live with HIP events (on the
context stream, which is torch's current stream) during the timed solves.
cpu_baseline: the FP64 C++ CPU oracle (OpenMP) on this host, same algorithm.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

SEED = 20261202
HBM_PEAK_GBS = 8000.0


def rhs_splitmix(dim: int, k: int, level: int, seed: int = SEED):
    """b_i = 2 u01(i) - 1, u01(i) = (splitmix64(seed ^ i) >> 11) 2^-53,
    zero on Dirichlet DoFs."""
    from gputensorproductsmoother_amd import problems
    return problems.rhs_random(dim, k, level, seed)


def smoother_colour_values(dim: int, k: int, level: int):
    """Algorithmic values moved by each colour launch of one multiplicative
    sweep: b read on the patch interiors plus u read on the union of the
    patch boxes minus the interiors (skeleton) plus u written on the
    interiors, i.e. prod_j p_j (2k-1) + prod_j (2k p_j + 1) for p_j patches
    of the colour along direction j."""
    n = 1 << level
    out = []
    for c in range(1 << dim):
        p = [(n - ((c >> j) & 1)) // 2 for j in range(dim)]
        inner, union = 1, 1
        for pj in p:
            inner *= pj * (2 * k - 1)
            union *= 2 * k * pj + 1
        out.append(inner + union if all(p) else 0)
    return out


def measured_traffic(kernel: str, dim: int, k: int, level: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (profiles/traffic.json: FETCH_SIZE and WRITE_SIZE runs, each
    scaled by the ratio a plain copy of known size reads for the same
    counter), or None when no summary matches this kernel and workload."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None
    for e in data.get("kernels", []):
        if e.get("kernel") == kernel and e.get("workload") == [dim, k, level]:
            return int(e["bytes_per_launch"])
    return None


def dominant_kernel_name(core, dim: int, k: int, level: int) -> str:
    """Name of the FP32 smoother kernel a finest-level colour launch runs
    with the current tuning (3D k <= 3: the patch-group kernel)."""
    if dim == 3 and k <= 3:
        n = 1 << level
        patches = (n // 2) ** 3
        npw = 1
        if patches > core.get_tuning("group_min_patches"):
            npw = core.get_tuning("patches_per_wave")
        waves = core.get_tuning("group_waves") if npw <= 2 else 1
        if waves == 0:
            waves = 4  # automatic: 4 wavefronts per work-group in FP32
        return f"patch3d_group_kernel<{k}, float, 0, {npw}, {waves}>"
    return f"patch3d_kernel<{k}, float, false>"


def _best_ms(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        best = min(best, (time.perf_counter() - t0) * 1e3)
    return best


def config_report(args):
    """One JSON line per workload of the parity configurations (not the
    contract line): iterations, time per solve, MDoF/s, and the relative L2
    distance of each solve to a tight FP64 solve of the same system."""
    import numpy as np
    import torch
    from gputensorproductsmoother_amd import core, problems

    ctx = core.Context(0)
    dev = "cuda:0"
    reps = max(1, args.steps)

    def rel_l2(a, b):
        return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))

    def solve_pair(dim, k, level, b, rtol, schedule=core.MultigridSolvers.SYMMETRIC,
                   coef=None, precs=("fp32", "fp64")):
        ref_mg = core.MultigridSolvers(ctx, dim, k, level, torch.float64)
        if coef is not None:
            ref_mg.set_coefficient(coef)
        xref, _, _, _ = ref_mg.solve_cg(b, rtol=1e-13, maxit=500, raise_on_failure=False)
        xref = xref.clone()
        del ref_mg
        res = {}
        n = b.numel()
        for prec in precs:
            mg = core.MultigridSolvers(ctx, dim, k, level,
                                       torch.float32 if prec == "fp32" else torch.float64,
                                       schedule)
            if coef is not None:
                mg.set_coefficient(coef)
            x = torch.empty_like(b)
            info = {}

            def run():
                _, it, rel, _ = mg.solve_cg(b, x, rtol=rtol, maxit=500)
                info.update(iterations=it, relres=rel)
            ms = _best_ms(run, reps)
            res[prec] = {"iterations": info["iterations"], "relres": info["relres"],
                         "ms_per_solve": round(ms, 3), "MDoF_s": round(n / ms / 1e3, 2),
                         "rel_l2_vs_fp64_tight": rel_l2(x, xref)}
            del mg
        return res

    def cell_mass(k):
        op = core.LaplaceOperator(ctx, 3, k, 1, torch.float64)
        return op.tensor_product_data()["cell_mass"]

    for name in args.configs.split(","):
        name = name.strip().lower()
        if name == "c1":
            dim, k, level = 2, 2, 7
            b = torch.tensor(problems.rhs_constant(dim, k, level, cell_mass(k)), device=dev)
            r = solve_pair(dim, k, level, b, 1e-10, core.MultigridSolvers.ADDITIVE,
                           precs=("fp64",))
            line = {"config": "C1", "workload": "2D Q2 Poisson f=1, 128^2 cells, additive "
                    "vertex-patch smoother (relaxation 0.4), all FP64, CG to 1e-10",
                    "n_dofs": b.numel(), **r}
        elif name == "c2":
            dim, k, level = 3, 3, 6
            b = torch.tensor(problems.rhs_random(dim, k, level), device=dev)
            r = solve_pair(dim, k, level, b, 1e-8)
            line = {"config": "C2", "workload": "3D Q3 Poisson, 64^3 cells, random rhs, "
                    "multiplicative smoother, CG to 1e-8", "n_dofs": b.numel(), **r}
        elif name == "c3":
            dim, k, level = 3, 7, 5
            b = torch.tensor(problems.rhs_constant(dim, k, level, cell_mass(k)), device=dev)
            r = solve_pair(dim, k, level, b, 1e-8)
            line = {"config": "C3", "workload": "3D Q7 Poisson f=1, 32^3 cells, "
                    "multiplicative smoother (FP32: 1D contractions on the matrix cores; "
                    "FP64: LDS/FMA kernel), CG to 1e-8", "n_dofs": b.numel(),
                    "mfma_utilisation": mfma_report(core, ctx, k, level, reps),
                    "hbm_per_kernel": hbm_report(core, ctx, k, level, reps), **r}
        elif name == "c4":
            line = c4_report(ctx, core, problems, reps, solve_pair, rel_l2)
        elif name == "c5":
            line = c5_report(ctx, core, problems, reps, rel_l2)
        else:
            raise SystemExit(f"unknown config {name}")
        print(json.dumps(line), flush=True)


def mfma_flops_per_patch(k: int) -> int:
    """FLOPs issued on the matrix cores per patch by the FP32 matrix-core
    smoother (v_mfma_f32_16x16x4f32 = 16*16*4 multiply-adds): z pass 2
    products on the NP non-zero tiles, y pass 3 and x pass 2 products plus
    S^T on the 13 stored planes, S^T along y, S^T and S along z on the NI
    interior tiles, S along y on 13 and along x on NI tiles; 4 k-steps
    each."""
    NP, NI, NZ = 2 * k + 1, 2 * k - 1, 13
    mfma = 4 * (2 * NP + 3 * NZ + 3 * NZ + NZ + 2 * NI + NZ + NI)
    return mfma * 16 * 16 * 4 * 2


def mfma_report(core, ctx, k, level, reps):
    """Finest-level colour pass of the FP32 smoother with the matrix-core
    kernel and with the LDS/FMA kernel (eager launches, best of reps x 8
    colours); MFMA utilisation = issued matrix-core FLOPs / time / the
    157.3 TFLOP/s dense FP32 matrix peak."""
    import torch
    op = core.LaplaceOperator(ctx, 3, k, level, torch.float32)
    g = torch.Generator(device="cuda:0").manual_seed(SEED)
    u = torch.rand(op.n_dofs, generator=g, device="cuda:0")
    b = torch.rand(op.n_dofs, generator=g, device="cuda:0")
    sm = core.PatchSmoother(op)
    n = 1 << level
    patches = (n - 1) ** 3
    res = {}
    for mf in (1, 0):
        prev = core.set_tuning("mfma_smoother", mf)
        try:
            ms = _best_ms(lambda: sm.step(u, b), max(3, reps))
        finally:
            core.set_tuning("mfma_smoother", prev)
        res["mfma" if mf else "fma"] = ms * 1e3 / 8
    fl = mfma_flops_per_patch(k) * patches
    ach = fl / (res["mfma"] * 8 * 1e-6) / 1e12
    waves = core.get_tuning("mfma_waves")  # wavefronts per patch
    return {"kernel": f"patch3d_mfma_kernel<{k}, {waves}>",
            "us_per_colour_mfma": round(res["mfma"], 2),
            "us_per_colour_lds_fma": round(res["fma"], 2), "issued_TFLOPs": round(ach, 1),
            "peak_TFLOPs": 157.3, "frac": round(ach / 157.3, 3),
            "note": "issued FLOPs include the zero padding of 13/15-wide lines to 16"}


def hbm_report(core, ctx, k, level, reps):
    """Finest-level kernels of the 3D solve at (k, level): mean time per
    launch (HIP events, best of reps) and algorithmic HBM bytes per launch
    over that time, against the 8 TB/s peak.  Bytes: smoother colour
    s (interior b + union of the patch boxes), vmult 2 N s, residual 3 N s,
    prolongate-add s (N_c + 2 N), restrict-add s (N + 2 N_c), FP64 CG
    matvec with the fused dot 2 N 8."""
    import torch
    g = torch.Generator(device="cuda:0").manual_seed(SEED)
    vals = smoother_colour_values(3, k, level)
    out = {}

    def rate(name, fn, nbytes, launches):
        ms = _best_ms(fn, max(3, reps))
        us = ms * 1e3 / launches
        gbs = nbytes / launches / (us * 1e-6) / 1e9
        out[name] = {"us": round(us, 2), "GB_s": round(gbs, 1),
                     "frac_of_hbm": round(gbs / HBM_PEAK_GBS, 3)}

    for dt, s in ((torch.float32, 4), (torch.float64, 8)):
        tag = "fp32" if s == 4 else "fp64"
        op = core.LaplaceOperator(ctx, 3, k, level, dt)
        sm = core.PatchSmoother(op)
        u = torch.rand(op.n_dofs, generator=g, device="cuda:0", dtype=dt)
        b = torch.rand(op.n_dofs, generator=g, device="cuda:0", dtype=dt)
        y = torch.empty_like(u)
        if s == 4:
            for mf in (1, 0):
                prev = core.set_tuning("mfma_smoother", mf)
                try:
                    rate(f"smoother_colour_fp32_{'mfma' if mf else 'lds_fma'}",
                         lambda: sm.step(u, b), s * sum(vals), len(vals))
                finally:
                    core.set_tuning("mfma_smoother", prev)
        else:
            rate("smoother_colour_fp64", lambda: sm.step(u, b), s * sum(vals), len(vals))
        rate(f"vmult_{tag}", lambda: op.vmult(y, u), 2 * s * op.n_dofs, 1)
        rate(f"residual_{tag}", lambda: op.residual(y, b, u), 3 * s * op.n_dofs, 1)
        tr = core.MGTransferCUDA(ctx, 3, k, 1, level, dt)
        nc = (k * (1 << (level - 1)) + 1) ** 3
        c = torch.rand(nc, generator=g, device="cuda:0", dtype=dt)
        rate(f"prolongate_add_{tag}", lambda: tr.prolongate_and_add(level, u, c),
             s * (nc + 2 * op.n_dofs), 1)
        rate(f"restrict_add_{tag}", lambda: tr.restrict_and_add(level, c, u),
             s * (op.n_dofs + 2 * nc), 1)
        del op, sm, tr
    return out


def c5_report(ctx, core, problems, reps, rel_l2):
    """SIPG-DG workload: Q4 DG on 24^3 cells (3^3 coarse mesh, 3 refinements),
    penalty 2.5 (k+1)^2, random f, FP32 V-cycle levels or FP64 everywhere,
    FP64 CG to 1e-8; then J = 1/2 sum_m ||U_m||^2 over 4 right-hand sides
    solved together (batched MG-CG) with dJ/dB = A^{-1} U through the
    autograd Function, checked by a central
    difference along a random direction (J is quadratic in B, so the
    difference is exact up to the solver tolerance)."""
    import torch
    dev = "cuda:0"
    k, n0, maxl = 4, 3, 3
    n = n0 << maxl
    B = torch.tensor(problems.rhs_dg_random(3, k, n, batch=4), device=dev)
    b = B[0].contiguous()
    ref = core.DGMultigrid(ctx, 3, k, n0, maxl, torch.float64)
    xref, _, _, _ = ref.solve_cg(b, rtol=1e-13, maxit=500, raise_on_failure=False)
    xref = xref.clone()
    res = {}
    for prec in ("fp32", "fp64"):
        mg = core.DGMultigrid(ctx, 3, k, n0, maxl,
                              torch.float32 if prec == "fp32" else torch.float64)
        x = torch.empty_like(b)
        info = {}

        def run():
            _, it, rel, _ = mg.solve_cg(b, x, rtol=1e-8, maxit=500)
            info.update(iterations=it, relres=rel)
        ms = _best_ms(run, reps)
        res[prec] = {"iterations": info["iterations"], "relres": info["relres"],
                     "ms_per_solve": round(ms, 3), "MDoF_s": round(b.numel() / ms / 1e3, 2),
                     "rel_l2_vs_fp64_tight": rel_l2(x, xref)}
        del mg
    mg = core.DGMultigrid(ctx, 3, k, n0, maxl, torch.float32)

    def fwd_bwd(rtol):
        Bv = B.clone().requires_grad_(True)
        U = mg.solve(Bv, rtol=rtol)
        (0.5 * (U * U).sum()).backward()
        return Bv.grad

    ms_fb = _best_ms(lambda: fwd_bwd(1e-8), reps)
    grad = fwd_bwd(1e-12)
    D = torch.tensor(problems.rhs_dg_random(3, k, n, batch=4, seed=problems.SEED + 1),
                     device=dev)
    h = 1e-3

    def J(Bv):
        U = core.DGSolve._solve_rows(mg, Bv.contiguous(), 1e-12, 500)
        return 0.5 * float((U * U).sum())

    fd = (J(B + h * D) - J(B - h * D)) / (2 * h)
    an = float((grad * D).sum())
    return {"config": "C5", "workload": "3D SIPG-DG Q4 Laplace, 24^3 cells (3^3 coarse, "
            "3 refinements), penalty 2.5 (k+1)^2, random f, 27-colour vertex-patch "
            "smoother, CG to 1e-8", "n_dofs": b.numel(), **res,
            "batched_adjoint": {"rhs": 4, "forward_plus_adjoint_ms": round(ms_fb, 3),
                                "solves": "the 4 right-hand sides together: one batched "
                                          "MG-CG forward, one for the adjoint",
                                "fd_direction_h": h,
                                "rel_dev_adjoint_vs_fd": abs(an - fd) / abs(fd)}}


def c4_report(ctx, core, problems, reps, solve_pair, rel_l2):
    """Variable-coefficient adjoint workload: forward MG-CG, adjoint solve and
    dJ/dc kernel through the autograd Function, checked against central
    finite differences (h = 1e-6) of J on 16 cells with FP64 solves to 1e-14."""
    import torch
    dev = "cuda:0"
    dim, k, level = 3, 4, 5
    coef_h, ud_h, cells = problems.coeff_problem(dim, k, level)
    op = core.LaplaceOperator(ctx, 3, k, 1, torch.float64)
    b = torch.tensor(problems.rhs_constant(dim, k, level,
                                           op.tensor_product_data()["cell_mass"]), device=dev)
    coef = torch.tensor(coef_h, device=dev)
    ud = torch.tensor(ud_h, device=dev)
    r = solve_pair(dim, k, level, b, 1e-8, coef=coef)
    mg = core.MultigridSolvers(ctx, dim, k, level, torch.float32)
    mask = mg.boundary_mask()

    def J(u):
        d = torch.where(mask, torch.zeros_like(u), u - ud)
        return 0.5 * (d * d).sum()

    def fwd_bwd(rtol):
        c = coef.clone().requires_grad_(True)
        u = mg.solve(b, c, rtol=rtol, maxit=500)
        J(u).backward()
        return c.grad

    ms_fb = _best_ms(lambda: fwd_bwd(1e-8), reps)
    grad = fwd_bwd(1e-12).cpu().numpy()
    mg64 = core.MultigridSolvers(ctx, dim, k, level, torch.float64)

    mg64.set_coefficient(coef)
    u0, _, _, _ = mg64.solve_cg(b, rtol=1e-14, maxit=500, raise_on_failure=False)
    u0 = u0.clone()

    def u64(cv):
        # u(c') = u0 + A(c')^{-1} (b - A(c') u0): the correction is O(h) and
        # solved to a relative 1e-12, so u(c+) - u(c-) keeps ~12 digits
        mg64.set_coefficient(cv)
        r = torch.empty_like(u0)
        mg64.do_matvec(r, u0)
        r = b - r
        d, _, _, _ = mg64.solve_cg(r, rtol=1e-12, maxit=500, raise_on_failure=False)
        return u0 + d

    # central difference J(c + h e_K) - J(c - h e_K), evaluated as
    # sum (u+ - u-) ((u+ + u-)/2 - u_d) over the interior: the same number
    # without the cancellation of two ~1e3-sized sums
    h = 1e-6
    dev_max = 0.0
    for K in cells.tolist():
        cp, cm = coef.clone(), coef.clone()
        cp[K] += h
        cm[K] -= h
        up, um = u64(cp), u64(cm)
        dj = torch.where(mask, torch.zeros_like(up), (up - um) * (0.5 * (up + um) - ud)).sum()
        fd = dj.item() / (2 * h)
        dev_max = max(dev_max, abs(grad[K] - fd) / abs(fd))
    return {"config": "C4", "workload": "3D Q4 -div(c grad u) = 1, 32^3 cells, c = exp(0.5 g), "
            "J = 1/2 ||u - u_d||^2, inexact-FD multiplicative smoother, CG to 1e-8",
            "n_dofs": b.numel(), **r,
            "forward_plus_adjoint_ms": round(ms_fb, 3),
            "fd_check": {"cells": len(cells), "h": h,
                         "max_rel_dev_adjoint_vs_fd": dev_max}}


def solve_trace(args):
    """Workload for a kernel trace: the configuration's FP32-V-cycle solve
    (C5: also the batch of 4 right-hand sides) three times, a 1 s pause,
    then --steps timed solves (one pause before each further group), so
    that --trace-summary can cut the timed dispatches out of the trace."""
    import torch
    from gputensorproductsmoother_amd import core, problems
    ctx = core.Context(0)
    dev = "cuda:0"
    name = args.solve_trace.lower()

    def cell_mass(k):
        op = core.LaplaceOperator(ctx, 3, k, 1, torch.float64)
        return op.tensor_product_data()["cell_mass"]

    groups = []
    if name in ("c2", "c3", "c4"):
        k, level = {"c2": (3, 6), "c3": (7, 5), "c4": (4, 5)}[name]
        if name == "c2":
            b = torch.tensor(problems.rhs_random(3, k, level), device=dev)
        else:
            b = torch.tensor(problems.rhs_constant(3, k, level, cell_mass(k)), device=dev)
        mg = core.MultigridSolvers(ctx, 3, k, level, torch.float32)
        if name == "c4":
            mg.set_coefficient(torch.tensor(problems.coeff_problem(3, k, level)[0], device=dev))
        x = torch.empty_like(b)
        groups.append((f"{name} fp32 solve", lambda: mg.solve_cg(b, x, rtol=1e-8)))
    elif name == "c5":
        k, n0, maxl = 4, 3, 3
        B = torch.tensor(problems.rhs_dg_random(3, k, n0 << maxl, batch=4), device=dev)
        mg = core.DGMultigrid(ctx, 3, k, n0, maxl, torch.float32)
        b = B[0].contiguous()
        x = torch.empty_like(b)
        X = torch.empty_like(B)
        groups.append(("c5 fp32 solve", lambda: mg.solve_cg(b, x, rtol=1e-8)))
        groups.append(("c5 fp32 batch-4 solve", lambda: mg.solve_cg_batch(B, X, rtol=1e-8)))
    else:
        raise SystemExit(f"unknown config {name}")
    for label, fn in groups:
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        time.sleep(1.0)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        print(json.dumps({"trace_group": label, "calls": args.steps,
                          "ms_per_call": round((time.perf_counter() - t0) * 1e3 / args.steps, 3)}),
              flush=True)
        time.sleep(1.0)


def trace_summary(csv_path: str, calls: int, top: int = 30):
    """Per-call kernel table of the timed groups of a --solve-trace run:
    the rocprofv3 kernel-trace CSV is cut at gaps over 0.5 s, the segments
    after the first (setup and warm-up) are reported, kernels grouped by
    name and grid size, times divided by the number of calls."""
    import collections
    import csv
    rows = []
    for r in csv.DictReader(open(csv_path)):
        name = r["Kernel_Name"].replace("(anonymous namespace)::", "").replace("void ", "")
        name = name.replace("gtps::", "").split("(")[0]
        grid = 1
        for ax in ("X", "Y", "Z"):
            grid *= int(r.get(f"Grid_Size_{ax}") or 1)
        rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), name, grid))
    rows.sort()
    segs, cur = [], [rows[0]]
    for a, b in zip(rows, rows[1:]):
        if b[0] - a[1] > 5e8:
            segs.append(cur)
            cur = []
        cur.append(b)
    segs.append(cur)
    for i, seg in enumerate(segs[1:], start=1):
        agg = collections.defaultdict(lambda: [0, 0.0])
        for s0, e0, n, g in seg:
            agg[(n, g)][0] += 1
            agg[(n, g)][1] += (e0 - s0) / 1e3
        busy = sum(v[1] for v in agg.values())
        span = (seg[-1][1] - seg[0][0]) / 1e3
        print(f"segment {i}: {len(seg) / calls:.1f} dispatches per call, kernel time "
              f"{busy / calls:.1f} us per call, span {span / calls:.1f} us per call")
        print("| us per call | launches per call | us per launch | grid | kernel |")
        print("|---:|---:|---:|---:|---|")
        for (n, g), (c, t) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
            print(f"| {t / calls:.1f} | {c / calls:.1f} | {t / c:.2f} | {g} | `{n}` |")


def traffic_summary(fetch_dir: str, write_dir: str, kernel: str, workload, out_path: str):
    """Turn two rocprofv3 --pmc runs of `bench.py --kernels` (FETCH_SIZE,
    then WRITE_SIZE, KiB per dispatch) into HBM bytes per launch of `kernel`:
    each counter is scaled by what it reads for the 256 MiB device copy of
    the same run (known bytes read and written), and the result is written
    to out_path (profiles/traffic.json)."""
    import csv
    import glob

    def rows(d, counter):
        path = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not path:
            raise SystemExit(f"no counter_collection.csv under {d}")
        with open(path[0]) as f:
            return [r for r in csv.DictReader(f) if r["Counter_Name"] == counter]

    copy_bytes = 256 * 2 ** 20
    res = {}
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rs = rows(d, counter)
        copies = [float(r["Counter_Value"]) for r in rs if "copyBuffer" in r["Kernel_Name"]]
        if not copies:
            raise SystemExit(f"no large copy dispatch in {d}")
        scale = copy_bytes / (max(copies) * 1024.0)
        vals = [float(r["Counter_Value"]) for r in rs
                if "::" + kernel + "(" in r["Kernel_Name"]]
        if not vals:
            raise SystemExit(f"no dispatch of {kernel} in {d}")
        res[counter] = {"kib_per_launch_raw": sum(vals) / len(vals), "scale": round(scale, 4),
                        "launches": len(vals),
                        "bytes_per_launch": sum(vals) / len(vals) * 1024.0 * scale}
    total = res["FETCH_SIZE"]["bytes_per_launch"] + res["WRITE_SIZE"]["bytes_per_launch"]
    entry = {"kernel": kernel, "workload": list(workload), "bytes_per_launch": int(round(total)),
             "fetch": res["FETCH_SIZE"], "write": res["WRITE_SIZE"],
             "calibration": "counter KiB x scale, scale = 256 MiB / counter of the "
                            "256 MiB device copy of the same run"}
    data = {"kernels": []}
    if os.path.exists(out_path):
        with open(out_path) as f:
            data = json.load(f)
    data["kernels"] = [e for e in data.get("kernels", [])
                       if not (e.get("kernel") == kernel and e.get("workload") == list(workload))]
    data["kernels"].append(entry)
    with open(out_path, "w") as f:
        json.dump(data, f, indent=1)
    print(json.dumps(entry), flush=True)


CPU_NOTE = ("the oracle: an FP64 C++/OpenMP restatement of the reference's algorithm in its "
            "own formulation (cell-quadrature operator with scatter, full-box local "
            "residual, valence-weighted transfer), not a tuned CPU code, so the GPU/CPU "
            "ratio says little about the kernels")


def cpu_baseline(args):
    """FP64 C++ oracle MG-PCG on the host cores (bounded sample): the full
    workload solved repeatedly for about 12 s (at most 5 solves, after a
    probe one level below), best solve reported; one level below when a
    full solve would not fit in about 40 s."""
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = O.set_threads(threads)
    dim, k = args.dim, args.degree
    lvl = args.level - 1
    b = O.rhs_random(dim, k, lvl, SEED)
    t0 = time.perf_counter()
    _, it, _, _ = O.pcg(dim, k, lvl, b, rtol=args.rtol, maxit=100)
    t_probe = time.perf_counter() - t0
    n = len(b)
    if t_probe * 8.5 < 40.0:
        lvl = args.level
        b = O.rhs_random(dim, k, lvl, SEED)
        n = len(b)
    times = []
    t_start = time.perf_counter()
    while len(times) < 5 and (not times or time.perf_counter() - t_start < 12.0):
        t0 = time.perf_counter()
        _, it, _, _ = O.pcg(dim, k, lvl, b, rtol=args.rtol, maxit=100)
        times.append(time.perf_counter() - t0)
        if lvl < args.level:
            break
    t = min(times)
    what = ("the full workload" if lvl == args.level
            else "one level below the workload (a full solve would exceed the time bound)")
    sample = (f"{what}: level {lvl} ({n} DoFs), FP64 MG-PCG to {args.rtol:g}, {it} iterations, "
              f"best of {len(times)} solve(s), {t:.2f} s; {CPU_NOTE}")
    return {"value": round(n / t / 1e6, 3), "unit": "MDoF/s", "cores": threads,
            "kind": "port", "sample": sample}


def cpu_baseline_one_core(args):
    """The same FP64 oracle MG-PCG on one host core, on a bounded sample: the
    solve one level below the workload (one eighth of the DoFs in 3D)."""
    import oracle as O
    prev = O.set_threads(0)
    O.set_threads(1)
    try:
        lvl = args.level - 1
        b = O.rhs_random(args.dim, args.degree, lvl, SEED)
        t0 = time.perf_counter()
        _, it, _, _ = O.pcg(args.dim, args.degree, lvl, b, rtol=args.rtol, maxit=100)
        t = time.perf_counter() - t0
    finally:
        O.set_threads(prev)
    n = len(b)
    return {"value": round(n / t / 1e6, 3), "unit": "MDoF/s", "cores": 1,
            "kind": "port",
            "sample": f"level {lvl} ({n} DoFs) full FP64 MG-PCG solve, {it} iterations, "
                      f"one solve, {t:.1f} s (the oracle, as for cpu_baseline)"}


def kernel_bench(args):
    """Per-kernel timing of the level-`level` building blocks (not the
    contract line): finest smoother sweep FP32/FP64, vmult FP64 and FP32
    residual, FP32 transfers, and a plain device copy of known size used to
    calibrate the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    from gputensorproductsmoother_amd import core

    dim, k, level, R = args.dim, args.degree, args.level, args.repeat
    ctx = core.Context(0)
    out = {}

    def timed(name, fn, nbytes, launches):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(R):
            fn()
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / (R * launches)
        out[name] = {"us_per_launch": round(us, 2), "launches": launches,
                     "GB_s_algorithmic": round(nbytes / launches / (us * 1e-6) / 1e9, 1)}

    g = torch.Generator(device="cuda:0").manual_seed(SEED)
    vals = smoother_colour_values(dim, k, level)
    for dt, s in ((torch.float32, 4), (torch.float64, 8)):
        op = core.LaplaceOperator(ctx, dim, k, level, dt)
        sm = core.PatchSmoother(op)
        u = torch.rand(op.n_dofs, generator=g, device="cuda:0", dtype=dt)
        b = torch.rand(op.n_dofs, generator=g, device="cuda:0", dtype=dt)
        tag = "fp32" if s == 4 else "fp64"
        timed(f"smoother_colour_{tag}", lambda: sm.step(u, b), s * sum(vals), len(vals))
        if dim == 3 and k <= 3:
            # patch-group kernel at 1, 2 and 4 patches per wavefront
            prev_min = core.set_tuning("group_min_patches", 0)
            for npw in (1, 2, 4):
                prev_npw = core.set_tuning("patches_per_wave", npw)
                timed(f"smoother_colour_{tag}_npw{npw}", lambda: sm.step(u, b),
                      s * sum(vals), len(vals))
                core.set_tuning("patches_per_wave", prev_npw)
            core.set_tuning("group_min_patches", prev_min)
        y = torch.empty_like(u)
        timed(f"vmult_{tag}", lambda: op.vmult(y, u), 2 * s * op.n_dofs, 1)
        timed(f"residual_{tag}", lambda: op.residual(y, b, u), 3 * s * op.n_dofs, 1)
        if level >= 2:
            tr = core.MGTransferCUDA(ctx, dim, k, 1, level, dt)
            nc = (k * (1 << (level - 1)) + 1) ** dim
            c = torch.rand(nc, generator=g, device="cuda:0", dtype=dt)
            timed(f"prolongate_add_{tag}", lambda: tr.prolongate_and_add(level, u, c),
                  s * (nc + 2 * op.n_dofs), 1)
            timed(f"restrict_add_{tag}", lambda: tr.restrict_and_add(level, c, u),
                  s * (op.n_dofs + 2 * nc), 1)
    src = torch.rand(1 << 26, generator=g, device="cuda:0")  # 256 MiB
    dst = torch.empty_like(src)
    timed("copy_256MiB", lambda: dst.copy_(src), 2 * src.numel() * 4, 1)
    if args.smoother_levels:
        # both smoother kernels on every level (colour-launch latency vs size)
        for lv in range(2, level + 1):
            vals_l = smoother_colour_values(dim, k, lv)
            for dt, s in ((torch.float32, 4), (torch.float64, 8)):
                op = core.LaplaceOperator(ctx, dim, k, lv, dt)
                sm = core.PatchSmoother(op)
                u = torch.rand(op.n_dofs, generator=g, device="cuda:0", dtype=dt)
                b = torch.rand(op.n_dofs, generator=g, device="cuda:0", dtype=dt)
                for path, lim in (("single", 1 << 40), ("multi", 0)):
                    prev = core.set_tuning("small_colour_patches", lim)
                    timed(f"smoother_l{lv}_{'fp32' if s == 4 else 'fp64'}_{path}",
                          lambda: sm.step(u, b), s * sum(vals_l), len(vals_l))
                    core.set_tuning("small_colour_patches", prev)
                yv = torch.empty_like(u)
                for path, lim in (("box", 1 << 20), ("march", 0)):
                    prev = core.set_tuning("box_vmult_max_n", lim)
                    timed(f"residual_l{lv}_{'fp32' if s == 4 else 'fp64'}_{path}",
                          lambda: op.residual(yv, b, u), 3 * s * op.n_dofs, 1)
                    core.set_tuning("box_vmult_max_n", prev)
                if lv >= 2:
                    trl = core.MGTransferCUDA(ctx, dim, k, 1, lv, dt)
                    ncl = (k * (1 << (lv - 1)) + 1) ** dim
                    cl = torch.rand(ncl, generator=g, device="cuda:0", dtype=dt)
                    timed(f"restrict_l{lv}_{'fp32' if s == 4 else 'fp64'}",
                          lambda: trl.restrict_and_add(lv, cl, u), s * (op.n_dofs + 2 * ncl), 1)
                    timed(f"prolongate_l{lv}_{'fp32' if s == 4 else 'fp64'}",
                          lambda: trl.prolongate_and_add(lv, u, cl), s * (ncl + 2 * op.n_dofs), 1)
    print(json.dumps({"kernels": out, "workload": f"{dim}D Q{k} level {level}",
                      "repeat": R}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--dim", type=int, default=3)
    ap.add_argument("--degree", type=int, default=3)
    ap.add_argument("--level", type=int, default=6)
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-everywhere run")
    ap.add_argument("--breakdown", action="store_true",
                    help="also time every stage group (extra solves, not in value)")
    ap.add_argument("--kernels", action="store_true",
                    help="per-kernel timings of the building blocks instead of the solve")
    ap.add_argument("--repeat", type=int, default=10)
    ap.add_argument("--smoother-levels", action="store_true",
                    help="with --kernels: both smoother kernels on every level")
    ap.add_argument("--configs", default="", metavar="c1,c2,c3,c4,c5",
                    help="per-configuration report (iterations, times, FP32-vs-FP64 "
                         "distance) instead of the contract line")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="library tuning knob (gtps_set_tuning) applied before the run")
    ap.add_argument("--traffic-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write profiles/traffic.json from two rocprofv3 --pmc runs of "
                         "--kernels (no GPU work)")
    ap.add_argument("--traffic-kernel", default="patch3d_group_kernel<3, float, 0, 2, 4>")
    ap.add_argument("--traffic-out", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="output of --traffic-summary")
    ap.add_argument("--solve-trace", metavar="c2|c3|c4|c5",
                    help="run a configuration's FP32 solve for a kernel trace (warm-up, "
                         "1 s pause, --steps timed solves)")
    ap.add_argument("--trace-summary", metavar="KERNEL_TRACE_CSV",
                    help="per-call kernel table of a --solve-trace rocprofv3 kernel trace "
                         "(--steps = calls per group; no GPU work)")
    args = ap.parse_args()

    if args.trace_summary:
        trace_summary(args.trace_summary, args.steps)
        return

    if args.traffic_summary:
        traffic_summary(args.traffic_summary[0], args.traffic_summary[1], args.traffic_kernel,
                        (args.dim, args.degree, args.level), args.traffic_out)
        return

    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build measures one GPU "
              "(the multi-GPU decomposition is a design in DESIGN.md, not a run)",
              file=sys.stderr)
        sys.exit(2)

    if args.tune:
        from gputensorproductsmoother_amd import core as _core
        for kv in args.tune:
            key, val = kv.split("=", 1)
            _core.set_tuning(key, int(val))

    if args.kernels:
        kernel_bench(args)
        return
    if args.solve_trace:
        solve_trace(args)
        return
    if args.configs:
        config_report(args)
        return

    import torch
    from gputensorproductsmoother_amd import core

    dim, k, level = args.dim, args.degree, args.level
    ctx = core.Context(0)
    b_host = rhs_splitmix(dim, k, level)
    N = b_host.size
    b = torch.tensor(b_host, device="cuda:0")
    x = torch.empty_like(b)

    def run(prec, steps, warmup, kernel_pass):
        """Timed solves (HIP-graph replay of the V-cycle, no stage timers),
        then, when kernel_pass, the same number of solves again with HIP
        events around every finest-level sweep (kernel-by-kernel launches)
        for the dominant kernel's mean launch duration."""
        t0 = time.perf_counter()
        mg = core.MultigridSolvers(ctx, dim, k, level, prec)
        # first preconditioner application: tabulates the coarse sub-cycle
        # (part of the setup, outside the timed solves)
        z = torch.empty_like(b)
        mg.vmult(z, torch.zeros_like(b))
        torch.cuda.synchronize()
        setup_s = time.perf_counter() - t0
        for _ in range(warmup):
            mg.solve_cg(b, x, rtol=args.rtol, maxit=100)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        its = []
        rel = None
        for _ in range(steps):
            _, it, rel, _ = mg.solve_cg(b, x, rtol=args.rtol, maxit=100)
            its.append(it)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        prof = None
        if kernel_pass:
            mg.set_profiling(("smooth_fine",))
            mg.profile_reset()
            for _ in range(steps):
                mg.solve_cg(b, x, rtol=args.rtol, maxit=100)
            prof = mg.profile()
            mg.set_profiling(())
        return mg, dict(ms=dt / steps * 1e3, iterations=its, relres=rel, setup_s=setup_s,
                        profile=prof)

    mg32, r32 = run(torch.float32, args.steps, args.warmup, True)
    value = N / (r32["ms"] / 1e3) / 1e6

    # dominant kernel: finest-level patch-smoother colour launch (FP32)
    cnt, tot_ms = r32["profile"]["smooth_fine"]
    vals = smoother_colour_values(dim, k, level)
    # each sweep launches every colour once; bytes per launch averaged over
    # the colours of a sweep
    bytes_per_launch = 4.0 * sum(vals) / len(vals)
    avg_launch_ms = tot_ms / max(cnt, 1)
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    kernel_name = dominant_kernel_name(core, dim, k, level)
    traffic = measured_traffic(kernel_name, dim, k, level)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic,
                "kernel": f"{kernel_name} (finest-level colour launch)",
                "avg_launch_us": round(avg_launch_ms * 1e3, 2), "launches": cnt,
                "algorithmic_bytes_per_launch": round(bytes_per_launch)}

    extra = {"fp32_vcycle": {"ms_per_solve": round(r32["ms"], 3),
                             "iterations": r32["iterations"][-1],
                             "relres": r32["relres"], "setup_s": round(r32["setup_s"], 3)}}
    if not args.no_fp64:
        _, r64 = run(torch.float64, max(2, args.steps // 2), 1, False)
        v64 = N / (r64["ms"] / 1e3) / 1e6
        extra["fp64_everywhere"] = {"value": round(v64, 2), "ms_per_solve": round(r64["ms"], 3),
                                    "iterations": r64["iterations"][-1],
                                    "relres": r64["relres"], "setup_s": round(r64["setup_s"], 3)}
        extra["speedup_fp32_vs_fp64"] = round(value / v64, 3)
    if args.breakdown:
        groups = tuple(core.L.PROF_GROUPS) if hasattr(core, "L") else ()
        mg32.set_profiling(groups)
        mg32.profile_reset()
        for _ in range(2):
            mg32.solve_cg(b, x, rtol=args.rtol, maxit=100)
        extra["breakdown_ms_per_solve"] = {g: round(v[1] / 2, 4)
                                           for g, v in mg32.profile().items()}
        mg32.set_profiling(())

    cpu = None if args.no_cpu_baseline else cpu_baseline(args)
    if cpu is not None:
        extra["cpu_baseline_one_core"] = cpu_baseline_one_core(args)

    line = {
        "metric": "MDoF/s of MG-CG to 1e-8 rel, FP32-smoother vs FP64, 3D Q3 Poisson, 1 GPU",
        "value": round(value, 2),
        "unit": "MDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycle / f64 CG",
        "data": "synthetic (SplitMix64 uniform[-1,1] per DoF, seed 20261202)",
        "config": {"workload": f"{dim}D Q{k} Poisson, {1 << level}^{dim} cells, "
                               f"N={N}, MG V(1,1) multiplicative vertex-patch FD smoother, "
                               f"FP64 CG to {args.rtol:g}",
                   "dim": dim, "degree": k, "level": level, "n_dofs": N},
        "roofline": roofline,
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
