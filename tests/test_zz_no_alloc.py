"""gtps_vec_dot_noalloc (the allocation-free variant of gtps_vec_dot, for
callers that cannot allocate device memory) against gtps_vec_dot, through
the C ABI: two calls in a row, the second with a smaller input, then the
empty vector, a second context with its own stream on the same device, and
the argument checks.  Both entry points run the same kernels with the same
grid, so the results must agree bit for bit."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from gputensorproductsmoother_amd import _lib as L
    return L


def _dot(L, ctx, x, y):
    out = ctypes.c_double()
    L.check(L.load().gtps_vec_dot(ctx.handle, x.numel(), ctypes.c_void_p(x.data_ptr()),
                                  ctypes.c_void_p(y.data_ptr()), ctypes.byref(out)), "dot")
    return out.value


def _dot_noalloc(L, ctx, x, y, n=None):
    res = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda:0")
    n = x.numel() if n is None else n
    xp = ctypes.c_void_p(x.data_ptr()) if n else None
    yp = ctypes.c_void_p(y.data_ptr()) if n else None
    L.check(L.load().gtps_vec_dot_noalloc(ctx.handle, n, xp, yp,
                                          ctypes.c_void_p(res.data_ptr())), "dot_noalloc")
    torch.cuda.synchronize()
    return res.item()


def test_dot_noalloc_matches_dot(lib):
    from gputensorproductsmoother_amd import core
    print("CASE dot-noalloc", flush=True)
    ctx = core.Context(0)
    g = np.random.default_rng(20261203)
    for n in (7_189_057, 4097):          # second call: smaller input
        x = torch.tensor(g.uniform(-1, 1, n), device="cuda:0")
        y = torch.tensor(g.uniform(-1, 1, n), device="cuda:0")
        ref = _dot(lib, ctx, x, y)
        got = _dot_noalloc(lib, ctx, x, y)
        print(f"n = {n}: dot {ref!r}, noalloc {got!r}", flush=True)
        assert np.float64(got).tobytes() == np.float64(ref).tobytes()
        assert abs(got - float(np.dot(x.cpu().numpy(), y.cpu().numpy()))) <= 1e-12 * n ** 0.5
    # the empty vector writes 0 (no stale partials from the calls above)
    e = torch.empty(0, dtype=torch.float64, device="cuda:0")
    assert _dot_noalloc(lib, ctx, e, e) == 0.0
    # another context with its own stream on the same device uses the same array
    side = torch.cuda.Stream(device="cuda:0")
    ctx2 = core.Context(0, stream=side.cuda_stream)
    x = torch.tensor(g.uniform(-1, 1, 300001), device="cuda:0")
    torch.cuda.synchronize()
    assert _dot_noalloc(lib, ctx2, x, x) == _dot(lib, ctx, x, x)


def test_dot_noalloc_rejects_bad_arguments(lib):
    from gputensorproductsmoother_amd import core
    print("CASE dot-noalloc-errors", flush=True)
    ctx = core.Context(0)
    L = lib.load()
    x = torch.zeros(8, dtype=torch.float64, device="cuda:0")
    res = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    xp, rp = ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(res.data_ptr())
    assert L.gtps_vec_dot_noalloc(ctx.handle, 8, xp, xp, None) == lib.GTPS_EINVAL
    assert L.gtps_vec_dot_noalloc(ctx.handle, -1, xp, xp, rp) == lib.GTPS_EINVAL
    assert L.gtps_vec_dot_noalloc(ctx.handle, (1 << 62) + 1, xp, xp, rp) == lib.GTPS_EINVAL
    assert b"GTPS_NOALLOC_MAX_N" in L.gtps_last_error()
    assert L.gtps_vec_dot_noalloc(None, 8, xp, xp, rp) == lib.GTPS_EINVAL
