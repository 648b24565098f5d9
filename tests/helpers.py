"""Shared helpers of the test-suite (GPU side talks to the C ABI through
gputensorproductsmoother_amd.core; the oracle is the FP64 CPU checker)."""
import numpy as np

SEED = 20261202


def rng(seed=SEED):
    return np.random.default_rng(seed)


def n_dofs(dim, k, level):
    return (k * (1 << level) + 1) ** dim


def boundary_mask(dim, k, level):
    n1 = k * (1 << level) + 1
    idx = np.arange(n1)
    b1 = (idx == 0) | (idx == n1 - 1)
    if dim == 2:
        return (b1[None, :] | b1[:, None]).ravel()
    return (b1[None, None, :] | b1[None, :, None] | b1[:, None, None]).ravel()


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


def rel_l2(a, b):
    """Relative L2 distance ||a - b|| / ||b||."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
