"""bench.py helpers and contract (no GPU needed)."""
import subprocess
import sys
import os

import numpy as np
import pytest

import bench
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("dim,k,level", [(2, 2, 4), (3, 3, 3), (3, 1, 2)])
def test_rhs_generator_matches_oracle(dim, k, level):
    assert np.array_equal(bench.rhs_splitmix(dim, k, level), O.rhs_random(dim, k, level))


def _count_by_enumeration(dim, k, level):
    """Values touched per colour: union of patch boxes (skeleton read +
    interior write) plus interior reads of b, by explicit enumeration."""
    n = 1 << level
    N1 = k * n + 1
    out = []
    for c in range(1 << dim):
        par = [(c >> j) & 1 for j in range(dim)]
        union = np.zeros((N1,) * dim, bool)
        inner = 0
        import itertools
        for cell in itertools.product(*[range(par[j], n - 1, 2) for j in range(dim)]):
            sl = tuple(slice(k * cell[dim - 1 - a], k * cell[dim - 1 - a] + 2 * k + 1)
                       for a in range(dim))
            union[sl] = True
            inner += (2 * k - 1) ** dim
        out.append(int(union.sum()) + inner if inner else 0)
    return out


@pytest.mark.parametrize("dim,k,level", [(3, 3, 3), (3, 2, 4), (2, 4, 4), (3, 1, 2)])
def test_smoother_byte_model(dim, k, level):
    assert bench.smoother_colour_values(dim, k, level) == _count_by_enumeration(dim, k, level)


def test_metric_config_sizes():
    v = bench.smoother_colour_values(3, 3, 6)
    assert sum(v) == 86_127_875          # 344.5 MB per FP32 sweep
    assert (3 * 64 + 1) ** 3 == 7_189_057


def test_multi_gpu_request_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
