import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "fgcs-openacc-advanced_amd")
for p in (REPO, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(REPO, "tests", "golden", "fold_ref_golden.npz")
GOLDEN_LARGE = os.path.join(REPO, "tests", "golden", "fold_ref_large.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI on HIP)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def golden():
    return np.load(GOLDEN)


@pytest.fixture(scope="session")
def golden_large():
    """Digests of the reference's 64^3 and 128^3 solves (tests/golden/make_golden.py)."""
    return np.load(GOLDEN_LARGE)


def matches_digest(T, g, key) -> bool:
    """Field T against the digest `key` of golden_large: sampled nodes, then
    per-plane sums, then the SHA-256 of the bytes (each names a different failure)."""
    import hashlib
    T = np.ascontiguousarray(T, dtype=np.float64)
    assert bits_equal(T[g[f"{key}_idx"]], g[f"{key}_val"]), f"{key}: sampled nodes differ"
    H = round(T.size ** (1 / 3))
    sums = np.abs(T.reshape(H, H, H)).clip(max=1e3).sum(axis=(1, 2))
    assert np.array_equal(sums, g[f"{key}_plane_sums"]), f"{key}: plane sums differ"
    return hashlib.sha256(T.tobytes()).digest() == g[f"{key}_sha256"].tobytes()


@pytest.fixture(scope="session")
def solver():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import fold3d
    s = fold3d.Solver(0)
    yield s
    s.close()


def bits_equal(a, b) -> bool:
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))
