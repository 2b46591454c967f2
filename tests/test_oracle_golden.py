"""Pin the oracle (CPU restatement) to the golden vectors written from the
reference's own solver (tests/golden/make_golden.py).  Bitwise."""
import numpy as np
import pytest

from conftest import bits_equal


def test_golden_digest():
    import hashlib
    import os
    from conftest import GOLDEN
    want = open(GOLDEN + ".sha256").read().split()[0]
    assert hashlib.sha256(open(GOLDEN, "rb").read()).hexdigest() == want
    assert os.path.getsize(GOLDEN) < 8 * 2 ** 20


def test_edge_kats(oracle, golden):
    E, want = golden["edge_in"], golden["edge_out"]
    got = np.array([oracle.edge(*row) for row in E])
    assert bits_equal(got, want)
    # the KATs exercise acceptance as well as rejection
    assert 0.2 < np.mean(want != E[:, 0]) < 0.8


def test_triangle_kats(oracle, golden):
    S, want = golden["surf_in"], golden["surf_out"]
    got = np.array([oracle.surf(*row) for row in S])
    assert bits_equal(got, want)
    assert 0.1 < np.mean(want != S[:, 0]) < 0.9


def test_point_source_kats(oracle, golden):
    P, want = golden["point_in"], golden["point_out"]
    got = np.array([oracle.fold_point(*row) for row in P])
    assert bits_equal(got, want)


@pytest.mark.parametrize("n", [16, 24])
def test_initialiser(oracle, golden, n):
    assert bits_equal(oracle.initialiser(n), golden[f"init_{n}"])


@pytest.mark.parametrize("d", range(6))
def test_plane_random_field(oracle, golden, d):
    planes = {0: 7, 1: 9, 2: 7, 3: 9, 4: 9, 5: 7}
    T = golden["plane_rand_in_16"].copy()
    oracle.plane(d, T, 16, planes[d])
    want = golden[f"plane_rand_{d}_16"]
    assert bits_equal(T, want)
    assert (T != golden["plane_rand_in_16"]).sum() > 50  # the step did work


@pytest.mark.parametrize("d", range(6))
def test_direction_initial_field(oracle, golden, d):
    T = golden["init_16"].copy()
    oracle.direction(d, T, 16)
    assert bits_equal(T, golden[f"direction_init_{d}_16"])


@pytest.mark.parametrize("n", [16, 24, 32])
def test_full_solves(oracle, golden, n):
    T = oracle.initialiser(n)
    oracle.sweeps(T, n, 1)
    assert bits_equal(T, golden[f"solve1_{n}"])
    oracle.sweeps(T, n, 7)
    assert bits_equal(T, golden[f"solve8_{n}"])
    # the first iteration that changes nothing, counted from the start
    its = int(golden[f"fixed_iters_{n}"][0])
    U = oracle.initialiser(n)
    oracle.sweeps(U, n, its - 1)
    V = U.copy()
    oracle.sweeps(V, n, 1)
    assert bits_equal(V, U) and bits_equal(U, golden[f"fixed_{n}"])
    W = oracle.initialiser(n)
    oracle.sweeps(W, n, its - 2)
    X = W.copy()
    oracle.sweeps(X, n, 1)
    assert not bits_equal(X, W)  # one iteration earlier it still moved
    if its <= 8:
        assert bits_equal(T, golden[f"fixed_{n}"])


@pytest.mark.slow
def test_solve_128_digest(oracle, golden):
    import hashlib
    T = oracle.initialiser(128)
    oracle.sweeps(T, 128, 8)
    assert np.array_equal(T[golden["solve8_128_idx"]], golden["solve8_128_val"])
    assert hashlib.sha256(T.tobytes()).digest() == golden["solve8_128_sha256"].tobytes()


def test_solves_64_digests(oracle, golden_large):
    """The oracle against the reference's 64^3 solve after 1 and 8 iterations
    and at its fixed point (18 iterations, the first that changes nothing)."""
    from conftest import matches_digest
    n = 64
    T = oracle.initialiser(n)
    oracle.sweeps(T, n, 1)
    assert matches_digest(T, golden_large, "solve1_64")
    oracle.sweeps(T, n, 7)
    assert matches_digest(T, golden_large, "solve8_64")
    its = int(golden_large["fixed_iters_64"][0])
    assert its == 18
    oracle.sweeps(T, n, its - 8)
    assert matches_digest(T, golden_large, "fixed_64")
    U = T.copy()
    oracle.sweeps(U, n, 1)
    assert bits_equal(U, T)


@pytest.mark.slow
def test_fixed_point_128_digest(oracle, golden_large):
    from conftest import matches_digest
    n = 128
    its = int(golden_large["fixed_iters_128"][0])
    T = oracle.initialiser(n)
    oracle.sweeps(T, n, its)
    assert matches_digest(T, golden_large, "fixed_128")
