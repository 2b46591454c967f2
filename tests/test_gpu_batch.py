"""Batched elliptic solves (many instances of uneven size in one launch
sequence) against the elliptic oracle, instance by instance.

The batched kernel builds the simplex constants per node from the lattice
form of the metric, so FP64 agrees with the oracle to rounding (held to
1e-12 relative), and the replayed hipGraph gives the same bits as direct
launches.  FP32 is held to 1e-5 relative of the FP64 oracle (nodes at least
8 cells from the source) and 1e-5 * max(T) absolute.
"""
import numpy as np
import pytest

from conftest import bits_equal

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import fold3d  # noqa: E402
from fold3d import workloads as wl  # noqa: E402


def dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def make_batch(sides, dtype_np=np.float64):
    Ws = []
    for k, n in enumerate(sides):
        if k % 3 == 0:
            W = wl.tilted_homogeneous(n)
            W.source = (0.37, 0.52, 0.61)
        elif k % 3 == 1:
            W = wl.blob_speed(n, seed=100 + k)
        else:
            W = wl.batch_instance(n, seed=200 + k, dtype=dtype_np)
        Ws.append(W)
    return Ws


def to_dev_workload(W, dtype):
    Wd = wl.Workload(W.name, W.n, W.h, W.M, W.speed, W.source, W.v_source, meta=W.meta)
    for k in ("speed_field", "theta", "phi", "ra", "rb"):
        f = getattr(W, k)
        setattr(Wd, k, None if f is None else dev(f, dtype))
    return Wd


def oracle_solve(oracle, W, T0, iters):
    T = T0.copy()
    f = {k: (None if getattr(W, k) is None else np.asarray(getattr(W, k), np.float64))
         for k in ("speed_field", "theta", "phi", "ra", "rb")}
    oracle.ellip_sweeps(T, W.n, iters, 0.0, W.h, M=W.M, speed=W.speed, **f)
    return T


SIDES = [17, 24, 9, 30, 12, 21, 26]


def test_batched_fp64_matches_oracle_per_instance(solver, oracle):
    print(f"batched FP64, sides {SIDES}", flush=True)
    Ws = make_batch(SIDES)
    T0s = [wl.point_source_start(W) for W in Ws]
    Wd = [to_dev_workload(W, torch.float64) for W in Ws]
    Ts = [dev(T0, torch.float64) for T0 in T0s]
    its, mds, ms = solver.solve_batched(Wd, Ts, max_iters=3, tol=0.0)
    assert its == [3] * len(SIDES)
    for W, T0, T in zip(Ws, T0s, Ts):
        want = oracle_solve(oracle, W, T0, 3)
        got = T.cpu().numpy()
        reached = want < 1e15
        assert np.array_equal(reached, got < 1e15)
        rel = np.abs(got[reached] - want[reached]) / want[reached]
        assert rel.max() < 1e-12, (W.name, W.n, rel.max())


def test_graph_and_direct_launches_agree(solver):
    Ws = make_batch(SIDES)
    T0s = [wl.point_source_start(W) for W in Ws]
    Wd = [to_dev_workload(W, torch.float64) for W in Ws]
    A = [dev(T0, torch.float64) for T0 in T0s]
    B = [dev(T0, torch.float64) for T0 in T0s]
    ia, _, _ = solver.solve_batched(Wd, A, max_iters=50, tol=1e-12, use_graph=True)
    ib, _, _ = solver.solve_batched(Wd, B, max_iters=50, tol=1e-12, use_graph=False)
    assert ia == ib
    for a, b in zip(A, B):
        assert torch.equal(a.view(torch.int64), b.view(torch.int64))


def test_converged_instances_stop(solver, oracle):
    """Per-instance convergence: iteration counts differ and each instance's
    field is the one its own tolerance stop gives."""
    Ws = make_batch([10, 28, 16])
    T0s = [wl.point_source_start(W) for W in Ws]
    Wd = [to_dev_workload(W, torch.float64) for W in Ws]
    Ts = [dev(T0, torch.float64) for T0 in T0s]
    its, mds, _ = solver.solve_batched(Wd, Ts, max_iters=60, tol=1e-10)
    assert all(m < 1e-10 for m in mds)
    for W, T0, T, k in zip(Ws, T0s, Ts, its):
        want = oracle_solve(oracle, W, T0, k)
        got = T.cpu().numpy()
        r = want < 1e15
        assert np.abs(got[r] - want[r]).max() < 1e-12 * want[r].max()


def test_batched_fp32_within_tolerance(solver, oracle):
    Ws = make_batch([34, 40, 28])
    T0s = [wl.point_source_start(W) for W in Ws]
    Wd = [to_dev_workload(W, torch.float32) for W in Ws]
    Ts = [dev(T0, torch.float32) for T0 in T0s]
    its, mds, _ = solver.solve_batched(Wd, Ts, max_iters=80, tol=1e-6)
    for W, T0, T, k in zip(Ws, T0s, Ts, its):
        want = T0.copy()
        f = {kk: (None if getattr(W, kk) is None else np.asarray(getattr(W, kk), np.float64))
             for kk in ("speed_field", "theta", "phi", "ra", "rb")}
        oracle.ellip_sweeps(want, W.n, 80, 1e-10, W.h, M=W.M, speed=W.speed, **f)
        got = T.double().cpu().numpy()
        r = want < 1e15
        far = r & (want > 8 * W.h / 3.6)
        rel = np.abs(got[far] - want[far]) / want[far]
        absd = np.abs(got[r] - want[r]).max() / want[r].max()
        print(f"batched FP32 {W.name} n={W.n}: {k} iterations, max rel {rel.max():.2e}, abs {absd:.2e}")
        assert rel.max() < 1e-5 and absd < 1e-5


def test_batch_rejects_bad_instances(solver):
    W = wl.tilted_homogeneous(12)
    Wd = to_dev_workload(W, torch.float64)
    T = dev(wl.point_source_start(W), torch.float64)
    Wd.M = np.array([1.0, 1.0, -1.0, 0, 0, 0])
    with pytest.raises(fold3d.Fold3dError) as e:
        solver.solve_batched([Wd], [T])
    assert e.value.code == fold3d.FOLD3D_E_INVALID


@pytest.mark.parametrize("groups", [1, 2, 4])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batched_per_node_metric_equals_single_solves(solver, groups, dtype):
    """Box-grid instances with per-node tilt and speed fields: the batch gives
    the same bits as one metric-engine solve per instance."""
    shapes = [(45, 70, 38), (20, 9, 33), (64, 64, 12), (7, 40, 25)]
    rng = np.random.default_rng(groups)
    Ws, T0s = [], []
    for n3 in shapes:
        h = tuple(1.0 / (x + 3) for x in n3)
        size = (n3[0] + 2) * (n3[1] + 2) * (n3[2] + 2)
        k, j, i = np.indices((n3[2] + 2, n3[1] + 2, n3[0] + 2)).reshape(3, -1)
        c = rng.uniform(0.2, 0.6, 3)
        X = np.stack([(i - 1) * h[0], (j - 1) * h[1], (k - 1) * h[2]]) - c[:, None]
        r = np.sqrt((X * X).sum(0))
        T0 = np.where(r < 2.5 * max(h), r, 1e16)
        ghost = (i == 0) | (j == 0) | (k == 0) | (i == n3[0] + 1) | (j == n3[1] + 1) | (k == n3[2] + 1)
        T0[ghost] = 1e16
        W = wl.Workload("box", n3, h, wl.CONFIG_METRIC, 1.0)
        W.speed_field = dev(1.0 + 0.5 * rng.random(size), dtype)
        W.theta = dev(rng.uniform(0, 1.0, size), dtype)
        W.phi = dev(rng.uniform(0, 6.2, size), dtype)
        W.ra = dev(rng.uniform(0.6, 1.0, size), dtype)
        W.rb = dev(rng.uniform(0.6, 1.0, size), dtype)
        Ws.append(W)
        T0s.append(T0)
    solver.set_tuning(groups, fold3d.MODEL_ELLIP)
    solver.set_batch_tuning(groups, 4, 2)
    try:
        Ts = [dev(T0, dtype) for T0 in T0s]
        its, _, _ = solver.solve_batched(Ws, Ts, max_iters=3, tol=0.0)
        assert its == [3] * len(shapes)
        for W, T0, T in zip(Ws, T0s, Ts):
            S = dev(T0, dtype)
            solver.solve(S, W.n, spacing=W.h, model=W.model(), max_iters=3,
                         engine=fold3d.ENGINE_SPLIT)
            assert torch.equal(T.double(), S.double()), W.n
    finally:
        solver.set_tuning(0, fold3d.MODEL_ELLIP)
        solver.set_batch_tuning(0, 0, 0)


def test_autotune_batched_keeps_fields_and_picks_a_schedule(solver):
    Ws = make_batch([17, 24, 30], dtype_np=np.float32)
    T0s = [wl.point_source_start(W) for W in Ws]
    Wd = [to_dev_workload(W, torch.float32) for W in Ws]
    Ts = [dev(T0, torch.float32) for T0 in T0s]
    before = [t.clone() for t in Ts]
    try:
        (g, b, r), ms = solver.autotune_batched(Wd, Ts, iterations=1)
        assert g in (1, 2, 4) and 1 <= b <= 64 and -1 <= r <= 64 and ms > 0
        for a, c in zip(Ts, before):
            assert torch.equal(a, c)
        its, _, _ = solver.solve_batched(Wd, Ts, max_iters=3, tol=0.0)
        ref = [c.clone() for c in before]
        solver.set_batch_tuning(2, 8, 8)  # the device queue
        solver.solve_batched(Wd, ref, max_iters=3, tol=0.0)
        for a, c in zip(Ts, ref):
            assert torch.equal(a, c)
    finally:
        solver.set_batch_tuning(0, 0, 0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_sub_batches_on_several_streams_give_the_same_bits(solver, dtype):
    """Cut into 2, 3 or 8 bands solved at once on their own streams, the batch
    gives every instance the bits, iteration count and max|dT| of one batch,
    with instances retiring at different iterations (tolerance stop)."""
    sides = [30, 9, 24, 17, 12, 28, 21, 26, 14]
    Ws = make_batch(sides, dtype_np=np.float64 if dtype == torch.float64 else np.float32)
    T0s = [wl.point_source_start(W) for W in Ws]
    Wd = [to_dev_workload(W, dtype) for W in Ws]
    tol = 1e-10 if dtype == torch.float64 else 1e-6
    results = {}
    try:
        for k in (1, 2, 3, 8):
            solver.set_batch_streams(k)
            assert solver.batch_streams() == k
            Ts = [dev(T0, dtype) for T0 in T0s]
            its, mds, ms = solver.solve_batched(Wd, Ts, max_iters=60, tol=tol)
            assert ms > 0
            results[k] = (its, mds, Ts)
        its1, mds1, T1 = results[1]
        assert len(set(its1)) > 1  # instances retire at different iterations
        for k in (2, 3, 8):
            its, mds, Ts = results[k]
            assert its == its1, k
            assert mds == mds1, k
            for a, b in zip(Ts, T1):
                assert torch.equal(a, b), k
    finally:
        solver.set_batch_streams(0)
    assert solver.batch_streams() == 3
