"""The tolerance stop's max|dT|: each engine keeps it in many words (one per
group of blocks) and the host takes the largest.  The reported value must be
exactly max |T_k - T_(k-1)| over the nodes, which the host computes here from
two fixed-count solves of the same start field (FP64, so both are exact)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import fold3d  # noqa: E402
from fold3d import workloads as wl  # noqa: E402


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", torch.float64)


def reported_vs_host(solver, T0, n, k, **kw):
    a, b, c = dev(T0), dev(T0), dev(T0)
    if k > 1:
        solver.solve(a, n, max_iters=k - 1, **kw)
    solver.solve(b, n, max_iters=k, **kw)
    st = solver.solve(c, n, max_iters=k, tol=1e-300, raise_not_converged=False, **kw)
    assert torch.equal(b.view(torch.int64), c.view(torch.int64))  # the stop rule changes no bits
    want = float((b - a).abs().max())
    return st, want


@pytest.mark.parametrize("engine", [fold3d.ENGINE_PLANE, fold3d.ENGINE_HOISTED,
                                    fold3d.ENGINE_TILED, fold3d.ENGINE_SPLIT,
                                    fold3d.ENGINE_LOCKED])
@pytest.mark.parametrize("k", [1, 3])
def test_fold_max_dT_is_the_largest_change(solver, engine, k):
    n = 96  # 95 x 2 row blocks per plane launch: every word is used
    print(f"fold max|dT| n={n} engine={engine} k={k}", flush=True)
    T0 = fold3d.implicit_initialiser(n)
    st, want = reported_vs_host(solver, T0, n, k, engine=engine)
    assert st.iterations == k
    assert want > 0 and st.max_dT == want


@pytest.mark.parametrize("name", ["blob_speed", "tilted_field"])
def test_ellip_max_dT_is_the_largest_change(solver, name):
    n = 80
    print(f"elliptic max|dT| {name} n={n}", flush=True)
    W = getattr(wl, name)(n)
    T0 = wl.point_source_start(W)
    Wd = wl.Workload(W.name, W.n, W.h, W.M, W.speed, W.source, W.v_source, meta=W.meta)
    for f in ("speed_field", "theta", "phi", "ra", "rb"):
        v = getattr(W, f)
        setattr(Wd, f, None if v is None else dev(v))
    st, want = reported_vs_host(solver, T0, n, 2, spacing=W.h, model=Wd.model())
    assert want > 0 and st.max_dT == want
