"""Parity of the HIP elliptic-model solve with the elliptic oracle.

FP64 with a constant metric is bit-for-bit equal to the oracle (same
constants, same candidate arithmetic; the kernel prunes simplices that cannot
win, the oracle evaluates all of them).  With a per-node metric the device
evaluates sin/cos itself, so FP64 is held to 1e-12 relative.  FP32 is held to
1e-5 relative of the FP64 oracle on nodes at least 8 cells from the source
and to 1e-5 * max(T) absolute everywhere.
"""
import numpy as np
import pytest

from conftest import bits_equal

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import fold3d  # noqa: E402
from fold3d import workloads as wl  # noqa: E402


def dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def oracle_fields(W):
    out = {}
    for k in ("speed_field", "theta", "phi", "ra", "rb"):
        f = getattr(W, k)
        out[k] = None if f is None else np.asarray(f, dtype=np.float64)
    return out


def to_device_workload(W, dtype):
    Wd = wl.Workload(W.name, W.n, W.h, W.M, W.speed, W.source, W.v_source, meta=W.meta)
    for k in ("speed_field", "theta", "phi", "ra", "rb"):
        f = getattr(W, k)
        setattr(Wd, k, None if f is None else dev(f, dtype))
    return Wd


@pytest.mark.parametrize("n", [16, 33])
def test_homogeneous_bitwise(solver, oracle, n):
    print(f"ellip homogeneous n={n}", flush=True)
    W = wl.tilted_homogeneous(n)
    T0 = wl.point_source_start(W)
    want = T0.copy()
    it_o, md_o = oracle.ellip_sweeps(want, n, 3, 0.0, W.h, M=W.M, speed=W.speed)
    T = dev(T0)
    st = solver.solve(T, n, spacing=W.h, model=W.model(), max_iters=3)
    assert bits_equal(T.cpu().numpy(), want)


def test_blob_speed_bitwise_and_converges(solver, oracle):
    n = 28
    W = wl.blob_speed(n)
    T0 = wl.point_source_start(W)
    want = T0.copy()
    it_o, md_o = oracle.ellip_sweeps(want, n, 40, 1e-10, W.h, M=W.M, speed_field=W.speed_field)
    Wd = to_device_workload(W, torch.float64)
    T = dev(T0)
    st = solver.solve(T, n, spacing=W.h, model=Wd.model(), max_iters=40, tol=1e-10)
    assert st.iterations == it_o and st.max_dT == md_o
    assert st.bytes_moved == st.updates_computed * 3 * 8  # target read + write, speed read
    assert bits_equal(T.cpu().numpy(), want)


def test_per_node_metric_fp64(solver, oracle):
    n = 20
    W = wl.tilted_field(n, dtype=np.float64)
    T0 = wl.point_source_start(W)
    want = T0.copy()
    f = oracle_fields(W)
    oracle.ellip_sweeps(want, n, 4, 0.0, W.h, M=W.M, **f)
    Wd = to_device_workload(W, torch.float64)
    T = dev(T0)
    solver.solve(T, n, spacing=W.h, model=Wd.model(), max_iters=4)
    got = T.cpu().numpy()
    reached = want < 1e15
    assert np.array_equal(reached, got < 1e15)
    rel = np.abs(got[reached] - want[reached]) / want[reached].clip(min=1e-300)
    print(f"per-node metric FP64: max rel {rel.max():.2e}")
    assert rel.max() < 1e-12


def fp32_error(got, want, n, h, vmax):
    reached = want < 1e15
    far = reached & (want > 8 * h / vmax)
    rel = np.abs(got[far] - want[far]) / want[far]
    absd = np.abs(got[reached] - want[reached]).max() / want[reached].max()
    return rel.max(), absd


@pytest.mark.parametrize("which", ["homogeneous", "blob", "field"])
def test_fp32_vs_fp64_oracle(solver, oracle, which):
    n = 40
    if which == "homogeneous":
        W = wl.tilted_homogeneous(n)
        vmax = 1.0
    elif which == "blob":
        W = wl.blob_speed(n)
        vmax = 2.3
    else:
        W = wl.tilted_field(n, dtype=np.float64)
        vmax = 3.6
    T0 = wl.point_source_start(W)
    want = T0.copy()
    it, md = oracle.ellip_sweeps(want, n, 60, 1e-10, W.h, M=W.M, **oracle_fields(W))
    Wd = to_device_workload(W, torch.float32)
    T = dev(T0, torch.float32)
    st = solver.solve(T, n, spacing=W.h, model=Wd.model(), max_iters=60, tol=1e-6)
    got = T.double().cpu().numpy()
    rel, absd = fp32_error(got, want, n, W.h, vmax)
    print(f"FP32 {which}: {st.iterations} iterations (FP64 {it}); max rel {rel:.2e}, max abs/maxT {absd:.2e}")
    assert rel < 1e-5 and absd < 1e-5


def test_closed_form_error_reported(solver):
    """Homogeneous tilted medium: converged PMM field vs sqrt(x^T M^-1 x)/v."""
    n = 64
    W = wl.tilted_homogeneous(n)
    T = dev(wl.point_source_start(W))
    st = solver.solve(T, n, spacing=W.h, model=W.model(), max_iters=100, tol=1e-10)
    got = T.cpu().numpy()
    exact = wl.closed_form(W)
    H = n + 2
    inner = np.zeros((H, H, H), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    inner = inner.reshape(-1)
    err = np.abs(got[inner] - exact[inner])
    print(f"closed form n={n}: {st.iterations} iterations, max err {err.max():.3e}, "
          f"mean {err.mean():.3e}, max rel {(err / exact[inner].clip(min=W.h)).max():.3e}")
    assert st.max_dT < 1e-10
    assert np.all(got[inner] >= exact[inner] - 1e-12)
    assert err.max() < 0.05


@pytest.mark.parametrize("groups", [1, 2, 4])
@pytest.mark.parametrize("with_speed", [False, True])
def test_split_engine_bitwise_on_box_grid(solver, oracle, groups, with_speed):
    """Layout changes + LDS rows + units over waves: same bits as the oracle,
    on a box grid with partial tiles, with and without a speed field."""
    n3 = (45, 70, 38)
    h = (1.0 / 44, 1.0 / 50, 1.0 / 41)
    W = wl.tilted_homogeneous(20)
    rng = np.random.default_rng(groups)
    size = (n3[0] + 2) * (n3[1] + 2) * (n3[2] + 2)
    # a smooth start: closed form from an interior point with a few nodes set
    k, j, i = np.indices((n3[2] + 2, n3[1] + 2, n3[0] + 2)).reshape(3, -1)
    X = np.stack([(i - 1) * h[0], (j - 1) * h[1], (k - 1) * h[2]]) - np.array([[0.4], [0.7], [0.3]])
    N = np.linalg.inv(wl.sym(W.M))
    r = np.sqrt(np.einsum("in,ij,jn->n", X, N, X))
    T0 = np.where(r < 0.06, r, 1e16)
    ghost = (i == 0) | (j == 0) | (k == 0) | (i == n3[0] + 1) | (j == n3[1] + 1) | (k == n3[2] + 1)
    T0[ghost] = 1e16
    v = 1.0 + 0.5 * rng.random(size) if with_speed else None
    want = T0.copy()
    oracle.ellip_sweeps(want, n3, 3, 0.0, h, M=W.M, speed=1.3, speed_field=v)
    m = W.model()
    m.speed = 1.3
    vd = dev(v) if with_speed else None
    m.speed_field = vd.data_ptr() if with_speed else None
    solver.set_tuning(groups, fold3d.MODEL_ELLIP)
    try:
        T = dev(T0)
        solver.solve(T, n3, spacing=h, model=m, max_iters=3, engine=fold3d.ENGINE_SPLIT)
        assert bits_equal(T.cpu().numpy(), want)
        T = dev(T0)
        solver.solve(T, n3, spacing=h, model=m, max_iters=3, engine=fold3d.ENGINE_PLANE)
        assert bits_equal(T.cpu().numpy(), want)
    finally:
        solver.set_tuning(0, fold3d.MODEL_ELLIP)


def test_bad_metric_is_rejected(solver):
    m = wl.tilted_homogeneous(16).model()
    m.M[0] = -1.0
    T = torch.zeros(18 ** 3, dtype=torch.float64, device="cuda")
    with pytest.raises(fold3d.Fold3dError) as e:
        solver.solve(T, 16, spacing=0.1, model=m)
    assert e.value.code == fold3d.FOLD3D_E_INVALID


@pytest.mark.parametrize("groups", [1, 2, 4])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_metric_engine_matches_plane_engine(solver, oracle, groups, dtype):
    """Per-node metric: hoisted metric fields + layouts + LDS rows + units over
    waves give the same bits as the one-thread-per-node plane kernel, on a box
    grid with partial rows; FP64 also within 1e-12 of the oracle."""
    n3 = (45, 70, 38)
    h = (1.0 / 44, 1.0 / 50, 1.0 / 41)
    W = wl.tilted_homogeneous(20)
    rng = np.random.default_rng(groups)
    size = (n3[0] + 2) * (n3[1] + 2) * (n3[2] + 2)
    k, j, i = np.indices((n3[2] + 2, n3[1] + 2, n3[0] + 2)).reshape(3, -1)
    X = np.stack([(i - 1) * h[0], (j - 1) * h[1], (k - 1) * h[2]]) - np.array([[0.4], [0.7], [0.3]])
    N = np.linalg.inv(wl.sym(W.M))
    r = np.sqrt(np.einsum("in,ij,jn->n", X, N, X))
    T0 = np.where(r < 0.06, r, 1e16)
    ghost = (i == 0) | (j == 0) | (k == 0) | (i == n3[0] + 1) | (j == n3[1] + 1) | (k == n3[2] + 1)
    T0[ghost] = 1e16
    fields = {"speed_field": 1.0 + 0.5 * rng.random(size),
              "theta": rng.uniform(0, 1.0, size), "phi": rng.uniform(0, 6.2, size),
              "ra": rng.uniform(0.6, 1.0, size), "rb": rng.uniform(0.6, 1.0, size)}
    tdt = torch.float64 if dtype == "f64" else torch.float32
    dfields = {k: dev(v, tdt) for k, v in fields.items()}
    m = W.model()
    for k, v in dfields.items():
        setattr(m, k, v.data_ptr())
    solver.set_tuning(groups, fold3d.MODEL_ELLIP)
    try:
        out = {}
        for name, eng in (("metric", fold3d.ENGINE_SPLIT), ("plane", fold3d.ENGINE_PLANE)):
            T = dev(T0, tdt)
            solver.solve(T, n3, spacing=h, model=m, max_iters=3, engine=eng)
            out[name] = T.double().cpu().numpy()
        assert bits_equal(out["metric"], out["plane"])
    finally:
        solver.set_tuning(0, fold3d.MODEL_ELLIP)
    if dtype == "f64":
        want = T0.copy()
        oracle.ellip_sweeps(want, n3, 3, 0.0, h, M=W.M, **fields)
        got = out["metric"]
        reached = want < 1e15
        assert np.array_equal(reached, got < 1e15)
        rel = np.abs(got[reached] - want[reached]) / want[reached].clip(min=1e-300)
        assert rel.max() < 1e-12


@pytest.mark.parametrize("n3", [(1, 1, 1), (2, 2, 2), (1, 5, 3), (3, 2, 9), (65, 1, 2)])
def test_tiny_grids(solver, oracle, n3):
    """Constant metric (split engine, bitwise vs the oracle) and per-node metric
    (metric engine, bitwise vs the plane kernel) with one or two nodes along an axis."""
    h = (0.1, 0.12, 0.09)
    size = (n3[0] + 2) * (n3[1] + 2) * (n3[2] + 2)
    rng = np.random.default_rng(sum(n3))
    T0 = np.where(rng.random(size) < 0.3, rng.uniform(0.0, 1.0, size), 1e16)
    W = wl.tilted_homogeneous(8)
    want = T0.copy()
    oracle.ellip_sweeps(want, n3, 2, 0.0, h, M=W.M, speed=1.0)
    T = dev(T0)
    solver.solve(T, n3, spacing=h, model=W.model(), max_iters=2)
    assert bits_equal(T.cpu().numpy(), want)
    fields = {"theta": rng.uniform(0, 1, size), "phi": rng.uniform(0, 6, size),
              "ra": rng.uniform(0.6, 1, size), "rb": rng.uniform(0.6, 1, size)}
    d = {k: dev(v) for k, v in fields.items()}
    m = W.model()
    for k, v in d.items():
        setattr(m, k, v.data_ptr())
    out = []
    for eng in (fold3d.ENGINE_SPLIT, fold3d.ENGINE_PLANE):
        T = dev(T0)
        solver.solve(T, n3, spacing=h, model=m, max_iters=2, engine=eng)
        out.append(T.cpu().numpy())
    assert bits_equal(out[0], out[1])


@pytest.mark.parametrize("name,n", [("tilted_homogeneous", 20), ("tilted_field", 24)])
def test_device_point_source_start(solver, name, n):
    """fold3d_point_source against the host start field: MAXT at the same nodes,
    values within 1e-14 relative (the host inverts the metric by LU, the library
    by cofactors).  FP32 is the FP64 value rounded, to one ulp."""
    print(f"point source {name} n={n}", flush=True)
    W = getattr(wl, name)(n)
    want = wl.point_source_start(W)
    M = W.meta.get("metric_at_source", W.M)
    Mv = [float(x) for x in np.asarray(M).ravel()]  # (xx, yy, zz, xy, xz, yz)
    for dtype in (torch.float64, torch.float32):
        T = torch.empty((n + 2) ** 3, dtype=dtype, device="cuda")
        solver.point_source(T, n, W.h, W.source, Mv, W.v_source)
        torch.cuda.synchronize()
        got = T.cpu().numpy().astype(np.float64)
        far = want >= fold3d.MAXT
        assert np.array_equal(got >= fold3d.MAXT * (1 - 1e-6), far)
        assert int((~far).sum()) >= 8
        ref = want[~far] if dtype == torch.float64 else want[~far].astype(np.float32).astype(np.float64)
        rtol = 1e-14 if dtype == torch.float64 else 1.2e-7
        np.testing.assert_allclose(got[~far], ref, rtol=rtol, atol=0)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("ratios", [(0.8, 0.6), (0.25, 0.1)])
def test_plane_bound_pruning_keeps_the_bits(solver, oracle, dtype, ratios):
    """The split engine skips simplices by the plane bound (smallest vertex +
    s times the N-distance to the upwind plane), the plane engine by the vertex
    test alone: the same bits in FP64 (both equal to the oracle) and in FP32,
    for a mild and a strongly anisotropic metric (condition numbers of N 2.8
    and 100), with a speed field."""
    torch_dtype = getattr(torch, dtype)
    n3 = (36, 41, 33)
    h = (1.0 / 35, 1.0 / 37, 1.0 / 30)
    print(f"plane bound {dtype} ratios={ratios}", flush=True)
    M = np.asarray(wl.tilted_metric(0.6, 2.2, 1.0, ratios[0], ratios[1]))
    k, j, i = np.indices((n3[2] + 2, n3[1] + 2, n3[0] + 2)).reshape(3, -1)
    X = np.stack([(i - 1) * h[0], (j - 1) * h[1], (k - 1) * h[2]]) - np.array([[0.45], [0.5], [0.4]])
    N = np.linalg.inv(wl.sym(M))
    r = np.sqrt(np.einsum("in,ij,jn->n", X, N, X))
    near = (np.abs(X) <= 1.6 * np.array(h)[:, None]).all(axis=0)  # closed form within 1.6 cells
    T0 = np.where(near, r, 1e16)
    ghost = (i == 0) | (j == 0) | (k == 0) | (i == n3[0] + 1) | (j == n3[1] + 1) | (k == n3[2] + 1)
    T0[ghost] = 1e16
    assert near.sum() >= 8
    v = 1.0 + 0.5 * np.random.default_rng(7).random(T0.size)
    m = fold3d.Model()
    m.kind = fold3d.MODEL_ELLIP
    for q in range(6):
        m.M[q] = float(M[q])
    vd = dev(v, torch_dtype)
    m.speed_field = vd.data_ptr()
    out = {}
    for name, eng in (("split", fold3d.ENGINE_SPLIT), ("plane", fold3d.ENGINE_PLANE)):
        T = dev(T0, torch_dtype)
        solver.solve(T, n3, spacing=h, model=m, max_iters=4, engine=eng)
        out[name] = T.cpu().numpy()
    assert bits_equal(out["split"], out["plane"])
    assert (out["split"][~ghost] < 1e15).all()  # every interior node reached
    if dtype == "float64":
        want = T0.copy()
        oracle.ellip_sweeps(want, n3, 4, 0.0, h, M=M, speed_field=v)
        assert bits_equal(out["split"], want)
