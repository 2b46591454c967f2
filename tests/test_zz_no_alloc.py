"""fold3d_solve_noalloc (through the C-ABI) against fold3d_solve: same bits,
two calls in a row with the second on a smaller grid (the scratch keeps the
previous call's contents), FP64 and FP32, with and without the tolerance stop,
and no device memory taken by a call."""
import ctypes as C

import numpy as np
import pytest

import fold3d
from conftest import bits_equal

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _opts(precision, max_iters=8, tol=0.0, engine=fold3d.ENGINE_AUTO):
    o = fold3d.Opts()
    fold3d.lib().fold3d_default_opts(C.byref(o))
    o.precision, o.max_iters, o.tol, o.engine = precision, max_iters, tol, engine
    o.stream = torch.cuda.current_stream().cuda_stream
    return o


def _noalloc(T, n3, h, model, o):
    st = fold3d.Stats()
    rc = fold3d.lib().fold3d_solve_noalloc(C.c_void_p(T.data_ptr()), *n3, *h, C.byref(model),
                                           C.byref(o), C.byref(st))
    return rc, st


def _cases():
    """(name, n3, h, model, start field): the reference workload and a box grid
    with a random signed field and another model; the second is smaller."""
    n = 64
    ref = ("reference 64^3", (n, n, n), (fold3d.reference_spacing(n),) * 3,
           fold3d.reference_model(), fold3d.implicit_initialiser(n))
    n3 = (37, 29, 23)
    rng = np.random.default_rng(23)
    size = (n3[0] + 2) * (n3[1] + 2) * (n3[2] + 2)
    T0 = np.where(rng.random(size) < 0.3, -1.0, 1.0) * rng.uniform(1.0, 5.0, size)
    T0[rng.random(size) < 0.05] = fold3d.MAXT
    box = ("box 37x29x23", n3, (0.31, 0.27, 0.35), fold3d.fold_model(1.2, (0.25, -0.4, 0.3)), T0)
    return [ref, box]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_noalloc_equals_solve_twice_in_a_row(solver, dtype):
    prec = fold3d.F64 if dtype == torch.float64 else fold3d.F32
    iview = torch.int64 if dtype == torch.float64 else torch.int32
    for name, n3, h, model, T0 in _cases():
        print(f"noalloc vs solve: {name}, {dtype}", flush=True)
        A = torch.from_numpy(T0).to("cuda", dtype)
        B = A.clone()
        sa = solver.solve(A, n3, spacing=h, model=model, max_iters=8, engine=fold3d.ENGINE_SPLIT)
        rc, sb = _noalloc(B, n3, h, model, _opts(prec))
        torch.cuda.synchronize()
        assert rc == fold3d.FOLD3D_OK, fold3d.last_error()
        assert sb.iterations == sa.iterations == 8
        assert sb.cell_updates == sa.cell_updates
        assert torch.equal(A.view(iview), B.view(iview)), name


def test_noalloc_tolerance_stop_matches_solve(solver, golden):
    n = 32
    T0 = fold3d.implicit_initialiser(n)
    A = torch.from_numpy(T0).cuda()
    rc, st = _noalloc(A, (n, n, n), (fold3d.reference_spacing(n),) * 3, fold3d.reference_model(),
                      _opts(fold3d.F64, max_iters=100, tol=1e-300))
    torch.cuda.synchronize()
    assert rc == fold3d.FOLD3D_OK, fold3d.last_error()
    assert st.iterations == int(golden[f"fixed_iters_{n}"][0]) and st.max_dT == 0.0
    assert bits_equal(A.cpu().numpy(), golden[f"fixed_{n}"])
    B = torch.from_numpy(T0).cuda()
    rc, st = _noalloc(B, (n, n, n), (fold3d.reference_spacing(n),) * 3, fold3d.reference_model(),
                      _opts(fold3d.F64, max_iters=2, tol=1e-300))
    assert rc == fold3d.FOLD3D_E_NOT_CONVERGED and st.iterations == 2


def test_noalloc_two_waves_per_row_on_a_large_grid(solver):
    """A grid large enough for 2 waves per row (side >= 448), one iteration."""
    n3 = (460, 450, 440)
    rng = np.random.default_rng(5)
    size = (n3[0] + 2) * (n3[1] + 2) * (n3[2] + 2)
    A = torch.where(torch.rand(size, device="cuda", dtype=torch.float64) < 0.3, -1.0, 1.0) * \
        (1.0 + 4.0 * torch.rand(size, device="cuda", dtype=torch.float64))
    B = A.clone()
    h = (0.02, 0.021, 0.019)
    m = fold3d.reference_model()
    solver.solve(A, n3, spacing=h, model=m, max_iters=1, engine=fold3d.ENGINE_SPLIT)
    rc, _ = _noalloc(B, n3, h, m, _opts(fold3d.F64, max_iters=1))
    torch.cuda.synchronize()
    assert rc == fold3d.FOLD3D_OK, fold3d.last_error()
    assert torch.equal(A.view(torch.int64), B.view(torch.int64))
    del A, B, rng


def test_noalloc_takes_no_device_memory(solver):
    n = 48
    T0 = torch.from_numpy(fold3d.implicit_initialiser(n)).cuda()
    h = (fold3d.reference_spacing(n),) * 3
    T = T0.clone()
    assert _noalloc(T, (n, n, n), h, fold3d.reference_model(), _opts(fold3d.F64))[0] == 0
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(3):
        T.copy_(T0)
        assert _noalloc(T, (n, n, n), h, fold3d.reference_model(), _opts(fold3d.F64))[0] == 0
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0


def test_noalloc_rejects_what_it_does_not_do(solver):
    n = 8
    T = torch.zeros((n + 2) ** 3, dtype=torch.float64, device="cuda")
    h = (0.1,) * 3
    m = fold3d.reference_model()
    host = np.zeros((n + 2) ** 3)
    st = fold3d.Stats()
    L = fold3d.lib()
    o = _opts(fold3d.F64)
    assert L.fold3d_solve_noalloc(C.c_void_p(host.ctypes.data), n, n, n, *h, C.byref(m),
                                  C.byref(o), C.byref(st)) == fold3d.FOLD3D_E_INVALID
    assert "device memory" in fold3d.last_error()
    assert _noalloc(T, (n, n, n), h, m, _opts(fold3d.F64, engine=fold3d.ENGINE_PLANE))[0] == \
        fold3d.FOLD3D_E_UNSUPPORTED
    me = fold3d.reference_model()
    me.kind = fold3d.MODEL_ELLIP
    assert _noalloc(T, (n, n, n), h, me, _opts(fold3d.F64))[0] == fold3d.FOLD3D_E_UNSUPPORTED
    assert _noalloc(T, (3_000_000,) * 3, h, m, _opts(fold3d.F64))[0] == fold3d.FOLD3D_E_NOMEM
    bad = fold3d.fold_model(1.1, (1.2, 0.0, 0.0))
    assert _noalloc(T, (n, n, n), h, bad, _opts(fold3d.F64))[0] == fold3d.FOLD3D_E_INVALID
