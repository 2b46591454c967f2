"""The C-ABI library without a GPU: it loads, exports every symbol the headers
declare, validates arguments before touching HIP, reports errors as codes,
and its host-side initializer (the reference's ImplicitInitialiser, served by
libfold3d.so) is bitwise equal to the reference's output."""
import ctypes as C
import subprocess

import numpy as np
import pytest

import fold3d
from conftest import bits_equal


def test_library_loads():
    L = fold3d.lib()
    assert b"gfx950" in L.fold3d_version()


def test_exports_every_header_symbol():
    names = fold3d.header_symbols()
    assert "fold3d_solve" in names and "ImplicitInitialiser" in names
    out = subprocess.run(["nm", "-D", "--defined-only", fold3d.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_code_object_targets_gfx950():
    blob = open(fold3d.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


@pytest.mark.parametrize("n", [16, 24])
def test_host_initialiser_matches_reference(golden, n):
    assert bits_equal(fold3d.implicit_initialiser(n), golden[f"init_{n}"])


def test_host_point_source_matches_reference(golden):
    L = fold3d.lib()
    P, want = golden["point_in"], golden["point_out"]
    got = np.array([L.InitFoldFromPoint(*map(float, row)) for row in P])
    assert bits_equal(got, want)


def test_explicit_grid_voxel_update_matches_implicit():
    """fold3d_update_voxel / fold3d_init_node on an explicit (nx,ny,nz) give the
    same bits as updateVoxel / InitNode after fold3d_set_grid on that grid, and
    ignore whatever grid was recorded last.  The implicit pair is the one the
    golden 16^3 and 24^3 initializer fields pin (ImplicitInitialiser calls it)."""
    L = fold3d.lib()
    dp = C.POINTER(C.c_double)
    nx, ny, nz = 7, 9, 11
    h = (0.6, 0.45, 0.8)
    rng = np.random.default_rng(5)
    base = np.where(rng.random((nx + 2) * (ny + 2) * (nz + 2)) < 0.5, -1.0, 1.0) * 1e5
    # voxels whose 44 neighbours (offsets -1..2) stay inside the (n+2)^3 box
    vox = [(int(rng.integers(1, nx - 1)), int(rng.integers(1, ny - 1)),
            int(rng.integers(1, nz - 1))) for _ in range(12)]
    pts = rng.random((len(vox), 3)) * np.array(h)

    implicit = base.copy()
    L.fold3d_set_grid(nx, ny, nz)
    got = (C.c_int(), C.c_int(), C.c_int())
    L.fold3d_get_grid(*map(C.byref, got))
    assert [g.value for g in got] == [nx, ny, nz]
    for (i, j, k), p in zip(vox, pts):
        L.updateVoxel(implicit.ctypes.data_as(dp), i, j, k, *map(float, p), *h)
    L.InitNode(0.1, -0.2, 0.3, 2, 3, 4, implicit.ctypes.data_as(dp))

    explicit = base.copy()
    L.fold3d_set_grid(30, 5, 2)  # a different recorded grid must not matter
    for (i, j, k), p in zip(vox, pts):
        L.fold3d_update_voxel(explicit.ctypes.data_as(dp), i, j, k, *map(float, p), *h,
                              nx, ny, nz)
    L.fold3d_init_node(0.1, -0.2, 0.3, 2, 3, 4, explicit.ctypes.data_as(dp), nx, ny, nz)
    L.fold3d_set_grid(16, 16, 16)

    assert bits_equal(explicit, implicit)
    assert (implicit != base).sum() > 100  # the updates did reach many nodes
    # the sign of every node is kept; only magnitudes shrink
    assert np.all(np.sign(implicit) == np.sign(base))
    assert np.all(np.abs(implicit) <= np.abs(base))


def _solve(handle, T, n=16, dx=0.5, model=None, opts=None):
    L = fold3d.lib()
    m = model if model is not None else fold3d.reference_model()
    o = opts if opts is not None else fold3d.Opts()
    if opts is None:
        L.fold3d_default_opts(C.byref(o))
    st = fold3d.Stats()
    return L.fold3d_solve(handle, T, n, n, n, dx, dx, dx, C.byref(m), C.byref(o), C.byref(st))


def test_invalid_arguments_are_codes_not_crashes():
    L = fold3d.lib()
    dummy = C.c_void_p(0x1000)  # never dereferenced: validation fails first
    assert _solve(None, dummy) == fold3d.FOLD3D_E_INVALID
    assert "NULL" in fold3d.last_error()
    assert L.fold3d_set_batch_streams(None, 2) == fold3d.FOLD3D_E_INVALID
    assert L.fold3d_get_batch_streams(None, None) == fold3d.FOLD3D_E_INVALID
    assert _solve(dummy, None) == fold3d.FOLD3D_E_INVALID
    assert _solve(dummy, dummy, n=0) == fold3d.FOLD3D_E_INVALID
    assert "too small" in fold3d.last_error()
    assert _solve(dummy, dummy, dx=-1.0) == fold3d.FOLD3D_E_INVALID
    assert _solve(dummy, dummy, dx=float("nan")) == fold3d.FOLD3D_E_INVALID
    o = fold3d.Opts()
    L.fold3d_default_opts(C.byref(o))
    o.max_iters = 0
    assert _solve(dummy, dummy, opts=o) == fold3d.FOLD3D_E_INVALID
    o.max_iters, o.tol = 8, -1.0
    assert _solve(dummy, dummy, opts=o) == fold3d.FOLD3D_E_INVALID
    o.tol, o.precision = 0.0, 7
    assert _solve(dummy, dummy, opts=o) == fold3d.FOLD3D_E_INVALID
    m = fold3d.reference_model()
    m.F = 0.0
    assert _solve(dummy, dummy, model=m) == fold3d.FOLD3D_E_INVALID
    m = fold3d.reference_model()
    m.kind = 9
    assert _solve(dummy, dummy, model=m) in (fold3d.FOLD3D_E_INVALID, fold3d.FOLD3D_E_UNSUPPORTED)
    # plane step: direction and plane range are checked
    assert L.fold3d_plane_step(dummy, 6, dummy, 16, 16, 16, 3, *([1.0] * 10), None) == fold3d.FOLD3D_E_INVALID
    assert L.fold3d_plane_step(dummy, 0, dummy, 16, 16, 16, 16, *([1.0] * 10), None) == fold3d.FOLD3D_E_INVALID
    assert L.fold3d_plane_step(dummy, 1, dummy, 16, 16, 16, 1, *([1.0] * 10), None) == fold3d.FOLD3D_E_INVALID
    # the per-direction entry points check the same way
    assert L.fold3d_function_x_down(None, dummy, 16, 16, 16, 3, *([1.0] * 10), None) == fold3d.FOLD3D_E_INVALID
    assert L.fold3d_function_k_up(dummy, dummy, 16, 16, 16, 16, *([1.0] * 10), None) == fold3d.FOLD3D_E_INVALID
    assert L.fold3d_function_y_up(dummy, dummy, 16, 16, 16, 1, *([1.0] * 10), None) == fold3d.FOLD3D_E_INVALID


def test_fold_model_needs_a_finite_wind_slower_than_F():
    """F|grad T| + a.grad T = 1 is an eikonal equation only for |a| < F."""
    dummy = C.c_void_p(0x1000)  # never dereferenced: validation fails first
    for a in ((float("nan"), 0.0, 0.0), (0.0, float("inf"), 0.0), (1.1, 0.0, 0.0),
              (0.8, 0.8, 0.0), (0.0, 0.0, -2.0)):
        m = fold3d.fold_model(1.1, a)
        assert _solve(dummy, dummy, model=m) == fold3d.FOLD3D_E_INVALID, a
        assert "|a| < F" in fold3d.last_error()


def test_huge_grid_is_nomem_before_any_allocation():
    """A field too large for any device is FOLD3D_E_NOMEM from the size check,
    with no allocation tried (and no 64-bit overflow in the size arithmetic)."""
    L = fold3d.lib()
    dummy = C.c_void_p(0x1000)
    m = fold3d.reference_model()
    o = fold3d.Opts()
    L.fold3d_default_opts(C.byref(o))
    st = fold3d.Stats()
    for nx, ny, nz in ((2_000_000, 2_000_000, 2_000_000), (1 << 30, 1 << 30, 4)):
        rc = L.fold3d_solve(dummy, dummy, nx, ny, nz, 0.1, 0.1, 0.1, C.byref(m), C.byref(o),
                            C.byref(st))
        assert rc == fold3d.FOLD3D_E_NOMEM
        assert "does not fit" in fold3d.last_error()


def test_batch_instances_are_checked_on_the_host_first():
    """Non-finite or indefinite metrics and oversized instances are rejected
    before any pointer is queried."""
    L = fold3d.lib()
    dummy = C.c_void_p(0x1000)
    o = fold3d.Opts()
    L.fold3d_default_opts(C.byref(o))
    inst = (fold3d.Instance * 1)()
    x = inst[0]
    x.T = 0x1000
    x.n[0] = x.n[1] = x.n[2] = 8
    x.h[0] = x.h[1] = x.h[2] = 0.1
    x.M[0] = x.M[1] = x.M[2] = 1.0
    x.speed = 1.0

    def call():
        return L.fold3d_solve_batched(dummy, 1, inst, C.byref(o), None, None, None)

    x.M[4] = float("nan")
    assert call() == fold3d.FOLD3D_E_INVALID
    assert "not finite positive definite" in fold3d.last_error()
    x.M[4] = 0.0
    x.n[0] = x.n[1] = x.n[2] = 3_000_000
    assert call() == fold3d.FOLD3D_E_NOMEM
    x.n[0] = x.n[1] = x.n[2] = 8
    x.speed = float("nan")
    assert call() == fold3d.FOLD3D_E_INVALID


def test_point_source_arguments_are_checked():
    L = fold3d.lib()
    dummy = C.c_void_p(0x1000)  # never dereferenced: validation fails first
    src = (C.c_double * 3)(1.0, 1.0, 1.0)
    eye = (C.c_double * 6)(1.0, 1.0, 1.0, 0.0, 0.0, 0.0)
    bad = (C.c_double * 6)(1.0, 1.0, 1.0, 2.0, 0.0, 0.0)  # not positive definite

    def call(h=dummy, T=dummy, n=8, d=0.5, M=eye, v=1.0, r=1.6, prec=fold3d.F64):
        return L.fold3d_point_source(h, T, n, n, n, d, d, d, src, M, v, r, prec, None)

    assert call(h=None) == fold3d.FOLD3D_E_INVALID
    assert call(T=None) == fold3d.FOLD3D_E_INVALID
    assert call(n=0) == fold3d.FOLD3D_E_INVALID
    assert call(d=0.0) == fold3d.FOLD3D_E_INVALID
    assert call(v=-1.0) == fold3d.FOLD3D_E_INVALID
    assert call(v=float("inf")) == fold3d.FOLD3D_E_INVALID
    assert call(r=-1.0) == fold3d.FOLD3D_E_INVALID
    assert call(prec=5) == fold3d.FOLD3D_E_INVALID
    assert call(M=bad) == fold3d.FOLD3D_E_INVALID
    assert "positive definite" in fold3d.last_error()


def test_batched_adjoint_arguments_are_checked():
    L = fold3d.lib()
    dummy = C.c_void_p(0x1000)  # never dereferenced: validation fails first
    m = fold3d.Model()
    m.kind = fold3d.MODEL_ELLIP
    m.speed = 1.0
    m.M[0] = m.M[1] = m.M[2] = 1.0
    o = fold3d.Opts()
    L.fold3d_default_opts(C.byref(o))
    src = (fold3d.AdjointSource * 2)()
    src[0].T = src[1].T = 0x1000

    def call(h=dummy, count=2, s=src, n=8, model=m, grad=dummy, gt=None):
        return L.fold3d_adjoint_batched(h, count, s, n, n, n, 0.1, 0.1, 0.1,
                                        None if model is None else C.byref(model), C.byref(o),
                                        grad, gt, None, None, None, None)

    assert call(h=None) == fold3d.FOLD3D_E_INVALID
    assert call(s=None) == fold3d.FOLD3D_E_INVALID
    assert call(count=0) == fold3d.FOLD3D_E_INVALID
    assert call(n=0) == fold3d.FOLD3D_E_INVALID
    assert call(model=None) == fold3d.FOLD3D_E_INVALID
    assert call(grad=None) == fold3d.FOLD3D_E_INVALID
    assert call(gt=dummy) == fold3d.FOLD3D_E_INVALID  # no tilt fields
    assert "tilt" in fold3d.last_error()
    m.theta = 0x1000  # tilt fields come as all four or none
    assert call() == fold3d.FOLD3D_E_INVALID
    assert "all four" in fold3d.last_error()
    m.theta = None
    m.M[3] = 2.0  # not positive definite
    assert call() == fold3d.FOLD3D_E_INVALID
    assert "positive definite" in fold3d.last_error()
    m.M[3] = 0.0
    src[1].nrec = 3  # receivers without arrays
    assert call() == fold3d.FOLD3D_E_INVALID
    assert "source 1" in fold3d.last_error()
    src[1].nrec = 0
    src[1].tol = -1.0
    assert call() == fold3d.FOLD3D_E_INVALID


def test_create_without_gpu_is_an_error_code():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = fold3d.lib().fold3d_create(C.byref(h), 0)
    assert rc in (fold3d.FOLD3D_E_HIP, fold3d.FOLD3D_E_INVALID)
    assert fold3d.last_error()


def test_defaults_follow_the_reference():
    L = fold3d.lib()
    o = fold3d.Opts()
    L.fold3d_default_opts(C.byref(o))
    assert (o.precision, o.memory, o.max_iters, o.tol) == (fold3d.F64, fold3d.DEVICE, 8, 0.0)
    m = fold3d.reference_model()
    assert (m.F, tuple(m.a)) == (1.1, (-0.35, 0.4, 0.7))


def test_python_layer_checks_host_arrays_and_tensors_without_a_gpu():
    """The wrapper checks what the C-ABI cannot (it sees raw pointers): size,
    contiguity and dtype of every array, before any library call."""
    import torch
    s = fold3d.Solver.__new__(fold3d.Solver)  # no handle: nothing may reach the library
    s.h, s.device = None, 0
    n = 6
    cnt = (n + 2) ** 3
    with pytest.raises(ValueError):
        s.solve(np.zeros(cnt - 1), n)
    with pytest.raises(ValueError):
        s.solve(np.zeros(2 * cnt)[::2], n)
    with pytest.raises(TypeError):
        s.solve(np.zeros(cnt, dtype=np.float16), n)
    with pytest.raises(TypeError):
        s.solve(np.zeros(cnt, dtype=np.int64), n)
    with pytest.raises(TypeError):
        s.solve(torch.zeros(cnt, dtype=torch.bfloat16), n)
    with pytest.raises(ValueError):
        s.solve(torch.zeros(cnt + 1), n)
    with pytest.raises(ValueError):
        s.solve(torch.zeros(cnt, 2)[:, 0], n)
    with pytest.raises(TypeError):  # right size and dtype, but not on a GPU
        s.solve(torch.zeros(cnt, dtype=torch.float64), n)
    with pytest.raises(TypeError):
        s.initialiser_device(torch.zeros(cnt, dtype=torch.float32), n)
    with pytest.raises(ValueError):
        s.point_source(torch.zeros(cnt - 3), n, 0.1, (0.3, 0.3, 0.3), (1, 1, 1, 0, 0, 0), 1.0)
