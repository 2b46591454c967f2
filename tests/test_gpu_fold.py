"""Parity of the HIP fold-model solve (through the C-ABI) with the reference.

FP64 is held bit-for-bit: against golden vectors written from the reference's
own solver and, at sizes the fixtures do not cover, against the oracle on the
same inputs.  FP32 is held to a stated relative tolerance of the FP64 result.
"""
import hashlib

import numpy as np
import pytest

from conftest import bits_equal

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import fold3d  # noqa: E402

# Driver arguments per direction for a cube with spacing h
# (reference AN_ISO_1_OpenACC.c:1198-1235): (dzz, dxx, dyy, dxy, dxz, dyz, F, ay, ax, az)
AXX, AYY, AZZ = fold3d.REF_A
FC = fold3d.REF_F


def driver_args(d, h):
    dg = np.sqrt(h * h + h * h)
    sp = (h, h, h, dg, dg, dg)
    model = {
        0: (FC, AYY, AZZ, AXX), 1: (FC, -AYY, -AZZ, -AXX),
        2: (FC, AXX, AYY, AZZ), 3: (FC, -AXX, -AYY, -AZZ),
        4: (FC, -AZZ, -AXX, -AYY), 5: (FC, AZZ, AXX, AYY),
    }[d]
    return sp + model


def planes_of(d, n):
    return range(1, n) if d in (0, 2, 5) else range(n, 1, -1)


def to_dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


@pytest.fixture(autouse=True)
def lock_from_the_start(solver):
    """Locked solves in this module lock from iteration 0 (every sweep after
    the first iteration may skip rows), the strictest test of exactness; the
    default start is tested on its own below."""
    solver.set_lock_start(0)
    yield
    solver.set_lock_start(-1)


@pytest.mark.parametrize("d", range(6))
def test_plane_step_random_field(solver, golden, d):
    print(f"plane step direction {d} on the 16^3 random field", flush=True)
    planes = {0: 7, 1: 9, 2: 7, 3: 9, 4: 9, 5: 7}
    T = to_dev(golden["plane_rand_in_16"])
    h = fold3d.reference_spacing(16)
    solver.plane_step(d, T, 16, planes[d], *driver_args(d, h))
    torch.cuda.synchronize()
    assert bits_equal(T.cpu().numpy(), golden[f"plane_rand_{d}_16"])


def test_per_direction_entry_points(solver, golden):
    """fold3d_function_x_down & co. (handle, grid and stream added to the
    reference's plane-step signature) on the 16^3 random field, each against
    the reference's own plane step."""
    import ctypes as C
    L = fold3d.lib()
    planes = {0: 7, 1: 9, 2: 7, 3: 9, 4: 9, 5: 7}
    h = fold3d.reference_spacing(16)
    for d, name in enumerate(fold3d.DIRECTIONS):
        T = to_dev(golden["plane_rand_in_16"])
        fold3d.check(getattr(L, "fold3d_function_" + name)(
            solver.h, C.c_void_p(T.data_ptr()), 16, 16, 16, planes[d], *driver_args(d, h),
            None))
        torch.cuda.synchronize()
        assert bits_equal(T.cpu().numpy(), golden[f"plane_rand_{d}_16"]), name
    # out-of-range plane is an error code, not a launch
    assert getattr(L, "fold3d_function_x_up")(solver.h, C.c_void_p(T.data_ptr()), 16, 16, 16,
                                              1, *driver_args(1, h), None) \
        == fold3d.FOLD3D_E_INVALID


@pytest.mark.parametrize("d", range(6))
def test_direction_on_initial_field(solver, golden, d):
    T = to_dev(golden["init_16"])
    h = fold3d.reference_spacing(16)
    for p in planes_of(d, 16):
        solver.plane_step(d, T, 16, p, *driver_args(d, h))
    torch.cuda.synchronize()
    assert bits_equal(T.cpu().numpy(), golden[f"direction_init_{d}_16"])


ENGINES = [fold3d.ENGINE_PLANE, fold3d.ENGINE_HOISTED, fold3d.ENGINE_TILED, fold3d.ENGINE_SPLIT,
           fold3d.ENGINE_LOCKED]


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("n", [16, 24, 32])
def test_full_solve_matches_reference(solver, golden, n, engine):
    print(f"full solve n={n} engine={engine}", flush=True)
    T = to_dev(fold3d.implicit_initialiser(n))
    solver.solve(T, n, max_iters=1, engine=engine)
    assert bits_equal(T.cpu().numpy(), golden[f"solve1_{n}"])
    solver.solve(T, n, max_iters=7, engine=engine)
    assert bits_equal(T.cpu().numpy(), golden[f"solve8_{n}"])


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("n3", [(19, 27, 23), (70, 9, 33), (5, 66, 130)])
def test_box_grid_matches_oracle(solver, oracle, n3, engine):
    """Unequal axes exercise every layout change and partial tiles."""
    print(f"box grid {n3} engine={engine}", flush=True)
    rng = np.random.default_rng(sum(n3))
    size = (n3[0] + 2) * (n3[1] + 2) * (n3[2] + 2)
    T0 = np.where(rng.random(size) < 0.25, -1.0, 1.0) * rng.uniform(1.0, 5.0, size)
    T0[rng.random(size) < 0.05] = fold3d.MAXT
    h = (0.31, 0.27, 0.35)
    a = (0.25, -0.4, 0.3)
    want = T0.copy()
    oracle.sweeps3(want, n3, 2, 1.2, a, h)
    T = to_dev(T0)
    solver.solve(T, n3, spacing=h, model=fold3d.fold_model(1.2, a), max_iters=2, engine=engine)
    assert bits_equal(T.cpu().numpy(), want)


@pytest.mark.parametrize("groups", [2, 4, 8])
def test_split_groups_are_bitwise_equal(solver, oracle, groups):
    n3 = (37, 70, 29)
    rng = np.random.default_rng(11)
    size = (n3[0] + 2) * (n3[1] + 2) * (n3[2] + 2)
    T0 = np.where(rng.random(size) < 0.3, -1.0, 1.0) * rng.uniform(1.0, 5.0, size)
    h, a = (0.3, 0.3, 0.3), (-0.35, 0.4, 0.7)
    want = T0.copy()
    oracle.sweeps3(want, n3, 2, 1.1, a, h)
    solver.set_tuning(groups, fold3d.MODEL_FOLD)
    try:
        T = to_dev(T0)
        solver.solve(T, n3, spacing=h, model=fold3d.fold_model(1.1, a), max_iters=2,
                     engine=fold3d.ENGINE_SPLIT)
        assert bits_equal(T.cpu().numpy(), want)
    finally:
        solver.set_tuning(0, fold3d.MODEL_FOLD)


def test_autotune_keeps_input_and_picks_a_choice(solver):
    n = 48
    T0 = to_dev(fold3d.implicit_initialiser(n))
    before = T0.clone()
    groups, ms = solver.autotune(T0, n, iterations=1)
    assert groups in (2, 4, 8) and ms > 0
    assert torch.equal(T0.view(torch.int64), before.view(torch.int64))
    solver.set_tuning(0, fold3d.MODEL_FOLD)
    assert solver.tuning() == 4


@pytest.mark.parametrize("n", [16, 24, 32])
def test_convergence_stop_reaches_fixed_point(solver, golden, n):
    T = to_dev(fold3d.implicit_initialiser(n))
    st = solver.solve(T, n, max_iters=100, tol=1e-300)
    assert st.iterations == int(golden[f"fixed_iters_{n}"][0])
    assert st.max_dT == 0.0
    assert bits_equal(T.cpu().numpy(), golden[f"fixed_{n}"])


def test_not_converged_is_reported(solver):
    n = 24
    T = to_dev(fold3d.implicit_initialiser(n))
    with pytest.raises(fold3d.Fold3dError) as e:
        solver.solve(T, n, max_iters=2, tol=1e-300)
    assert e.value.code == fold3d.FOLD3D_E_NOT_CONVERGED
    st = solver.solve(T, n, max_iters=1, tol=1e-300, raise_not_converged=False)
    assert st.iterations == 1 and st.max_dT > 0


def test_solve_128_matches_reference_digest(solver, golden):
    print("full solve n=128, 8 iterations", flush=True)
    T = to_dev(fold3d.implicit_initialiser(128))
    st = solver.solve(T, 128)
    out = T.cpu().numpy()
    assert np.array_equal(out[golden["solve8_128_idx"]], golden["solve8_128_val"])
    assert hashlib.sha256(out.tobytes()).digest() == golden["solve8_128_sha256"].tobytes()
    assert st.cell_updates == 8 * 6 * 127 * 128 * 128
    assert st.bytes_moved == st.updates_computed * 16  # read and write each target, FP64


@pytest.mark.parametrize("engine", ENGINES)
def test_solves_64_match_reference_digests(solver, golden_large, engine):
    """64^3 after 1 and 8 iterations and at the fixed point (the convergence
    stop), each against the digest of the reference's own solve."""
    from conftest import matches_digest
    n = 64
    print(f"64^3 digests, engine {engine}", flush=True)
    for iters, key in ((1, "solve1_64"), (8, "solve8_64")):
        T = to_dev(fold3d.implicit_initialiser(n))
        solver.solve(T, n, max_iters=iters, engine=engine)
        assert matches_digest(T.cpu().numpy(), golden_large, key)
    T = to_dev(fold3d.implicit_initialiser(n))
    st = solver.solve(T, n, max_iters=100, tol=1e-300, engine=engine)
    assert st.iterations == int(golden_large["fixed_iters_64"][0])
    assert matches_digest(T.cpu().numpy(), golden_large, "fixed_64")


@pytest.mark.parametrize("engine", [fold3d.ENGINE_SPLIT, fold3d.ENGINE_LOCKED,
                                    fold3d.ENGINE_AUTO])
def test_fixed_point_128_matches_reference_digest(solver, golden_large, engine):
    """The reference's fixed point at 128^3 (34 iterations).  ENGINE_AUTO runs
    the split engine and, on a solve allowed past 24 iterations, locks from
    iteration 24 on: the same bits, with rows skipped in the last ten."""
    from conftest import matches_digest
    n = 128
    print(f"128^3 to the fixed point, engine {engine}", flush=True)
    T = to_dev(fold3d.implicit_initialiser(n))
    st = solver.solve(T, n, max_iters=100, tol=1e-300, engine=engine)
    assert st.iterations == int(golden_large["fixed_iters_128"][0])
    assert matches_digest(T.cpu().numpy(), golden_large, "fixed_128")
    if engine == fold3d.ENGINE_AUTO:
        # the first 24 iterations compute every update, the locked ones fewer
        w = st.cell_updates / st.iterations
        assert 24 * w <= st.updates_computed < st.cell_updates
        T = to_dev(fold3d.implicit_initialiser(n))
        st = solver.solve(T, n, max_iters=24, engine=engine)  # no locking within 24
        assert st.updates_computed == st.cell_updates


def test_solve_64_matches_oracle(solver, oracle):
    n = 64
    T0 = fold3d.implicit_initialiser(n)
    want = T0.copy()
    oracle.sweeps(want, n, 8)
    T = to_dev(T0)
    solver.solve(T, n)
    assert bits_equal(T.cpu().numpy(), want)


def test_random_field_and_model_match_oracle(solver, oracle):
    """A model other than the reference's (F, a) and an unstructured field."""
    n = 20
    rng = np.random.default_rng(7)
    H = n + 2
    T0 = np.where(rng.random(H ** 3) < 0.3, -1.0, 1.0) * rng.uniform(1.0, 4.0, H ** 3)
    F, a, h = 1.3, (0.2, -0.5, 0.3), (0.3, 0.3, 0.3)
    want = T0.copy()
    oracle.sweeps(want, n, 2, F=F, a=a, h=h)
    T = to_dev(T0)
    solver.solve(T, n, spacing=h, model=fold3d.fold_model(F, a), max_iters=2)
    assert bits_equal(T.cpu().numpy(), want)


def test_host_memory_solve(solver, golden):
    T = fold3d.implicit_initialiser(24)
    solver.solve(T, 24, max_iters=8)
    assert bits_equal(T, golden["solve8_24"])


def test_reference_signature_plane_steps(golden):
    """function_x_down & co. with the reference's signatures on the grid set
    by ImplicitInitialiser()."""
    import ctypes as C
    L = fold3d.lib()
    n = 16
    T = to_dev(fold3d.implicit_initialiser(n))  # sets the grid to 16^3
    h = fold3d.reference_spacing(n)
    for d, name in enumerate(fold3d.DIRECTIONS):
        for p in planes_of(d, n):
            getattr(L, "function_" + name)(C.c_void_p(T.data_ptr()), p, *driver_args(d, h))
    torch.cuda.synchronize()
    want = golden["init_16"].copy()
    from oracle.oracle import Oracle
    o = Oracle()
    for d in range(6):
        o.direction(d, want, n)
    assert bits_equal(T.cpu().numpy(), want)
    assert bits_equal(want, golden["solve1_16"])


def test_deterministic(solver):
    n = 48
    T0 = fold3d.implicit_initialiser(n)
    A, B = to_dev(T0), to_dev(T0)
    solver.solve(A, n, max_iters=3)
    solver.solve(B, n, max_iters=3)
    assert torch.equal(A.view(torch.int64), B.view(torch.int64))


# The FP32 bound, stated for the reference workload (8 iterations) up to 512^3,
# over reached nodes, relative to the FP64 field (which equals the reference's
# bits): signs and the reached set identical; max < 1e-2; at most 0.5 % of
# nodes above 1e-3, 10 % above 1e-4, 25 % above 1e-5.  Measured at 512^3:
# max 4.8e-3, 0.25 %, 8.4 %, 19.2 %; at 128^3: 4.1e-3, 0.08 %, 2.6 %, 6.2 %.
FP32_BOUND = {"max": 1e-2, 1e-3: 0.005, 1e-4: 0.10, 1e-5: 0.25}


@pytest.mark.parametrize("n", [32, 64, 128])
def test_fp32_within_stated_bound_of_fp64(solver, oracle, n):
    """FP32 field vs the FP64 oracle on the same input.  The FP32 engines give
    the literal kernel's FP32 bits (test_fp32_engines_equal_the_fp32_literal_kernel),
    so this is the reference's arithmetic run in float.  Its acceptance tests
    are discontinuous (the gate ignores the -x^/-y^ neighbours; candidates are
    accepted on strict foot-point bounds): where FP32 rounding flips one, a
    node takes another candidate and the difference is carried downstream, so
    the 1e-5 target of the elliptic model does not hold here for every node."""
    print(f"FP32 vs the FP64 oracle at {n}^3", flush=True)
    T0 = fold3d.implicit_initialiser(n)
    want = T0.copy()
    oracle.sweeps(want, n, 8)
    T = to_dev(T0, torch.float32)
    solver.solve(T, n, max_iters=8)
    got = T.cpu().double().numpy()
    assert np.array_equal(np.sign(got), np.sign(want))
    reached = np.abs(want) < 1e15
    assert np.array_equal(reached, np.abs(got) < 1e15)
    rel = np.abs(got[reached] - want[reached]) / np.abs(want[reached])
    fr = {t: float(np.mean(rel > t)) for t in (1e-3, 1e-4, 1e-5)}
    print(f"FP32 n={n}: max rel err {rel.max():.3e}, median {np.median(rel):.3e}, "
          f"above 1e-3/1e-4/1e-5: {fr[1e-3]:.2e} {fr[1e-4]:.2e} {fr[1e-5]:.2e}")
    assert rel.max() < FP32_BOUND["max"]
    for t in (1e-3, 1e-4, 1e-5):
        assert fr[t] <= FP32_BOUND[t], (t, fr[t])


FAST_ENGINES = [fold3d.ENGINE_HOISTED, fold3d.ENGINE_TILED, fold3d.ENGINE_SPLIT,
                fold3d.ENGINE_LOCKED]


# waves per row (groups) apply to the split and locked engines only
FP32_ENGINE_CASES = [(e, g) for e in FAST_ENGINES for g in (2, 4, 8)
                     if g == 2 or e in (fold3d.ENGINE_SPLIT, fold3d.ENGINE_LOCKED)]


@pytest.mark.parametrize("engine,groups", FP32_ENGINE_CASES)
def test_fp32_engines_equal_the_fp32_literal_kernel(solver, engine, groups):
    """FP32: hoisted constants (computed in float with the literal kernel's
    expressions), pruning by the plane and edge-line bounds (shrunk to 90 %),
    LDS tiles, units split over waves and locking change no bit against the
    literal FP32 kernel (fold_local.h run in float, which prunes nothing).
    On the reference workload at 64^3 (8 iterations and to the fixed point),
    and on a box grid with a random signed field and another model."""
    print(f"FP32 engine {engine}, {groups} waves per row, vs the literal FP32 kernel", flush=True)
    solver.set_tuning(groups, fold3d.MODEL_FOLD)
    try:
        n = 64
        T0 = to_dev(fold3d.implicit_initialiser(n), torch.float32)
        for iters, tol in ((8, 0.0), (200, 1e-30)):
            A, B = T0.clone(), T0.clone()
            sa = solver.solve(A, n, max_iters=iters, tol=tol, engine=fold3d.ENGINE_PLANE)
            sb = solver.solve(B, n, max_iters=iters, tol=tol, engine=engine)
            assert sa.iterations == sb.iterations
            assert torch.equal(A.view(torch.int32), B.view(torch.int32)), (iters, tol)
        n3 = (45, 70, 29)
        rng = np.random.default_rng(groups + 10 * engine)
        size = (n3[0] + 2) * (n3[1] + 2) * (n3[2] + 2)
        R0 = np.where(rng.random(size) < 0.3, -1.0, 1.0) * rng.uniform(1.0, 5.0, size)
        R0[rng.random(size) < 0.05] = fold3d.MAXT
        h, a = (0.31, 0.27, 0.35), (0.25, -0.4, 0.3)
        m = fold3d.fold_model(1.2, a)
        A, B = to_dev(R0, torch.float32), to_dev(R0, torch.float32)
        solver.solve(A, n3, spacing=h, model=m, max_iters=3, engine=fold3d.ENGINE_PLANE)
        solver.solve(B, n3, spacing=h, model=m, max_iters=3, engine=engine)
        assert torch.equal(A.view(torch.int32), B.view(torch.int32))
    finally:
        solver.set_tuning(0, fold3d.MODEL_FOLD)


@pytest.mark.parametrize("groups", [2, 4, 8])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_locking_gives_the_same_bits(solver, groups, dtype):
    """Exact locking: the reference workload at 64^3 for 8 iterations and a
    box grid with a random field, locked and unlocked, same bits, with fewer
    updates computed once the field settles."""
    solver.set_tuning(groups, fold3d.MODEL_FOLD)
    try:
        n = 64
        T0 = to_dev(fold3d.implicit_initialiser(n), dtype)
        A, B = T0.clone(), T0.clone()
        sa = solver.solve(A, n, max_iters=8, engine=fold3d.ENGINE_SPLIT)
        sb = solver.solve(B, n, max_iters=8, engine=fold3d.ENGINE_LOCKED)
        assert torch.equal(A.view(torch.int64 if dtype == torch.float64 else torch.int32),
                           B.view(torch.int64 if dtype == torch.float64 else torch.int32))
        assert sa.updates_computed == sa.cell_updates
        # a block is skipped when no 64-node line chunk of its source patch
        # changed; at 64^3 after 8 iterations about 82 % of the updates remain
        assert sb.updates_computed < 0.9 * sb.cell_updates
        print(f"locked {dtype} G={groups}: {sb.updates_computed / sb.cell_updates:.3f} of the updates computed")
        n3 = (37, 70, 29)
        rng = np.random.default_rng(groups)
        size = (n3[0] + 2) * (n3[1] + 2) * (n3[2] + 2)
        R0 = np.where(rng.random(size) < 0.3, -1.0, 1.0) * rng.uniform(1.0, 5.0, size)
        h, a = (0.3, 0.25, 0.35), (-0.35, 0.4, 0.7)
        A, B = to_dev(R0, dtype), to_dev(R0, dtype)
        m = fold3d.fold_model(1.1, a)
        solver.solve(A, n3, spacing=h, model=m, max_iters=6, engine=fold3d.ENGINE_SPLIT)
        sb = solver.solve(B, n3, spacing=h, model=m, max_iters=60, tol=1e-300,
                          engine=fold3d.ENGINE_LOCKED, raise_not_converged=False)
        C = to_dev(R0, dtype)
        sc = solver.solve(C, n3, spacing=h, model=m, max_iters=60, tol=1e-300,
                          engine=fold3d.ENGINE_SPLIT, raise_not_converged=False)
        assert sb.iterations == sc.iterations
        assert torch.equal(B.double(), C.double())
    finally:
        solver.set_tuning(0, fold3d.MODEL_FOLD)


def test_default_lock_start(solver, golden_large):
    """With the default start (iteration 8) the first 8 iterations run
    unlocked and the locked ones after them still give the reference's
    fixed point (64^3, 18 iterations)."""
    from conftest import matches_digest
    solver.set_lock_start(-1)
    assert solver.lock_start() == 8
    n = 64
    T = to_dev(fold3d.implicit_initialiser(n))
    st = solver.solve(T, n, max_iters=8, engine=fold3d.ENGINE_LOCKED)
    assert st.updates_computed == st.cell_updates
    T = to_dev(fold3d.implicit_initialiser(n))
    st = solver.solve(T, n, max_iters=100, tol=1e-300, engine=fold3d.ENGINE_LOCKED)
    assert st.iterations == int(golden_large["fixed_iters_64"][0])
    assert matches_digest(T.cpu().numpy(), golden_large, "fixed_64")
    assert st.updates_computed < st.cell_updates
    with pytest.raises(fold3d.Fold3dError):
        solver.set_lock_start(-2)


def test_locking_is_rejected_for_the_elliptic_model(solver):
    from fold3d import workloads as wl
    W = wl.tilted_homogeneous(16)
    T = torch.zeros(18 ** 3, dtype=torch.float64, device="cuda")
    with pytest.raises(fold3d.Fold3dError) as e:
        solver.solve(T, 16, spacing=W.h, model=W.model(), engine=fold3d.ENGINE_LOCKED)
    assert e.value.code == fold3d.FOLD3D_E_UNSUPPORTED


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("n3", [(1, 1, 1), (2, 2, 2), (1, 5, 3), (2, 7, 1), (3, 2, 9), (65, 1, 2)])
def test_tiny_grids_match_oracle(solver, oracle, n3, engine):
    """Grids with one or two nodes along an axis (no or one plane step)."""
    rng = np.random.default_rng(sum(n3) + engine)
    size = (n3[0] + 2) * (n3[1] + 2) * (n3[2] + 2)
    T0 = np.where(rng.random(size) < 0.3, -1.0, 1.0) * rng.uniform(1.0, 5.0, size)
    h, a = (0.3, 0.25, 0.35), (-0.35, 0.4, 0.7)
    want = T0.copy()
    oracle.sweeps3(want, n3, 2, 1.1, a, h)
    T = to_dev(T0)
    solver.solve(T, n3, spacing=h, model=fold3d.fold_model(1.1, a), max_iters=2, engine=engine)
    assert bits_equal(T.cpu().numpy(), want)


def test_host_pointers_are_rejected_not_dereferenced(solver):
    """A plain host pointer where a device array is required is an error code
    from the C-ABI, not a kernel fault: solve with memory = DEVICE, a batch
    instance, an adjoint receiver array (called through the C-ABI directly;
    the Python layer would stop these earlier)."""
    import ctypes as C
    from fold3d import workloads as wl
    n = 8
    host = np.zeros((n + 2) ** 3)
    L = fold3d.lib()
    o = fold3d.Opts()
    L.fold3d_default_opts(C.byref(o))
    st = fold3d.Stats()
    m = fold3d.reference_model()
    rc = L.fold3d_solve(solver.h, C.c_void_p(host.ctypes.data), n, n, n, 0.1, 0.1, 0.1,
                        C.byref(m), C.byref(o), C.byref(st))
    assert rc == fold3d.FOLD3D_E_INVALID and "device memory" in fold3d.last_error()
    W = wl.tilted_homogeneous(n)
    T = torch.zeros((n + 2) ** 3, dtype=torch.float64, device="cuda")
    inst = (fold3d.Instance * 1)()
    x = inst[0]
    x.T = T.data_ptr()
    for a in range(3):
        x.n[a], x.h[a] = n, W.h
    for a in range(6):
        x.M[a] = float(W.M[a])
    x.speed = 1.0
    x.speed_field = host.ctypes.data  # host memory
    assert L.fold3d_solve_batched(solver.h, 1, inst, C.byref(o), None, None, None) \
        == fold3d.FOLD3D_E_INVALID
    assert "device memory" in fold3d.last_error()
    me = W.model()
    idx = np.zeros(4, dtype=np.int32)
    res = np.zeros(4)
    g = torch.zeros_like(T)
    rc = L.fold3d_adjoint(solver.h, C.c_void_p(T.data_ptr()), n, n, n, W.h, W.h, W.h,
                          C.byref(me), 4, C.c_void_p(idx.ctypes.data),
                          C.c_void_p(res.ctypes.data), C.byref(o), C.c_void_p(g.data_ptr()),
                          None, C.byref(st))
    assert rc == fold3d.FOLD3D_E_INVALID


def test_python_layer_rejects_bad_device_tensors(solver):
    """Wrong size, layout, dtype or device of a device tensor is a TypeError /
    ValueError in the Python layer, before the C-ABI sees a raw pointer."""
    from fold3d import workloads as wl
    n = 8
    cnt = (n + 2) ** 3
    ok = torch.zeros(cnt, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        solver.solve(torch.zeros(cnt - 1, dtype=torch.float64, device="cuda"), n)
    with pytest.raises(ValueError):
        solver.solve(torch.zeros(2 * cnt, dtype=torch.float64, device="cuda")[::2], n)
    with pytest.raises(TypeError):
        solver.solve(torch.zeros(cnt, dtype=torch.float16, device="cuda"), n)
    with pytest.raises(TypeError):
        solver.solve(torch.zeros(cnt, dtype=torch.float64), n)  # host tensor
    W = wl.tilted_homogeneous(n)
    Wd = wl.Workload(W.name, n, W.h, W.M, W.speed)
    Wd.speed_field = torch.ones(cnt, dtype=torch.float32, device="cuda")  # dtype differs from T
    with pytest.raises(TypeError):
        solver.solve(ok, n, spacing=W.h, model=Wd.model())
    with pytest.raises(TypeError):
        solver.solve_batched([Wd], [ok])
    me = W.model()
    idx = torch.zeros(4, dtype=torch.int32, device="cuda")
    with pytest.raises(TypeError):
        solver.adjoint(ok, n, W.h, me, idx, torch.zeros(4, dtype=torch.float32, device="cuda"),
                       torch.zeros_like(ok))
    with pytest.raises(ValueError):
        solver.adjoint(ok, n, W.h, me, idx, torch.zeros(3, dtype=torch.float64, device="cuda"),
                       torch.zeros_like(ok))
    with pytest.raises(ValueError):
        solver.adjoint(ok, n, W.h, me, idx, torch.zeros(4, dtype=torch.float64, device="cuda"),
                       torch.zeros(cnt + 1, dtype=torch.float64, device="cuda"))
