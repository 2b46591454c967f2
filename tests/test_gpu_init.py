"""The reference's input field built on the device: the same bits as the host
initializer (itself bit-identical to the reference's, see the golden vectors),
on cubic and box grids."""
import numpy as np
import pytest

from conftest import bits_equal

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import fold3d  # noqa: E402


@pytest.mark.parametrize("n", [16, 24, 64, 128])
def test_device_initialiser_matches_host(solver, golden, n):
    print(f"device initialiser n={n}", flush=True)
    host = fold3d.implicit_initialiser(n)
    if f"init_{n}" in golden.files:
        assert bits_equal(host, golden[f"init_{n}"])
    T = torch.empty((n + 2) ** 3, dtype=torch.float64, device="cuda")
    ne = solver.initialiser_device(T, n)
    assert 0 < ne < (n + 2) ** 3 // 2
    assert bits_equal(T.cpu().numpy(), host)


@pytest.mark.parametrize("n3", [(30, 17, 41), (9, 50, 12)])
def test_device_initialiser_box_grid(solver, n3):
    import ctypes as C
    h = (10.0 / (n3[0] - 1), 10.0 / (n3[1] - 1), 10.0 / (n3[2] - 1))
    size = (n3[0] + 2) * (n3[1] + 2) * (n3[2] + 2)
    host = np.empty(size)
    fold3d.lib().ImplicitInitialiser(host.ctypes.data_as(C.POINTER(C.c_double)), *n3, *h)
    T = torch.empty(size, dtype=torch.float64, device="cuda")
    solver.initialiser_device(T, n3, spacing=h)
    assert bits_equal(T.cpu().numpy(), host)


def test_device_initialiser_then_solve(solver, golden):
    n = 32
    T = torch.empty((n + 2) ** 3, dtype=torch.float64, device="cuda")
    solver.initialiser_device(T, n)
    solver.solve(T, n)
    assert bits_equal(T.cpu().numpy(), golden["solve8_32"])
