"""Pin the elliptic-model oracle (no reference code exists for it) by what the
mathematics fixes:
  * brute-force minimisation of T_lin(Y) + s|X - Y|_N over the upwind patch;
  * exact reproduction of plane waves (linear fields are fixed points);
  * the closed-form point-source traveltime sqrt(x^T M^-1 x)/v, with
    first-order convergence under grid refinement;
  * the isotropic case against the reference fold model with a = 0
    (F |grad T| = 1), an independent local solver;
  * metric assembly R diag(a)^2 R^T against numpy.
"""
import numpy as np
import pytest
from scipy.optimize import minimize

RING = [(1, 0), (1, 1), (0, 1), (-1, 1), (-1, 0), (-1, -1), (0, -1), (1, -1)]
FR = [(0, 1, 2, 1, 1, 1), (0, -1, 2, -1, 1, -1), (2, 1, 1, 1, 0, 1),
      (2, -1, 1, -1, 0, -1), (1, -1, 0, -1, 2, -1), (1, 1, 0, 1, 2, 1)]


def sym(M6):
    a, b, c, d, e, f = M6
    return np.array([[a, d, e], [d, b, f], [e, f, c]])


def random_spd(rng, lo=0.4, hi=1.6):
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    A = Q @ np.diag(rng.uniform(lo, hi, 3)) @ Q.T
    return np.array([A[0, 0], A[1, 1], A[2, 2], A[0, 1], A[0, 2], A[1, 2]])


def local_N(d, N6):
    m, s, ux, sx, uy, sy = FR[d]
    u, sg = (ux, uy, m), (sx, sy, s)
    Ng = sym(N6)
    return np.array([[sg[i] * sg[j] * Ng[u[i], u[j]] for j in range(3)] for i in range(3)])


def brute_node(p9, s, NL, h3):
    """min over the 8 closed triangles of T_lin(Y) + s |X - Y|_N."""
    dxx, dyy, dzz = h3
    X = np.array([0.0, 0.0, dzz])
    pos = [np.zeros(3)] + [np.array([a * dxx, b * dyy, 0.0]) for a, b in RING]

    def f(uv, k):
        u, v = uv
        i1, i2 = 1 + k, 1 + (k + 1) % 8
        Y = pos[0] + u * pos[i1] + v * pos[i2]
        t = p9[0] + u * (p9[i1] - p9[0]) + v * (p9[i2] - p9[0])
        w = X - Y
        return t + s * np.sqrt(w @ NL @ w)

    best = np.inf
    g = np.linspace(0, 1, 41)
    for k in range(8):
        U, V = np.meshgrid(g, g)
        mask = U + V <= 1 + 1e-12
        vals = np.array([f((u, v), k) for u, v in zip(U[mask], V[mask])])
        j = np.argmin(vals)
        x0 = (U[mask][j], V[mask][j])
        r = minimize(lambda z: f(z, k), x0, method="SLSQP", bounds=[(0, 1), (0, 1)],
                     constraints=[{"type": "ineq", "fun": lambda z: 1 - z[0] - z[1]}],
                     options={"ftol": 1e-15, "maxiter": 200})
        best = min(best, vals[j], r.fun if r.success else np.inf)
    return best


def test_metric_assembly(oracle):
    rng = np.random.default_rng(3)
    for _ in range(20):
        th, ph = rng.uniform(0, np.pi / 3), rng.uniform(0, 2 * np.pi)
        a = rng.uniform(0.6, 1.0, 3)
        Rz = np.array([[np.cos(ph), -np.sin(ph), 0], [np.sin(ph), np.cos(ph), 0], [0, 0, 1]])
        Ry = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
        R = Rz @ Ry
        want = R @ np.diag(a ** 2) @ R.T
        got = sym(oracle.tilted_metric(th, ph, *a))
        assert np.allclose(got, want, rtol=0, atol=1e-15)
        N = sym(oracle.sym3_inverse(oracle.tilted_metric(th, ph, *a)))
        assert np.allclose(N @ want, np.eye(3), atol=1e-13)


@pytest.mark.parametrize("seed", range(12))
def test_node_update_equals_brute_force_minimum(oracle, seed):
    rng = np.random.default_rng(100 + seed)
    d = seed % 6
    M6 = random_spd(rng)
    N6 = oracle.sym3_inverse(M6)
    h = rng.uniform(0.5, 1.5, 3)
    s = rng.uniform(0.5, 2.0)
    K = oracle.ellip_constants(d, N6, h)
    m, _, ux, _, uy, _ = FR[d]
    h3 = (h[ux], h[uy], h[m])
    # smooth patch (plane wave + curvature + noise) or rough patch
    base = rng.uniform(1, 3)
    g = rng.normal(size=2) * s * 0.5
    p9 = [base] + [base + g[0] * a * h3[0] + g[1] * b * h3[1] + 0.1 * (a * a + b * b)
                   for a, b in RING]
    p9 = np.array(p9) + (rng.normal(size=9) * 0.3 if seed % 3 == 2 else 0.0)
    got = oracle.ellip_node(1e16, p9, s, K)
    want = brute_node(p9, s, local_N(d, N6), h3)
    assert got == pytest.approx(want, rel=2e-9, abs=1e-12)


@pytest.mark.parametrize("d", range(6))
def test_plane_wave_is_reproduced(oracle, d):
    rng = np.random.default_rng(50 + d)
    n = 10
    M6 = random_spd(rng, 0.7, 1.3)
    M = sym(M6)
    h = np.array([0.1, 0.12, 0.09])
    v = 1.7
    # direction of travel along this sweep's march axis, tilted a little
    m, sgn = FR[d][0], FR[d][1]
    w = rng.normal(size=3) * 0.15
    w[m] = sgn
    # gradient p with M p parallel to w and |p|_M = 1/v
    p = np.linalg.solve(M, w)
    p /= np.sqrt(p @ M @ p) * v
    H = n + 2
    idx = np.indices((H, H, H)).reshape(3, -1)  # k, j, i order of a C array
    x = np.stack([idx[2] * h[0], idx[1] * h[1], idx[0] * h[2]])
    T = (p @ x + 5.0).astype(np.float64)
    target = 4 + sgn
    exact = T.copy()
    # knock the target plane up and let one plane step restore it
    stride = [1, H, H * H][m]
    sel = (x[m] / h[m]).round().astype(int) == target
    T[sel] += 1.0
    oracle.ellip_plane(d, T, n, 4, h, M=M6, speed=v)
    interior = sel.copy()
    for ax in range(3):
        if ax != m:
            c = (x[ax] / h[ax]).round().astype(int)
            interior &= (c >= 1) & (c <= n)
    assert np.allclose(T[interior], exact[interior], rtol=0, atol=1e-12)


def test_point_source_closed_form_first_order(oracle):
    """Homogeneous tilted medium (the first configuration's metric): error
    against sqrt(x^T M^-1 x)/v shrinks about linearly with h."""
    M6 = oracle.tilted_metric(np.radians(30), np.radians(45), 1.0, 0.8, 0.6)
    N = sym(oracle.sym3_inverse(M6))
    v = 1.0
    errs = []
    for n in (17, 33):
        h = 1.0 / (n - 1)
        H = n + 2
        k, j, i = np.indices((H, H, H)).reshape(3, -1)
        X = np.stack([(i - 1) * h, (j - 1) * h, (k - 1) * h]) - 0.5
        exact = np.sqrt(np.einsum("in,ij,jn->n", X, N, X)) / v
        T = np.full(H ** 3, 1e16)
        near = np.max(np.abs(X), axis=0) <= 1.01 * h
        T[near] = exact[near]
        ghost = (i == 0) | (j == 0) | (k == 0) | (i == H - 1) | (j == H - 1) | (k == H - 1)
        T[ghost] = 1e16
        it, md = oracle.ellip_sweeps(T, n, 50, 1e-12, h, M=M6, speed=v)
        assert md < 1e-12 and it < 50
        inner = ~ghost
        assert np.all(T[inner] < 1e15)
        assert np.all(T[inner] >= exact[inner] - 1e-12)  # upwind PMM overestimates
        errs.append(np.max(np.abs(T[inner] - exact[inner])))
    assert errs[1] < 0.75 * errs[0]
    assert errs[1] < 0.05


def test_isotropic_matches_fold_model_without_drift(oracle):
    """M = I, v = F against the reference solver with a = 0 on a smooth field."""
    n = 20
    h = 0.25
    F = 1.3
    H = n + 2
    k, j, i = np.indices((H, H, H)).reshape(3, -1)
    X = np.stack([(i - 1) * h, (j - 1) * h, (k - 1) * h]) - np.array([[2.1], [1.7], [2.6]])
    r = np.sqrt((X ** 2).sum(0)) / F
    near = r < 1.2 * h / F * np.sqrt(3)
    T0 = np.where(near, r, 1e16)
    ghost = (i == 0) | (j == 0) | (k == 0) | (i == H - 1) | (j == H - 1) | (k == H - 1)
    T0[ghost] = 1e16
    A = T0.copy()
    oracle.sweeps(A, n, 6, F=F, a=(0.0, 0.0, 0.0), h=(h, h, h))
    B = T0.copy()
    oracle.ellip_sweeps(B, n, 6, 0.0, h, M=[1, 1, 1, 0, 0, 0], speed=F)
    reached = (A < 1e15) & (B < 1e15)
    assert reached.sum() > 0.9 * (~ghost).sum()
    assert np.max(np.abs(A[reached] - B[reached]) / B[reached]) < 1e-12
