"""Pin the FP64 CPU adjoint of the elliptic model (no reference code exists
for it) by what the mathematics fixes: its gradient is the derivative of the
misfit J = 1/2 sum (T[rec] - T_obs)^2 of the oracle's own converged forward
solve, so central differences of that solve (small step, where the discrete
solution is smooth) must reproduce it, for the slowness and for the tilt
angle; and it is linear in the receiver residuals."""
import numpy as np
import pytest

from fold3d.tomography import tomography_fields, top_face_survey
from fold3d.workloads import Workload, point_source_start, tilted_metric

N = 12


def _solve(oracle, S, fl):
    n, H = S.n, S.n + 2
    src = S.sources[0]
    ijk = [int(round(c / S.h)) + 1 for c in src]
    flat = ijk[0] + H * (ijk[1] + H * ijk[2])
    W = Workload("t", n, S.h, np.zeros(6), 1.0, src, 1.0 / fl["slowness"][flat])
    W.meta = {"metric_at_source": tilted_metric(fl["theta"][flat], fl["phi"][flat], 1.0,
                                                fl["ra"][flat], fl["rb"][flat])}
    T = point_source_start(W)
    _, md = oracle.ellip_sweeps(T, n, 300, 1e-300, S.h, speed_field=1.0 / fl["slowness"],
                                theta=fl["theta"], phi=fl["phi"], ra=fl["ra"], rb=fl["rb"])
    assert md == 0.0
    return T, flat


@pytest.fixture(scope="module")
def case(oracle):
    f = tomography_fields(N, xp=np, dtype=np.float64)
    ft = tomography_fields(N, xp=np, dtype=np.float64, anomalies=True)
    S = top_face_survey(N, src_side=1, rec_side=N)
    obs = _solve(oracle, S, ft)[0][S.rec_idx]
    T, src = _solve(oracle, S, f)
    res = T[S.rec_idx] - obs
    return f, S, obs, T, src, res


def _adjoint(oracle, f, S, T, res, local="closure"):
    return oracle.ellip_adjoint(T, S.n, S.h, S.rec_idx.astype(np.int64), res,
                                speed_field=1.0 / f["slowness"], theta=f["theta"],
                                phi=f["phi"], ra=f["ra"], rb=f["rb"], local=local)


def _cells(g, src, H, count, seed):
    """Seeded cells with a gradient above 1e-3 of its maximum, 3+ cells from
    the source node."""
    si, sj, sk = src % H, (src // H) % H, src // (H * H)
    idx = np.flatnonzero(np.abs(g) > 1e-3 * np.abs(g).max())
    i, j, k = idx % H, (idx // H) % H, idx // (H * H)
    far = np.maximum(np.maximum(abs(i - si), abs(j - sj)), abs(k - sk)) >= 3
    idx = idx[far]
    rng = np.random.default_rng(seed)
    return rng.choice(idx, min(count, idx.size), replace=False)


@pytest.mark.parametrize("local", ["closure", "solver"])
@pytest.mark.parametrize("name", ["slowness", "theta"])
def test_adjoint_matches_central_differences(oracle, case, name, local):
    """Both definitions of the node derivatives (the winning simplex with its
    faces; the node's whole local solver) against central differences of the
    oracle's own forward solve."""
    f, S, obs, T, src, res = case
    gs, gt, passes = _adjoint(oracle, f, S, T, res, local)
    assert passes >= 1
    g = gs if name == "slowness" else gt
    cells = _cells(g, src, S.n + 2, 10, seed=7)
    assert cells.size >= 5

    def J(fl):
        return 0.5 * np.sum((_solve(oracle, S, fl)[0][S.rec_idx] - obs) ** 2)

    for c in cells:
        d = 1e-6 * abs(f[name][c])
        fp = {k: v.copy() for k, v in f.items()}
        fm = {k: v.copy() for k, v in f.items()}
        fp[name][c] += d
        fm[name][c] -= d
        fd = (J(fp) - J(fm)) / (2 * d)
        assert abs(g[c] - fd) <= 1e-6 * abs(fd), (name, c, g[c], fd)


def test_adjoint_zero_outside_the_fan_and_at_start(oracle, case):
    f, S, obs, T, src, res = case
    gs, gt, _ = _adjoint(oracle, f, S, T, res)
    H = S.n + 2
    # ghosts carry nothing
    g3 = gs.reshape(H, H, H)
    assert not g3[0].any() and not g3[-1].any() and not g3[:, 0].any()
    assert not g3[:, -1].any() and not g3[:, :, 0].any() and not g3[:, :, -1].any()
    # the closed-form start node nearest the source is fixed: no gradient
    assert gs[src] == 0.0 and gt[src] == 0.0


def test_adjoint_is_linear_in_residuals(oracle, case):
    f, S, obs, T, src, res = case
    rng = np.random.default_rng(3)
    ra_, rb_ = rng.normal(size=res.size), rng.normal(size=res.size)
    ga = _adjoint(oracle, f, S, T, ra_)[0]
    gb = _adjoint(oracle, f, S, T, rb_)[0]
    gab = _adjoint(oracle, f, S, T, 2.0 * ra_ + rb_)[0]
    scale = np.abs(gab).max()
    assert np.abs(gab - (2.0 * ga + gb)).max() <= 1e-12 * scale


@pytest.mark.parametrize("local", ["closure", "solver"])
def test_single_receiver_gradient_is_a_path_sensitivity(oracle, case, local):
    """With residual 1 at one receiver, dJ/ds_x = dT_r/ds_x >= 0 (a slower
    node can only delay the arrival), and since every candidate is
    homogeneous of degree 1 in (upwind values, slowness),
        T_r = sum_x s_x dT_r/ds_x + sum_{y held fixed} mu_y T_y
    (Euler's relation, unrolled down the upwind graph to the start nodes)."""
    f, S, obs, T, src, res = case
    for r in (S.nrec // 2 + S.n // 2, 0, S.nrec - 1):
        gs, _, _, mu, fixed = oracle.ellip_adjoint(
            T, S.n, S.h, S.rec_idx[r:r + 1].astype(np.int64), np.ones(1),
            speed_field=1.0 / f["slowness"], theta=f["theta"], phi=f["phi"],
            ra=f["ra"], rb=f["rb"], state=True, local=local)
        assert (gs >= 0).all()
        assert fixed[src] and fixed.sum() <= 64  # within the 4^3 closed-form start
        euler = float(np.sum(f["slowness"] * gs) + np.sum(mu[fixed] * T[fixed]))
        Tr = T[S.rec_idx[r]]
        # 1e-9: the oracle's local dT/ds are central differences (step 1e-6 s)
        assert abs(euler - Tr) <= 1e-9 * Tr, (r, euler, Tr)
