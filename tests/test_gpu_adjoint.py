"""Adjoint-state gradients on the GPU (fold3d_adjoint through the C-ABI and
the torch.autograd harness), against
  * the FP64 CPU adjoint of the oracle on the same converged field;
  * central finite differences of the GPU forward solve (FP64, relative
    steps 1e-3 and 5e-4, Richardson-extrapolated) at 32 seeded cells;
  * itself under checkpoint budgets and batch sizes, and in FP32."""
import numpy as np
import pytest

import fold3d
from fold3d.tomography import Tomography, adjoint_tol, tomography_fields, top_face_survey

pytestmark = pytest.mark.gpu


def _fields(n, dtype, anomalies=False):
    import torch
    return tomography_fields(n, xp=torch, dtype=dtype, device="cuda", anomalies=anomalies)


def _tomo(solver, n, dtype, src_side, rec_side, **kw):
    f = _fields(n, dtype)
    ft = _fields(n, dtype, anomalies=True)
    S = top_face_survey(n, src_side=src_side, rec_side=rec_side)
    tomo = Tomography(solver, S, f["phi"], f["ra"], f["rb"], **kw)
    tomo.observe(ft["slowness"], ft["theta"])
    return tomo, f


def _host(f):
    return {k: v.double().cpu().numpy() for k, v in f.items()}


def test_adjoint_matches_cpu_fp64(solver, oracle):
    import torch
    n = 20
    tomo, f = _tomo(solver, n, torch.float64, 1, n)
    S = tomo.survey
    T = tomo.solve_group([0], f["slowness"], f["theta"])[0]
    res = (tomo.receiver_times(T) - tomo.T_obs[0]).contiguous()
    gs = torch.zeros_like(T)
    gt = torch.zeros_like(T)
    m = tomo.model((1.0 / f["slowness"]).contiguous(), f["theta"])
    idx = torch.as_tensor(S.rec_idx, device="cuda")
    st = solver.adjoint(T, n, S.h, m, idx, res, gs, gt, max_iters=200, tol=adjoint_tol(res))
    assert 1 <= st.iterations < 200
    h = _host(f)
    cs, ct, _ = oracle.ellip_adjoint(T.cpu().numpy(), n, S.h, S.rec_idx.astype(np.int64),
                                     res.cpu().numpy(), speed_field=1.0 / h["slowness"],
                                     theta=h["theta"], phi=h["phi"], ra=h["ra"], rb=h["rb"])
    g_s, g_t = gs.cpu().numpy(), gt.cpu().numpy()
    assert np.abs(cs).max() > 0 and np.abs(ct).max() > 0
    # 1e-8: the oracle's local derivatives are central differences (step 1e-6)
    assert np.abs(g_s - cs).max() <= 1e-8 * np.abs(cs).max()
    assert np.abs(g_t - ct).max() <= 1e-8 * np.abs(ct).max()
    assert (np.abs(g_s - cs) <= 1e-8 * np.abs(cs).max()).all()


def test_adjoint_matches_cpu_fp64_on_the_survey_layout_at_64(solver, oracle):
    """The tomography survey's layout (16 sources on a 4x4 grid below the top
    face, 64x64 top-face receivers: every top-face node at 64^3), FP64, per-node
    speed and tilt: the GPU gradient of the misfit (batched forward, batched
    adjoint of the 16 sources) against the sum over the sources of the CPU
    adjoint on the same converged fields, to 1e-8 of the largest value (the
    oracle's node derivatives are central differences, step 1e-6, of the
    winning simplex with its faces; those of the node's whole local solver
    average two slopes where candidates of two patches tie within the step,
    which moves the tilt gradient of one node per source by up to 1e-6 of
    the maximum).  The reached nodes held fixed are only those of the
    closed-form starts."""
    import torch
    n = 64
    print("adjoint vs CPU adjoint, 16 sources, 64x64 receivers, 64^3 FP64", flush=True)
    tomo, f = _tomo(solver, n, torch.float64, 4, 64, batch=16)
    S = tomo.survey
    assert len(S.sources) == 16 and S.nrec == 64 * 64
    s0 = f["slowness"].clone().requires_grad_(True)
    t0 = f["theta"].clone().requires_grad_(True)
    J, res, ckpt = tomo.forward_pass(s0.detach(), t0.detach())
    tomo.misfit(s0, t0).backward()
    held = list(solver.last_held)
    h = _host(f)
    cs = np.zeros((n + 2) ** 3)
    ct = np.zeros((n + 2) ** 3)
    starts = []
    for k in range(16):
        T = ckpt[k].cpu().numpy()
        a, b, _ = oracle.ellip_adjoint(T, n, S.h, S.rec_idx.astype(np.int64),
                                       res[k].cpu().numpy(), speed_field=1.0 / h["slowness"],
                                       theta=h["theta"], phi=h["phi"], ra=h["ra"], rb=h["rb"])
        cs += a
        ct += b
        starts.append(int((tomo.start_field(k, f["slowness"], f["theta"]) < 1e15).sum()))
    g_s, g_t = s0.grad.cpu().numpy(), t0.grad.cpu().numpy()
    es = np.abs(g_s - cs).max() / np.abs(cs).max()
    et = np.abs(g_t - ct).max() / np.abs(ct).max()
    print(f"misfit {float(J):.6e}; max |GPU - CPU| / max: slowness {es:.2e}, theta {et:.2e}; "
          f"held per source {held}, start nodes {starts}")
    assert np.abs(cs).max() > 0 and np.abs(ct).max() > 0
    assert es <= 1e-8 and et <= 1e-8
    assert all(0 <= hk <= sk for hk, sk in zip(held, starts))


def _seeded_cells(g, srcs, H, count, seed):
    idx = np.flatnonzero(np.abs(g) > 1e-3 * np.abs(g).max())
    i, j, k = idx % H, (idx // H) % H, idx // (H * H)
    far = np.ones(idx.size, bool)
    for s in srcs:
        si, sj, sk = s % H, (s // H) % H, s // (H * H)
        far &= np.maximum(np.maximum(abs(i - si), abs(j - sj)), abs(k - sk)) >= 3
    return np.random.default_rng(seed).choice(idx[far], count, replace=False)


@pytest.mark.parametrize("name", ["slowness", "theta"])
def test_gradient_vs_finite_differences(solver, name):
    """32 seeded cells with a non-negligible gradient, 3+ cells from every
    source.  Central differences at relative step 1e-3 carry an O(step^2)
    error of order 1e-2 here (the discrete traveltime map is strongly curved
    at grazing simplices), so the comparison uses the Richardson combination
    of steps 1e-3 and 5e-4; at least 80 % of the cells must agree to 1e-4
    relative (a cell whose perturbation moves a foot point across a simplex
    boundary sees a kink) and every cell to 1e-3 of the largest gradient
    checked."""
    import torch
    n = 32
    tomo, f = _tomo(solver, n, torch.float64, 2, 16, batch=4)
    S = tomo.survey
    s0 = f["slowness"].clone().requires_grad_(True)
    t0 = f["theta"].clone().requires_grad_(True)
    J = tomo.misfit(s0, t0)
    J.backward()
    g = (s0.grad if name == "slowness" else t0.grad).cpu().numpy()
    H = n + 2
    srcs = [min(n, max(1, int(round(c / S.h)) + 1)) for src in S.sources for c in src]
    srcs = [srcs[3 * q] + H * (srcs[3 * q + 1] + H * srcs[3 * q + 2]) for q in range(len(S.sources))]
    cells = _seeded_cells(g, srcs, H, 32, seed=20261207)

    def Jof(field, c, d):
        s, t = f["slowness"].clone(), f["theta"].clone()
        (s if field == "slowness" else t)[c] += d
        with torch.no_grad():
            return float(tomo.forward_pass(s, t)[0])

    rows = []
    for c in cells:
        base = float(f[name][c])
        d = 1e-3 * abs(base)
        d1 = (Jof(name, c, d) - Jof(name, c, -d)) / (2 * d)
        d2 = (Jof(name, c, d / 2) - Jof(name, c, -d / 2)) / d
        rows.append((g[c], d1, (4 * d2 - d1) / 3))
    R = np.array(rows)
    rel = np.abs(R[:, 0] - R[:, 2]) / np.abs(R[:, 2])
    print(f"{name}: raw central 1e-3 median rel {np.median(np.abs(R[:, 0] - R[:, 1]) / np.abs(R[:, 1])):.2e}; "
          f"Richardson median {np.median(rel):.2e} max {rel.max():.2e}")
    assert (rel < 1e-4).mean() >= 0.8
    assert np.abs(R[:, 0] - R[:, 2]).max() <= 1e-3 * np.abs(R[:, 2]).max()


def test_checkpoint_budget_and_batching_give_the_same_gradient(solver):
    import torch
    n = 24
    grads = {}
    for key, kw in (("all", dict(batch=4)), ("none", dict(batch=4, budget_bytes=0)),
                    ("half", dict(batch=1, budget_bytes=2 * 26 ** 3 * 8))):
        tomo, f = _tomo(solver, n, torch.float64, 2, 12, **kw)
        s0 = f["slowness"].clone().requires_grad_(True)
        t0 = f["theta"].clone().requires_grad_(True)
        tomo.misfit(s0, t0).backward()
        grads[key] = (s0.grad.cpu().numpy(), t0.grad.cpu().numpy(), tomo.stats["recomputed"])
    assert grads["all"][2] == 0 and grads["none"][2] == 4 and grads["half"][2] == 2
    for q in (0, 1):
        a, b = grads["all"][q], grads["none"][q]
        assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
        a, b = grads["all"][q], grads["half"][q]
        assert np.abs(a - b).max() <= 1e-10 * np.abs(a).max()


def test_fp32_gradient_close_to_fp64(solver):
    import torch
    n = 32
    out = {}
    for dt in (torch.float64, torch.float32):
        tomo, f = _tomo(solver, n, dt, 2, 16, batch=4)
        s0 = f["slowness"].clone().requires_grad_(True)
        t0 = f["theta"].clone().requires_grad_(True)
        J = tomo.misfit(s0, t0)
        J.backward()
        out[dt] = (float(J.detach()), s0.grad.double().cpu().numpy(), t0.grad.double().cpu().numpy())
    J64, s64, t64 = out[torch.float64]
    J32, s32, t32 = out[torch.float32]
    assert abs(J32 - J64) <= 1e-2 * J64
    for a, b in ((s64, s32), (t64, t32)):
        err = np.linalg.norm(a - b) / np.linalg.norm(a)
        print("fp32 gradient relative L2 error", err)
        assert err < 1e-2


def test_zero_residual_and_invalid_arguments(solver):
    import torch
    n = 16
    tomo, f = _tomo(solver, n, torch.float64, 1, 8)
    S = tomo.survey
    T = tomo.solve_group([0], f["slowness"], f["theta"])[0]
    m = tomo.model((1.0 / f["slowness"]).contiguous(), f["theta"])
    idx = torch.as_tensor(S.rec_idx, device="cuda")
    gs = torch.zeros_like(T)
    gt = torch.zeros_like(T)
    solver.adjoint(T, n, S.h, m, idx, torch.zeros(S.nrec, dtype=T.dtype, device="cuda"), gs, gt)
    assert not gs.any() and not gt.any()
    bad = idx.clone()
    bad[3] = 0  # a ghost node
    with pytest.raises(fold3d.Fold3dError):
        solver.adjoint(T, n, S.h, m, bad, torch.ones(S.nrec, dtype=T.dtype, device="cuda"), gs, gt)
    m2 = tomo.model((1.0 / f["slowness"]).contiguous(), f["theta"])
    m2.theta = m2.phi = m2.ra = m2.rb = None
    with pytest.raises(fold3d.Fold3dError):
        solver.adjoint(T, n, S.h, m2, idx, torch.ones(S.nrec, dtype=T.dtype, device="cuda"), gs, gt)
    with pytest.raises(fold3d.Fold3dError):
        solver.adjoint(T, n, S.h, fold3d.reference_model(), idx,
                       torch.ones(S.nrec, dtype=T.dtype, device="cuda"), gs, None)
    assert not gs.any()


def test_gradient_vs_finite_differences_at_full_size(solver):
    """The tomography grid at its full 256^3 size, FP64, one source and 16x16
    top-face receivers: the adjoint gradient at 4 seeded cells per parameter
    against Richardson-extrapolated central differences (relative steps 1e-3,
    5e-4) of the forward solve, to 1e-4 relative."""
    import torch
    n = 256
    tomo, f = _tomo(solver, n, torch.float64, 1, 16, batch=1)
    S = tomo.survey
    s0 = f["slowness"].clone().requires_grad_(True)
    t0 = f["theta"].clone().requires_grad_(True)
    tomo.misfit(s0, t0).backward()
    H = n + 2
    src = [min(n, max(1, int(round(c / S.h)) + 1)) for c in S.sources[0]]
    srcs = [src[0] + H * (src[1] + H * src[2])]
    for name, g in (("slowness", s0.grad), ("theta", t0.grad)):
        g = g.cpu().numpy()
        cells = _seeded_cells(g, srcs, H, 4, seed=20261208)
        for c in cells:
            base = float(f[name][c])
            d = 1e-3 * abs(base)

            def Jof(delta):
                s, t = f["slowness"].clone(), f["theta"].clone()
                (s if name == "slowness" else t)[c] += delta
                with torch.no_grad():
                    return float(tomo.forward_pass(s, t)[0])

            d1 = (Jof(d) - Jof(-d)) / (2 * d)
            d2 = (Jof(d / 2) - Jof(-d / 2)) / d
            fd = (4 * d2 - d1) / 3
            print(f"{name} cell {c}: adjoint {g[c]:.6e} fd {fd:.6e} rel {abs(g[c] - fd) / abs(fd):.2e}")
            assert abs(g[c] - fd) <= 1e-4 * abs(fd)


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
def test_batched_adjoint_same_bits_as_one_source_at_a_time(solver, dtype_name):
    """fold3d_adjoint_batched over 5 sources (own residuals and tolerances) adds
    exactly the gradients of 5 fold3d_adjoint calls in the same order, with the
    same sweep count per source; per-node speed and tilt."""
    import torch
    dtype = getattr(torch, dtype_name)
    n = 24
    print(f"batched adjoint {dtype_name} n={n}", flush=True)
    tomo, f = _tomo(solver, n, dtype, 3, 12)
    S = tomo.survey
    ks = [0, 4, 2, 7, 8]
    Ts = tomo.solve_group(ks, f["slowness"], f["theta"])
    rs = [(tomo.receiver_times(T) - tomo.T_obs[k]).contiguous() for k, T in zip(ks, Ts)]
    tols = [adjoint_tol(r) for r in rs]
    m = tomo.model((1.0 / f["slowness"]).contiguous(), f["theta"])
    idx = torch.as_tensor(S.rec_idx, device="cuda", dtype=torch.int32)
    gs1, gt1 = torch.zeros_like(Ts[0]), torch.zeros_like(Ts[0])
    its1 = []
    for T, r, tol in zip(Ts, rs, tols):
        st = solver.adjoint(T, n, S.h, m, idx, r, gs1, gt1, max_iters=200, tol=tol)
        its1.append(st.iterations)
    gs2, gt2 = torch.zeros_like(Ts[0]), torch.zeros_like(Ts[0])
    its2, mds, ms = solver.adjoint_batched(Ts, n, S.h, m, idx, rs, tols, gs2, gt2, max_iters=200)
    assert its2 == its1
    assert all(md <= tol for md, tol in zip(mds, tols)) and ms > 0
    assert float(gs1.abs().max()) > 0 and float(gt1.abs().max()) > 0
    assert torch.equal(gs1, gs2) and torch.equal(gt1, gt2)
    # a second call adds the same amounts again: the handle's scratch is reused
    solver.adjoint_batched(Ts, n, S.h, m, idx, rs, tols, gs2, gt2, max_iters=200)
    for T, r, tol in zip(Ts, rs, tols):
        solver.adjoint(T, n, S.h, m, idx, r, gs1, gt1, max_iters=200, tol=tol)
    assert torch.equal(gs1, gs2) and torch.equal(gt1, gt2)


def test_batched_adjoint_rejects_bad_sources(solver):
    import ctypes as C
    import torch
    n = 12
    T = torch.zeros((n + 2) ** 3, dtype=torch.float32, device="cuda")
    g = torch.zeros_like(T)
    m = fold3d.Model()
    m.kind = fold3d.MODEL_ELLIP
    m.speed = 1.0
    m.M[0] = m.M[1] = m.M[2] = 1.0
    idx = torch.tensor([0], dtype=torch.int32, device="cuda")  # a ghost node
    r = torch.ones(1, dtype=torch.float32, device="cuda")
    with pytest.raises(fold3d.Fold3dError, match="interior"):
        solver.adjoint_batched([T, T], n, 1.0 / n, m, idx, [r, r], [0.0, 0.0], g)
    with pytest.raises(fold3d.Fold3dError, match="tilt"):
        solver.adjoint_batched([T], n, 1.0 / n, m, idx, [r], [0.0], g, grad_theta=g)
    L = fold3d.lib()
    o = fold3d.Opts()
    L.fold3d_default_opts(C.byref(o))
    o.precision = fold3d.F32
    assert L.fold3d_adjoint_batched(solver.h, 0, (fold3d.AdjointSource * 1)(), n, n, n, 0.1, 0.1,
                                    0.1, C.byref(m), C.byref(o), C.c_void_p(g.data_ptr()), None,
                                    None, None, None, None) == fold3d.FOLD3D_E_INVALID
