"""Traveltime tomography on the elliptic model: the receiver misfit
    J = 1/2 sum_sources sum_receivers (T_s[rec] - T_obs[s, rec])^2
and its gradient with respect to per-node slowness s = 1/v and tilt angle
theta, through the adjoint-state entry point fold3d_adjoint, wrapped in a
torch.autograd.Function.

Forward fields are solved in groups of `batch` sources with the batched
solver (one launch sequence per group, per-instance convergence), or one
source at a time with batch=1.  Each solved field is kept as a checkpoint
for the backward pass while the checkpoints fit `budget_bytes`; the rest
are recomputed in the backward pass, group by group.

Sources sit a few cells below the top face z = 1; receivers are top-face
nodes.  The start of every solve is the closed form sqrt(d^T N d) / v at
nodes within 1.6 cells (max norm) of the source, with N and v taken at the
node nearest the source (fold3d_point_source); those nodes are held fixed by the adjoint (they do
not depend on the parameters except through that one node).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import (F32, F64, MODEL_ELLIP, Model, Solver, check, lib)  # noqa: F401
from .workloads import _axis, _gauss_sum, tilted_metric


@dataclass
class Survey:
    n: int
    h: float
    sources: list          # (x, y, z) in [0, 1]^3
    rec_idx: np.ndarray    # int32 linear indices on the padded grid

    @property
    def nrec(self) -> int:
        return int(self.rec_idx.size)


def top_face_survey(n: int, src_side: int = 4, rec_side: int = 64) -> Survey:
    """src_side x src_side sources 2.5 cells below the top face and
    rec_side x rec_side receivers on the top-face nodes (k = n)."""
    h = 1.0 / (n - 1)
    rec_side = min(rec_side, n)
    sources = [((a + 0.5) / src_side, (b + 0.5) / src_side, 1.0 - 2.5 * h)
               for b in range(src_side) for a in range(src_side)]
    nodes = [1 + int(round((c + 0.5) * (n - 1) / rec_side)) for c in range(rec_side)]
    if len(set(nodes)) != rec_side:
        raise ValueError("receivers do not fall on distinct nodes")
    H = n + 2
    idx = [i + H * (j + H * n) for j in nodes for i in nodes]
    return Survey(n, h, sources, np.asarray(idx, dtype=np.int32))


def tomography_fields(n: int = 256, seed: int = 20261206, xp=np, dtype=np.float32,
                      device=None, anomalies: bool = False):
    """Fields of the tomography workload on the padded grid.

    Speed: v = 1 + 6 d + sum_8 A_b exp(-|x-c_b|^2/(2*0.1^2)), A_b ~ U(+-.3),
    c_b ~ U[0,1]^3, with d = 1 - z the depth below the top face (the vertical
    gradient turns the first arrivals between top-face points back up through
    the volume, down to d ~ 0.35 at unit offset; without it they run along
    the face).  With `anomalies`, v is multiplied by 1 + 3 seeded Gaussian
    anomalies of +-5 %, width 0.08, centred at x, y ~ U(.25, .75), d ~ U(.1, .3).
    Tilt: theta = 30 deg varying by up to +-10 deg over 4 seeded Gaussian
    modes, phi = 45 deg, ra = 0.8, rb = 0.6.
    Returns dict(slowness, theta, phi, ra, rb)."""
    rng = np.random.default_rng(seed)
    h = 1.0 / (n - 1)
    amps = rng.uniform(-0.3, 0.3, 8)
    centres = rng.uniform(0.0, 1.0, (8, 3))
    z = _axis(xp, n, dtype, device) * h
    v = (1.0 + 6.0 * (1.0 - z))[:, None, None] + _gauss_sum(xp, n, h, centres, [0.1] * 8, amps,
                                                             dtype, device)
    an_amp = rng.choice([-0.05, 0.05], 3)
    an_c = np.column_stack([rng.uniform(0.25, 0.75, (3, 2)), 1.0 - rng.uniform(0.1, 0.3, 3)])
    if anomalies:
        v = v * (1.0 + _gauss_sum(xp, n, h, an_c, [0.08] * 3, an_amp, dtype, device))
    th_amp = rng.uniform(-1.0, 1.0, 4)
    th_c = rng.uniform(0.0, 1.0, (4, 3))
    g = _gauss_sum(xp, n, h, th_c, [0.2] * 4, th_amp, dtype, device)
    theta = math.radians(30) + math.radians(10) * g / max(1.0, float(abs(th_amp).sum()))
    ones = v * 0 + 1
    out = {"slowness": 1.0 / v, "theta": theta, "phi": ones * math.radians(45),
           "ra": ones * 0.8, "rb": ones * 0.6}
    for k, f in out.items():
        if xp is np:
            out[k] = np.ascontiguousarray(f, dtype=dtype).reshape(-1)
        else:
            out[k] = f.to(dtype).contiguous().reshape(-1)
    return out


def adjoint_tol(res) -> float:
    """Stop rule of the adjoint sweeps: no adjoint value changes by more than a
    few units in the last place of sum|res| (an upper bound of every adjoint
    value along a single path; sweeps reach the exact fixed point unless equal
    traveltimes close a loop of stencils, where the last bits can cycle)."""
    import torch
    eps = torch.finfo(res.dtype).eps
    return 4.0 * eps * float(res.abs().sum())


def _node_of(pos, n, h):
    return [min(n, max(1, int(round(c / h)) + 1)) for c in pos]


class Tomography:
    """Misfit and gradient of one survey on one device.

    phi, ra, rb: fixed per-node tensors (device, the field dtype);
    T_obs: observed traveltimes, tensor (nsrc, nrec) of the field dtype;
    autotune: before the first batched forward solve, time the batched
    solver's schedules on one iteration of that group (fold3d_autotune_batched)
    and keep the fastest in the solver's handle.  Off by default: from the
    point-source start the first iteration is not representative (at 256^3 it
    picks 2 waves per row, 0.64 s per 16 solves, against 0.54 s for the
    default schedule)."""

    def __init__(self, solver: Solver, survey: Survey, phi, ra, rb, T_obs=None,
                 batch: int = 16, budget_bytes: float = float("inf"),
                 max_iters: int = 400, tol: float = 1e-30, adj_iters: int = 400,
                 autotune: bool = False):
        self.solver, self.survey = solver, survey
        self.phi, self.ra, self.rb = phi, ra, rb
        self.T_obs = T_obs
        self.batch = max(1, int(batch))
        self.budget = budget_bytes
        self.max_iters, self.tol, self.adj_iters = max_iters, tol, adj_iters
        self.stats = {}
        self._tuned = not autotune

    # model -----------------------------------------------------------------
    def model(self, speed, theta) -> Model:
        m = Model()
        m.kind = MODEL_ELLIP
        m.F = 1.0
        m.speed = 1.0
        m.speed_field = speed.data_ptr()
        m.theta, m.phi = theta.data_ptr(), self.phi.data_ptr()
        m.ra, m.rb = self.ra.data_ptr(), self.rb.data_ptr()
        m._refs = (speed, theta, self.phi, self.ra, self.rb)  # the pointers stay valid with m
        return m

    def start_field(self, k: int, slowness, theta):
        """Closed-form start of source k (device tensor, the field dtype),
        written by the library's point-source kernel."""
        import torch
        S = self.survey
        n, h = S.n, S.h
        src = S.sources[k]
        i, j, kk = _node_of(src, n, h)
        H = n + 2
        flat = i + H * (j + H * kk)
        M = tilted_metric(float(theta[flat]), float(self.phi[flat]), 1.0,
                          float(self.ra[flat]), float(self.rb[flat]))
        T = torch.empty(H ** 3, dtype=slowness.dtype, device=slowness.device)
        self.solver.point_source(T, n, h, src, M, 1.0 / float(slowness[flat]), radius=1.6)
        return T

    def solve_group(self, ks, slowness, theta):
        """Solved fields of the sources in `ks` (list of tensors)."""
        S = self.survey
        speed = (1.0 / slowness).contiguous()
        Ts = [self.start_field(k, slowness, theta) for k in ks]
        if len(ks) == 1:
            st = self.solver.solve(Ts[0], S.n, S.h, self.model(speed, theta),
                                   max_iters=self.max_iters, tol=self.tol)
            self.stats.setdefault("forward_iters", []).append(st.iterations)
        else:
            from .workloads import Workload
            Ws = []
            for _ in ks:
                W = Workload("tomography", S.n, S.h, np.zeros(6), 1.0)
                W.speed_field, W.theta = speed, theta
                W.phi, W.ra, W.rb = self.phi, self.ra, self.rb
                Ws.append(W)
            if not self._tuned:
                self.solver.autotune_batched(Ws, Ts, iterations=1)
                self._tuned = True
            its, _, _ = self.solver.solve_batched(Ws, Ts, max_iters=self.max_iters,
                                                  tol=self.tol)
            self.stats.setdefault("forward_iters", []).extend(its)
        del speed
        return Ts

    def receiver_times(self, T):
        import torch
        idx = torch.as_tensor(self.survey.rec_idx, device=T.device, dtype=torch.long)
        return T[idx]

    # forward / backward -----------------------------------------------------
    def forward_pass(self, slowness, theta):
        """J, residuals per source and the checkpoints kept (None = dropped)."""
        import torch
        nsrc = len(self.survey.sources)
        res, ckpt = [None] * nsrc, [None] * nsrc
        used = 0.0
        J = torch.zeros((), dtype=torch.float64, device=slowness.device)
        for g0 in range(0, nsrc, self.batch):
            ks = list(range(g0, min(nsrc, g0 + self.batch)))
            Ts = self.solve_group(ks, slowness, theta)
            for k, T in zip(ks, Ts):
                r = self.receiver_times(T) - self.T_obs[k]
                res[k] = r
                J = J + 0.5 * (r.double() ** 2).sum()
                nbytes = T.numel() * T.element_size()
                if used + nbytes <= self.budget:
                    ckpt[k] = T
                    used += nbytes
            del Ts
        self.stats["checkpoint_bytes"] = used
        self.stats["recomputed"] = sum(c is None for c in ckpt)
        return J, res, ckpt

    def backward_pass(self, slowness, theta, res, ckpt, scale=1.0):
        import torch
        S = self.survey
        gs = torch.zeros_like(slowness)
        gt = torch.zeros_like(theta)
        speed = (1.0 / slowness).contiguous()
        m = self.model(speed, theta)
        idx = torch.as_tensor(S.rec_idx, device=slowness.device, dtype=torch.int32)
        nsrc = len(S.sources)
        missing = [k for k in range(nsrc) if ckpt[k] is None]
        fields = dict((k, ckpt[k]) for k in range(nsrc) if ckpt[k] is not None)
        adj_iters, adj_ms = [], []

        def adjoint_group(ks, Ts):
            rs = [(res[k] * scale).contiguous() for k in ks]
            its, _, ms = self.solver.adjoint_batched(Ts, S.n, S.h, m, idx, rs,
                                                     [adjoint_tol(r) for r in rs], gs, gt,
                                                     max_iters=self.adj_iters)
            adj_iters.extend(its)
            adj_ms.append(ms)

        # checkpointed sources first, in groups of `batch`, then the recomputed ones
        order = [k for k in range(nsrc) if ckpt[k] is not None]
        for g0 in range(0, len(order), self.batch):
            ks = order[g0:g0 + self.batch]
            adjoint_group(ks, [fields[k] for k in ks])
        for g0 in range(0, len(missing), self.batch):
            ks = missing[g0:g0 + self.batch]
            Ts = self.solve_group(ks, slowness, theta)
            adjoint_group(ks, Ts)
            del Ts
        self.stats["adjoint_iters"] = adj_iters
        self.stats["adjoint_ms"] = adj_ms
        return gs, gt

    def misfit(self, slowness, theta):
        """J(slowness, theta) as a differentiable torch scalar (float64)."""
        return _Misfit.apply(slowness, theta, self)

    def observe(self, slowness, theta):
        """Set T_obs to the receiver times of the model (slowness, theta)."""
        import torch
        nsrc = len(self.survey.sources)
        obs = []
        for g0 in range(0, nsrc, self.batch):
            ks = list(range(g0, min(nsrc, g0 + self.batch)))
            for T in self.solve_group(ks, slowness, theta):
                obs.append(self.receiver_times(T).clone())
        self.T_obs = torch.stack(obs)
        return self.T_obs


def _function():
    import torch

    class Misfit(torch.autograd.Function):
        @staticmethod
        def forward(ctx, slowness, theta, tomo):
            J, res, ckpt = tomo.forward_pass(slowness, theta)
            ctx.tomo, ctx.res, ctx.ckpt = tomo, res, ckpt
            ctx.save_for_backward(slowness, theta)
            return J

        @staticmethod
        def backward(ctx, gJ):
            slowness, theta = ctx.saved_tensors
            gs, gt = ctx.tomo.backward_pass(slowness, theta, ctx.res, ctx.ckpt,
                                            scale=float(gJ))
            ctx.res = ctx.ckpt = None
            return gs, gt, None

    return Misfit


class _Lazy:
    _f = None

    def apply(self, *a):
        if _Lazy._f is None:
            _Lazy._f = _function()
        return _Lazy._f.apply(*a)


_Misfit = _Lazy()
