"""Synthetic inputs of the elliptic-model workloads (no datasets exist).

All fields live on the padded grid (n+2)^3 of the reference layout
T[i + H*(j + H*k)], node (i,j,k) at ((i-1)h, (j-1)h, (k-1)h), i.e. a C array
of shape (H, H, H) indexed [k, j, i].  Every generator takes ``xp`` = numpy
or torch and builds the field from separable Gaussians (outer products), so
1280^3 fields are cheap on the GPU.  Random parameters come from numpy's
PCG64 seeded generator, so both backends see identical parameters.

Workloads:
  tilted_homogeneous  v = 1, M = R(30 deg, 45 deg) diag(1, .8, .6)^2 R^T, source at the
                      centre of [0,1]^3, h = 1/(n-1)
  blob_speed          v = 2 + sum_8 A_b exp(-|x-c_b|^2 / (2 w^2)), A_b ~ U(-.3,.3),
                      c_b ~ U[0,1]^3, w = 0.1, same metric, seeded interior source
  tilted_field        theta in [0, 60 deg], phi in [0, 360 deg), ra, rb in [.6, 1], each a
                      16-mode Gaussian sum rescaled to its range; v = 1.5 + 2z + 5% blobs;
                      source near the top face
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MAXT = 10e15 * 1.0


def tilted_metric(theta, phi, a1, a2, a3):
    """M = R diag(a1^2, a2^2, a3^2) R^T, R = Rz(phi) Ry(theta); (xx,yy,zz,xy,xz,yz)."""
    ct, st, cp, sp = math.cos(theta), math.sin(theta), math.cos(phi), math.sin(phi)
    R = np.array([[cp * ct, -sp, cp * st], [sp * ct, cp, sp * st], [-st, 0.0, ct]])
    A = R @ np.diag([a1 * a1, a2 * a2, a3 * a3]) @ R.T
    return np.array([A[0, 0], A[1, 1], A[2, 2], A[0, 1], A[0, 2], A[1, 2]])


def sym(M6):
    a, b, c, d, e, f = M6
    return np.array([[a, d, e], [d, b, f], [e, f, c]])


CONFIG_METRIC = tilted_metric(math.radians(30), math.radians(45), 1.0, 0.8, 0.6)


@dataclass
class Workload:
    name: str
    n: int
    h: float
    M: np.ndarray                 # constant metric (6), used when theta is None
    speed: float = 1.0            # constant speed when speed_field is None
    source: tuple = (0.5, 0.5, 0.5)
    v_source: float = 1.0         # speed at the source (for the closed-form start)
    speed_field: object = None    # padded fields (numpy or torch), or None
    theta: object = None
    phi: object = None
    ra: object = None
    rb: object = None
    meta: dict = field(default_factory=dict)

    def model(self):
        """fold3d Model for this workload (per-node pointers for torch fields)."""
        from . import Model, MODEL_ELLIP
        m = Model()
        m.kind = MODEL_ELLIP
        m.F = 1.0
        for k in range(6):
            m.M[k] = float(self.M[k])
        m.speed = float(self.speed)
        refs = []
        for name in ("speed_field", "theta", "phi", "ra", "rb"):
            f = getattr(self, name)
            setattr(m, name, None if f is None else f.data_ptr())
            refs.append(f)
        m._refs = tuple(refs)  # the pointers stay valid with m
        return m


def _axis(xp, n, dtype, device):
    if xp is np:
        return (np.arange(n + 2, dtype=np.float64) - 1.0)
    import torch
    return torch.arange(n + 2, dtype=torch.float64, device=device) - 1.0


def _gauss_sum(xp, n, h, centres, widths, amps, dtype, device):
    """sum_b amps[b] * exp(-|x - c_b|^2 / (2 w_b^2)) on the padded grid, [k,j,i]."""
    x = _axis(xp, n, dtype, device) * h
    out = None
    for c, w, a in zip(centres, widths, amps):
        ex = xp.exp(-((x - c[0]) ** 2) / (2 * w * w))
        ey = xp.exp(-((x - c[1]) ** 2) / (2 * w * w))
        ez = xp.exp(-((x - c[2]) ** 2) / (2 * w * w))
        term = a * ez[:, None, None] * ey[None, :, None] * ex[None, None, :]
        out = term if out is None else out + term
    return out


def _cast(xp, a, dtype):
    if xp is np:
        return np.ascontiguousarray(a, dtype=dtype)
    return a.to(dtype).contiguous()


def tilted_homogeneous(n=128, xp=np, dtype=np.float64, device=None) -> Workload:
    return Workload("tilted_homogeneous", n, 1.0 / (n - 1), CONFIG_METRIC, 1.0,
                    (0.5, 0.5, 0.5), 1.0)


def blob_speed(n=512, seed=20261203, xp=np, dtype=np.float64, device=None) -> Workload:
    rng = np.random.default_rng(seed)
    amps = rng.uniform(-0.3, 0.3, 8)
    centres = rng.uniform(0.0, 1.0, (8, 3))
    src = tuple(rng.uniform(0.2, 0.8, 3))
    h = 1.0 / (n - 1)
    v = 2.0 + _gauss_sum(xp, n, h, centres, [0.1] * 8, amps, dtype, device)
    vs = 2.0 + sum(a * math.exp(-sum((s - c) ** 2 for s, c in zip(src, cc)) / (2 * 0.01))
                   for a, cc in zip(amps, centres))
    return Workload("blob_speed", n, h, CONFIG_METRIC, 2.0, src, vs,
                    speed_field=_cast(xp, v, dtype).reshape(-1),
                    meta={"seed": seed, "amps": amps.tolist(), "centres": centres.tolist()})


def _rescaled_modes(xp, n, h, rng, lo, hi, dtype, device, modes=16):
    centres = rng.uniform(0.0, 1.0, (modes, 3))
    widths = rng.uniform(0.15, 0.35, modes)
    amps = rng.normal(size=modes)
    f = _gauss_sum(xp, n, h, centres, widths, amps, dtype, device)
    fmin, fmax = float(f.min()), float(f.max())
    return lo + (f - fmin) * ((hi - lo) / (fmax - fmin))


def tilted_field(n=1280, seed=20261204, xp=np, dtype=np.float32, device=None) -> Workload:
    rng = np.random.default_rng(seed)
    h = 1.0 / (n - 1)
    theta = _rescaled_modes(xp, n, h, rng, 0.0, math.radians(60), dtype, device)
    phi = _rescaled_modes(xp, n, h, rng, 0.0, 2 * math.pi * (1 - 1e-6), dtype, device)
    ra = _rescaled_modes(xp, n, h, rng, 0.6, 1.0, dtype, device)
    rb = _rescaled_modes(xp, n, h, rng, 0.6, 1.0, dtype, device)
    amps = rng.uniform(-0.05, 0.05, 8)
    centres = rng.uniform(0.0, 1.0, (8, 3))
    z = _axis(xp, n, dtype, device) * h
    blobs = _gauss_sum(xp, n, h, centres, [0.1] * 8, amps, dtype, device)
    v = (1.5 + 2.0 * z)[:, None, None] * (1.0 + blobs)
    src = (float(rng.uniform(0.3, 0.7)), float(rng.uniform(0.3, 0.7)), 1.0 - 2.5 * h)
    vs = (1.5 + 2.0 * src[2]) * (1.0 + sum(
        a * math.exp(-sum((s - c) ** 2 for s, c in zip(src, cc)) / (2 * 0.01))
        for a, cc in zip(amps, centres)))
    W = Workload("tilted_field", n, h, CONFIG_METRIC, 1.0, src, vs,
                 speed_field=_cast(xp, v, dtype).reshape(-1))
    W.theta = _cast(xp, theta, dtype).reshape(-1)
    W.phi = _cast(xp, phi, dtype).reshape(-1)
    W.ra = _cast(xp, ra, dtype).reshape(-1)
    W.rb = _cast(xp, rb, dtype).reshape(-1)
    # metric at the source node, for the closed-form start
    ijk = [int(round(c / h)) + 1 for c in src]
    flat = ijk[0] + (n + 2) * (ijk[1] + (n + 2) * ijk[2])
    tv = [float(f[flat]) for f in (W.theta, W.phi, W.ra, W.rb)]
    W.meta = {"seed": seed, "metric_at_source": tilted_metric(tv[0], tv[1], 1.0, tv[2], tv[3]).tolist()}
    return W


def batch_instance(n, seed, xp=np, dtype=np.float32, device=None) -> Workload:
    """One instance of the batched workload: blob speed field (as blob_speed),
    tilt and axis-ratio fields (as tilted_field) and a seeded interior source."""
    rng = np.random.default_rng(seed)
    h = 1.0 / (n - 1)
    amps = rng.uniform(-0.3, 0.3, 8)
    centres = rng.uniform(0.0, 1.0, (8, 3))
    v = 2.0 + _gauss_sum(xp, n, h, centres, [0.1] * 8, amps, dtype, device)
    theta = _rescaled_modes(xp, n, h, rng, 0.0, math.radians(60), dtype, device)
    phi = _rescaled_modes(xp, n, h, rng, 0.0, 2 * math.pi * (1 - 1e-6), dtype, device)
    ra = _rescaled_modes(xp, n, h, rng, 0.6, 1.0, dtype, device)
    rb = _rescaled_modes(xp, n, h, rng, 0.6, 1.0, dtype, device)
    src = tuple(float(c) for c in rng.uniform(0.2, 0.8, 3))
    vs = 2.0 + sum(a * math.exp(-sum((s - c) ** 2 for s, c in zip(src, cc)) / (2 * 0.01))
                   for a, cc in zip(amps, centres))
    W = Workload("batch_instance", n, h, CONFIG_METRIC, 2.0, src, vs,
                 speed_field=_cast(xp, v, dtype).reshape(-1))
    W.theta = _cast(xp, theta, dtype).reshape(-1)
    W.phi = _cast(xp, phi, dtype).reshape(-1)
    W.ra = _cast(xp, ra, dtype).reshape(-1)
    W.rb = _cast(xp, rb, dtype).reshape(-1)
    ijk = [int(round(c / h)) + 1 for c in src]
    flat = ijk[0] + (n + 2) * (ijk[1] + (n + 2) * ijk[2])
    tv = [float(f[flat]) for f in (W.theta, W.phi, W.ra, W.rb)]
    W.meta = {"seed": seed, "metric_at_source": tilted_metric(tv[0], tv[1], 1.0, tv[2], tv[3]).tolist()}
    return W


def batch_sides(count=96, seed=20261205, choices=tuple(range(64, 321, 32))):
    """Instance sides drawn uniformly from {64, 96, ..., 320}."""
    rng = np.random.default_rng(seed)
    return [int(x) for x in rng.choice(choices, count)]


def point_source_start(W: Workload, xp=np, dtype=np.float64, device=None, radius=1.6):
    """T = MAXT except the nodes within `radius` cells (max-norm) of the source,
    which get the closed form sqrt(d^T M^-1 d)/v(source).  Ghosts stay MAXT."""
    n, h = W.n, W.h
    M = np.array(W.meta.get("metric_at_source", W.M))
    N = np.linalg.inv(sym(M))
    x = _axis(xp, n, dtype, device) * h
    dx = (x - W.source[0])[None, None, :]
    dy = (x - W.source[1])[None, :, None]
    dz = (x - W.source[2])[:, None, None]
    q = (N[0, 0] * dx * dx + N[1, 1] * dy * dy + N[2, 2] * dz * dz
         + 2 * N[0, 1] * dx * dy + 2 * N[0, 2] * dx * dz + 2 * N[1, 2] * dy * dz)
    T = xp.sqrt(q) / W.v_source
    r = h * radius
    near = (abs(dx) <= r) & (abs(dy) <= r) & (abs(dz) <= r)
    T = xp.where(near, T, xp.full_like(T, MAXT) if xp is not np else MAXT)
    idx = _axis(xp, n, dtype, device)
    inside = (idx >= 1) & (idx <= n)
    mask = inside[:, None, None] & inside[None, :, None] & inside[None, None, :]
    T = xp.where(mask, T, xp.full_like(T, MAXT) if xp is not np else MAXT)
    return _cast(xp, T, dtype).reshape(-1)


def closed_form(W: Workload, xp=np, device=None):
    """sqrt(d^T M^-1 d)/v for a homogeneous medium (the first workload)."""
    n, h = W.n, W.h
    N = np.linalg.inv(sym(W.M))
    x = _axis(xp, n, np.float64, device) * h
    dx = (x - W.source[0])[None, None, :]
    dy = (x - W.source[1])[None, :, None]
    dz = (x - W.source[2])[:, None, None]
    q = (N[0, 0] * dx * dx + N[1, 1] * dy * dy + N[2, 2] * dz * dz
         + 2 * N[0, 1] * dx * dy + 2 * N[0, 2] * dx * dz + 2 * N[1, 2] * dy * dz)
    return (xp.sqrt(q) / W.speed).reshape(-1)
