"""Write the golden vectors of the fold-model solve from the reference's own
solver, compiled in this container from its sources (oracle/Makefile `ref`).

The reference ships no tests, fixtures or golden data, so every fixture here
comes from running its code:
  Stencil-OpenACC/AN_ISO_1_OpenACC.c   FoldEdge2Points :59, FoldSurf3Points :119,
                                       plane kernels :180-1092, driver order :1192-1237
  Stencil-OpenACC/Fold3dPMM.c          InitFoldFromPoint :9, ImplicitInitialiser :129

Run:  make -C oracle ref && python tests/golden/make_golden.py [--large-only]
Output: tests/golden/fold_ref_golden.npz and fold_ref_large.npz (+ .sha256 each).
The second holds digests (SHA-256 of the field's bytes, 4,096 seeded sample
nodes, per-plane sums) of the 64^3 solve after 1 and 8 iterations and at the
fixed point, and of the 128^3 solve at the fixed point.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from oracle.oracle import MAXT, REF_A, REF_F, RefBuild  # noqa: E402

SEED = 20261203


def edge_inputs(rng: np.random.Generator, count: int) -> np.ndarray:
    """Tuples (tnew, st, xt, F, ax, ay, az, dxx, dzz) covering accept/reject
    on every branch: equal values, MAXT, a = 0, |a| close to and above F."""
    X = np.empty((count, 9))
    for r in range(count):
        kind = r % 8
        F = REF_F if kind < 4 else rng.uniform(0.6, 2.0)
        a = rng.uniform(-0.8, 0.8, 3) * F
        if kind == 5:
            a[:] = 0.0
        if kind == 6:  # |a| close to or beyond F
            a = a / np.linalg.norm(a) * F * rng.uniform(0.95, 1.05)
        st, xt = rng.uniform(0.0, 12.0, 2)
        if kind == 7:
            xt = st
        dxx = rng.uniform(0.05, 1.0)
        dzz = rng.uniform(0.05, 1.0) if kind != 3 else dxx
        tnew = rng.choice([MAXT, st + rng.uniform(-0.5, 2.0) * dzz, xt + dzz])
        X[r] = (tnew, st, xt, F, a[0], a[1], a[2], dxx, dzz)
    return X


def surf_inputs(rng: np.random.Generator, count: int) -> np.ndarray:
    """Tuples (tnew, st, xt, yt, F, ax, ay, az, dxx, dyy, dzz, dxz)."""
    X = np.empty((count, 12))
    for r in range(count):
        kind = r % 8
        F = REF_F if kind < 4 else rng.uniform(0.6, 2.0)
        a = rng.uniform(-0.8, 0.8, 3) * F
        if kind == 5:
            a[:] = 0.0
        if kind == 6:
            a = a / np.linalg.norm(a) * F * rng.uniform(0.95, 1.05)
        base = rng.uniform(0.0, 12.0)
        st, xt, yt = base + rng.uniform(-0.6, 0.6, 3)
        if kind == 7:
            yt = xt
        dxx, dyy = rng.uniform(0.05, 1.0, 2)
        dzz = rng.uniform(0.05, 1.0)
        if kind == 2:
            dxx = dyy = dzz
        dxz = np.sqrt(dxx * dxx + dzz * dzz)
        tnew = rng.choice([MAXT, min(st, xt, yt) + rng.uniform(0.0, 2.0) * dzz])
        X[r] = (tnew, st, xt, yt, F, a[0], a[1], a[2], dxx, dyy, dzz, dxz)
    return X


def fold_point_inputs(rng: np.random.Generator, count: int) -> np.ndarray:
    X = np.empty((count, 7))
    for r in range(count):
        F = REF_F if r % 2 == 0 else rng.uniform(0.6, 2.0)
        a = rng.uniform(-0.9, 0.9, 3) * F
        d = rng.uniform(-3, 3, 3)
        if r % 17 == 0:
            d[:] = 0.0
        X[r] = (F, a[0], a[1], a[2], d[0], d[1], d[2])
    return X


def random_field(rng: np.random.Generator, n: int) -> np.ndarray:
    """Signed field with magnitudes in a narrow band, so most nodes take the
    full local-solver path; a sprinkle of MAXT (unreached) nodes."""
    H = n + 2
    mag = rng.uniform(2.0, 6.0, H ** 3)
    mag[rng.random(H ** 3) < 0.05] = MAXT
    sgn = np.where(rng.random(H ** 3) < 0.3, -1.0, 1.0)
    return sgn * mag


def main() -> None:
    rng = np.random.default_rng(SEED)
    out: dict[str, np.ndarray] = {}
    r16 = RefBuild(16)

    E = edge_inputs(rng, 6000)
    out["edge_in"] = E
    out["edge_out"] = np.array([r16.edge(*row) for row in E])
    S = surf_inputs(rng, 6000)
    out["surf_in"] = S
    out["surf_out"] = np.array([r16.surf(*row) for row in S])
    P = fold_point_inputs(rng, 2000)
    out["point_in"] = P
    out["point_out"] = np.array([r16.fold_point(*row) for row in P])

    # initializer, full fields
    for n in (16, 24):
        out[f"init_{n}"] = RefBuild(n).initialiser()

    # one plane step per direction on a random field and on the initial field
    planes = {0: 7, 1: 9, 2: 7, 3: 9, 4: 9, 5: 7}
    Trand = random_field(rng, 16)
    out["plane_rand_in_16"] = Trand
    Tinit = out["init_16"]
    for d, p in planes.items():
        A = Trand.copy()
        r16.plane(d, A, p)
        out[f"plane_rand_{d}_16"] = A
        B = Tinit.copy()
        r16.direction(d, B)
        out[f"direction_init_{d}_16"] = B

    # full solves: after 1 and 8 iterations, and at the fixed point.
    # fixed_iters = index of the first iteration that changes no node.
    for n in (16, 24, 32):
        ref = RefBuild(n)
        A = ref.initialiser()
        it, fixed_at = 0, None
        while it < 8 or fixed_at is None:
            B = A.copy()
            ref.sweeps(A, 1)
            it += 1
            if it == 1:
                out[f"solve1_{n}"] = A.copy()
            if it == 8:
                out[f"solve8_{n}"] = A.copy()
            if fixed_at is None and np.array_equal(A, B):
                fixed_at = it
                out[f"fixed_{n}"] = A.copy()
        out[f"fixed_iters_{n}"] = np.array([fixed_at])

    # 128^3: digests only (the full field is 18 MB)
    ref = RefBuild(128)
    T = ref.initialiser()
    ref.sweeps(T, 8)
    out["solve8_128_sha256"] = np.frombuffer(hashlib.sha256(T.tobytes()).digest(), np.uint8)
    idx = np.sort(rng.choice(T.size, 4096, replace=False))
    out["solve8_128_idx"] = idx
    out["solve8_128_val"] = T[idx]
    H = 130
    out["solve8_128_plane_sums"] = np.abs(T.reshape(H, H, H)).clip(max=1e3).sum(axis=(1, 2))

    save(out, "fold_ref_golden.npz")


def save(out: dict, name: str) -> None:
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **out)
    digest = hashlib.sha256(open(path, "rb").read()).hexdigest()
    with open(path + ".sha256", "w") as f:
        f.write(f"{digest}  {name}\n")
    print(path, os.path.getsize(path), "bytes", digest)


def field_digest(out: dict, key: str, T: np.ndarray, n: int, rng: np.random.Generator) -> None:
    out[f"{key}_sha256"] = np.frombuffer(hashlib.sha256(T.tobytes()).digest(), np.uint8)
    idx = np.sort(rng.choice(T.size, 4096, replace=False))
    out[f"{key}_idx"] = idx
    out[f"{key}_val"] = T[idx]
    H = n + 2
    out[f"{key}_plane_sums"] = np.abs(T.reshape(H, H, H)).clip(max=1e3).sum(axis=(1, 2))


def large() -> None:
    """Digests of the 64^3 and 128^3 solves (the full fields are 2.3 and 18 MB)."""
    rng = np.random.default_rng(SEED + 1)
    out: dict[str, np.ndarray] = {}
    for n in (64, 128):
        ref = RefBuild(n)
        A = ref.initialiser()
        it, fixed_at = 0, None
        while fixed_at is None:
            B = A.copy()
            ref.sweeps(A, 1)
            it += 1
            if n == 64 and it == 1:
                field_digest(out, "solve1_64", A, n, rng)
            if n == 64 and it == 8:
                field_digest(out, "solve8_64", A, n, rng)
            if np.array_equal(A, B):
                fixed_at = it
                field_digest(out, f"fixed_{n}", A, n, rng)
            print(f"n={n} iteration {it}", flush=True)
        out[f"fixed_iters_{n}"] = np.array([fixed_at])
    save(out, "fold_ref_large.npz")


if __name__ == "__main__":
    if "--large-only" not in sys.argv:
        main()
    large()
