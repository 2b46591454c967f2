"""ctypes view of the oracle (oracle/liboracle.so) and of the reference checker
builds (oracle/_ref/libfoldref_<N>.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, ``__graft_entry__.smoke()`` and
the cpu_baseline leg of bench.py, always as the checker; the shipped library
never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

# Reference model constants (reference Stencil-OpenACC/Fold3dPMM.h:49-52).
REF_F = 1.1
REF_A = (-0.35, 0.4, 0.7)
MAXT = 10e15 * 1.0  # Fold3dPMM.h:6

DIRECTIONS = ("x_down", "x_up", "k_up", "k_down", "y_up", "y_down")

_d = C.c_double
_dp = C.POINTER(C.c_double)
_i = C.c_int


def build(quiet: bool = True) -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


class Oracle:
    """The CPU restatement (FP64, OpenMP over plane cells)."""

    def __init__(self, path: str = LIB_PATH):
        if not os.path.exists(path):
            build()
        lib = C.CDLL(path)
        lib.orc_edge.restype = _d
        lib.orc_edge.argtypes = [_d] * 9
        lib.orc_surf.restype = _d
        lib.orc_surf.argtypes = [_d] * 12
        lib.orc_fold_point.restype = _d
        lib.orc_fold_point.argtypes = [_d] * 7
        lib.orc_impsurf.restype = _d
        lib.orc_impsurf.argtypes = [_d] * 3
        lib.orc_initialiser.argtypes = [_dp, _i]
        lib.orc_plane.argtypes = [_i, _dp, _i, _i, _d, _dp, _dp]
        lib.orc_direction.argtypes = [_i, _dp, _i, _d, _dp, _dp]
        lib.orc_sweeps.argtypes = [_dp, _i, _i, _d, _dp, _dp]
        lib.orc_set_threads.argtypes = [_i]
        self.lib = lib

    # local solvers -------------------------------------------------------
    def edge(self, *args):
        return self.lib.orc_edge(*map(float, args))

    def surf(self, *args):
        return self.lib.orc_surf(*map(float, args))

    def fold_point(self, *args):
        return self.lib.orc_fold_point(*map(float, args))

    # grids ---------------------------------------------------------------
    @staticmethod
    def ref_spacing(n: int):
        h = (10.0 - 0.0) / (n - 1.0)
        return (h, h, h)

    def initialiser(self, n: int) -> np.ndarray:
        T = np.empty((n + 2) ** 3, dtype=np.float64)
        self.lib.orc_initialiser(_ptr(T), n)
        return T

    def _model(self, n, F, a, h):
        F = REF_F if F is None else F
        a = np.ascontiguousarray(REF_A if a is None else a, dtype=np.float64)
        h = np.ascontiguousarray(self.ref_spacing(n) if h is None else h, dtype=np.float64)
        return F, a, h

    def plane(self, d: int, T: np.ndarray, n: int, p: int, F=None, a=None, h=None):
        F, a, h = self._model(n, F, a, h)
        self.lib.orc_plane(d, _ptr(T), n, p, F, _ptr(a), _ptr(h))

    def direction(self, d: int, T: np.ndarray, n: int, F=None, a=None, h=None):
        F, a, h = self._model(n, F, a, h)
        self.lib.orc_direction(d, _ptr(T), n, F, _ptr(a), _ptr(h))

    def sweeps(self, T: np.ndarray, n: int, iterations: int, F=None, a=None, h=None):
        F, a, h = self._model(n, F, a, h)
        self.lib.orc_sweeps(_ptr(T), n, iterations, F, _ptr(a), _ptr(h))

    def sweeps3(self, T: np.ndarray, n3, iterations: int, F, a, h):
        """Fold-model sweeps on a box grid n3 = (nx, ny, nz)."""
        lib = self.lib
        ip = C.POINTER(C.c_int)
        lib.orc_sweeps3.argtypes = [_dp, ip, _i, _d, _dp, _dp]
        nn = (C.c_int * 3)(*n3)
        a = np.ascontiguousarray(a, dtype=np.float64)
        h = np.ascontiguousarray(h, dtype=np.float64)
        lib.orc_sweeps3(_ptr(T), nn, iterations, F, _ptr(a), _ptr(h))

    def set_threads(self, k: int):
        self.lib.orc_set_threads(k)
        self.lib.orc_ellip_set_threads(k)

    # tilted elliptic model -------------------------------------------------
    def _ellip_proto(self):
        lib = self.lib
        if getattr(self, "_ellip_ready", False):
            return lib
        lib.orc_tilted_metric.argtypes = [_d] * 5 + [_dp]
        lib.orc_sym3_inverse.argtypes = [_dp, _dp]
        lib.orc_ellip_constants.argtypes = [_i, _dp, _dp, C.c_void_p]
        lib.orc_ellip_edge.restype = _d
        lib.orc_ellip_edge.argtypes = [_d, _d, _d, _dp]
        lib.orc_ellip_tri.restype = _d
        lib.orc_ellip_tri.argtypes = [_d, _d, _d, _d, _dp]
        lib.orc_ellip_node.restype = _d
        lib.orc_ellip_node.argtypes = [_d, _dp, _d, C.c_void_p]
        lib.orc_ellip_set_threads.argtypes = [_i]
        ip = C.POINTER(C.c_int)
        lib.orc_ellip_plane.restype = _d
        lib.orc_ellip_plane.argtypes = [_i, _dp, ip, _i, _dp, _dp, _d] + [_dp] * 5
        lib.orc_ellip_sweeps.restype = _i
        lib.orc_ellip_sweeps.argtypes = [_dp, ip, _i, _d, _dp, _dp, _d] + [_dp] * 5 + [_dp]
        lp = C.POINTER(C.c_long)
        lib.orc_ellip_adjoint_local.argtypes = [_i]
        lib.orc_ellip_adjoint_local.restype = None
        lib.orc_ellip_adjoint.restype = _i
        lib.orc_ellip_adjoint.argtypes = ([_dp, ip, _dp, _dp, _d] + [_dp] * 5 +
                                          [_i, lp, _dp, _dp, _dp, _d, _i, _dp, C.c_void_p])
        self._ellip_ready = True
        return lib

    def tilted_metric(self, theta, phi, a1, a2, a3) -> np.ndarray:
        M = np.zeros(6)
        self._ellip_proto().orc_tilted_metric(theta, phi, a1, a2, a3, _ptr(M))
        return M

    def sym3_inverse(self, M) -> np.ndarray:
        M = np.ascontiguousarray(M, dtype=np.float64)
        N = np.zeros(6)
        self._ellip_proto().orc_sym3_inverse(_ptr(M), _ptr(N))
        return N

    def ellip_constants(self, d: int, N, h) -> np.ndarray:
        """(9 + 48 + 48) doubles: vert[9], edge[16][3], tri[8][6]."""
        K = np.zeros(9 + 48 + 48)
        N = np.ascontiguousarray(N, dtype=np.float64)
        h = np.ascontiguousarray(h, dtype=np.float64)
        self._ellip_proto().orc_ellip_constants(d, _ptr(N), _ptr(h), K.ctypes.data)
        return K

    def ellip_edge(self, t0, t1, s2, E):
        E = np.ascontiguousarray(E, dtype=np.float64)
        return self._ellip_proto().orc_ellip_edge(t0, t1, s2, _ptr(E))

    def ellip_tri(self, t0, t1, t2, s2, Tr):
        Tr = np.ascontiguousarray(Tr, dtype=np.float64)
        return self._ellip_proto().orc_ellip_tri(t0, t1, t2, s2, _ptr(Tr))

    def ellip_node(self, told, p9, s, K):
        p9 = np.ascontiguousarray(p9, dtype=np.float64)
        K = np.ascontiguousarray(K, dtype=np.float64)
        return self._ellip_proto().orc_ellip_node(told, _ptr(p9), s, K.ctypes.data)

    @staticmethod
    def _fields(n, speed_field, theta, phi, ra, rb):
        out = []
        for f in (speed_field, theta, phi, ra, rb):
            if f is None:
                out.append(None)
            else:
                f = np.ascontiguousarray(f, dtype=np.float64).ravel()
                assert f.size == (n[0] + 2) * (n[1] + 2) * (n[2] + 2)
                out.append(f)
        return out

    def ellip_plane(self, d, T, n, p, h, M=None, speed=1.0, speed_field=None,
                    theta=None, phi=None, ra=None, rb=None) -> float:
        n3 = (C.c_int * 3)(*((n,) * 3 if np.isscalar(n) else n))
        nn = tuple(n3)
        M = np.ascontiguousarray(np.eye(3)[[0, 1, 2, 0, 0, 1], [0, 1, 2, 1, 2, 2]]
                                 if M is None else M, dtype=np.float64)
        h = np.ascontiguousarray((h,) * 3 if np.isscalar(h) else h, dtype=np.float64)
        fs = self._fields(nn, speed_field, theta, phi, ra, rb)
        ptrs = [None if f is None else _ptr(f) for f in fs]
        return self._ellip_proto().orc_ellip_plane(d, _ptr(T), n3, p, _ptr(h), _ptr(M),
                                                   speed, *ptrs)

    def ellip_sweeps(self, T, n, max_iters, tol, h, M=None, speed=1.0, speed_field=None,
                     theta=None, phi=None, ra=None, rb=None):
        """Returns (iterations run, max|dT| of the last iteration)."""
        n3 = (C.c_int * 3)(*((n,) * 3 if np.isscalar(n) else n))
        nn = tuple(n3)
        M = np.ascontiguousarray(np.eye(3)[[0, 1, 2, 0, 0, 1], [0, 1, 2, 1, 2, 2]]
                                 if M is None else M, dtype=np.float64)
        h = np.ascontiguousarray((h,) * 3 if np.isscalar(h) else h, dtype=np.float64)
        fs = self._fields(nn, speed_field, theta, phi, ra, rb)
        ptrs = [None if f is None else _ptr(f) for f in fs]
        md = np.zeros(1)
        it = self._ellip_proto().orc_ellip_sweeps(_ptr(T), n3, max_iters, tol, _ptr(h),
                                                  _ptr(M), speed, *ptrs, _ptr(md))
        return it, float(md[0])

    def ellip_adjoint(self, T, n, h, rec_idx, rec_res, M=None, speed=1.0, speed_field=None,
                      theta=None, phi=None, ra=None, rb=None, rel_tol=1e-12, max_passes=100000,
                      state=False, local="closure"):
        """FP64 adjoint of the converged field T: returns (dJ/ds, dJ/dtheta or
        None, Jacobi passes) for J = 1/2 sum rec_res^2 at nodes rec_idx; with
        `state`, also the adjoint state mu and the held-fixed flags.  `local`:
        the node derivatives are central differences (step 1e-6) of the
        winning simplex with its faces ("closure": the candidate the node
        follows) or of the node's whole local solver ("solver": averages the
        slopes of two near-tied candidates of different patches)."""
        n3 = (C.c_int * 3)(*((n,) * 3 if np.isscalar(n) else n))
        nn = tuple(n3)
        M = np.ascontiguousarray(np.eye(3)[[0, 1, 2, 0, 0, 1], [0, 1, 2, 1, 2, 2]]
                                 if M is None else M, dtype=np.float64)
        h = np.ascontiguousarray((h,) * 3 if np.isscalar(h) else h, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64).ravel()
        fs = self._fields(nn, speed_field, theta, phi, ra, rb)
        ptrs = [None if f is None else _ptr(f) for f in fs]
        idx = np.ascontiguousarray(rec_idx, dtype=np.int64)
        res = np.ascontiguousarray(rec_res, dtype=np.float64)
        gs = np.zeros(T.size)
        gt = np.zeros(T.size) if theta is not None else None
        mu = np.zeros(T.size) if state else None
        fixed = np.zeros(T.size, dtype=np.uint8) if state else None
        if local not in ("closure", "solver"):
            raise ValueError("local must be 'closure' or 'solver'")
        self._ellip_proto().orc_ellip_adjoint_local(1 if local == "closure" else 0)
        passes = self._ellip_proto().orc_ellip_adjoint(
            _ptr(T), n3, _ptr(h), _ptr(M), speed, *ptrs, int(idx.size),
            idx.ctypes.data_as(C.POINTER(C.c_long)), _ptr(res), _ptr(gs),
            None if gt is None else _ptr(gt), rel_tol, max_passes,
            None if mu is None else _ptr(mu), None if fixed is None else fixed.ctypes.data)
        if state:
            return gs, gt, passes, mu, fixed.astype(bool)
        return gs, gt, passes


class RefBuild:
    """The reference's own solver compiled for an n^3 grid (oracle/_ref)."""

    def __init__(self, n: int):
        path = os.path.join(REF_DIR, f"libfoldref_{n}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        lib = C.CDLL(path)
        lib.ref_edge.restype = _d
        lib.ref_edge.argtypes = [_d] * 9
        lib.ref_surf.restype = _d
        lib.ref_surf.argtypes = [_d] * 12
        lib.ref_init_fold.restype = _d
        lib.ref_init_fold.argtypes = [_d] * 7
        lib.ref_impsurf.restype = _d
        lib.ref_impsurf.argtypes = [_d] * 3
        lib.ref_initialiser.argtypes = [_dp]
        lib.ref_plane.argtypes = [_i, _dp, _i]
        lib.ref_direction.argtypes = [_i, _dp]
        lib.ref_sweeps.argtypes = [_dp, _i]
        assert lib.ref_n() == n
        self.lib, self.n = lib, n

    def edge(self, *args):
        return self.lib.ref_edge(*map(float, args))

    def surf(self, *args):
        return self.lib.ref_surf(*map(float, args))

    def fold_point(self, *args):
        return self.lib.ref_init_fold(*map(float, args))

    def initialiser(self) -> np.ndarray:
        T = np.empty((self.n + 2) ** 3, dtype=np.float64)
        self.lib.ref_initialiser(_ptr(T))
        return T

    def plane(self, d: int, T: np.ndarray, p: int):
        self.lib.ref_plane(d, _ptr(T), p)

    def direction(self, d: int, T: np.ndarray):
        self.lib.ref_direction(d, _ptr(T))

    def sweeps(self, T: np.ndarray, iterations: int):
        self.lib.ref_sweeps(_ptr(T), iterations)


def ref_available(n: int) -> bool:
    return os.path.exists(os.path.join(REF_DIR, f"libfoldref_{n}.so"))
