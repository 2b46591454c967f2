"""Python view of libfold3d.so (the MI355X Fold3dPMM solver) through its C-ABI.

This is a thin ctypes layer over include/fold3d.h and include/Fold3dPMM.h;
it performs no computation of its own.  Device buffers are torch tensors on
``cuda`` (torch supplies memory and streams only); host buffers are numpy
arrays.  Loading fails loudly when the shared library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libfold3d.so")
REPO_DIR = os.path.dirname(PKG_DIR)
INCLUDE_DIR = os.path.join(REPO_DIR, "include")

FOLD3D_OK = 0
FOLD3D_E_INVALID = -1
FOLD3D_E_HIP = -2
FOLD3D_E_NOMEM = -3
FOLD3D_E_UNSUPPORTED = -4
FOLD3D_E_NOT_CONVERGED = -5
FOLD3D_E_IO = -6

F64, F32 = 0, 1
DEVICE, HOST = 0, 1
MODEL_FOLD, MODEL_ELLIP = 0, 1
ENGINE_AUTO, ENGINE_PLANE, ENGINE_HOISTED, ENGINE_TILED, ENGINE_SPLIT = 0, 1, 2, 3, 4
ENGINE_LOCKED = 5

MAXT = 10e15 * 1.0
REF_F = 1.1                      # reference Fold3dPMM.h:52
REF_A = (-0.35, 0.4, 0.7)        # reference Fold3dPMM.h:49-51
DIRECTIONS = ("x_down", "x_up", "k_up", "k_down", "y_up", "y_down")


class Fold3dError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"fold3d error {code}: {msg}")
        self.code = code


class Model(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("F", C.c_double),
        ("a", C.c_double * 3),
        ("M", C.c_double * 6),
        ("speed", C.c_double),
        ("speed_field", C.c_void_p),
        ("theta", C.c_void_p),
        ("phi", C.c_void_p),
        ("ra", C.c_void_p),
        ("rb", C.c_void_p),
    ]


class Opts(C.Structure):
    _fields_ = [
        ("precision", C.c_int),
        ("memory", C.c_int),
        ("max_iters", C.c_int),
        ("tol", C.c_double),
        ("stream", C.c_void_p),
        ("engine", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int),
        ("max_dT", C.c_double),
        ("cell_updates", C.c_double),
        ("sweep_ms", C.c_double),
        ("updates_computed", C.c_double),
        ("bytes_moved", C.c_double),
        ("nodes_held", C.c_double),
    ]


class AdjointSource(C.Structure):
    _fields_ = [
        ("T", C.c_void_p),
        ("nrec", C.c_int),
        ("rec_idx", C.c_void_p),
        ("rec_res", C.c_void_p),
        ("tol", C.c_double),
    ]


class Instance(C.Structure):
    _fields_ = [
        ("T", C.c_void_p),
        ("n", C.c_int * 3),
        ("h", C.c_double * 3),
        ("M", C.c_double * 6),
        ("speed", C.c_double),
        ("speed_field", C.c_void_p),
        ("theta", C.c_void_p),
        ("phi", C.c_void_p),
        ("ra", C.c_void_p),
        ("rb", C.c_void_p),
    ]


_lib = None


def lib() -> C.CDLL:
    """Load libfold3d.so once; raise if it is missing (no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C csrc)")
    L = C.CDLL(LIB_PATH)
    d, i, vp = C.c_double, C.c_int, C.c_void_p
    dp = C.POINTER(C.c_double)
    L.fold3d_version.restype = C.c_char_p
    L.fold3d_last_error.restype = C.c_char_p
    L.fold3d_create.argtypes = [C.POINTER(vp), i]
    L.fold3d_destroy.argtypes = [vp]
    L.fold3d_default_opts.argtypes = [C.POINTER(Opts)]
    L.fold3d_reference_model.argtypes = [C.POINTER(Model)]
    L.fold3d_solve.argtypes = [vp, vp, i, i, i, d, d, d, C.POINTER(Model),
                               C.POINTER(Opts), C.POINTER(Stats)]
    L.fold3d_plane_step.argtypes = [vp, i, vp, i, i, i, i] + [d] * 10 + [vp]
    L.fold3d_solve_noalloc.argtypes = [vp, i, i, i, d, d, d, C.POINTER(Model),
                                       C.POINTER(Opts), C.POINTER(Stats)]
    L.fold3d_measure_peaks.argtypes = [vp, dp, dp, dp]
    L.fold3d_solve_batched.argtypes = [vp, i, C.POINTER(Instance), C.POINTER(Opts),
                                       C.POINTER(i), dp, dp]
    L.fold3d_set_tuning.argtypes = [vp, i, i]
    L.fold3d_set_lock_start.argtypes = [vp, i]
    L.fold3d_get_lock_start.argtypes = [vp, C.POINTER(i)]
    L.fold3d_get_tuning.argtypes = [vp, i, C.POINTER(i)]
    L.fold3d_set_batch_tuning.argtypes = [vp, i, i, i]
    L.fold3d_set_batch_streams.argtypes = [vp, i]
    L.fold3d_get_batch_streams.argtypes = [vp, C.POINTER(i)]
    L.fold3d_get_batch_tuning.argtypes = [vp, C.POINTER(i), C.POINTER(i), C.POINTER(i)]
    L.fold3d_autotune_batched.argtypes = [vp, i, C.POINTER(Instance), C.POINTER(Opts), i,
                                          C.POINTER(C.c_float)]
    L.fold3d_autotune.argtypes = [vp, vp, i, i, i, d, d, d, C.POINTER(Model),
                                  C.POINTER(Opts), i, C.POINTER(C.c_float)]
    L.fold3d_adjoint.argtypes = [vp, vp, i, i, i, d, d, d, C.POINTER(Model), i, vp, vp,
                                 C.POINTER(Opts), vp, vp, C.POINTER(Stats)]
    L.fold3d_adjoint_batched.argtypes = [vp, i, C.POINTER(AdjointSource), i, i, i, d, d, d,
                                         C.POINTER(Model), C.POINTER(Opts), vp, vp,
                                         C.POINTER(C.c_int), C.POINTER(C.c_double),
                                         C.POINTER(C.c_longlong), C.POINTER(C.c_double)]
    L.fold3d_synchronize.argtypes = [vp]
    L.fold3d_point_source.argtypes = [vp, vp, i, i, i, d, d, d, dp, dp, d, d, i, vp]
    L.fold3d_initialiser_device.argtypes = [vp, vp, i, i, i, d, d, d, vp, C.POINTER(C.c_longlong)]
    cp = C.c_char_p
    L.fold3d_write_text.argtypes = [cp, dp, i, i, i]
    L.fold3d_read_text.argtypes = [cp, dp, i, i, i]
    L.fold3d_write_vtk.argtypes = [cp, dp, i, i, i, d, d, d, i]
    L.ImplicitInitialiser.argtypes = [dp, i, i, i, d, d, d]
    L.InitFoldFromPoint.restype = d
    L.InitFoldFromPoint.argtypes = [d] * 7
    L.ImpSurfF.restype = d
    L.ImpSurfF.argtypes = [d] * 3
    L.fold3d_set_grid.argtypes = [i, i, i]
    L.fold3d_get_grid.argtypes = [C.POINTER(i)] * 3
    L.InitNode.argtypes = [d, d, d, i, i, i, dp]
    L.InitNode.restype = None
    L.fold3d_init_node.argtypes = [d, d, d, i, i, i, dp, i, i, i]
    L.fold3d_init_node.restype = None
    L.updateVoxel.argtypes = [dp, i, i, i] + [d] * 6
    L.updateVoxel.restype = None
    L.fold3d_update_voxel.argtypes = [dp, i, i, i] + [d] * 6 + [i, i, i]
    L.fold3d_update_voxel.restype = None
    for name in DIRECTIONS:
        fn = getattr(L, "function_" + name)
        fn.argtypes = [vp, i] + [d] * 10
        fn.restype = None
        getattr(L, "fold3d_function_" + name).argtypes = L.fold3d_plane_step.argtypes[:1] + \
            L.fold3d_plane_step.argtypes[2:]
    _lib = L
    return L


def last_error() -> str:
    return lib().fold3d_last_error().decode()


def check(rc: int) -> int:
    if rc != FOLD3D_OK:
        raise Fold3dError(rc, last_error())
    return rc


def reference_spacing(n: int) -> float:
    """Node spacing of the reference's [0,10] domain (AN_ISO_1_OpenACC.c:1154)."""
    return (10.0 - 0.0) / (n - 1.0)


def reference_model() -> Model:
    m = Model()
    lib().fold3d_reference_model(C.byref(m))
    return m


def fold_model(F: float = REF_F, a=REF_A) -> Model:
    m = reference_model()
    m.F = F
    for k in range(3):
        m.a[k] = a[k]
    return m


def measure_peaks(stream=None) -> dict:
    """Sustained FP64/FP32 FMA rate and device copy bandwidth of this GPU."""
    f64, f32, cp = C.c_double(), C.c_double(), C.c_double()
    check(lib().fold3d_measure_peaks(stream, C.byref(f64), C.byref(f32), C.byref(cp)))
    return {"fp64_tflops": f64.value / 1e12, "fp32_tflops": f32.value / 1e12,
            "copy_gbs": cp.value / 1e9}


def _dims(n):
    return (n, n, n) if np.isscalar(n) else tuple(n)


def write_text(path: str, T: np.ndarray, n) -> None:
    """The reference's "%f" solution dump of a host FP64 field."""
    nx, ny, nz = _dims(n)
    T = np.ascontiguousarray(T, dtype=np.float64)
    check(lib().fold3d_write_text(os.fsencode(path), T.ctypes.data_as(C.POINTER(C.c_double)),
                                  nx, ny, nz))


def read_text(path: str, n) -> np.ndarray:
    nx, ny, nz = _dims(n)
    T = np.empty((nx + 2) * (ny + 2) * (nz + 2))
    check(lib().fold3d_read_text(os.fsencode(path), T.ctypes.data_as(C.POINTER(C.c_double)),
                                 nx, ny, nz))
    return T


def write_vtk(path: str, T: np.ndarray, n, spacing, binary: bool = True) -> None:
    nx, ny, nz = _dims(n)
    dx, dy, dz = (spacing,) * 3 if np.isscalar(spacing) else spacing
    T = np.ascontiguousarray(T, dtype=np.float64)
    check(lib().fold3d_write_vtk(os.fsencode(path), T.ctypes.data_as(C.POINTER(C.c_double)),
                                 nx, ny, nz, dx, dy, dz, int(binary)))


def implicit_initialiser(n: int) -> np.ndarray:
    """The reference's input field on an n^3 grid (host C, Fold3dPMM.c:129)."""
    h = reference_spacing(n)
    T = np.empty((n + 2) ** 3, dtype=np.float64)
    lib().ImplicitInitialiser(T.ctypes.data_as(C.POINTER(C.c_double)), n, n, n, h, h, h)
    return T


def padded_count(nx: int, ny: int, nz: int) -> int:
    """Nodes of the padded grid, ghosts included: (nx+2)(ny+2)(nz+2)."""
    return (nx + 2) * (ny + 2) * (nz + 2)


def _check_tensor(x, count, what, device, dtype=None):
    """A torch tensor the C-ABI may read and write as `count` contiguous FP64
    or FP32 values on GPU `device` (of `dtype` when given).  The library sees
    only a raw pointer, so a wrong size, layout, dtype or device is rejected
    here, before any launch."""
    import torch
    if not isinstance(x, torch.Tensor):
        raise TypeError(f"{what} must be a torch tensor, not {type(x).__name__}")
    if x.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"{what} must be float32 or float64, not {x.dtype}")
    if dtype is not None and x.dtype != dtype:
        raise TypeError(f"{what} is {x.dtype}; it must be {dtype}, like T")
    if not x.is_contiguous():
        raise ValueError(f"{what} must be contiguous")
    if x.numel() != count:
        raise ValueError(f"{what} has {x.numel()} elements; the padded grid has {count}")
    if not x.is_cuda or (device is not None and x.device.index != device):
        raise TypeError(f"{what} must be on cuda:{device}, not {x.device}")


def _check_host_array(x, count, what):
    if x.dtype not in (np.float64, np.float32):
        raise TypeError(f"{what} must be float64 or float32, not {x.dtype}")
    if not x.flags.c_contiguous or not x.flags.writeable:
        raise ValueError(f"{what} must be a writeable C-contiguous array")
    if x.size != count:
        raise ValueError(f"{what} has {x.size} elements; the padded grid has {count}")


def _check_model_fields(m, count, device, dtype):
    """Per-node fields a Model was built from (Workload.model() and the
    tomography harness keep them in m._refs): same dtype, device and size as T."""
    for name, f in zip(("speed_field", "theta", "phi", "ra", "rb"), getattr(m, "_refs", ())):
        if f is not None:
            _check_tensor(f, count, f"model {name}", device, dtype)


def _check_receivers(rec_idx, rec_res, device, dtype, what=""):
    import torch
    if not isinstance(rec_idx, torch.Tensor) or rec_idx.dtype != torch.int32:
        raise TypeError(f"rec_idx{what} must be an int32 torch tensor")
    if not isinstance(rec_res, torch.Tensor) or rec_res.dtype != dtype:
        raise TypeError(f"rec_res{what} must be a torch tensor of T's dtype {dtype}")
    for name, x in (("rec_idx", rec_idx), ("rec_res", rec_res)):
        if not x.is_cuda or x.device.index != device:
            raise TypeError(f"{name}{what} must be on cuda:{device}")
        if not x.is_contiguous():
            raise ValueError(f"{name}{what} must be contiguous")
    if rec_res.numel() != rec_idx.numel():
        raise ValueError(f"rec_res{what} has {rec_res.numel()} values for {rec_idx.numel()} receivers")


class Solver:
    """One fold3d handle bound to a device."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().fold3d_create(C.byref(h), device))
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            lib().fold3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, T, n, spacing=None, model: Model | None = None,
              max_iters: int = 8, tol: float = 0.0, stream=None,
              engine: int = ENGINE_AUTO, raise_not_converged: bool = True) -> Stats:
        """In-place solve of T (torch cuda tensor or numpy host array)."""
        nx, ny, nz = (n, n, n) if np.isscalar(n) else n
        if spacing is None:
            spacing = reference_spacing(nx)
        dx, dy, dz = (spacing,) * 3 if np.isscalar(spacing) else spacing
        o = Opts()
        lib().fold3d_default_opts(C.byref(o))
        o.max_iters, o.tol, o.engine = max_iters, tol, engine
        count = padded_count(nx, ny, nz)
        m = model if model is not None else reference_model()
        if isinstance(T, np.ndarray):
            _check_host_array(T, count, "T")
            if m.kind == MODEL_ELLIP and any(f is not None for f in getattr(m, "_refs", ())):
                raise TypeError("per-node model fields need a device-resident T")
            o.memory = HOST
            o.precision = F64 if T.dtype == np.float64 else F32
            ptr = T.ctypes.data
        else:
            import torch
            _check_tensor(T, count, "T", self.device)
            _check_model_fields(m, count, self.device, T.dtype)
            o.memory = DEVICE
            o.precision = F64 if T.dtype == torch.float64 else F32
            ptr = T.data_ptr()
            if stream is None:
                stream = torch.cuda.current_stream(T.device).cuda_stream
        o.stream = stream
        st = Stats()
        rc = lib().fold3d_solve(self.h, C.c_void_p(ptr), nx, ny, nz, dx, dy, dz,
                                C.byref(m), C.byref(o), C.byref(st))
        if rc == FOLD3D_E_NOT_CONVERGED and not raise_not_converged:
            return st
        check(rc)
        return st

    def _instances(self, workloads, Ts):
        count = len(workloads)
        if count < 1 or len(Ts) != count:
            raise ValueError(f"{count} workloads for {len(Ts)} fields")
        arr = (Instance * count)()
        for k, (W, T) in enumerate(zip(workloads, Ts)):
            x = arr[k]
            nn = (W.n,) * 3 if np.isscalar(W.n) else W.n
            cnt = padded_count(*map(int, nn))
            _check_tensor(T, cnt, f"T of instance {k}", self.device, Ts[0].dtype)
            for name in ("speed_field", "theta", "phi", "ra", "rb"):
                f = getattr(W, name)
                if f is not None:
                    _check_tensor(f, cnt, f"{name} of instance {k}", self.device, T.dtype)
            x.T = T.data_ptr()
            hh = (W.h,) * 3 if np.isscalar(W.h) else W.h
            for a in range(3):
                x.n[a] = int(nn[a])
                x.h[a] = float(hh[a])
            for a in range(6):
                x.M[a] = float(W.M[a])
            x.speed = float(W.speed)
            for name in ("speed_field", "theta", "phi", "ra", "rb"):
                f = getattr(W, name)
                setattr(x, name, None if f is None else f.data_ptr())
        return arr

    def _batch_opts(self, Ts, max_iters, tol, use_graph, stream):
        import torch
        o = Opts()
        lib().fold3d_default_opts(C.byref(o))
        o.precision = F64 if Ts[0].dtype == torch.float64 else F32
        o.max_iters, o.tol = max_iters, tol
        o.engine = ENGINE_AUTO if use_graph else ENGINE_PLANE
        o.stream = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        return o

    def solve_batched(self, workloads, Ts, max_iters=200, tol=1e-6, use_graph=True,
                      stream=None, raise_not_converged=True):
        """Solve many elliptic-model instances (fold3d.workloads.Workload with
        torch fields on the device, and their torch T tensors) in one batch.
        Returns (iterations per instance, max|dT| per instance, device ms)."""
        count = len(workloads)
        arr = self._instances(workloads, Ts)
        o = self._batch_opts(Ts, max_iters, tol, use_graph, stream)
        its = (C.c_int * count)()
        mds = (C.c_double * count)()
        ms = C.c_double()
        rc = lib().fold3d_solve_batched(self.h, count, arr, C.byref(o), its, mds, C.byref(ms))
        if not (rc == FOLD3D_E_NOT_CONVERGED and not raise_not_converged):
            check(rc)
        return list(its), list(mds), ms.value

    def autotune_batched(self, workloads, Ts, iterations=1, stream=None):
        """Time the batched solve's schedules on copies of the fields; keep the
        fastest.  Returns ((groups, blocks per CU, rows per pull), device ms)."""
        arr = self._instances(workloads, Ts)
        o = self._batch_opts(Ts, 1, 0.0, True, stream)
        ms = C.c_float()
        check(lib().fold3d_autotune_batched(self.h, len(workloads), arr, C.byref(o),
                                            iterations, C.byref(ms)))
        return self.batch_tuning(), ms.value

    def adjoint(self, T, n, spacing, model: Model, rec_idx, rec_res, grad_slowness,
                grad_theta=None, max_iters: int = 200, tol: float = 0.0, stream=None,
                raise_not_converged: bool = True) -> Stats:
        """Adds dJ/d(slowness) (and dJ/d(theta) with per-node tilt fields) of
        J = 1/2 sum (T[rec_idx] - T_obs)^2 to the gradient tensors, given the
        converged field T and the residuals rec_res = T[rec_idx] - T_obs (torch
        device tensors; rec_idx int32 linear indices on the padded layout)."""
        import torch
        nx, ny, nz = (n, n, n) if np.isscalar(n) else n
        dx, dy, dz = (spacing,) * 3 if np.isscalar(spacing) else spacing
        count = padded_count(nx, ny, nz)
        _check_tensor(T, count, "T", self.device)
        _check_model_fields(model, count, self.device, T.dtype)
        _check_receivers(rec_idx, rec_res, self.device, T.dtype)
        _check_tensor(grad_slowness, count, "grad_slowness", self.device, T.dtype)
        if grad_theta is not None:
            _check_tensor(grad_theta, count, "grad_theta", self.device, T.dtype)
        o = Opts()
        lib().fold3d_default_opts(C.byref(o))
        o.precision = F64 if T.dtype == torch.float64 else F32
        o.max_iters, o.tol = max_iters, tol
        o.stream = stream if stream is not None else torch.cuda.current_stream(T.device).cuda_stream
        st = Stats()
        rc = lib().fold3d_adjoint(
            self.h, C.c_void_p(T.data_ptr()), nx, ny, nz, dx, dy, dz, C.byref(model),
            int(rec_idx.numel()), C.c_void_p(rec_idx.data_ptr()), C.c_void_p(rec_res.data_ptr()),
            C.byref(o), C.c_void_p(grad_slowness.data_ptr()),
            None if grad_theta is None else C.c_void_p(grad_theta.data_ptr()), C.byref(st))
        if not (rc == FOLD3D_E_NOT_CONVERGED and not raise_not_converged):
            check(rc)
        return st

    def adjoint_batched(self, Ts, n, spacing, model: Model, rec_idx, rec_res, tols,
                        grad_slowness, grad_theta=None, max_iters: int = 200, stream=None,
                        raise_not_converged: bool = True):
        """adjoint() for several sources on one grid and model at once: Ts, rec_idx,
        rec_res and tols are per-source lists (rec_idx may be one tensor shared by
        all).  Gradients of all sources are added in list order.  Returns
        (sweeps per source, last max change per source, device ms); the reached
        nodes held fixed per source are left in self.last_held."""
        import torch
        count = len(Ts)
        nx, ny, nz = (n, n, n) if np.isscalar(n) else n
        dx, dy, dz = (spacing,) * 3 if np.isscalar(spacing) else spacing
        idx = rec_idx if isinstance(rec_idx, (list, tuple)) else [rec_idx] * count
        if count < 1 or len(idx) != count or len(rec_res) != count or len(tols) != count:
            raise ValueError("Ts, rec_idx, rec_res and tols must have one entry per source")
        cnt = padded_count(nx, ny, nz)
        _check_model_fields(model, cnt, self.device, Ts[0].dtype)
        _check_tensor(grad_slowness, cnt, "grad_slowness", self.device, Ts[0].dtype)
        if grad_theta is not None:
            _check_tensor(grad_theta, cnt, "grad_theta", self.device, Ts[0].dtype)
        arr = (AdjointSource * count)()
        for k in range(count):
            _check_tensor(Ts[k], cnt, f"T of source {k}", self.device, Ts[0].dtype)
            _check_receivers(idx[k], rec_res[k], self.device, Ts[0].dtype, f" of source {k}")
            arr[k].T = Ts[k].data_ptr()
            arr[k].nrec = int(idx[k].numel())
            arr[k].rec_idx = idx[k].data_ptr()
            arr[k].rec_res = rec_res[k].data_ptr()
            arr[k].tol = float(tols[k])
        o = Opts()
        lib().fold3d_default_opts(C.byref(o))
        o.precision = F64 if Ts[0].dtype == torch.float64 else F32
        o.max_iters = max_iters
        o.stream = stream if stream is not None else torch.cuda.current_stream(Ts[0].device).cuda_stream
        its = (C.c_int * count)()
        mds = (C.c_double * count)()
        held = (C.c_longlong * count)()
        ms = C.c_double()
        rc = lib().fold3d_adjoint_batched(
            self.h, count, arr, nx, ny, nz, dx, dy, dz, C.byref(model), C.byref(o),
            C.c_void_p(grad_slowness.data_ptr()),
            None if grad_theta is None else C.c_void_p(grad_theta.data_ptr()), its, mds,
            held, C.byref(ms))
        if not (rc == FOLD3D_E_NOT_CONVERGED and not raise_not_converged):
            check(rc)
        self.last_held = list(held)
        return list(its), list(mds), ms.value

    def set_lock_start(self, iteration: int = -1):
        """First iteration the locked engine runs locked (-1: the default, 8)."""
        check(lib().fold3d_set_lock_start(self.h, iteration))

    def lock_start(self) -> int:
        k = C.c_int()
        check(lib().fold3d_get_lock_start(self.h, C.byref(k)))
        return k.value

    def set_tuning(self, groups: int, kind: int = MODEL_FOLD):
        check(lib().fold3d_set_tuning(self.h, kind, groups))

    def point_source(self, T, n, spacing, source, M, speed, radius=1.6, stream=None):
        """Closed-form start of an elliptic solve around a point source, on the
        device into the torch tensor T (its dtype sets the precision)."""
        import torch
        nx, ny, nz = _dims(n)
        _check_tensor(T, padded_count(nx, ny, nz), "T", self.device)
        dx, dy, dz = (spacing,) * 3 if np.isscalar(spacing) else spacing
        src = (C.c_double * 3)(*[float(v) for v in source])
        Mv = (C.c_double * 6)(*[float(v) for v in M])
        if stream is None:
            stream = torch.cuda.current_stream(T.device).cuda_stream
        prec = F64 if T.dtype == torch.float64 else F32
        check(lib().fold3d_point_source(self.h, C.c_void_p(T.data_ptr()), nx, ny, nz, dx, dy,
                                        dz, src, Mv, float(speed), float(radius), prec, stream))

    def initialiser_device(self, T, n, spacing=None, stream=None) -> int:
        """The reference's input field built on the device into the FP64 torch
        tensor T (same bits as implicit_initialiser); returns the number of
        nodes whose surface function the host evaluated."""
        import torch
        nx, ny, nz = _dims(n)
        if spacing is None:
            spacing = reference_spacing(nx)
        dx, dy, dz = (spacing,) * 3 if np.isscalar(spacing) else spacing
        _check_tensor(T, padded_count(nx, ny, nz), "T", self.device, torch.float64)
        if stream is None:
            stream = torch.cuda.current_stream(T.device).cuda_stream
        ne = C.c_longlong()
        check(lib().fold3d_initialiser_device(self.h, C.c_void_p(T.data_ptr()), nx, ny, nz,
                                              dx, dy, dz, stream, C.byref(ne)))
        return ne.value

    def set_batch_tuning(self, groups: int = 0, blocks_per_cu: int = 0, rows_per_pull: int = 0):
        check(lib().fold3d_set_batch_tuning(self.h, groups, blocks_per_cu, rows_per_pull))

    def set_batch_streams(self, streams: int = 0):
        """Sub-batches the batched solve runs at once (0: the default, 3)."""
        check(lib().fold3d_set_batch_streams(self.h, streams))

    def batch_streams(self) -> int:
        k = C.c_int()
        check(lib().fold3d_get_batch_streams(self.h, C.byref(k)))
        return k.value

    def batch_tuning(self):
        g, b, r = C.c_int(), C.c_int(), C.c_int()
        check(lib().fold3d_get_batch_tuning(self.h, C.byref(g), C.byref(b), C.byref(r)))
        return g.value, b.value, r.value

    def tuning(self, kind: int = MODEL_FOLD) -> int:
        g = C.c_int()
        check(lib().fold3d_get_tuning(self.h, kind, C.byref(g)))
        return g.value

    def autotune(self, T, n, spacing=None, model: Model | None = None,
                 iterations: int = 1, stream=None):
        """Time 2/4/8 waves per node row on a copy of T; keep the fastest.
        Returns (groups, sweep ms of the winner)."""
        import torch
        nx, ny, nz = (n, n, n) if np.isscalar(n) else n
        if spacing is None:
            spacing = reference_spacing(nx)
        dx, dy, dz = (spacing,) * 3 if np.isscalar(spacing) else spacing
        m = model if model is not None else reference_model()
        _check_tensor(T, padded_count(nx, ny, nz), "T", self.device)
        _check_model_fields(m, padded_count(nx, ny, nz), self.device, T.dtype)
        o = Opts()
        lib().fold3d_default_opts(C.byref(o))
        o.precision = F64 if T.dtype == torch.float64 else F32
        o.stream = stream if stream is not None else torch.cuda.current_stream(T.device).cuda_stream
        ms = C.c_float()
        check(lib().fold3d_autotune(self.h, C.c_void_p(T.data_ptr()), nx, ny, nz, dx, dy, dz,
                                    C.byref(m), C.byref(o), iterations, C.byref(ms)))
        return self.tuning(m.kind), ms.value

    def plane_step(self, direction: int, T, n, plane: int, dzz, dxx, dyy, dxy,
                   dxz, dyz, F, ay, ax, az, stream=None):
        import torch
        nx, ny, nz = (n, n, n) if np.isscalar(n) else n
        _check_tensor(T, padded_count(nx, ny, nz), "T", self.device, torch.float64)
        if stream is None:
            stream = torch.cuda.current_stream(T.device).cuda_stream
        check(lib().fold3d_plane_step(self.h, direction, C.c_void_p(T.data_ptr()),
                                      nx, ny, nz, plane, dzz, dxx, dyy, dxy, dxz,
                                      dyz, F, ay, ax, az, stream))


def header_symbols() -> list[str]:
    """Function names declared in include/*.h (for export checks)."""
    import re
    names = []
    for fn in ("fold3d.h", "Fold3dPMM.h"):
        src = open(os.path.join(INCLUDE_DIR, fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", src)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", src):
            name = m.group(1)
            if name in ("if", "sizeof", "defined", "return"):
                continue
            # a declaration: preceded by a type on the same statement
            names.append(name)
    # drop macro-ish and type names
    skip = {"_DOUBLE_", "sign", "fmin"}
    out = []
    for nm in names:
        if nm in skip or nm in out:
            continue
        out.append(nm)
    return out
