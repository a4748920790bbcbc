"""ctypes binding of the engine's C ABI (include/lc2e.h).

The shared library is built in-tree by ``lc2registration_amd.build`` into
``lc2registration_amd/_lib/liblc2e.so``. There is no fallback: if the library is
missing or the device cannot be opened, every call raises.

Arrays follow numpy (Z, Y, X) order (x fastest); geometry tuples are (x, y, z).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "liblc2e.so")
_lib = None

OK, OK_DEGENERATE = 0, 1
ERR = {-1: "INVALID", -2: "HIP", -3: "BUDGET", -4: "UNSUPPORTED", -5: "SINGULAR"}
U8, I16, F16, F32, STORE_AUTO = 0, 1, 2, 3, 100
STAGE_WARP_I, STAGE_WARP_US, STAGE_GRAD, STAGE_W, STAGE_MEASURE = range(5)
_DTYPES = {np.dtype(np.uint8): U8, np.dtype(np.int16): I16, np.dtype(np.float16): F16,
           np.dtype(np.float32): F32}


class LC2Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lc2e error {ERR.get(code, code)}: {msg}")
        self.code = code


class Geom(C.Structure):
    _fields_ = [("dims", C.c_int * 3), ("spacing", C.c_float * 3), ("origin", C.c_float * 3)]


class PairInfo(C.Structure):
    _fields_ = [("virt", Geom), ("fixed_center", C.c_float * 3), ("shift", C.c_float * 3),
                ("grid_ndim", C.c_int), ("src_ndim", C.c_int), ("high_res_is_moving", C.c_int),
                ("patch_radius", C.c_int), ("n_frames", C.c_int),
                ("per_frame_matrices", C.c_int)]


def library_path() -> str:
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"{_LIB_PATH} is not built; run python -m lc2registration_amd.build")
        L = C.CDLL(_LIB_PATH)
        vp, fp, dp = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_double)
        L.lc2e_last_error.restype = C.c_char_p
        L.lc2e_create.argtypes = [C.c_int, C.c_size_t, C.POINTER(vp)]
        L.lc2e_destroy.argtypes = [vp]
        L.lc2e_destroy.restype = None
        L.lc2e_memory.argtypes = [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        L.lc2e_volume_upload.argtypes = [vp, vp, C.c_int, C.POINTER(Geom), C.c_int, C.POINTER(vp)]
        L.lc2e_volume_destroy.argtypes = [vp]
        L.lc2e_volume_destroy.restype = None
        L.lc2e_pair_create.argtypes = [vp, vp, vp, fp, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(vp)]
        L.lc2e_pair_destroy.argtypes = [vp]
        L.lc2e_pair_destroy.restype = None
        L.lc2e_pair_get_info.argtypes = [vp, C.POINTER(PairInfo)]
        L.lc2e_pair_set_per_frame.argtypes = [vp, C.c_int]
        L.lc2e_pose_to_matrix.argtypes = [vp, dp, C.c_int, fp]
        L.lc2e_eval.argtypes = [vp, fp, dp]
        L.lc2e_eval_batch.argtypes = [vp, C.c_int, fp, dp, dp, dp]
        L.lc2e_eval_batch_noalloc.argtypes = [vp, C.c_int, fp, dp, dp, dp]
        L.lc2e_eval_batch_async.argtypes = [vp, C.c_int, vp, vp, vp]
        L.lc2e_sync.argtypes = [vp]
        L.lc2e_debug_fetch.argtypes = [vp, C.c_int, fp, vp]
        L.lc2e_eval_mi_batch.argtypes = [vp, C.c_int, fp, C.c_int, C.c_int,
                                         C.POINTER(C.c_double)]
        L.lc2e_pair_set_tuning.argtypes = [vp, C.c_int]
        L.lc2e_pair_set_tuning_n.argtypes = [vp, C.c_int, C.c_int]
        L.lc2e_pair_get_tuning.argtypes = [vp, C.POINTER(C.c_int)]
        L.lc2e_pair_use_graph.argtypes = [vp, C.c_int]
        L.lc2e_pair_autotune.argtypes = [vp, C.c_int, fp, C.c_int, C.POINTER(C.c_int),
                                         C.POINTER(C.c_double)]
        L.lc2e_pair_profile.argtypes = [vp, C.c_int]
        L.lc2e_pair_profile_read.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise LC2Error(rc, lib().lc2e_last_error().decode())
    return rc


def make_geom(dims, spacing, origin) -> Geom:
    d = list(dims) + [1] * (3 - len(dims))
    s = list(spacing) + [1.0] * (3 - len(spacing))
    o = list(origin) + [0.0] * (3 - len(origin))
    return Geom((C.c_int * 3)(*d), (C.c_float * 3)(*s), (C.c_float * 3)(*o))


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Engine:
    """One device, one HIP stream, an HBM budget (bytes, 0 = unlimited)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = C.c_void_p()
        _check(lib().lc2e_create(device, hbm_budget, C.byref(h)))
        self.h = h
        self._children = []

    def memory(self):
        used, budget = C.c_size_t(), C.c_size_t()
        _check(lib().lc2e_memory(self.h, C.byref(used), C.byref(budget)))
        return used.value, budget.value

    def upload(self, data: np.ndarray, spacing, origin, storage: int = STORE_AUTO) -> "Volume":
        data = np.ascontiguousarray(data)
        if data.dtype not in _DTYPES:
            raise TypeError(f"unsupported dtype {data.dtype}")
        dims = tuple(reversed(data.shape)) + (1,) * (3 - data.ndim)
        g = make_geom(dims, spacing, origin)
        h = C.c_void_p()
        _check(lib().lc2e_volume_upload(self.h, data.ctypes.data_as(C.c_void_p),
                                        _DTYPES[data.dtype], C.byref(g), storage, C.byref(h)))
        return Volume(self, h, g)

    def pair(self, fixed: "Volume", moving: "Volume", resampled_spacing, high_res_is_moving: bool,
             center_overlaid: bool, patch_radius: int, dim: int) -> "Pair":
        rs = np.zeros(3, np.float32)
        rs[: len(resampled_spacing)] = resampled_spacing
        h = C.c_void_p()
        _check(lib().lc2e_pair_create(self.h, fixed.h, moving.h, _fptr(rs), int(high_res_is_moving),
                                      int(center_overlaid), patch_radius, dim, C.byref(h)))
        return Pair(self, h, fixed, moving)

    def sync(self):
        _check(lib().lc2e_sync(self.h))

    def close(self):
        if self.h:
            lib().lc2e_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Volume:
    def __init__(self, engine: Engine, h, geom: Geom):
        self.engine, self.h, self.geom = engine, h, geom

    def close(self):
        if self.h:
            lib().lc2e_volume_destroy(self.h)
            self.h = None


@dataclass
class BatchResult:
    lc2: np.ndarray
    sum_measure: np.ndarray
    sum_w: np.ndarray
    status: int


class Pair:
    def __init__(self, engine: Engine, h, fixed: Volume, moving: Volume):
        self.engine, self.h, self.fixed, self.moving = engine, h, fixed, moving
        info = PairInfo()
        _check(lib().lc2e_pair_get_info(h, C.byref(info)))
        self.info = info

    @property
    def virt_dims(self):
        return tuple(self.info.virt.dims)

    @property
    def virt_shape(self):
        d = self.virt_dims
        return (d[2], d[1], d[0])

    def pose_to_matrix(self, params) -> np.ndarray:
        p = np.ascontiguousarray(params, np.float64)
        out = np.zeros(12, np.float32)
        _check(lib().lc2e_pose_to_matrix(self.h, _dptr(p), p.size, _fptr(out)))
        return out

    def set_per_frame(self, on: bool):
        _check(lib().lc2e_pair_set_per_frame(self.h, int(on)))
        _check(lib().lc2e_pair_get_info(self.h, C.byref(self.info)))

    def eval(self, m12) -> float:
        m = np.ascontiguousarray(m12, np.float32).ravel()
        out = C.c_double()
        _check(lib().lc2e_eval(self.h, _fptr(m), C.byref(out)))
        return out.value

    def eval_batch(self, mats) -> BatchResult:
        m = np.ascontiguousarray(mats, np.float32)
        per = 12 * (self.info.n_frames if self.info.per_frame_matrices else 1)
        n = m.size // per
        if n * per != m.size:
            raise ValueError("matrix buffer size is not a multiple of the per-pose size")
        l, sm, sw = np.zeros(n), np.zeros(n), np.zeros(n)
        rc = _check(lib().lc2e_eval_batch(self.h, n, _fptr(m.ravel()), _dptr(l), _dptr(sm),
                                          _dptr(sw)))
        return BatchResult(l, sm, sw, rc)

    def eval_batch_noalloc(self, mats) -> BatchResult:
        """eval_batch through the allocation-free entry point (device scratch array)."""
        m = np.ascontiguousarray(mats, np.float32)
        per = 12 * (self.info.n_frames if self.info.per_frame_matrices else 1)
        n = m.size // per
        if n * per != m.size:
            raise ValueError("matrix buffer size is not a multiple of the per-pose size")
        l, sm, sw = np.zeros(n), np.zeros(n), np.zeros(n)
        rc = _check(lib().lc2e_eval_batch_noalloc(self.h, n, _fptr(m.ravel()), _dptr(l),
                                                  _dptr(sm), _dptr(sw)))
        return BatchResult(l, sm, sw, rc)

    def debug_fetch(self, stage: int, m12) -> np.ndarray:
        m = np.ascontiguousarray(m12, np.float32).ravel()
        dt = np.float64 if stage in (STAGE_W, STAGE_MEASURE) else np.float32
        out = np.zeros(self.virt_shape, dt)
        _check(lib().lc2e_debug_fetch(self.h, stage, _fptr(m), out.ctypes.data_as(C.c_void_p)))
        return out

    def eval_mi(self, mats, bins_us: int = 256, bins_hr: int = 256) -> np.ndarray:
        """Normalized mutual information (H_US + H_HR) / H_joint for each pose."""
        m = np.ascontiguousarray(mats, np.float32)
        per = 12 * (self.info.n_frames if self.info.per_frame_matrices else 1)
        n = m.size // per
        out = np.zeros(n, np.float64)
        _check(lib().lc2e_eval_mi_batch(self.h, n, _fptr(m.ravel()), bins_us, bins_hr,
                                        out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def set_tuning(self, zchunk: int):
        """Force the 3D z-chunk for every launch (0: automatic)."""
        _check(lib().lc2e_pair_set_tuning(self.h, zchunk))

    def set_tuning_n(self, n: int, zchunk: int):
        """z-chunk for launches of exactly n poses (0: automatic)."""
        _check(lib().lc2e_pair_set_tuning_n(self.h, n, zchunk))

    def get_tuning(self) -> int:
        z = C.c_int()
        _check(lib().lc2e_pair_get_tuning(self.h, C.byref(z)))
        return z.value

    def profile(self, on: bool):
        """HIP-event timing of every fused-kernel launch (resets the counters)."""
        _check(lib().lc2e_pair_profile(self.h, int(on)))

    def profile_read(self):
        ms, n = C.c_double(), C.c_longlong()
        _check(lib().lc2e_pair_profile_read(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def autotune(self, mats, iters: int = 3):
        """Pick the fastest z-chunk for batches shaped like `mats`; returns (zchunk, ms)."""
        m = np.ascontiguousarray(mats, np.float32)
        per = 12 * (self.info.n_frames if self.info.per_frame_matrices else 1)
        z, ms = C.c_int(), C.c_double()
        _check(lib().lc2e_pair_autotune(self.h, m.size // per, _fptr(m.ravel()), iters,
                                        C.byref(z), C.byref(ms)))
        return z.value, ms.value

    def use_graph(self, on: bool):
        _check(lib().lc2e_pair_use_graph(self.h, int(on)))

    def close(self):
        if self.h:
            lib().lc2e_pair_destroy(self.h)
            self.h = None
