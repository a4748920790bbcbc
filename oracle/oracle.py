"""ctypes front end of the LC2 oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module; the engine (``lc2registration_amd``) never does.
The C restatement lives in ``oracle/lc2_oracle.c`` and each function there cites
the reference file:line it follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liblc2oracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


class Geom(C.Structure):
    _fields_ = [("dims", C.c_int * 3), ("spacing", C.c_float * 3), ("origin", C.c_float * 3)]


class Result(C.Structure):
    _fields_ = [("sum_measure", C.c_double), ("sum_w", C.c_double), ("lc2", C.c_double),
                ("n_valid", C.c_longlong)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        fp = C.POINTER(C.c_float)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        gp = C.POINTER(Geom)
        L.lc2o_center.argtypes = [gp, C.c_int, fp]
        L.lc2o_virtual_dims.argtypes = [gp, fp, C.c_int, ip]
        L.lc2o_affine_init_3d.argtypes = [fp, fp, fp, fp, fp, fp, fp]
        L.lc2o_affine_init_2d.argtypes = [C.c_float, fp, fp, C.c_float, fp, fp, fp]
        L.lc2o_inverse_4x4.argtypes = [fp, fp]
        L.lc2o_inverse_3x3.argtypes = [fp, fp]
        L.lc2o_warp_3d.argtypes = [fp, gp, gp, fp, fp]
        L.lc2o_warp_2d.argtypes = [fp, gp, gp, fp, fp]
        for n in ("lc2o_sobel_3d_ref", "lc2o_sobel_3d_sep", "lc2o_sobel_2d_ref",
                  "lc2o_sobel_2d_sep"):
            getattr(L, n).argtypes = [fp, ip, fp]
        L.lc2o_lc2_exact.argtypes = [fp, fp, fp, ip, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp,
                                     C.POINTER(Result)]
        L.lc2o_lc2_cg.argtypes = [fp, fp, fp, ip, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  fp, fp, C.POINTER(Result)]
        L.lc2o_eval.argtypes = [fp, gp, fp, gp, gp, C.c_int, C.c_int, C.c_int, C.c_int, fp,
                                C.c_int, C.c_int, C.POINTER(Result)]
        L.lc2o_num_threads.restype = C.c_int
        L.lc2o_set_num_threads.argtypes = [C.c_int]
        L.lc2o_set_num_threads.restype = C.c_int
        L.lc2o_nmi.argtypes = [fp, fp, C.c_longlong, C.c_int, C.c_int, C.c_int, ip, ip, ip]
        L.lc2o_nmi.restype = C.c_double
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def geom(dims, spacing, origin) -> Geom:
    d = list(dims) + [1] * (3 - len(dims))
    s = list(spacing) + [1.0] * (3 - len(spacing))
    o = list(origin) + [0.0] * (3 - len(origin))
    return Geom((C.c_int * 3)(*d), (C.c_float * 3)(*s), (C.c_float * 3)(*o))


@dataclass
class LC2Result:
    sum_measure: float
    sum_w: float
    lc2: float
    n_valid: int


def _res(r: Result) -> LC2Result:
    return LC2Result(r.sum_measure, r.sum_w, r.lc2, r.n_valid)


# ---- geometry ----
def center(g: Geom, ndim: int) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().lc2o_center(C.byref(g), ndim, _f(out))
    return out[:ndim]


def virtual_dims(fixed: Geom, resampled, ndim: int):
    rs = _f32(list(resampled) + [1.0] * (3 - len(resampled)))
    out = (C.c_int * 3)()
    lib().lc2o_virtual_dims(C.byref(fixed), _f(rs), ndim, out)
    return tuple(out)


def affine_3d(trans, rot, scale, shear, center_, shift) -> np.ndarray:
    """Parameter order of the reference's pose vector: T, R (degrees), S, Sh."""
    out = np.zeros(16, np.float32)
    args = [_f32(v) for v in (rot, trans, scale, shear, center_, shift)]
    lib().lc2o_affine_init_3d(*[_f(a) for a in args], _f(out))
    return out.reshape(4, 4)


def affine_2d(rot, trans, scale, shear, center_, shift) -> np.ndarray:
    out = np.zeros(9, np.float32)
    t, s, c, sh = _f32(trans), _f32(scale), _f32(center_), _f32(shift)
    lib().lc2o_affine_init_2d(float(rot), _f(t), _f(s), float(shear), _f(c), _f(sh), _f(out))
    return out.reshape(3, 3)


def inverse(m: np.ndarray) -> np.ndarray:
    m = _f32(m)
    out = np.zeros_like(m)
    fn = lib().lc2o_inverse_4x4 if m.size == 16 else lib().lc2o_inverse_3x3
    if not fn(_f(m.ravel()), _f(out.ravel())):
        raise ValueError("singular transform")
    return out


# ---- warp / sobel ----
def warp_3d(src, src_geom: Geom, virt: Geom, m12) -> np.ndarray:
    src = _f32(src)
    out = np.zeros(tuple(virt.dims)[::-1], np.float32)
    m = _f32(np.asarray(m12, np.float32).ravel()[:12])
    lib().lc2o_warp_3d(_f(src), C.byref(src_geom), C.byref(virt), _f(m), _f(out))
    return out


def warp_2d(src, src_geom: Geom, virt: Geom, m6) -> np.ndarray:
    src = _f32(src)
    out = np.zeros((virt.dims[1], virt.dims[0]), np.float32)
    m = _f32(np.asarray(m6, np.float32).ravel()[:6])
    lib().lc2o_warp_2d(_f(src), C.byref(src_geom), C.byref(virt), _f(m), _f(out))
    return out


def sobel(I, order: str = "sep") -> np.ndarray:
    """|Sobel| with zero padding; I is (Z,Y,X) for 3D or (Y,X) for 2D."""
    I = _f32(I)
    G = np.zeros_like(I)
    if I.ndim == 3:
        dims = (C.c_int * 3)(I.shape[2], I.shape[1], I.shape[0])
        fn = lib().lc2o_sobel_3d_ref if order == "ref" else lib().lc2o_sobel_3d_sep
    else:
        dims = (C.c_int * 3)(I.shape[1], I.shape[0], 1)
        fn = lib().lc2o_sobel_2d_ref if order == "ref" else lib().lc2o_sobel_2d_sep
    fn(_f(I), dims, _f(G))
    return G


def _dims_of(U):
    if U.ndim == 3:
        return (C.c_int * 3)(U.shape[2], U.shape[1], U.shape[0]), 3, U.shape[0]
    return (C.c_int * 3)(U.shape[1], U.shape[0], 1), 2, U.shape[0]


def lc2_exact(U, I, G, r: int, zlo: int = 0, zhi: int | None = None, maps: bool = False):
    U, I, G = _f32(U), _f32(I), _f32(G)
    dims, nd, outer = _dims_of(U)
    res = Result()
    wm = np.zeros(U.shape, np.float64) if maps else None
    mm = np.zeros(U.shape, np.float64) if maps else None
    rc = lib().lc2o_lc2_exact(_f(U), _f(I), _f(G), dims, nd, r, zlo,
                              outer if zhi is None else zhi,
                              _d(wm) if maps else None, _d(mm) if maps else None, C.byref(res))
    if rc:
        raise ValueError("lc2o_lc2_exact failed")
    return (_res(res), wm, mm) if maps else _res(res)


def lc2_cg(U, I, G, r: int, fp64: bool = False, zlo: int = 0, zhi: int | None = None,
           maps: bool = False):
    U, I, G = _f32(U), _f32(I), _f32(G)
    dims, nd, outer = _dims_of(U)
    res = Result()
    wm = np.zeros(U.shape, np.float32) if maps else None
    mm = np.zeros(U.shape, np.float32) if maps else None
    rc = lib().lc2o_lc2_cg(_f(U), _f(I), _f(G), dims, nd, r, zlo, outer if zhi is None else zhi,
                           int(fp64), _f(wm) if maps else None, _f(mm) if maps else None,
                           C.byref(res))
    if rc:
        raise ValueError("lc2o_lc2_cg failed")
    return (_res(res), wm, mm) if maps else _res(res)


def evaluate(fixed, fixed_geom: Geom, moving, moving_geom: Geom, virt: Geom, grid_ndim: int,
             src_ndim: int, high_res_is_moving: bool, r: int, m12, solver: str = "exact",
             sobel_order: str = "sep") -> LC2Result:
    """One LC2 evaluation in the reference's data flow (warp both, Sobel, LC2)."""
    fixed, moving = _f32(fixed), _f32(moving)
    m = np.zeros(12, np.float32)
    mm = np.asarray(m12, np.float32).ravel()
    m[: min(12, mm.size)] = mm[:12]
    res = Result()
    solver_code = {"exact": 0, "cg": 1, "cg64": 2}[solver]
    rc = lib().lc2o_eval(_f(fixed), C.byref(fixed_geom), _f(moving), C.byref(moving_geom),
                         C.byref(virt), grid_ndim, src_ndim, int(high_res_is_moving), r, _f(m),
                         solver_code, 0 if sobel_order == "ref" else 1, C.byref(res))
    if rc:
        raise ValueError("lc2o_eval failed")
    return _res(res)


def nmi(U, I, bins_u: int = 256, bins_i: int = 256, fp32: bool = True, hists: bool = False):
    """Normalized mutual information of the patchSize 0 path ((H_U + H_I) / H_joint) of
    two images of equal size; with hists also the U, I and joint histograms."""
    U, I = _f32(U).ravel(), _f32(I).ravel()
    if U.size != I.size:
        raise ValueError("images differ in size")
    hu = np.zeros(bins_u, np.int32)
    hi = np.zeros(bins_i, np.int32)
    hj = np.zeros((bins_i, bins_u), np.int32)
    ip = C.POINTER(C.c_int)
    v = lib().lc2o_nmi(_f(U), _f(I), U.size, bins_u, bins_i, int(fp32), hu.ctypes.data_as(ip),
                       hi.ctypes.data_as(ip), hj.ctypes.data_as(ip))
    return (v, hu, hi, hj) if hists else v


def num_threads() -> int:
    return lib().lc2o_num_threads()


def set_num_threads(n: int) -> int:
    """OpenMP threads of the oracle's loops from now on; returns the count in force."""
    return lib().lc2o_set_num_threads(n)
