"""The reference's LC2 classes (LC2_3D_class / LC2_2D_class, cu_lc2_3d.cuh:14-55,
cu_lc2_2d.cuh:14-47) as the facade provides them, through lc2reg_lc2_class_metric: two
images that already share a grid, GetSimilarityMetric, the |Sobel| buffer and the
weighted map, against the oracle on the same images."""
import ctypes as C

import numpy as np
import pytest

from lc2registration_amd import engine as E
from lc2registration_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

_F = C.POINTER(C.c_float)


def _lib():
    lib = C.CDLL(E.library_path())
    fn = lib.lc2reg_lc2_class_metric
    fn.argtypes = [C.c_int, C.c_int, C.c_int, _F, _F, C.c_int, C.POINTER(C.c_double), _F, _F]
    fn.restype = C.c_int
    lib.lc2reg_last_error.restype = C.c_char_p
    return lib


def _metric(U, I, r):
    lib = _lib()
    U = np.ascontiguousarray(U, np.float32)
    I = np.ascontiguousarray(I, np.float32)
    dims = U.shape[::-1] + (1,) * (3 - U.ndim)
    l = C.c_double()
    g = np.zeros_like(U)
    wm = np.zeros_like(U)
    rc = lib.lc2reg_lc2_class_metric(dims[0], dims[1], dims[2], U.ctypes.data_as(_F),
                                     I.ctypes.data_as(_F), r, C.byref(l), g.ctypes.data_as(_F),
                                     wm.ctypes.data_as(_F))
    assert rc == 0, lib.lc2reg_last_error()
    return l.value, g, wm


def _images():
    """US and MR of the small 3D workload resampled onto the US grid at the generating pose."""
    w = synth.small3d()
    f, m = w.fixed, w.moving
    fg, mg = O.geom(f.dims, f.spacing, f.origin), O.geom(m.dims, m.spacing, m.origin)
    vd = O.virtual_dims(fg, w.resampled_spacing, 3)
    vg = O.geom(vd, w.resampled_spacing, f.origin)
    e = E.Engine(0)
    fv = e.upload(f.data, f.spacing, f.origin)
    mv = e.upload(m.data, m.spacing, m.origin)
    p = e.pair(fv, mv, w.resampled_spacing, True, False, w.patch_radius, 3)
    M = p.pose_to_matrix(w.true_pose)
    p.close()
    fv.close()
    mv.close()
    e.close()
    U = O.warp_3d(f.data.astype(np.float32), fg, vg, [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0])
    I = O.warp_3d(m.data, mg, vg, M)
    return U, I


def _check(U, I, r, l, g, wm, min_valid):
    G = O.sobel(I, "sep")
    res, ow, om = O.lc2_exact(U, I, G, r, maps=True)
    assert res.n_valid > min_valid
    assert np.array_equal(g, G)
    assert abs(l - res.lc2) < 1e-9, (l, res.lc2)
    ratio = np.where(ow > 0, om / np.where(ow > 0, ow, 1.0), 0.0).astype(np.float32)
    assert np.count_nonzero(ratio) > min_valid // 2
    assert np.allclose(wm, ratio, rtol=1e-6, atol=1e-9)


def test_lc2_3d_class_matches_oracle():
    U, I = _images()
    l, g, wm = _metric(U, I, 3)
    _check(U, I, 3, l, g, wm, 5000)


def test_lc2_2d_class_matches_oracle():
    U, I = _images()
    z = U.shape[0] // 2
    U2, I2 = np.ascontiguousarray(U[z]), np.ascontiguousarray(I[z])
    l, g, wm = _metric(U2, I2, 9)  # PatchSize 9: the reference's 2D default
    _check(U2, I2, 9, l, g, wm, 100)


def test_lc2_class_rejects_bad_arguments():
    lib = _lib()
    buf = np.zeros(8, np.float32)
    l = C.c_double()
    assert lib.lc2reg_lc2_class_metric(2, 2, 2, buf.ctypes.data_as(_F), buf.ctypes.data_as(_F),
                                       0, C.byref(l), None, None) == -1
    assert b"invalid argument" in lib.lc2reg_last_error()
