"""Pins the product's host geometry (lc2e_geometry.hpp, reached through the C ABI's
host-only entry points lc2reg_pair_geometry / lc2reg_pose_to_matrix /
lc2reg_moving_to_fixed) against the oracle and against closed forms, on the CPU.

These are the functions behind lc2e_pose_to_matrix and the facade's
GetMovingToFixedTransform, so a slip in the product's shear order, scale, centre or
CenterOverlaid shift fails here even though the parity tests feed the product's own
matrix to both sides.
  * virtual grid, fixed centre, shift: equal to the oracle's (interpolator.cu:320,
    :337-339, :495-529);
  * VirtualToMoving for 12-DoF poses (rotation, translation, scale, shear) with and
    without CenterOverlaid: bit-identical to the oracle's affine_3d/affine_2d
    (spatialMappers.cuh:637-850, :999-1033), and within FP32 rounding of the closed form
    Rz Ry Rx Sh S about the fixed centre composed in numpy;
  * MovingToFixed: bit-identical to the oracle's inverse, and an inverse of the matrix
    to FP32 rounding.
"""
import ctypes as C

import numpy as np
import pytest

from lc2registration_amd import engine as E
from oracle import oracle as O


def _lib():
    if not __import__("os").path.exists(E.library_path()):
        from lc2registration_amd import build
        build.build()
    L = C.CDLL(E.library_path())
    G = C.POINTER(E.Geom)
    L.lc2reg_pair_geometry.argtypes = [G, G, C.POINTER(C.c_float), C.c_int, C.c_int,
                                       C.POINTER(E.PairInfo)]
    L.lc2reg_pose_to_matrix.argtypes = [C.POINTER(E.PairInfo), C.POINTER(C.c_double), C.c_int,
                                        C.POINTER(C.c_float)]
    L.lc2reg_moving_to_fixed.argtypes = L.lc2reg_pose_to_matrix.argtypes
    return L


CASES_3D = [
    # fixed geometry, moving geometry, resampled spacing, center_overlaid
    (((192, 160, 128), (0.5, 0.5, 0.5), (32.0, 48.0, 64.0)),
     ((256, 256, 192), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)), (0.5, 0.5, 0.5), False),
    (((394, 466, 378), (0.5, 0.5, 0.5), (-98.0, -134.0, -72.0)),
     ((323, 366, 371), (0.3, 0.3, 0.3), (-42.2, -52.7, -59.4)), (1.3, 1.3, 1.3), True),
    (((100, 90, 80), (0.8, 0.7, 1.1), (3.0, -5.0, 7.5)),
     ((64, 70, 50), (1.2, 0.9, 1.5), (-11.0, 4.0, 2.0)), (1.0, 0.9, 1.2), True),
]


def _info(L, fg, mg, rs, dim, co):
    out = E.PairInfo()
    rsf = (C.c_float * 3)(*(list(rs) + [1.0] * (3 - len(rs))))
    assert L.lc2reg_pair_geometry(C.byref(E.make_geom(*fg)), C.byref(E.make_geom(*mg)), rsf, dim,
                                  int(co), C.byref(out)) == 0
    return out


def _pose_matrix(L, info, p):
    p = np.ascontiguousarray(p, np.float64)
    m = np.zeros(12, np.float32)
    assert L.lc2reg_pose_to_matrix(C.byref(info), p.ctypes.data_as(C.POINTER(C.c_double)), p.size,
                                   m.ctypes.data_as(C.POINTER(C.c_float))) == 0
    return m.reshape(3, 4)


def _m2f(L, info, p, n_out):
    p = np.ascontiguousarray(p, np.float64)
    m = np.zeros(n_out, np.float32)
    assert L.lc2reg_moving_to_fixed(C.byref(info), p.ctypes.data_as(C.POINTER(C.c_double)), p.size,
                                    m.ctypes.data_as(C.POINTER(C.c_float))) == 0
    return m


def _poses_3d(rng, k=12):
    out = [np.r_[0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0.0]]
    for _ in range(k):
        out.append(np.concatenate([rng.uniform(-20, 20, 3), rng.uniform(-25, 25, 3),
                                   rng.uniform(0.8, 1.2, 3), rng.uniform(-0.2, 0.2, 3)]))
    return out


def _closed_form_3d(p, C_, shift):
    """Rz Ry Rx Sh S about the fixed centre, in numpy float64 (from float32 parameters)."""
    p = np.asarray(p, np.float32).astype(np.float64)
    t, r, s, sh = p[:3], np.deg2rad(p[3:6]), p[6:9], p[9:12]
    cx, sx, cy, sy, cz, sz = np.cos(r[0]), np.sin(r[0]), np.cos(r[1]), np.sin(r[1]), np.cos(r[2]), np.sin(r[2])
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    Sh = np.array([[1, sh[0], sh[1]], [0, 1, sh[2]], [0, 0, 1]])
    A = Rz @ Ry @ Rx @ Sh @ np.diag(s)
    C64, sh64 = np.asarray(C_, np.float64), np.asarray(shift, np.float64)
    return np.hstack([A, (A @ (sh64 - C64) + C64 + t)[:, None]])


@pytest.mark.parametrize("case", range(len(CASES_3D)))
def test_3d_geometry_matches_oracle_and_closed_form(case):
    L = _lib()
    fg, mg, rs, co = CASES_3D[case]
    info = _info(L, fg, mg, rs, 3, co)
    ofg, omg = O.geom(*fg), O.geom(*mg)
    vd = O.virtual_dims(ofg, rs, 3)
    assert tuple(info.virt.dims) == tuple(vd)
    c_f = O.center(ofg, 3)
    c_v = O.center(O.geom(vd, rs, fg[2]), 3)
    shift = (O.center(omg, 3) - c_v) if co else np.zeros(3, np.float32)
    assert np.array_equal(np.array(info.fixed_center, np.float32), c_f)
    assert np.array_equal(np.array(info.shift, np.float32), shift.astype(np.float32))
    if co:
        assert np.abs(shift).max() > 1.0  # the case exercises a real shift
    rng = np.random.default_rng(case)
    for p in _poses_3d(rng):
        m = _pose_matrix(L, info, p)
        om = O.affine_3d(p[:3], p[3:6], p[6:9], p[9:12], c_f, shift)
        assert np.array_equal(m, om[:3]), (p, m - om[:3])
        cf = _closed_form_3d(p, c_f, shift)
        scale = np.abs(cf).max(axis=1, keepdims=True)
        assert np.all(np.abs(m - cf) <= 4 * np.finfo(np.float32).eps * np.maximum(scale, 1.0))
        # rigid parameter vector = the affine one with S = 1, Sh = 0
        if np.allclose(p[6:], [1, 1, 1, 0, 0, 0]):
            assert np.array_equal(_pose_matrix(L, info, p[:6]), m)
        inv = _m2f(L, info, p, 16).reshape(4, 4)
        oinv = O.inverse(om)
        assert np.array_equal(inv, oinv)
        full = np.vstack([m.astype(np.float64), [0, 0, 0, 1]])
        assert np.allclose(inv.astype(np.float64) @ full, np.eye(4), atol=2e-5)
        assert np.array_equal(inv[3], [0, 0, 0, 1])


@pytest.mark.parametrize("co", [False, True])
def test_2d_geometry_matches_oracle(co):
    L = _lib()
    fg = ((164, 151), (1.0, 1.0), (0.0, 0.0))
    mg = ((419, 310), (0.310571, 0.310571), (5.0, -3.0))
    info = _info(L, fg, mg, (0.9, 0.9), 2, co)
    ofg, omg = O.geom(*fg), O.geom(*mg)
    vd = O.virtual_dims(ofg, (0.9, 0.9), 2)
    assert tuple(info.virt.dims[:2]) == tuple(vd[:2])
    c_f = O.center(ofg, 2)
    c_v = O.center(O.geom(vd, (0.9, 0.9, 1.0), (0.0, 0.0, 0.0)), 2)
    shift = (O.center(omg, 2) - c_v) if co else np.zeros(2, np.float32)
    rng = np.random.default_rng(11)
    for _ in range(10):
        p = np.r_[rng.uniform(-15, 15, 2), rng.uniform(-30, 30), rng.uniform(0.8, 1.2, 2),
                  rng.uniform(-0.2, 0.2)]
        m = _pose_matrix(L, info, p)
        om = O.affine_2d(p[2], p[:2], p[3:5], p[5], c_f, shift)
        assert np.array_equal(m[:2, [0, 1, 3]], om[:2])
        assert np.array_equal(m[2], [0, 0, 1, 0]) and m[0, 2] == 0 and m[1, 2] == 0
        inv = _m2f(L, info, p, 9).reshape(3, 3)
        assert np.array_equal(inv, O.inverse(om))


def test_host_geometry_rejects_bad_input():
    L = _lib()
    info = _info(L, *CASES_3D[0][:3], 3, False)
    p = np.zeros(5)
    m = np.zeros(12, np.float32)
    assert L.lc2reg_pose_to_matrix(C.byref(info), p.ctypes.data_as(C.POINTER(C.c_double)), 5,
                                   m.ctypes.data_as(C.POINTER(C.c_float))) == -1
    sing = np.r_[0, 0, 0, 0, 0, 0, 0, 1, 1, 0, 0, 0.0]  # scale 0: singular
    out = np.zeros(16, np.float32)
    assert L.lc2reg_moving_to_fixed(C.byref(info), sing.ctypes.data_as(C.POINTER(C.c_double)), 12,
                                    out.ctypes.data_as(C.POINTER(C.c_float))) == -1
