"""Pins the oracle's geometry and warp (CPU only).

Anchors: the reference's own virtual-grid dumps (config/Data, copied to
tests/golden/kidney) and closed forms of the affine conventions.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = []


def _kgeom(man, key):
    m = man[key]
    return O.geom(m["dims"], m["spacing"], m["origin"])


def test_identity_dumps_bit_exact(kidney):
    """fixed.bin / CT_KidneyPhantom_virtual.bin are the CT resampled at identity
    (interpolator.cu:1364-1368 debug dump): reproduced bit for bit."""
    man = kidney["manifest"]
    g_ct = _kgeom(man, "fixed")
    for name, sp in man["identity_dumps"].items():
        vd = O.virtual_dims(g_ct, (sp, sp), 2)
        assert vd[:2] == tuple(reversed(kidney[name].shape))
        out = O.warp_2d(kidney["ct"], g_ct, O.geom(vd[:2], (sp, sp), (0, 0)), [1, 0, 0, 0, 1, 0])
        assert np.array_equal(out, kidney[name]), name


def test_rigid_dumps(kidney):
    """moving.bin / US_KidneyPhantom_virtual.bin: the US warped by a rigid pose with
    CenterOverlaid, rotation about the fixed centre (interpolator.cu:495-529)."""
    man = kidney["manifest"]
    g_ct, g_us = _kgeom(man, "fixed"), _kgeom(man, "moving")
    for name, rec in man["rigid_dumps"].items():
        sp = rec["virtual_spacing"]
        g_v = O.geom(rec["virtual_dims"], (sp, sp), (0, 0))
        C = O.center(g_ct, 2)
        shift = O.center(g_us, 2) - O.center(g_v, 2)
        tx, ty, r = rec["pose_tx_ty_rdeg"]
        M = O.affine_2d(r, (tx, ty), (1, 1), 0.0, C, shift)
        out = O.warp_2d(kidney["us"], g_us, g_v, M.ravel()[:6])
        err = np.abs(out - kidney[name])
        assert err.max() <= 1.5e-3, (name, err.max())
        # a wrong centre convention (rotation about the moving centre) does not fit
        M_bad = O.affine_2d(r, (tx, ty), (1, 1), 0.0, O.center(g_us, 2), shift)
        bad = O.warp_2d(kidney["us"], g_us, g_v, M_bad.ravel()[:6])
        assert np.abs(bad - kidney[name]).max() > 1.0


def test_virtual_dims_ceil_rule(kidney):
    g = _kgeom(kidney["manifest"], "fixed")
    assert O.virtual_dims(g, (1.0, 1.0), 2)[:2] == (164, 151)
    assert O.virtual_dims(g, (0.9, 0.9), 2)[:2] == (182, 168)


def test_warp3d_equals_warp2d_on_single_slice(kidney):
    """A one-slice volume warped trilinearly by an in-plane transform equals the
    (golden-pinned) bilinear 2D warp bit for bit."""
    man = kidney["manifest"]
    us = kidney["us"]
    g2 = _kgeom(man, "moving")
    g3 = O.geom(man["moving"]["dims"] + [1], man["moving"]["spacing"] + [1.0], [0, 0, 0])
    gv2 = O.geom((160, 150), (0.7, 0.7), (3.0, -2.0))
    gv3 = O.geom((160, 150, 1), (0.7, 0.7, 1.0), (3.0, -2.0, 0.0))
    M2 = O.affine_2d(7.5, (4.0, -3.0), (1.03, 0.97), 0.02, (60.0, 50.0), (0.5, -0.5))
    M3 = np.array([[M2[0, 0], M2[0, 1], 0, M2[0, 2]], [M2[1, 0], M2[1, 1], 0, M2[1, 2]],
                   [0, 0, 1, 0]], np.float32)
    a = O.warp_2d(us, g2, gv2, M2.ravel()[:6])
    b = O.warp_3d(us[None], g3, gv3, M3)[0]
    assert np.array_equal(a, b)


def test_warp3d_exact_on_affine_field():
    """Trilinear interpolation reproduces an affine intensity field; outside the
    source's [0, dim-1] box the sample is 0."""
    rng = np.random.default_rng(3)
    X, Y, Z = 17, 13, 11
    zz, yy, xx = np.mgrid[0:Z, 0:Y, 0:X].astype(np.float64)
    a, b, c, d = 0.5, -0.25, 0.125, 3.0
    src = (a * xx + b * yy + c * zz + d).astype(np.float32)
    gs = O.geom((X, Y, Z), (1.5, 1.25, 2.0), (-3.0, 1.0, 2.0))
    gv = O.geom((20, 16, 12), (1.2, 1.1, 1.9), (-5.0, -1.0, 0.0))
    M = O.affine_3d((4, -3, 5), (0.8, -0.4, 0.7), (1, 1, 1), (0, 0, 0), (2, 3, 4), (0, 0, 0))
    out = O.warp_3d(src, gs, gv, M[:3].ravel())
    vz, vy, vx = np.mgrid[0:12, 0:16, 0:20].astype(np.float64)
    L = np.stack([vx * 1.2 - 5.0, vy * 1.1 - 1.0, vz * 1.9])
    P = np.tensordot(M[:3, :3].astype(np.float64), L, 1) + M[:3, 3].astype(np.float64)[:, None, None, None]
    v = np.stack([(P[0] + 3.0) / 1.5, (P[1] - 1.0) / 1.25, (P[2] - 2.0) / 2.0])
    inside = ((v[0] >= 0) & (v[0] <= X - 1) & (v[1] >= 0) & (v[1] <= Y - 1) & (v[2] >= 0) &
              (v[2] <= Z - 1))
    margin = ((v[0] > 1e-3) & (v[0] < X - 1 - 1e-3) & (v[1] > 1e-3) & (v[1] < Y - 1 - 1e-3) &
              (v[2] > 1e-3) & (v[2] < Z - 1 - 1e-3))
    expect = a * v[0] + b * v[1] + c * v[2] + d
    assert inside.sum() > 500 and (~inside).sum() > 100
    assert np.allclose(out[margin], expect[margin], rtol=0, atol=2e-5)
    far = ~inside & ~((v[0] > -1e-3) & (v[0] < X - 1 + 1e-3) & (v[1] > -1e-3) & (v[1] < Y - 1 + 1e-3) &
                      (v[2] > -1e-3) & (v[2] < Z - 1 + 1e-3))
    assert np.all(out[far] == 0.0)


# ---- affine conventions (spatialMappers.cuh:579-850, :977-1033) ----

def _apply(M, x):
    return M[:3, :3].astype(np.float64) @ np.asarray(x, np.float64) + M[:3, 3]


def test_affine3d_pure_translation_and_centre():
    C = (10.0, -4.0, 7.0)
    M = O.affine_3d((1, 2, 3), (0, 0, 0), (1, 1, 1), (0, 0, 0), C, (0, 0, 0))
    assert np.allclose(M[:3, :3], np.eye(3)) and np.allclose(M[:3, 3], (1, 2, 3))
    M = O.affine_3d((0, 0, 0), (20, -35, 50), (1, 1, 1), (0, 0, 0), C, (0, 0, 0))
    assert np.allclose(_apply(M, C), C, atol=1e-5)           # rotation about C
    R = M[:3, :3].astype(np.float64)
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-6) and abs(np.linalg.det(R) - 1) < 1e-6


@pytest.mark.parametrize("axis,frm,to", [(2, 0, 1), (0, 1, 2), (1, 2, 0)])
def test_affine3d_right_handed_axes(axis, frm, to):
    rot = [0, 0, 0]
    rot[axis] = 90.0
    M = O.affine_3d((0, 0, 0), rot, (1, 1, 1), (0, 0, 0), (0, 0, 0), (0, 0, 0))
    e = np.zeros(3)
    e[frm] = 1
    expect = np.zeros(3)
    expect[to] = 1
    assert np.allclose(M[:3, :3] @ e, expect, atol=1e-6)


def test_affine3d_rotation_order_and_scale_shear_first():
    # A = Rz Ry Rx Sh S: scale and shear act before the rotation, Rx first of the rotations
    M = O.affine_3d((0, 0, 0), (0, 0, 90), (2, 1, 1), (0, 0, 0), (0, 0, 0), (0, 0, 0))
    assert np.allclose(M[:3, :3] @ [1, 0, 0], [0, 2, 0], atol=1e-6)
    M = O.affine_3d((0, 0, 0), (0, 0, 90), (1, 1, 1), (0.3, 0, 0), (0, 0, 0), (0, 0, 0))
    assert np.allclose(M[:3, :3] @ [0, 1, 0], [-1, 0.3, 0], atol=1e-6)
    M = O.affine_3d((0, 0, 0), (90, 0, 90), (1, 1, 1), (0, 0, 0), (0, 0, 0), (0, 0, 0))
    # Rx(90) e_y = e_z, then Rz leaves e_z
    assert np.allclose(M[:3, :3] @ [0, 1, 0], [0, 0, 1], atol=1e-6)


def test_affine3d_reduces_to_2d():
    """Rotation about z with in-plane scale/shear equals the golden-pinned 2D form."""
    C, sh = (30.0, 40.0, 5.0), (1.5, -2.0, 0.0)
    M3 = O.affine_3d((3, -4, 0), (0, 0, 12.5), (1.05, 0.95, 1), (0.04, 0, 0), C, sh)
    M2 = O.affine_2d(12.5, (3, -4), (1.05, 0.95), 0.04, C[:2], sh[:2])
    assert np.allclose(M3[:2, [0, 1, 3]], M2[:2], atol=1e-5)


def test_center_overlaid_maps_centres():
    """With CenterOverlaid the identity pose maps the virtual centre onto the moving centre."""
    gf = O.geom((100, 80, 60), (0.8, 0.8, 1.0), (-10, 5, 2))
    gm = O.geom((70, 90, 50), (1.1, 0.9, 1.3), (4, -6, 1))
    vd = O.virtual_dims(gf, (1.2, 1.2, 1.2), 3)
    gv = O.geom(vd, (1.2, 1.2, 1.2), (-10, 5, 2))
    shift = O.center(gm, 3) - O.center(gv, 3)
    M = O.affine_3d((0, 0, 0), (0, 0, 0), (1, 1, 1), (0, 0, 0), O.center(gf, 3), shift)
    assert np.allclose(_apply(M, O.center(gv, 3)), O.center(gm, 3), atol=1e-4)


def test_inverse_round_trip():
    M = O.affine_3d((5, -3, 2), (10, 20, -30), (1.1, 0.9, 1.05), (0.05, -0.02, 0.03),
                    (12, 13, 14), (1, 2, 3))
    Mi = O.inverse(M)
    assert np.allclose(Mi.astype(np.float64) @ M.astype(np.float64), np.eye(4), atol=1e-5)
    M2 = O.affine_2d(17.0, (3, 4), (1.1, 0.95), 0.1, (5, 6), (0.5, 0.25))
    assert np.allclose(O.inverse(M2).astype(np.float64) @ M2.astype(np.float64), np.eye(3),
                       atol=1e-5)
    with pytest.raises(ValueError):
        O.inverse(np.zeros((4, 4), np.float32))
