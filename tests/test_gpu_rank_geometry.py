"""GPU cases for the two pieces the host tests pin: the rank-deficient branch of the
projection, and the product's pose -> matrix / MovingToFixed geometry.

  * Plateau label maps (with and without noise), linear ramps (constant |Sobel|) and
    exponential ramps, whose |Sobel| is collinear with I (positively for e^{kx}, negatively
    for -e^{kx}), through the reference's LC2 class interface (LC2_3D_class /
    LC2_2D_class via lc2reg_lc2_class_metric): LC2 and weighted map against the oracle.
  * lc2e_pose_to_matrix on a device pair equals the host geometry (pinned against the
    oracle and closed forms in test_geometry_product.py) and the oracle, bit for bit. (The
    matrix a YAML registration reports is checked against the oracle's inverse in
    test_gpu_registration.py.)
"""
import ctypes as C

import numpy as np
import pytest

from lc2registration_amd import engine as E
from lc2registration_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

_F = C.POINTER(C.c_float)


def _metric(U, I, r):
    lib = C.CDLL(E.library_path())
    fn = lib.lc2reg_lc2_class_metric
    fn.argtypes = [C.c_int, C.c_int, C.c_int, _F, _F, C.c_int, C.POINTER(C.c_double), _F, _F]
    U = np.ascontiguousarray(U, np.float32)
    I = np.ascontiguousarray(I, np.float32)
    dims = U.shape[::-1] + (1,) * (3 - U.ndim)
    l = C.c_double()
    g = np.zeros_like(U)
    wm = np.zeros_like(U)
    assert fn(dims[0], dims[1], dims[2], U.ctypes.data_as(_F), I.ctypes.data_as(_F), r,
              C.byref(l), g.ctypes.data_as(_F), wm.ctypes.data_as(_F)) == 0
    return l.value, g, wm


def _label_map(shape, seed, noise=0.0):
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    lab = np.zeros(shape, np.int32)
    for k in (1, 2):
        nrm = rng.normal(size=len(shape))
        off = rng.uniform(0.3, 0.7) * sum(n * abs(c) for n, c in zip(shape, nrm))
        lab[sum(c * g for c, g in zip(nrm, grids)) + 0.5 * off > off] = k
    I = np.array([-1000.0, 40.0, 1000.0])[lab]
    if noise:
        I = I + rng.normal(0.0, noise, size=shape)
    return I.astype(np.float32)


def _images(kind, ndim):
    shape = (40, 44, 48) if ndim == 3 else (70, 76)
    U = np.random.default_rng(3).uniform(10, 200, size=shape).astype(np.float32)
    g = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    x = g[-1]
    if kind == "plateau":
        I = _label_map(shape, 1)
    elif kind == "plateau_noise":
        I = _label_map(shape, 2, noise=20.0)
    elif kind == "ramp":
        I = (3.0 * x - 2.0 * g[-2] + (0.5 * g[0] if ndim == 3 else 0.0)).astype(np.float32)
    elif kind == "exp_pos":  # |Sobel| of e^{kx} is c e^{kx}: collinear with I
        I = np.exp(0.25 * x).astype(np.float32)
    else:  # exp_neg: I = -e^{kx}, |Sobel| = c e^{kx} = -c I
        I = (-np.exp(0.25 * x)).astype(np.float32)
    # (k = 0.25: a patch's spread of I is a large fraction of its mean, so the f32 rounding
    # of I and G leaves 1 - rho^2 near 1e-13, far below the 1e-10 collinearity threshold)
    if not kind.startswith("exp"):
        # U follows I and |grad I| loosely so the fit has something to explain
        U = (U + 0.05 * np.abs(I)).astype(np.float32)
    # U = 0 on the faces: patches that reach the zero-padded |Sobel| of the border rows are
    # excluded, so every patch is clearly inside or clearly outside a rank class (the rule
    # is discontinuous at its threshold, where rounding may decide either way)
    for ax in range(U.ndim):
        idx = [slice(None)] * U.ndim
        for e in (0, -1):
            idx[ax] = e
            U[tuple(idx)] = 0.0
    return U, I


@pytest.mark.parametrize("kind", ["plateau", "plateau_noise", "ramp", "exp_pos", "exp_neg"])
@pytest.mark.parametrize("ndim", [2, 3])
def test_rank_deficient_patches_match_oracle(kind, ndim):
    U, I = _images(kind, ndim)
    r = 3 if ndim == 3 else 5
    print(f"rank case {kind} {ndim}D r={r} shape={U.shape}", flush=True)
    l, g, wm = _metric(U, I, r)
    G = O.sobel(I, "sep")
    assert np.array_equal(g, G)
    res, ow, om = O.lc2_exact(U, I, G, r, maps=True)
    assert res.n_valid > 1000
    assert abs(l - res.lc2) < 1e-9, (l, res.lc2)
    ratio = np.where(ow > 0, om / np.where(ow > 0, ow, 1.0), 0.0).astype(np.float32)
    assert np.allclose(wm, ratio, rtol=1e-6, atol=1e-9)


def test_pair_pose_to_matrix_equals_host_geometry(engine):
    """lc2e_pose_to_matrix (device pair) == lc2reg_pose_to_matrix (host), 12-DoF with
    CenterOverlaid, and both equal the oracle's affine matrix."""
    w = synth.small3d()
    f, m = w.fixed, w.moving
    fv = engine.upload(f.data, f.spacing, f.origin)
    mv = engine.upload(m.data, m.spacing, m.origin)
    L = C.CDLL(E.library_path())
    L.lc2reg_pose_to_matrix.argtypes = [C.POINTER(E.PairInfo), C.POINTER(C.c_double), C.c_int, _F]
    rng = np.random.default_rng(4)
    for co in (False, True):
        p = engine.pair(fv, mv, w.resampled_spacing, True, co, w.patch_radius, 3)
        info = p.info
        for _ in range(6):
            q = np.concatenate([rng.uniform(-10, 10, 3), rng.uniform(-20, 20, 3),
                                rng.uniform(0.85, 1.15, 3), rng.uniform(-0.15, 0.15, 3)])
            a = p.pose_to_matrix(q)
            b = np.zeros(12, np.float32)
            assert L.lc2reg_pose_to_matrix(C.byref(info), q.ctypes.data_as(C.POINTER(C.c_double)),
                                           12, b.ctypes.data_as(_F)) == 0
            assert np.array_equal(a, b)
            om = O.affine_3d(q[:3], q[3:6], q[6:9], q[9:12], np.array(info.fixed_center, np.float32),
                             np.array(info.shift, np.float32))
            assert np.array_equal(a.reshape(3, 4), om[:3])
        p.close()
    fv.close()
    mv.close()
