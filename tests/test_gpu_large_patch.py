"""Patch radii beyond the fused kernels (3D r > 4, 2D r > 10) run the separable path (the
reference's large-patch global-memory kernel, cu_lc2_3d.cu:769-827; its 2D LDS path admits
r up to 24, cu_lc2_2d.cu:609-638): LC2 and the per-voxel maps against the oracle in both
roles, and batch = single bit for bit."""
import numpy as np
import pytest

from lc2registration_amd import engine as E
from lc2registration_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LC2_TOL = 1e-9


def _oracle(w, M, r, src_ndim, grid_ndim, maps=False):
    f, m = w.fixed, w.moving
    fg, mg = O.geom(f.dims, f.spacing, f.origin), O.geom(m.dims, m.spacing, m.origin)
    vd = O.virtual_dims(fg, w.resampled_spacing, len(f.dims))
    vg = O.geom(vd, w.resampled_spacing, f.origin)
    fix, mov = f.data.astype(np.float32), m.data.astype(np.float32)
    if src_ndim == 3:
        Wf = O.warp_3d(fix, fg, vg, [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0])
        Wm = O.warp_3d(mov, mg, vg, np.asarray(M).ravel()[:12])
    else:
        Wf = O.warp_2d(fix, fg, vg, [1, 0, 0, 0, 1, 0])
        M3 = np.asarray(M).reshape(3, 4)
        Wm = O.warp_2d(mov, mg, vg, [M3[0, 0], M3[0, 1], M3[0, 3], M3[1, 0], M3[1, 1], M3[1, 3]])
    I, U = (Wm, Wf) if w.high_res_is_moving else (Wf, Wm)
    if grid_ndim == 2 and I.ndim == 3:
        I, U = I[0], U[0]
    G = O.sobel(I, "sep")
    return O.lc2_exact(U, I, G, r, maps=maps)


def _pair(engine, w, r, dim):
    f, m = w.fixed, w.moving
    fv = engine.upload(f.data, f.spacing, f.origin)
    mv = engine.upload(m.data, m.spacing, m.origin)
    return engine.pair(fv, mv, w.resampled_spacing, w.high_res_is_moving, w.center_overlaid, r,
                       dim), (fv, mv)


@pytest.mark.parametrize("r", [5, 6])
@pytest.mark.parametrize("moving_high_res", [True, False])
def test_3d_large_radius_matches_oracle(engine, r, moving_high_res):
    w = synth.small3d()
    if not moving_high_res:  # the MR fixed (static |Sobel|), the ultrasound warped
        w = synth.Workload("swap", w.moving, w.fixed, (1.0, 1.0, 1.0), r, False, True, 3,
                           np.r_[0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0])
    print(f"3D separable path r={r} moving_high_res={moving_high_res}", flush=True)
    p, vols = _pair(engine, w, r, 3)
    rng = np.random.default_rng(r)
    poses = [w.true_pose, synth.perturb(w.true_pose, rng, 3), synth.perturb(w.true_pose, rng, 3)]
    mats = np.stack([p.pose_to_matrix(q) for q in poses])
    batch = p.eval_batch(mats)
    for k, M in enumerate(mats):
        res = _oracle(w, M, r, 3, 3)
        assert res.n_valid > 500
        assert abs(batch.lc2[k] - res.lc2) < LC2_TOL, (r, batch.lc2[k], res.lc2)
        assert p.eval(M) == batch.lc2[k]  # batch = single, bit for bit
    res, wm, mm = _oracle(w, mats[1], r, 3, 3, maps=True)
    gw = p.debug_fetch(E.STAGE_W, mats[1])
    gm = p.debug_fetch(E.STAGE_MEASURE, mats[1])
    assert np.array_equal(gw > 0, wm > 0)
    assert np.allclose(gw, wm, rtol=1e-9, atol=1e-12)
    assert np.allclose(gm, mm, rtol=1e-7, atol=1e-9)
    p.close()
    for v in vols:
        v.close()


@pytest.mark.parametrize("r", [11, 16, 24])
def test_2d_large_radius_matches_oracle(engine, w3_slice, r):
    """2D affine (the ultrasound warped onto a static image, 2D source) and 2D->3D (the MR
    warped onto a static ultrasound frame, 3D source) at the reference's 2D radii up to 24."""
    w = synth.config0(n=96, patch_radius=r)
    print(f"2D separable path r={r}", flush=True)
    p, vols = _pair(engine, w, r, 2)
    M = p.pose_to_matrix(w.true_pose)
    res = _oracle(w, M, r, 2, 2)
    assert res.n_valid > 100
    assert abs(p.eval(M) - res.lc2) < LC2_TOL, r
    p.close()
    for v in vols:
        v.close()
    ws = w3_slice(r)
    p, vols = _pair(engine, ws, r, 2)
    M = p.pose_to_matrix(ws.true_pose)
    res = _oracle(ws, M, r, 3, 2)
    if res.n_valid:
        assert abs(p.eval(M) - res.lc2) < LC2_TOL, r
    p.close()
    for v in vols:
        v.close()


@pytest.fixture(scope="module")
def w3_slice():
    w3 = synth.small3d(us_dims=(64, 60, 32))

    def make(r):
        us = w3.fixed
        sl = us.data[us.data.shape[0] // 2][None].copy()
        fixed = synth.Image(sl, (us.dims[0], us.dims[1], 1), us.spacing,
                            (us.origin[0], us.origin[1], us.origin[2] + us.spacing[2] * (us.dims[2] // 2)))
        return synth.Workload("2dto3d", fixed, w3.moving, us.spacing, r, True, False, 2, w3.true_pose)
    return make


def test_3d_source_with_big_slices_takes_separable_path(engine):
    """A 3D source whose slices hold 2^24 voxels or more is beyond the fused samplers'
    24-bit voxel offsets: such a pair is evaluated by the separable path at any radius (here
    r = 2, which the fused kernel would take otherwise). LC2 against the oracle."""
    X = Y = 4096
    x = np.arange(X, dtype=np.float32)
    plane = 100.0 + 40.0 * np.sin(x * 0.02)[None, :] * np.cos(x * 0.015)[:, None]
    mov = np.stack([plane, 0.5 * plane + 30.0]).astype(np.float32)           # (Z=2, Y, X)
    moving = synth.Image(mov, (X, Y, 2), (0.01, 0.01, 4.0), (0.0, 0.0, 0.0))
    rng = np.random.default_rng(12)
    us = rng.integers(0, 256, size=(8, 40, 48), dtype=np.uint8)
    fixed = synth.Image(us, (48, 40, 8), (0.5, 0.5, 0.5), (0.0, 0.0, 0.0))
    pose = np.r_[0.3, -0.2, 0.1, 1.0, -2.0, 0.5, 1, 1, 1, 0, 0, 0]
    w = synth.Workload("big-slices", fixed, moving, (0.5, 0.5, 0.5), 2, True, True, 3, pose)
    print("3D source 4096x4096x2 (f32), r=2: separable path", flush=True)
    p, vols = _pair(engine, w, 2, 3)
    pose2 = np.r_[-0.4, 0.3, -0.2, -1.5, 1.0, 2.0, 1.02, 0.98, 1.0, 0.01, 0.0, 0.0]
    mats = np.stack([p.pose_to_matrix(pose), p.pose_to_matrix(pose2)])
    batch = p.eval_batch(mats)
    for k, M in enumerate(mats):
        res = _oracle(w, M, 2, 3, 3)
        assert res.n_valid > 1000
        assert abs(batch.lc2[k] - res.lc2) < LC2_TOL, (k, batch.lc2[k], res.lc2)
    p.close()
    for v in vols:
        v.close()


def test_radius_limits(engine):
    w = synth.small3d()
    f, m = w.fixed, w.moving
    fv = engine.upload(f.data, f.spacing, f.origin)
    mv = engine.upload(m.data, m.spacing, m.origin)
    with pytest.raises(E.LC2Error) as ei:
        engine.pair(fv, mv, w.resampled_spacing, True, False, 65, 3)
    assert ei.value.code == -4
    # a radius no patch fits: every sum zero, LC2 0, degenerate status
    p = engine.pair(fv, mv, w.resampled_spacing, True, False, 30, 3)
    r = p.eval_batch(p.pose_to_matrix(w.true_pose)[None])
    assert r.status == 1 and r.lc2[0] == 0.0 and r.sum_w[0] == 0.0
    p.close()
    fv.close()
    mv.close()
