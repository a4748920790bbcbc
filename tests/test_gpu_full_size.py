"""GPU checks at the workloads' full sizes, through properties that do not depend on size
and through the oracle on a slab of output planes:

- 3D affine: MR 256x256x192 f32 (moving) against US 192x192x160 u8 (fixed grid), 7^3
  patch. The w and measure maps of a slab of output planes match the oracle; LC2 is
  bit-identical between a 32-pose batch, single evaluations and other z-chunks.
- Tracked sweep: 128 frames of 512x384 u8 against MR 320x320x256 stored f16, 7x7 patch,
  one matrix per frame. Two frames' maps match the oracle; a pose alone equals the same
  pose inside a 16-pose batch.
- CT: int16 CT 512x512x400 (HU) against US 256^3, 7^3 patch: f16 storage (lossless for
  integers in [-2048, 2048]) gives LC2 bit-identical to f32 storage.
"""
import numpy as np
import pytest

from lc2registration_amd import engine as E
from lc2registration_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def test_3d_affine_full_size(engine):
    w = synth.config1()
    f, m = w.fixed, w.moving
    fv = engine.upload(f.data, f.spacing, f.origin)
    mv = engine.upload(m.data, m.spacing, m.origin)
    p = engine.pair(fv, mv, w.resampled_spacing, True, False, w.patch_radius, 3)
    assert tuple(p.virt_dims) == (192, 192, 160)
    rng = np.random.default_rng(21)
    poses = [w.true_pose] + [synth.perturb(w.true_pose, rng, 3) for _ in range(31)]
    mats = np.stack([p.pose_to_matrix(q) for q in poses])
    base = p.eval_batch(mats)
    _log(f"3D full size: LC2 at the true pose {base.lc2[0]:.6f}")
    assert base.status == E.OK
    assert base.lc2[0] > np.median(base.lc2[1:])  # the generating pose scores high

    # oracle on a slab of 6 output planes around the middle, for the true pose and one other
    r = w.patch_radius
    fg, mg = O.geom(f.dims, f.spacing, f.origin), O.geom(m.dims, m.spacing, m.origin)
    vg = O.geom((192, 192, 160), w.resampled_spacing, f.origin)
    U = O.warp_3d(f.data.astype(np.float32), fg, vg, [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0])
    zlo, zhi = 77, 83
    for k in (0, 5):
        I = O.warp_3d(m.data, mg, vg, mats[k])
        G = O.sobel(I, "sep")
        _, wm, mm = O.lc2_exact(U, I, G, r, zlo=zlo, zhi=zhi, maps=True)
        gw = p.debug_fetch(E.STAGE_W, mats[k])[zlo:zhi]
        gm = p.debug_fetch(E.STAGE_MEASURE, mats[k])[zlo:zhi]
        wm, mm = wm[zlo:zhi], mm[zlo:zhi]
        assert np.count_nonzero(wm) > 100000
        assert np.array_equal(gw > 0, wm > 0), k
        assert np.allclose(gw, wm, rtol=1e-9, atol=1e-12), k
        assert np.allclose(gm, mm, rtol=1e-7, atol=1e-9), (k, np.abs(gm - mm).max())

    # the decomposition never changes the value
    for i in (0, 7, 31):
        assert p.eval(mats[i]) == base.lc2[i], i
    for zchunk in (16, 48):
        p.set_tuning(zchunk)
        assert np.array_equal(p.eval_batch(mats[:8]).lc2, base.lc2[:8]), zchunk
    p.set_tuning(0)
    assert np.array_equal(p.eval_batch(mats).lc2, base.lc2)  # and it is deterministic
    p.close()
    fv.close()
    mv.close()


def test_tracked_sweep_full_size(engine):
    w = synth.config2()
    f, m = w.fixed, w.moving
    fv = engine.upload(f.data, f.spacing, f.origin)
    mv = engine.upload(m.data, m.spacing, m.origin, E.F16)
    p = engine.pair(fv, mv, w.resampled_spacing, True, False, w.patch_radius, 2)
    p.set_per_frame(True)
    rng = np.random.default_rng(22)
    poses = [w.true_pose] + [w.true_pose + np.r_[rng.uniform(-2, 2, 3), rng.uniform(-2, 2, 3)]
                             for _ in range(15)]
    mats = np.stack([synth.sweep_matrices(q, w.extra["frames"], w.extra["centre"]) for q in poses])
    assert mats.shape[:2] == (16, 128) and mats[0, 0].size == 12
    b = p.eval_batch(mats)
    _log(f"sweep full size: LC2 at the true pose {b.lc2[0]:.6f}")
    assert b.status == E.OK
    for i in (0, 9):
        one = p.eval_batch(mats[i][None])
        assert one.lc2[0] == b.lc2[i] and one.sum_w[0] == b.sum_w[i], i

    # oracle on two frames (the first and one inside the second frame run)
    X, Y = f.dims[0], f.dims[1]
    mr16 = m.data.astype(np.float16).astype(np.float32)
    mg = O.geom(m.dims, m.spacing, m.origin)
    vg = O.geom((X, Y, 1), (f.spacing[0], f.spacing[1], 1.0), (0, 0, 0))
    fg = O.geom((X, Y, 1), (f.spacing[0], f.spacing[1], 1.0), (f.origin[0], f.origin[1], 0))
    gw = p.debug_fetch(E.STAGE_W, mats[3])
    gm = p.debug_fetch(E.STAGE_MEASURE, mats[3])
    for k in (0, 21):
        U = O.warp_2d(f.data[k].astype(np.float32), fg, vg, [1, 0, 0, 0, 1, 0])
        I = O.warp_3d(mr16, mg, vg, mats[3, k].ravel())[0]
        _, wm, mm = O.lc2_exact(U, I, O.sobel(I, "sep"), w.patch_radius, maps=True)
        assert np.count_nonzero(wm) > 10000, k
        assert np.array_equal(gw[k] > 0, wm > 0), k
        assert np.allclose(gw[k], wm, rtol=1e-9, atol=1e-12), k
        assert np.allclose(gm[k], mm, rtol=1e-6, atol=1e-7), (k, np.abs(gm[k] - mm).max())
    p.close()
    fv.close()
    mv.close()


def test_ct_full_size_half_storage_is_bit_identical(engine):
    w = synth.config3()
    f, m = w.fixed, w.moving
    assert m.data.dtype == np.int16 and m.data.min() >= -2048 and m.data.max() <= 2048
    fv = engine.upload(f.data, f.spacing, f.origin)
    out = {}
    for storage in (E.F16, E.F32):
        mv = engine.upload(m.data, m.spacing, m.origin, storage)
        p = engine.pair(fv, mv, w.resampled_spacing, True, False, w.patch_radius, 3)
        rng = np.random.default_rng(23)
        mats = np.stack([p.pose_to_matrix(w.true_pose)] +
                        [p.pose_to_matrix(synth.perturb(w.true_pose, rng, 3)) for _ in range(7)])
        out[storage] = p.eval_batch(mats)
        p.close()
        mv.close()
    _log(f"CT full size: LC2 at the true pose {out[E.F16].lc2[0]:.6f}")
    assert out[E.F16].status == E.OK
    assert np.array_equal(out[E.F16].lc2, out[E.F32].lc2)
    assert np.array_equal(out[E.F16].sum_w, out[E.F32].sum_w)
    fv.close()
