"""GPU parity: the HIP engine (through the C ABI) against the oracle on the same inputs.

Bars: FP32 stages (warps, |Sobel|) bit-exact; per-voxel w / measure within 1e-9
relative (FP64 moment order differs); LC2 within 1e-9 absolute here, against the
1e-5 the workload requires.
"""
import numpy as np
import pytest

from lc2registration_amd import engine as E
from lc2registration_amd import geometry as geo
from lc2registration_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LC2_TOL = 1e-9


def _ogeom(img):
    return O.geom(img.dims, img.spacing, img.origin)


def _pair_for(engine, w, storage=E.STORE_AUTO, moving_data=None, fixed_data=None, r=None):
    f, m = w.fixed, w.moving
    fv = engine.upload(f.data if fixed_data is None else fixed_data, f.spacing, f.origin)
    mv = engine.upload(m.data if moving_data is None else moving_data, m.spacing, m.origin,
                       storage)
    return engine.pair(fv, mv, w.resampled_spacing, w.high_res_is_moving, w.center_overlaid,
                       r or w.patch_radius, w.ndim)


def _oracle_eval(w, M, r=None, moving_data=None, src_ndim=None, grid_ndim=None, maps=False):
    f, m = w.fixed, w.moving
    fg, mg = _ogeom(f), _ogeom(m)
    nd = len(f.dims)
    vd = O.virtual_dims(fg, w.resampled_spacing, nd)
    vg = O.geom(vd, w.resampled_spacing, f.origin)
    mov = (m.data if moving_data is None else moving_data).astype(np.float32)
    fix = f.data.astype(np.float32)
    src_ndim = src_ndim or len(m.dims)
    grid_ndim = grid_ndim or w.ndim
    if src_ndim == 3:
        Wf = O.warp_3d(fix, fg, vg, [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0])
        Wm = O.warp_3d(mov, mg, vg, np.asarray(M).ravel()[:12])
    else:
        Wf = O.warp_2d(fix, fg, vg, [1, 0, 0, 0, 1, 0])
        M3 = np.asarray(M).reshape(3, 4)
        Wm = O.warp_2d(mov, mg, vg, [M3[0, 0], M3[0, 1], M3[0, 3], M3[1, 0], M3[1, 1], M3[1, 3]])
    I, U = (Wm, Wf) if w.high_res_is_moving else (Wf, Wm)
    G = O.sobel(I, "sep")
    res = O.lc2_exact(U, I, G, r or w.patch_radius, maps=maps)
    return (res, I, U, G) if not maps else (res[0], I, U, G, res[1], res[2])


@pytest.fixture(scope="module")
def w3():
    return synth.small3d()


def test_kidney_2d_stages_and_lc2(engine, kidney):
    """Reference data: static CT warp bit-exact vs the reference's dump; US warp at the
    recovered pose bit-exact vs the oracle and within 1e-3 of the reference's dump;
    LC2 and per-voxel maps vs the oracle for the reference's 2D patch radii 3 and 9."""
    man = kidney["manifest"]
    ct = kidney["ct"]
    us = kidney["us"]
    fv = engine.upload(ct, man["fixed"]["spacing"], man["fixed"]["origin"])
    mv = engine.upload(us, man["moving"]["spacing"], man["moving"]["origin"])
    rec = man["rigid_dumps"]["moving.bin"]
    for r in (3, 9):
        p = engine.pair(fv, mv, (1.0, 1.0), False, True, r, 2)
        assert p.virt_dims[:2] == (164, 151)
        M = p.pose_to_matrix(rec["pose_tx_ty_rdeg"])
        wi = p.debug_fetch(E.STAGE_WARP_I, M)[0]
        assert np.array_equal(wi, kidney["fixed.bin"])
        wu = p.debug_fetch(E.STAGE_WARP_US, M)[0]
        vg = O.geom((164, 151), (1.0, 1.0), (0, 0))
        ow = O.warp_2d(us, O.geom(man["moving"]["dims"], man["moving"]["spacing"], (0, 0)), vg,
                       [M[0], M[1], M[3], M[4], M[5], M[7]])
        assert np.array_equal(wu, ow)
        assert np.abs(wu - kidney["moving.bin"]).max() <= 1.5e-3
        g = p.debug_fetch(E.STAGE_GRAD, M)[0]
        assert np.array_equal(g, O.sobel(wi, "sep"))
        res, wm, mm = O.lc2_exact(wu, wi, g, r, maps=True)
        l = p.eval(M)
        assert abs(l - res.lc2) < LC2_TOL, (r, l, res.lc2)
        gw = p.debug_fetch(E.STAGE_W, M)[0]
        gm = p.debug_fetch(E.STAGE_MEASURE, M)[0]
        assert np.array_equal(gw > 0, wm > 0)
        assert np.allclose(gw, wm, rtol=1e-9, atol=1e-12)
        assert np.allclose(gm, mm, rtol=1e-7, atol=1e-9)
        p.close()
    fv.close()
    mv.close()


@pytest.mark.parametrize("r", [1, 2, 3, 4])
def test_3d_moving_highres_parity(engine, w3, r):
    p = _pair_for(engine, w3, r=r)
    rng = np.random.default_rng(r)
    poses = [w3.true_pose] + [synth.perturb(w3.true_pose, rng, 3) for _ in range(2)]
    for pose in poses:
        M = p.pose_to_matrix(pose)
        res, I, U, G = _oracle_eval(w3, M, r=r)
        assert np.array_equal(p.debug_fetch(E.STAGE_WARP_I, M), I)
        assert np.array_equal(p.debug_fetch(E.STAGE_WARP_US, M), U)
        assert np.array_equal(p.debug_fetch(E.STAGE_GRAD, M), G)
        l = p.eval(M)
        assert res.n_valid > 1000
        assert abs(l - res.lc2) < LC2_TOL, (r, l, res.lc2)
    p.close()


@pytest.mark.parametrize("spacing", [0.8, 0.7, 1.93696105])
def test_3d_warp_bit_exact_at_non_power_of_two_spacing(engine, w3, spacing):
    """(p - origin) / spacing must be the reference's rounded division: at spacings that are
    not powers of two the fused kernels' samplers divide (the kernel instantiation without
    the power-of-two product). The warp and |Sobel| are bit-exact against the oracle, and
    LC2 within 1e-9."""
    m = w3.moving
    mv = synth.Image(m.data, m.dims, (spacing, spacing * 1.1, spacing * 0.9), m.origin)
    w = synth.Workload("spacing", w3.fixed, mv, w3.resampled_spacing, 2, True, True, 3,
                       w3.true_pose)
    print(f"3D warp at MR spacing {spacing}", flush=True)
    p = _pair_for(engine, w)
    rng = np.random.default_rng(21)
    for pose in [w3.true_pose, synth.perturb(w3.true_pose, rng, 3)]:
        M = p.pose_to_matrix(pose)
        res, I, U, G = _oracle_eval(w, M)
        assert np.array_equal(p.debug_fetch(E.STAGE_WARP_I, M), I)
        assert np.array_equal(p.debug_fetch(E.STAGE_GRAD, M), G)
        assert res.n_valid > 1000
        assert abs(p.eval(M) - res.lc2) < LC2_TOL
    p.close()


def test_3d_maps_match_oracle(engine, w3):
    p = _pair_for(engine, w3)
    M = p.pose_to_matrix(w3.true_pose)
    res, I, U, G, wm, mm = _oracle_eval(w3, M, maps=True)
    gw = p.debug_fetch(E.STAGE_W, M)
    gm = p.debug_fetch(E.STAGE_MEASURE, M)
    assert np.array_equal(gw > 0, wm > 0)
    assert np.allclose(gw, wm, rtol=1e-9, atol=1e-12)
    assert np.allclose(gm, mm, rtol=1e-7, atol=1e-9)
    p.close()


def test_3d_fixed_highres_parity(engine, w3):
    """HighResolutionModality: Fixed in 3D -- the fixed MR is I (static |Sobel|), the
    moving US is warped per pose."""
    f, m = w3.moving, w3.fixed  # roles swapped
    w = synth.Workload("swap", f, m, (1.0, 1.0, 1.0), 2, False, True, 3, w3.true_pose)
    p = _pair_for(engine, w)
    rng = np.random.default_rng(9)
    for pose in [np.r_[0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0], synth.perturb(np.r_[0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0], rng, 3)]:
        M = p.pose_to_matrix(pose)
        res, I, U, G = _oracle_eval(w, M)
        assert np.array_equal(p.debug_fetch(E.STAGE_WARP_US, M), U)
        assert np.array_equal(p.debug_fetch(E.STAGE_GRAD, M), G)
        l = p.eval(M)
        assert res.n_valid > 100
        assert abs(l - res.lc2) < LC2_TOL
    p.close()


def test_2d_affine_cfg0_parity(engine):
    w = synth.config0(n=96, patch_radius=3)
    p = _pair_for(engine, w)
    rng = np.random.default_rng(1)
    for pose in [w.true_pose, synth.perturb(w.true_pose, rng, 2)]:
        M = p.pose_to_matrix(pose)
        assert np.allclose(M.reshape(3, 4), geo.embed_2d(geo.pose_to_matrix_2d(
            pose, w.fixed.dims, w.fixed.spacing, w.fixed.origin, w.moving.dims,
            w.moving.spacing, w.moving.origin, w.resampled_spacing)), atol=1e-5)
        res, I, U, G = _oracle_eval(w, M)
        assert np.array_equal(p.debug_fetch(E.STAGE_WARP_US, M)[0], U)
        l = p.eval(M)
        assert abs(l - res.lc2) < LC2_TOL
    p.close()


def test_2d_to_3d_slice_parity(engine, w3):
    """2D->3D (Interpolator2DTo3D): a one-slice fixed US, the 3D MR moving, 2D LC2."""
    us = w3.fixed
    sl = us.data[us.data.shape[0] // 2][None].copy()
    fixed = synth.Image(sl, (us.dims[0], us.dims[1], 1), us.spacing,
                        (us.origin[0], us.origin[1], us.origin[2] + us.spacing[2] * (us.dims[2] // 2)))
    w = synth.Workload("2dto3d", fixed, w3.moving, us.spacing, 3, True, False, 2, w3.true_pose)
    p = _pair_for(engine, w)
    assert p.info.grid_ndim == 2 and p.info.src_ndim == 3
    M = p.pose_to_matrix(w3.true_pose)
    res, I, U, G = _oracle_eval(w, M, src_ndim=3, grid_ndim=2)
    # the oracle's 2D LC2 works on (Y, X) planes
    res2 = O.lc2_exact(U[0], I[0], O.sobel(I[0], "sep"), 3)
    assert np.array_equal(p.debug_fetch(E.STAGE_WARP_I, M), I)
    l = p.eval(M)
    assert res2.n_valid > 100
    assert abs(l - res2.lc2) < LC2_TOL
    p.close()


def test_frame_stack_per_frame_matrices(engine, w3):
    """A stack of tracked frames with one matrix per frame: sums equal the per-frame sums."""
    us = w3.fixed
    zs = [8, 14, 20]
    stack = np.ascontiguousarray(us.data[zs])
    fv = engine.upload(stack, us.spacing, us.origin)
    mv = engine.upload(w3.moving.data, w3.moving.spacing, w3.moving.origin)
    p = engine.pair(fv, mv, us.spacing, True, False, 2, 2)
    p.set_per_frame(True)
    base = p.pose_to_matrix(w3.true_pose).reshape(3, 4).astype(np.float64)
    mats = []
    for z in zs:  # frame z lives at height origin_z + z*spacing_z: fold into its matrix
        T = np.eye(4)
        T[2, 3] = z * us.spacing[2]
        mats.append((np.vstack([base, [0, 0, 0, 1]]) @ T)[:3].astype(np.float32))
    mats = np.stack(mats)
    br = p.eval_batch(mats[None])
    sm = sw = 0.0
    for k, z in enumerate(zs):
        fk = engine.upload(np.ascontiguousarray(us.data[z][None]), us.spacing, us.origin)
        pk = engine.pair(fk, mv, us.spacing, True, False, 2, 2)
        r = pk.eval_batch(mats[k][None])
        sm += r.sum_measure[0]
        sw += r.sum_w[0]
        pk.close()
        fk.close()
    assert abs(br.sum_measure[0] - sm) < 1e-9 * max(1.0, abs(sm))
    assert abs(br.sum_w[0] - sw) < 1e-9 * max(1.0, abs(sw))
    p.close()


def test_batch_equals_single_and_deterministic(engine, w3):
    p = _pair_for(engine, w3)
    rng = np.random.default_rng(5)
    mats = np.stack([p.pose_to_matrix(synth.perturb(w3.true_pose, rng, 3)) for _ in range(7)])
    b = p.eval_batch(mats)
    singles = np.array([p.eval(m) for m in mats])
    assert np.array_equal(b.lc2, singles)
    b2 = p.eval_batch(mats)
    assert np.array_equal(b.lc2, b2.lc2)
    p.close()


def test_fp16_storage_matches_oracle_on_rounded_input(engine, w3):
    m16 = w3.moving.data.astype(np.float16)
    p = _pair_for(engine, w3, moving_data=m16)
    M = p.pose_to_matrix(w3.true_pose)
    res, I, U, G = _oracle_eval(w3, M, moving_data=m16.astype(np.float32))
    assert abs(p.eval(M) - res.lc2) < LC2_TOL
    p.close()


def test_degenerate_constant_us(engine, w3):
    w = synth.Workload("const", synth.Image(np.full_like(w3.fixed.data, 7), w3.fixed.dims,
                                             w3.fixed.spacing, w3.fixed.origin), w3.moving,
                       w3.resampled_spacing, 2, True, False, 3, w3.true_pose)
    p = _pair_for(engine, w)
    M = p.pose_to_matrix(w3.true_pose)
    b = p.eval_batch(M[None])
    assert b.status == E.OK_DEGENERATE and b.lc2[0] == 0.0 and b.sum_w[0] == 0.0
    p.close()


def test_upload_rejects_f32_volume_beyond_32_bit_offsets(engine):
    """An f32 volume of 2^30 voxels would need byte offsets beyond 32 bits in the samplers:
    refused before any allocation or read of the host buffer."""
    import ctypes as C
    g = E.make_geom((1024, 1024, 1024), (1, 1, 1), (0, 0, 0))
    h = C.c_void_p()
    small = np.zeros(16, np.float32)
    rc = E.lib().lc2e_volume_upload(engine.h, small.ctypes.data_as(C.c_void_p), E.F32,
                                    C.byref(g), E.F32, C.byref(h))
    assert E.ERR.get(rc) == "UNSUPPORTED" and not h.value
    assert b"F16" in E.lib().lc2e_last_error()


def test_upload_rejects_non_finite_voxels(engine):
    """The samplers rely on finite voxels (a corner of weight 0 contributes exactly 0):
    NaN or infinite values are refused at upload, as f32 or half input, before any
    allocation."""
    used0, _ = engine.memory()
    for bad in (np.nan, np.inf, -np.inf):
        vol = np.ones((6, 5, 4), np.float32)
        vol[2, 3, 1] = bad
        for dt in (np.float32, np.float16):
            with pytest.raises(E.LC2Error) as ei:
                engine.upload(vol.astype(dt), (1, 1, 1), (0, 0, 0))
            assert ei.value.code == -1 and b"finite" in E.lib().lc2e_last_error()
    assert engine.memory()[0] == used0
    ok = engine.upload(np.ones((6, 5, 4), np.float32), (1, 1, 1), (0, 0, 0))
    ok.close()


def test_budget_rejects_oversized_pair(w3):
    e = E.Engine(0, hbm_budget=200_000)
    with pytest.raises(E.LC2Error) as ei:
        e.upload(w3.moving.data, w3.moving.spacing, w3.moving.origin)
    assert ei.value.code == -3
    e.close()


def test_graph_replay_matches_eager(engine, w3):
    p = _pair_for(engine, w3)
    rng = np.random.default_rng(8)
    mats = np.stack([p.pose_to_matrix(synth.perturb(w3.true_pose, rng, 3)) for _ in range(4)])
    eager = p.eval_batch(mats).lc2
    p.use_graph(True)
    for k in range(3):
        g = p.eval_batch(mats).lc2
        assert np.array_equal(g, eager), k
        assert p.eval(mats[k]) == eager[k]  # other work between replays
    # kernel timing with replay on: timed launches run eagerly, replay resumes afterwards
    p.profile(True)
    for _ in range(3):
        assert np.array_equal(p.eval_batch(mats).lc2, eager)
    ms, n = p.profile_read()
    p.profile(False)
    assert n == 3 and ms > 0.0
    assert np.array_equal(p.eval_batch(mats).lc2, eager)
    p.use_graph(False)
    p.close()


def test_zchunk_and_batch_size_do_not_change_result(engine, w3):
    """The z-chunk (rounded up to a multiple of 8 planes) and the batch size change how the
    3D grid is split over work-groups, never the value: LC2 is bit-identical."""
    p = _pair_for(engine, w3)
    rng = np.random.default_rng(11)
    mats = np.stack([p.pose_to_matrix(w3.true_pose)] +
                    [p.pose_to_matrix(synth.perturb(w3.true_pose, rng, 3)) for _ in range(4)])
    ref = _oracle_eval(w3, mats[0])[0].lc2
    base = p.eval_batch(mats).lc2
    assert abs(base[0] - ref) < LC2_TOL
    for zchunk, expect in ((4, 8), (9, 16), (32, 32), (1000, 1000 + (-1000) % 8)):
        p.set_tuning(zchunk)
        assert p.get_tuning() == expect
        assert np.array_equal(p.eval_batch(mats).lc2, base), zchunk
        assert p.eval(mats[2]) == base[2]
    p.set_tuning(0)
    p.set_tuning_n(5, 16)
    assert np.array_equal(p.eval_batch(mats).lc2, base)
    assert p.get_tuning() == 16
    single = [p.eval(m) for m in mats]     # single poses: the model's chunk
    assert np.array_equal(np.array(single), base)
    p.close()


def test_tracked_sweep_2d_to_3d_matches_oracle(engine):
    """A small tracked sweep (per-frame matrices, MR stored f16) stage by stage against the
    oracle's slice warp (f16-rounded MR), 2D Sobel and exact LC2, frame by frame."""
    w = synth.config2(seed=3, mr_dims=(96, 96, 80), n_frames=5, frame_dims=(96, 64),
                      frame_spacing=0.5, sweep_mm=20.0, patch_radius=3)
    f, m = w.fixed, w.moving
    fv = engine.upload(f.data, f.spacing, f.origin)
    mv = engine.upload(m.data, m.spacing, m.origin, E.F16)
    p = engine.pair(fv, mv, w.resampled_spacing, True, False, 3, 2)
    p.set_per_frame(True)
    mats = synth.sweep_matrices(w.true_pose, w.extra["frames"], w.extra["centre"])
    wi = p.debug_fetch(E.STAGE_WARP_I, mats)
    gr = p.debug_fetch(E.STAGE_GRAD, mats)
    wu = p.debug_fetch(E.STAGE_WARP_US, mats)
    wm_gpu = p.debug_fetch(E.STAGE_W, mats)
    mm_gpu = p.debug_fetch(E.STAGE_MEASURE, mats)
    r = p.eval_batch(mats[None])
    mr16 = m.data.astype(np.float16).astype(np.float32)
    mg = O.geom(m.dims, m.spacing, m.origin)
    X, Y = f.dims[0], f.dims[1]
    vg = O.geom((X, Y, 1), (f.spacing[0], f.spacing[1], 1.0), (0, 0, 0))
    sm = sw = 0.0
    for k in range(len(mats)):
        I = O.warp_3d(mr16, mg, vg, mats[k].ravel())[0]
        U = f.data[k].astype(np.float32)
        assert np.array_equal(wi[k], I), k
        assert np.array_equal(wu[k], U), k
        G = O.sobel(I, "sep")
        assert np.array_equal(gr[k], G), k
        res, wm, mm = O.lc2_exact(U, I, G, 3, maps=True)
        assert np.array_equal(wm_gpu[k] > 0, wm > 0), k
        assert np.allclose(wm_gpu[k], wm, rtol=1e-9, atol=1e-12), k
        assert np.allclose(mm_gpu[k], mm, rtol=1e-6, atol=1e-7), (k, np.abs(mm_gpu[k] - mm).max())
        sm += res.sum_measure
        sw += res.sum_w
    assert sw > 0
    assert abs(r.sum_w[0] - sw) < 1e-9 * sw
    assert abs(r.sum_measure[0] - sm) < 1e-8 * sm
    assert abs(r.lc2[0] - sm / sw) < 1e-8
    p.close()
    fv.close()
    mv.close()


@pytest.fixture(scope="module")
def wct():
    return synth.config3(ct_dims=(72, 72, 60), ct_spacing=2.0, us_dims=(40, 40, 36),
                         us_spacing=1.5, patch_radius=2)


def test_ct_int16_half_storage_is_lossless(engine, wct):
    """CT in HU (int16 within [-2048, 2048]) goes to half storage under LC2E_STORE_AUTO,
    which is exact: the warp is bit-exact against the oracle on the int16 values, LC2 is
    bit-identical to f32 storage and within 1e-9 of the FP64 oracle. A CT value outside
    that range keeps f32 storage."""
    ct, us = wct.moving, wct.fixed
    assert ct.data.dtype == np.int16 and ct.data.max() <= 2048 and ct.data.min() >= -2048
    n = int(np.prod(ct.dims))
    u0, _ = engine.memory()
    mv16 = engine.upload(ct.data, ct.spacing, ct.origin)             # AUTO
    u1, _ = engine.memory()
    mv32 = engine.upload(ct.data, ct.spacing, ct.origin, E.F32)
    u2, _ = engine.memory()
    pad = ct.dims[0] * ct.dims[1] + ct.dims[0] + 64                   # zeroed tail, in voxels
    assert (u2 - u1) - (u1 - u0) == 2 * (n + pad)                    # 2 B vs 4 B per voxel
    hot = ct.data.copy()
    hot.flat[n // 2] = 3000
    mvh = engine.upload(hot, ct.spacing, ct.origin)                  # AUTO, out of range
    u3, _ = engine.memory()
    assert u3 - u2 == u2 - u1
    mvh.close()
    fv = engine.upload(us.data, us.spacing, us.origin)
    args = (wct.resampled_spacing, True, False, wct.patch_radius, 3)
    p16, p32 = engine.pair(fv, mv16, *args), engine.pair(fv, mv32, *args)
    rng = np.random.default_rng(30)
    poses = [wct.true_pose] + [synth.perturb(wct.true_pose, rng, 3) for _ in range(3)]
    mats = np.stack([p16.pose_to_matrix(q) for q in poses])
    r16, r32 = p16.eval_batch(mats), p32.eval_batch(mats)
    assert np.array_equal(r16.lc2, r32.lc2)
    assert np.array_equal(r16.sum_measure, r32.sum_measure)
    for k in range(len(mats)):
        res, I, U, G = _oracle_eval(wct, mats[k])
        assert np.array_equal(p16.debug_fetch(E.STAGE_WARP_I, mats[k]), I)
        assert res.n_valid > 1000
        assert abs(r16.lc2[k] - res.lc2) < LC2_TOL, (k, r16.lc2[k], res.lc2)
    for h in (p16, p32, fv, mv16, mv32):
        h.close()


@pytest.mark.parametrize("r", [1, 2, 4, 5, 6, 7, 10])
def test_2d_radius_classes_match_oracle(engine, w3, r):
    """Every radius class of the 2D kernel (r <= 3, <= 6, <= 10), in both roles: the
    ultrasound warped onto a static image (2D source) and the MR warped onto a static
    ultrasound frame (3D source)."""
    w = synth.config0(n=96, patch_radius=r)
    p = _pair_for(engine, w)
    M = p.pose_to_matrix(w.true_pose)
    res, I, U, G = _oracle_eval(w, M)
    assert res.n_valid > 100
    assert abs(p.eval(M) - res.lc2) < LC2_TOL, r
    p.close()
    us = w3.fixed
    sl = us.data[us.data.shape[0] // 2][None].copy()
    fixed = synth.Image(sl, (us.dims[0], us.dims[1], 1), us.spacing,
                        (us.origin[0], us.origin[1], us.origin[2] + us.spacing[2] * (us.dims[2] // 2)))
    ws = synth.Workload("2dto3d", fixed, w3.moving, us.spacing, r, True, False, 2, w3.true_pose)
    p = _pair_for(engine, ws)
    M = p.pose_to_matrix(w3.true_pose)
    res, I, U, G = _oracle_eval(ws, M, src_ndim=3, grid_ndim=2)
    res2 = O.lc2_exact(U[0], I[0], O.sobel(I[0], "sep"), r)
    assert np.array_equal(p.debug_fetch(E.STAGE_WARP_I, M), I)
    if res2.n_valid:
        assert abs(p.eval(M) - res2.lc2) < LC2_TOL, r
    p.close()


def test_autotune_keeps_results_and_caches(engine, w3, tmp_path):
    """The z-chunk autotuner picks a multiple of 8 planes, leaves LC2 bit-identical, and its
    JSON cache answers the second call without timing."""
    from lc2registration_amd import autotune
    p = _pair_for(engine, w3)
    rng = np.random.default_rng(12)
    mats = np.stack([p.pose_to_matrix(synth.perturb(w3.true_pose, rng, 3)) for _ in range(6)])
    before = p.eval_batch(mats).lc2
    cache = str(tmp_path / "tune.json")
    z = autotune.tune(p, mats, iters=2, path=cache)
    assert z > 0 and z % 8 == 0
    assert np.array_equal(p.eval_batch(mats).lc2, before)
    assert p.get_tuning() == z            # the chunk the last launch used
    entry = autotune.load(cache)
    assert len(entry) == 1 and list(entry.values())[0]["zchunk"] == z
    p2 = _pair_for(engine, w3)
    assert autotune.tune(p2, mats, iters=2, path=cache) == z      # from the cache
    assert np.array_equal(p2.eval_batch(mats).lc2, before)
    assert p2.get_tuning() == z
    p.close()
    p2.close()


# --- edge cases of the tiling --------------------------------------------------------------

@pytest.mark.parametrize("moving_high_res", [True, False])
def test_3d_ragged_grid_matches_oracle(engine, moving_high_res):
    """A grid far from the tile and chunk sizes: 29 x 13 x 7 patch centres at r = 3 (one
    partial 32 x 16 tile column, fewer planes than one partial-sum block), in both roles."""
    w = synth.config1(seed=11, mr_dims=(40, 36, 30), us_dims=(35, 19, 13), us_spacing=1.0,
                      patch_radius=3)
    if not moving_high_res:
        w = synth.Workload("ragged_swap", w.moving, w.fixed, (1.0, 1.0, 1.0), 3, False, True, 3,
                           np.r_[0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0])
    p = _pair_for(engine, w)
    if moving_high_res:
        assert tuple(p.virt_dims) == (35, 19, 13)
    rng = np.random.default_rng(4)
    poses = [w.true_pose, synth.perturb(w.true_pose, rng, 3, 0.5)]
    mats = np.stack([p.pose_to_matrix(q) for q in poses])
    b = p.eval_batch(mats)
    for k, M in enumerate(mats):
        res = _oracle_eval(w, M)[0]
        assert res.n_valid > 50
        assert abs(b.lc2[k] - res.lc2) < LC2_TOL, (k, b.lc2[k], res.lc2)
        assert abs(b.sum_w[k] - res.sum_w) < 1e-9 * res.sum_w
    p.close()


def test_grids_without_a_patch_centre_are_degenerate(engine):
    """No patch fits the grid (3D: 4 planes at r = 2; 2D: 4 columns at r = 2): every sum is
    0, LC2 is 0 and the status says degenerate, for one pose and for a batch."""
    rng = np.random.default_rng(2)
    M = np.eye(3, 4, dtype=np.float32).ravel()  # identity pose
    for us_shape, mr_shape, dim in (((4, 12, 12), (20, 20, 20), 3), ((30, 4), (20, 20), 2)):
        sp, org = (1.0,) * dim, (0.0,) * dim
        us = rng.integers(1, 255, us_shape).astype(np.uint8)
        mr = rng.uniform(1, 100, mr_shape).astype(np.float32)
        fv = engine.upload(us, sp, org)
        mv = engine.upload(mr, sp, org)
        p = engine.pair(fv, mv, sp, True, False, 2, dim)
        b = p.eval_batch(np.stack([M, M, M]))
        assert b.status == E.OK_DEGENERATE
        assert np.all(b.lc2 == 0.0) and np.all(b.sum_w == 0.0) and np.all(b.sum_measure == 0.0)
        assert p.eval(M) == 0.0
        p.close()
        fv.close()
        mv.close()


def test_sweep_with_partial_last_frame_run(engine):
    """18 tracked frames: one full run of 16 frames and a partial run of 2. The sums equal
    the oracle's per-frame sums; one pose alone equals the same pose in a batch. At 0.8 mm
    the identity resampling of a frame into the virtual grid is not exact (x * 0.8 / 0.8
    rounds), so U is the oracle's resampled frame, as in the reference, not the raw one."""
    w = synth.config2(seed=5, mr_dims=(64, 64, 56), n_frames=18, frame_dims=(40, 24),
                      frame_spacing=0.8, sweep_mm=24.0, patch_radius=2)
    f, m = w.fixed, w.moving
    fv = engine.upload(f.data, f.spacing, f.origin)
    mv = engine.upload(m.data, m.spacing, m.origin, E.F16)
    p = engine.pair(fv, mv, w.resampled_spacing, True, False, 2, 2)
    p.set_per_frame(True)
    rng = np.random.default_rng(6)
    poses = [w.true_pose, w.true_pose + np.r_[rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)]]
    mats = np.stack([synth.sweep_matrices(q, w.extra["frames"], w.extra["centre"]) for q in poses])
    b = p.eval_batch(mats)
    mr16 = m.data.astype(np.float16).astype(np.float32)
    mg = O.geom(m.dims, m.spacing, m.origin)
    X, Y = f.dims[0], f.dims[1]
    vg = O.geom((X, Y, 1), (f.spacing[0], f.spacing[1], 1.0), (0, 0, 0))
    fg = O.geom((X, Y, 1), (f.spacing[0], f.spacing[1], 1.0), (f.origin[0], f.origin[1], 0))
    Us = [O.warp_2d(f.data[k].astype(np.float32), fg, vg, [1, 0, 0, 0, 1, 0]) for k in range(18)]
    assert np.array_equal(p.debug_fetch(E.STAGE_WARP_US, mats[0]), np.stack(Us))
    for i in range(len(poses)):
        sm = sw = 0.0
        for k in range(18):
            I = O.warp_3d(mr16, mg, vg, mats[i, k].ravel())[0]
            res = O.lc2_exact(Us[k], I, O.sobel(I, "sep"), 2)
            sm += res.sum_measure
            sw += res.sum_w
        assert sw > 0
        assert abs(b.sum_w[i] - sw) < 1e-9 * sw, i
        assert abs(b.sum_measure[i] - sm) < 1e-8 * sm, i
        one = p.eval_batch(mats[i][None])
        assert one.lc2[0] == b.lc2[i] and one.sum_w[0] == b.sum_w[i]
    p.close()
    fv.close()
    mv.close()
