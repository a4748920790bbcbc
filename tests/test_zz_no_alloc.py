"""The allocation-free entry point (lc2e_eval_batch_noalloc) against lc2e_eval_batch.

Both run the same fused kernels, so the results must agree bit for bit: on two calls in a
row (the second with fewer poses, so the device scratch still holds the first call's
larger contents), for a 3D pair (moving MR, static ultrasound), a 2D pair (moving
ultrasound) and a 3D pair whose patch radius takes the separable path. The engine's allocated bytes must not move during the calls, and a call
larger than the scratch array must fail cleanly before any device work.
"""
import numpy as np
import pytest

from lc2registration_amd import engine as E
from lc2registration_amd import synth

pytestmark = pytest.mark.gpu


def _pair(engine, w):
    f, m = w.fixed, w.moving
    fv = engine.upload(f.data, f.spacing, f.origin)
    mv = engine.upload(m.data, m.spacing, m.origin)
    p = engine.pair(fv, mv, w.resampled_spacing, w.high_res_is_moving, w.center_overlaid,
                    w.patch_radius, w.ndim)
    return p, (fv, mv)


def _mats(p, w, n, seed):
    rng = np.random.default_rng(seed)
    poses = [w.true_pose] + [synth.perturb(w.true_pose, rng, w.ndim) for _ in range(n - 1)]
    return np.stack([p.pose_to_matrix(q) for q in poses])


@pytest.mark.parametrize("case", ["3d", "2d", "3d_r5"])
def test_noalloc_matches_eval_batch_on_two_calls(engine, case):
    w = synth.small3d() if case.startswith("3d") else synth.config0(n=96)
    if case == "3d_r5":  # beyond the fused kernel: the separable path
        w.patch_radius = 5
    p, vols = _pair(engine, w)
    print(f"no-alloc case {case}: virtual grid {p.virt_dims}", flush=True)
    big, small = _mats(p, w, 7, 1), _mats(p, w, 3, 2)
    used0 = engine.memory()[0]
    a1 = p.eval_batch_noalloc(big)
    a2 = p.eval_batch_noalloc(small)
    assert engine.memory()[0] == used0  # nothing allocated on the engine's budget
    b1 = p.eval_batch(big)
    b2 = p.eval_batch(small)
    for a, b in ((a1, b1), (a2, b2)):
        assert a.status == b.status
        assert np.array_equal(a.lc2, b.lc2)
        assert np.array_equal(a.sum_measure, b.sum_measure)
        assert np.array_equal(a.sum_w, b.sum_w)
        assert (a.lc2 > 0).all()
    # the first pose of each call is the generating pose: same value in both calls
    assert a1.lc2[0] == a2.lc2[0]
    p.close()
    for v in vols:
        v.close()


def test_noalloc_rejects_calls_beyond_the_scratch(engine):
    w = synth.small3d()
    p, vols = _pair(engine, w)
    m = p.pose_to_matrix(w.true_pose)
    many = np.tile(m, (200000, 1))  # partial sums alone exceed the 16 MiB array
    with pytest.raises(E.LC2Error) as ei:
        p.eval_batch_noalloc(many)
    assert ei.value.code == -1 and "scratch" in str(ei.value)
    # the pair still evaluates afterwards
    assert p.eval_batch_noalloc(m[None]).lc2[0] == p.eval(m)
    p.close()
    for v in vols:
        v.close()
