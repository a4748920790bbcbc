"""The serving loop (lc2serve.h) on the mixed serving workload at its own sizes: six
resident pairs in three request classes against shared resident volumes,
  2D affine 256x256 and 512x512 (u8 ultrasound warped onto an f32 image);
  2D-to-3D tracked sweeps of 32 and 64 frames 512x384 u8 against one MR 256x256x192
      stored f16 (both pairs share the MR volume);
  3D: US 192x192x160 u8 against MR 256x256x192 f32, and US 256^3 u8 against the int16
      CT 512x512x400 stored f16.
Requests of 1-8 candidate poses from four client threads: every result equals the
request's evaluation alone bit for bit, requests of a pair are merged. At these sizes one
request's launch fills the device, so launches run one after another; launches whose
work-groups fit the device together do overlap (second test, small pairs)."""
import threading
import time

import numpy as np
import pytest

from lc2registration_amd import engine as E
from lc2registration_amd import serve as S
from lc2registration_amd import synth

pytestmark = pytest.mark.gpu


def _mix(engine):
    t0 = time.time()
    keep, classes = [], []
    rng = np.random.default_rng(5)

    def add(name, p, mats, weight):
        classes.append((name, p, mats, weight))

    for seed, n in ((21, 256), (23, 512)):
        w = synth.config0(seed=seed, n=n)
        f, m = w.fixed, w.moving
        fv, mv = engine.upload(f.data, f.spacing, f.origin), engine.upload(m.data, m.spacing, m.origin)
        p = engine.pair(fv, mv, w.resampled_spacing, w.high_res_is_moving, w.center_overlaid,
                        w.patch_radius, 2)
        keep += [p, fv, mv]
        add(f"2d_affine_{n}", p, np.stack([p.pose_to_matrix(synth.perturb(w.true_pose, rng, 2))
                                           for _ in range(16)]), 0.2)
    mr = None
    for frames in (32, 64):
        w = synth.config2(seed=22, mr_dims=(256, 256, 192), n_frames=frames)
        if mr is None:
            mr = engine.upload(w.moving.data, w.moving.spacing, w.moving.origin, E.F16)
            keep.append(mr)
        f = w.fixed
        fv = engine.upload(f.data, f.spacing, f.origin)
        p = engine.pair(fv, mr, w.resampled_spacing, True, False, w.patch_radius, 2)
        p.set_per_frame(True)
        keep += [p, fv]
        mats = np.stack([synth.sweep_matrices(w.true_pose + np.r_[rng.uniform(-2, 2, 3),
                                                                 rng.uniform(-2, 2, 3)],
                                              w.extra["frames"], w.extra["centre"])
                         for _ in range(16)])
        add(f"sweep_{frames}_frames", p, mats, 0.2)
    for name, w in (("3d_mr_us192", synth.config1()), ("3d_ct_us256", synth.config3())):
        f, m = w.fixed, w.moving
        fv, mv = engine.upload(f.data, f.spacing, f.origin), engine.upload(m.data, m.spacing, m.origin)
        p = engine.pair(fv, mv, w.resampled_spacing, True, False, w.patch_radius, 3)
        keep += [p, fv, mv]
        add(name, p, np.stack([p.pose_to_matrix(synth.perturb(w.true_pose, rng, 3))
                               for _ in range(16)]), 0.1)
    print(f"serving mix resident after {time.time() - t0:.1f}s", flush=True)
    return classes, keep


def test_serving_mix_at_full_sizes_equals_alone(engine):
    classes, keep = _mix(engine)
    assert len(classes) == 6
    rng = np.random.default_rng(9)
    weights = np.array([c[3] for c in classes])
    reqs = []
    for k in range(48):
        c = int(rng.choice(len(classes), p=weights / weights.sum()))
        n = int(rng.integers(1, 9))
        reqs.append((c, classes[c][2][rng.integers(0, 16, n)]))
    # every class at least twice
    for c in range(len(classes)):
        reqs.append((c, classes[c][2][:3]))
        reqs.append((c, classes[c][2][3:4]))
    alone = [classes[c][1].eval_batch(m) for c, m in reqs]
    srv = S.Server(engine, max_poses=64, window_us=1000.0, slots=4)
    results = [None] * len(reqs)

    def client(ids):
        tickets = [(j, srv.submit(classes[reqs[j][0]][1], reqs[j][1])) for j in ids]
        for j, t in tickets:
            results[j] = srv.wait(t)

    threads = [threading.Thread(target=client, args=(list(range(c, len(reqs), 4)),))
               for c in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    st = srv.stats()
    srv.close()
    print("serving counters", st, flush=True)
    for j, (a, r) in enumerate(zip(alone, results)):
        assert r is not None and r.status == a.status == E.OK, j
        assert np.array_equal(r.lc2, a.lc2), (j, classes[reqs[j][0]][0])
        assert np.array_equal(r.sum_measure, a.sum_measure) and np.array_equal(r.sum_w, a.sum_w), j
    assert st["requests"] == len(reqs) and st["queued"] == 0
    assert st["poses"] == sum(len(m) for _, m in reqs)
    assert st["launches"] < len(reqs)
    for h in keep:
        h.close()


def test_small_launches_of_different_pairs_overlap(engine):
    """Four small 2D pairs (one pose = 36 work-groups): their launches fit the device
    together, so the worker keeps several on the device at once."""
    pairs, keep = [], []
    for seed in range(4):
        w = synth.config0(seed=30 + seed, n=96)
        f, m = w.fixed, w.moving
        fv, mv = engine.upload(f.data, f.spacing, f.origin), engine.upload(m.data, m.spacing, m.origin)
        p = engine.pair(fv, mv, w.resampled_spacing, w.high_res_is_moving, w.center_overlaid,
                        w.patch_radius, 2)
        keep += [p, fv, mv]
        pairs.append((p, p.pose_to_matrix(w.true_pose)[None]))
    alone = [p.eval_batch(M) for p, M in pairs]
    # a 20 ms batching window holds the four requests until all are queued
    srv = S.Server(engine, max_poses=64, window_us=20000.0, slots=4)
    tickets = [srv.submit(p, M) for p, M in pairs]
    res = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    print("serving counters", st, flush=True)
    for a, r in zip(alone, res):
        assert r.status == a.status and np.array_equal(r.lc2, a.lc2)
    assert st["max_in_flight"] >= 2
    for h in keep:
        h.close()
