"""Whole registrations as serving requests (lc2s_registrar, lc2serve.h): 2D affine, 3D rigid
and tracked-sweep (2D-to-3D, one matrix per frame) registrations submitted together end at
the same pose, LC2 and evaluation count, bit for bit, as each run alone; admission by HBM
budget runs registrations whose pairs do not fit together one after another, with the
same results."""
import numpy as np
import pytest

from lc2registration_amd import engine as E
from lc2registration_amd import serve as S
from lc2registration_amd import synth

pytestmark = pytest.mark.gpu


def _jobs(engine):
    """(volumes to close, [(name, submit kwargs, truth)])."""
    keep, jobs = [], []
    rng = np.random.default_rng(17)
    for seed in (4, 5):
        w = synth.config0(seed=seed)
        f, m = w.fixed, w.moving
        fv, mv = engine.upload(f.data, f.spacing, f.origin), engine.upload(m.data, m.spacing, m.origin)
        keep += [fv, mv]
        x0 = w.true_pose + np.r_[rng.uniform(-2, 2, 3), rng.uniform(-0.02, 0.02, 2), 0.0]
        half = np.r_[5, 5, 5, 0.08, 0.08, 0.05]
        jobs.append((f"2d_affine_256_{seed}",
                     dict(fixed=fv, moving=mv, resampled_spacing=w.resampled_spacing,
                          high_res_is_moving=w.high_res_is_moving, center_overlaid=w.center_overlaid,
                          patch_radius=w.patch_radius, dim=2, x0=x0, lower=x0 - half,
                          upper=x0 + half), w.true_pose))
    w = synth.small3d(seed=6)
    f, m = w.fixed, w.moving
    fv, mv = engine.upload(f.data, f.spacing, f.origin), engine.upload(m.data, m.spacing, m.origin)
    keep += [fv, mv]
    truth = np.asarray(w.true_pose, np.float64)[:6]
    x0 = truth + rng.uniform(-1.5, 1.5, 6)
    jobs.append(("3d_rigid", dict(fixed=fv, moving=mv, resampled_spacing=w.resampled_spacing,
                                  high_res_is_moving=True, center_overlaid=False,
                                  patch_radius=w.patch_radius, dim=3, x0=x0, lower=x0 - 4,
                                  upper=x0 + 4), truth))
    ws = synth.config2(seed=3, mr_dims=(96, 96, 80), n_frames=6, frame_dims=(96, 64),
                       frame_spacing=0.5, sweep_mm=20.0, patch_radius=3)
    f, m = ws.fixed, ws.moving
    fv, mv = engine.upload(f.data, f.spacing, f.origin), engine.upload(m.data, m.spacing, m.origin, E.F16)
    keep += [fv, mv]
    truth = np.asarray(ws.true_pose, np.float64)[:6]
    x0 = truth + rng.uniform(-1.5, 1.5, 6)
    jobs.append(("sweep_6_frames", dict(fixed=fv, moving=mv, resampled_spacing=ws.resampled_spacing,
                                        high_res_is_moving=True, center_overlaid=False,
                                        patch_radius=3, dim=2, x0=x0, lower=x0 - 4, upper=x0 + 4,
                                        frames=ws.extra["frames"],
                                        frames_centre=ws.extra["centre"]), truth))
    return keep, jobs


def _run(engine, jobs, sessions, together):
    srv = S.Server(engine, max_poses=64, slots=4)
    reg = S.Registrar(srv, engine, max_sessions=sessions)
    out = {}
    if together:
        tickets = [(name, reg.submit(**kw)) for name, kw, _ in jobs]
        for name, t in tickets:
            out[name] = reg.wait(t)
    else:
        for name, kw, _ in jobs:
            out[name] = reg.wait(reg.submit(**kw))
    st = reg.stats()
    reg.close()
    srv.close()
    return out, st


def test_registrations_in_a_mix_equal_alone(engine):
    keep, jobs = _jobs(engine)
    alone, _ = _run(engine, jobs, 1, False)
    mix, st = _run(engine, jobs, 4, True)
    print("registrar counters", st, flush=True)
    assert st["completed"] == len(jobs) and st["failed"] == 0 and st["max_running"] >= 2
    for name, _, truth in jobs:
        xa, la, na, _, _ = alone[name]
        xm, lm, nm, _, _ = mix[name]
        print(f"{name}: error {np.round(xa - truth, 3)} LC2 {la:.6f} after {na} evaluations",
              flush=True)
        assert np.array_equal(xa, xm) and la == lm and na == nm, name
        assert la > 0 and 20 < na < 5000
        # every registration moved towards its generating pose
        assert np.abs(xa - truth)[:3].max() < 1.0, name
    for v in keep:
        v.close()


def test_admission_by_budget_serialises_registrations():
    """Three 2D registrations over the same volumes under a budget that holds the volumes
    and one pair but not two: they run one at a time (the later ones wait for the pair of
    the one before to be freed) and end as they do alone."""
    w = synth.config0(seed=4)
    f, m = w.fixed, w.moving
    rng = np.random.default_rng(3)
    starts = [w.true_pose + np.r_[rng.uniform(-2, 2, 3), 0, 0, 0] for _ in range(3)]
    half = np.r_[5, 5, 5, 0.08, 0.08, 0.05]

    def kw(fv, mv, x0):
        return dict(fixed=fv, moving=mv, resampled_spacing=w.resampled_spacing,
                    high_res_is_moving=w.high_res_is_moving, center_overlaid=w.center_overlaid,
                    patch_radius=w.patch_radius, dim=2, x0=x0, lower=x0 - half, upper=x0 + half)

    probe = E.Engine(0)
    fv, mv = probe.upload(f.data, f.spacing, f.origin), probe.upload(m.data, m.spacing, m.origin)
    alone = {}
    srv = S.Server(probe, max_poses=64, slots=4)
    reg = S.Registrar(srv, probe, max_sessions=1)
    for k, x0 in enumerate(starts):
        alone[k] = reg.wait(reg.submit(**kw(fv, mv, x0)))
    reg.close()
    base, _ = probe.memory()  # volumes + the server's staging (its pairs are freed)
    p = probe.pair(fv, mv, w.resampled_spacing, w.high_res_is_moving, w.center_overlaid,
                   w.patch_radius, 2)
    created, _ = probe.memory()
    mats = np.stack([p.pose_to_matrix(starts[0])] * 64)
    p.eval_batch(mats)
    grown, _ = probe.memory()
    srv.close()
    for h in (p, fv, mv, probe):
        h.close()
    static, pair_bytes = created - base, grown - base   # at creation; with 64-pose buffers
    # room for one pair with its batch buffers, not for the static part of a second
    budget = base + pair_bytes + static // 4
    e = E.Engine(0, hbm_budget=budget)
    fv, mv = e.upload(f.data, f.spacing, f.origin), e.upload(m.data, m.spacing, m.origin)
    srv = S.Server(e, max_poses=64, slots=4)
    reg = S.Registrar(srv, e, max_sessions=3)
    tickets = [reg.submit(**kw(fv, mv, x0)) for x0 in starts]
    res = [reg.wait(t) for t in tickets]
    st = reg.stats()
    reg.close()
    srv.close()
    print("budget", budget, "pair bytes", pair_bytes, "counters", st, flush=True)
    assert st["completed"] == 3 and st["max_running"] == 1 and st["budget_waits"] >= 1
    for k, r in enumerate(res):
        assert np.array_equal(r[0], alone[k][0]) and r[1] == alone[k][1] and r[2] == alone[k][2]
    for h in (fv, mv, e):
        h.close()
