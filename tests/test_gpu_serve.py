"""Serving loop (lc2serve.h) on the GPU: requests from several client threads for two
resident pairs are merged into shared launches, and every request's values and status
equal its evaluation alone, bit for bit, also next to a poisoned (NaN) request."""
import threading

import numpy as np
import pytest

from lc2registration_amd import engine as E
from lc2registration_amd import serve as S
from lc2registration_amd import synth

pytestmark = pytest.mark.gpu


def _pairs(engine):
    w3 = synth.small3d(seed=6)
    w2 = synth.config0(seed=3, n=96)
    out = []
    for w, dim in ((w3, 3), (w2, 2)):
        f, m = w.fixed, w.moving
        fv = engine.upload(f.data, f.spacing, f.origin)
        mv = engine.upload(m.data, m.spacing, m.origin)
        p = engine.pair(fv, mv, w.resampled_spacing, w.high_res_is_moving, w.center_overlaid,
                        w.patch_radius, dim)
        out.append((w, p, fv, mv))
    return out


def test_mixed_requests_equal_alone_and_are_merged(engine):
    pairs = _pairs(engine)
    rng = np.random.default_rng(0)
    reqs = []  # (pair index, matrices)
    for k in range(24):
        i = k % 2
        w, p = pairs[i][0], pairs[i][1]
        n = int(rng.integers(1, 6))
        mats = np.stack([p.pose_to_matrix(synth.perturb(w.true_pose, rng, w.ndim)) for _ in range(n)])
        if k in (5, 10):   # poisoned requests: a NaN pose must not perturb its neighbours
            mats[0] = np.nan
        reqs.append((i, mats))
    alone = [pairs[i][1].eval_batch(m) for i, m in reqs]
    srv = S.Server(engine, max_poses=32, window_us=2000.0)
    results = [None] * len(reqs)

    def client(ids):
        tickets = [(j, srv.submit(pairs[reqs[j][0]][1], reqs[j][1])) for j in ids]
        for j, t in tickets:
            results[j] = srv.wait(t)

    threads = [threading.Thread(target=client, args=(list(range(c, len(reqs), 4)),)) for c in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    st = srv.stats()
    srv.close()
    for j, (a, r) in enumerate(zip(alone, results)):
        assert r is not None and r.status == a.status, j
        assert np.array_equal(r.lc2, a.lc2), j
        assert np.array_equal(r.sum_measure, a.sum_measure) and np.array_equal(r.sum_w, a.sum_w), j
    assert st["requests"] == len(reqs)
    assert st["poses"] == sum(len(m) for _, m in reqs)
    assert st["launches"] < len(reqs)          # requests of a pair were merged
    assert st["queued"] == 0 and st["max_in_flight"] >= 1
    print("serving counters", st)
    # request 5 (2D pair, ultrasound warped): its NaN pose leaves no weight; request 10 (3D
    # pair, MR warped, static ultrasound): its NaN pose samples nothing and measures 0
    assert alone[5].status == E.OK_DEGENERATE and alone[5].sum_w[0] == 0.0
    assert alone[10].status == E.OK and alone[10].lc2[0] == 0.0 and alone[10].sum_w[0] > 0.0
    assert all(a.status == E.OK for j, a in enumerate(alone) if j not in (5, 10))
    for _, p, fv, mv in pairs:
        p.close()
        fv.close()
        mv.close()


@pytest.mark.parametrize("graphs", [True, False])
def test_graph_replay_equals_eager_evaluation(engine, graphs):
    """Same-size requests of a pair one after another: from the second launch on each
    slot replays a captured HIP graph; every result equals lc2e_eval_batch bit for bit."""
    pairs = _pairs(engine)
    rng = np.random.default_rng(11)
    work = []
    for i, (w, p, _, _) in enumerate(pairs):
        for _ in range(5):
            mats = np.stack([p.pose_to_matrix(synth.perturb(w.true_pose, rng, w.ndim))
                             for _ in range(4)])
            work.append((i, mats, p.eval_batch(mats)))
    srv = S.Server(engine, max_poses=8, slots=1)
    srv.use_graphs(graphs)
    for i, mats, alone in work:
        r = srv.wait(srv.submit(pairs[i][1], mats))
        assert r.status == alone.status
        assert np.array_equal(r.lc2, alone.lc2) and np.array_equal(r.sum_w, alone.sum_w)
    st = srv.stats()
    srv.close()
    print("serving counters", st)
    # per pair: launch 1 eager, launch 2 captured and replayed, then replays
    assert st["graph_launches"] == (8 if graphs else 0)
    for _, p, fv, mv in pairs:
        p.close()
        fv.close()
        mv.close()


def test_time_out_and_unknown_ticket(engine):
    """A request held by the batching window is not done after a zero time-out (the
    time-out code, the request stays queued) and completes on the next wait; an unknown
    ticket is refused."""
    w = synth.small3d(seed=7)
    f, m = w.fixed, w.moving
    fv = engine.upload(f.data, f.spacing, f.origin)
    mv = engine.upload(m.data, m.spacing, m.origin)
    p = engine.pair(fv, mv, w.resampled_spacing, w.high_res_is_moving, w.center_overlaid,
                    w.patch_radius, w.ndim)
    M = p.pose_to_matrix(w.true_pose)[None]
    alone = p.eval_batch(M)
    srv = S.Server(engine, max_poses=64, window_us=300000.0)   # a 0.3 s window
    t = srv.submit(p, M)
    rc = S.lib().lc2s_wait(srv.h, t, 0.0, None, None, None)
    assert E.ERR.get(rc) == "UNSUPPORTED"
    r = srv.wait(t)
    assert r.status >= 0 and np.array_equal(r.lc2, alone.lc2)
    rc = S.lib().lc2s_wait(srv.h, t, 0.0, None, None, None)      # collected: no longer known
    assert E.ERR.get(rc) == "INVALID"
    assert b"unknown ticket" in E.lib().lc2e_last_error()
    srv.close()
    p.close()
    fv.close()
    mv.close()


def test_merged_launch_over_budget_falls_back_to_one_per_request():
    """Admission by bytes: with an HBM budget that holds the pair and a one-pose batch but
    not a merged one, the merged launch is refused by the budget and every request is then
    evaluated alone, with its alone values, instead of failing."""
    w = synth.small3d(seed=8)
    f, m = w.fixed, w.moving

    def make(e):
        fv = e.upload(f.data, f.spacing, f.origin)
        mv = e.upload(m.data, m.spacing, m.origin)
        return fv, mv, e.pair(fv, mv, w.resampled_spacing, w.high_res_is_moving,
                              w.center_overlaid, w.patch_radius, w.ndim)

    rng = np.random.default_rng(4)
    probe = E.Engine(0)
    fv, mv, p = make(probe)
    mats = np.stack([p.pose_to_matrix(synth.perturb(w.true_pose, rng, 3)) for _ in range(4)])
    p.eval_batch(mats[:1])
    used, _ = probe.memory()
    for h in (p, fv, mv, probe):
        h.close()

    staging = 4 * (12 * 4 + 3 * 8)              # the server's matrices and results, 4 poses
    e = E.Engine(0, hbm_budget=used + staging + 256)
    fv, mv, p = make(e)
    alone = [p.eval_batch(mats[k:k + 1]) for k in range(4)]
    with pytest.raises(E.LC2Error) as ei:       # the merged size does not fit
        p.eval_batch(mats)
    assert ei.value.code == -3
    srv = S.Server(e, max_poses=64, window_us=200000.0, slots=1)
    tickets = [srv.submit(p, mats[k:k + 1]) for k in range(4)]
    res = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    for a, r in zip(alone, res):
        assert r.status == E.OK and np.array_equal(r.lc2, a.lc2)
    assert st["launches"] == 5 and st["requests"] == 4   # one refused merged launch, four alone
    for h in (p, fv, mv, e):
        h.close()
