"""The mutual-information path (patchSize 0) on the GPU against the oracle, and a
YAML-driven registration that maximises it."""
import os

import numpy as np
import pytest
import yaml

from lc2registration_amd import engine as E
from lc2registration_amd import registration as REG
from lc2registration_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _warps(w, M):
    f, m = w.fixed, w.moving
    fg, mg = O.geom(f.dims, f.spacing, f.origin), O.geom(m.dims, m.spacing, m.origin)
    vg = O.geom(O.virtual_dims(fg, w.resampled_spacing, 3), w.resampled_spacing, f.origin)
    stat = O.warp_3d(f.data.astype(np.float32), fg, vg, [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0])
    mov = O.warp_3d(m.data.astype(np.float32), mg, vg, np.asarray(M).ravel()[:12])
    return stat, mov


@pytest.mark.parametrize("high_res_is_moving,bins", [(True, (256, 256)), (True, (64, 32)),
                                                      (False, (256, 128))])
def test_mi_matches_oracle(engine, high_res_is_moving, bins):
    """Histograms are integer-exact and the FP64 entropies agree with the oracle's FP64
    sums to 1e-12; the static image is the ultrasound exactly when the moving image is the
    high-resolution one."""
    w = synth.small3d(seed=4)
    f, m = w.fixed, w.moving
    fv = engine.upload(f.data, f.spacing, f.origin)
    mv = engine.upload(m.data, m.spacing, m.origin)
    p = engine.pair(fv, mv, w.resampled_spacing, high_res_is_moving, False, 2, 3)
    rng = np.random.default_rng(2)
    poses = [w.true_pose] + [synth.perturb(w.true_pose, rng, 3, 2.0) for _ in range(4)]
    mats = np.stack([p.pose_to_matrix(q) for q in poses])
    got = p.eval_mi(mats, *bins)
    for k, M in enumerate(mats):
        stat, mov = _warps(w, M)
        U, I = (stat, mov) if high_res_is_moving else (mov, stat)
        ref = O.nmi(U, I, bins[0], bins[1], fp32=False)
        assert abs(got[k] - ref) < 1e-12, (k, got[k], ref)
        assert 1.0 < got[k] < 2.0
        assert p.eval_mi(M[None], *bins)[0] == got[k]          # batch == single
    p.close()
    fv.close()
    mv.close()


def test_mi_yaml_registration_maximises_nmi(tmp_path):
    """PatchSize 0 in the configuration: both BOBYQA rounds maximise the normalized mutual
    information of an LC2-consistent pair, to within 10% of its gain at the generating pose."""
    import test_gpu_registration as TR
    w, U, _, _ = TR.consistent_pair()
    f, m = w.fixed, w.moving
    U.tofile(tmp_path / "fixed.raw")
    m.data.astype(np.float32).tofile(tmp_path / "moving.raw")
    truth = w.true_pose
    init = np.r_[truth[:6] + np.array([1.5, -1.0, 1.0, 1.5, -1.0, 1.0]), 1, 1, 1, 0, 0, 0]
    cfg = REG.write_parameters_yaml(
        str(tmp_path / "parameters3d.yaml"), 3,
        (tmp_path / "fixed.raw", "float32", f.dims, f.spacing, f.origin),
        (tmp_path / "moving.raw", "float32", m.dims, m.spacing, m.origin),
        "Moving", w.resampled_spacing, 0, tmp_path, False, init,
        [-4] * 6 + [0.9] * 3 + [-0.1] * 3, [4] * 6 + [1.1] * 3 + [0.1] * 3,
        rounds=[((True, True, False, False), 1e-4), ((True, True, True, True), 1e-4)])
    n_init = REG.evaluate_yaml("3D", cfg, init)
    x, nmi, nev, _ = REG.register_yaml("3D", cfg, write_images=False)
    print(f"MI registration: NMI {n_init:.6f} -> {nmi:.6f} in {nev} evaluations, "
          f"error {np.round(x - truth, 4)}")
    n_true = REG.evaluate_yaml("3D", cfg, truth)
    assert nmi > n_init
    # the 256 x 256-bin NMI of this 46k-voxel grid is rough: BOBYQA (as NLopt's) settles in a
    # local maximum next to the generating pose
    assert nmi > n_init + 0.9 * (n_true - n_init)
    assert np.abs(x - truth)[:3].max() < 0.5
    out = yaml.safe_load(open(tmp_path / "parameters3d_result.yaml"))
    assert abs(out["Outputs"]["LastSimilarity"] - nmi) < 1e-12
    log = open(tmp_path / "RegistrationInfo.log").read().splitlines()
    assert len(log) == nev + 2
