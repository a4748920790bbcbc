"""Writes the kidney-phantom golden fixtures used by the warp parity tests.

Run in a container that has the reference checkout; the committed outputs are
data only:
  * the six float32 image dumps of the reference's ``config/Data`` directory,
    copied byte for byte (inputs and the reference's own virtual-grid outputs);
  * ``kidney.json``: their geometry (from the reference's 2D configuration
    ``config/config_backup/parameters2d.yaml``) and the rigid poses at which the
    moving-image dumps were produced. The dumps carry no pose, so the pose is
    recovered here by Nelder-Mead on the mean squared difference between the
    oracle's warp and the dump; the tests then require the oracle to reproduce
    the dump at that pose to FP32 rounding.

Usage: python tests/golden/make_kidney_fixtures.py [/path/to/reference]
"""
import json
import os
import shutil
import sys

import numpy as np
from scipy.optimize import minimize

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))
from oracle import oracle as O  # noqa: E402

FILES = ["CT_KidneyPhantom.raw", "US_KidneyPhantom.raw", "fixed.bin", "moving.bin",
         "CT_KidneyPhantom_virtual.bin", "US_KidneyPhantom_virtual.bin"]

CT_DIMS, CT_SP = (321, 297), 0.508378906
US_DIMS, US_SP = (419, 310), 0.310571


def rigid_matrix(pose, fixed_g, moving_g, virt_g):
    """VirtualToMoving for (tx, ty, rot_deg) with CenterOverlaid: true."""
    c_f = O.center(fixed_g, 2)
    shift = O.center(moving_g, 2) - O.center(virt_g, 2)
    return O.affine_2d(pose[2], pose[:2], (1.0, 1.0), 0.0, c_f, shift)


def fit(us, fixed_g, moving_g, virt_g, target, x0):
    def cost(p):
        m = rigid_matrix(p, fixed_g, moving_g, virt_g)
        w = O.warp_2d(us, moving_g, virt_g, m.ravel()[:6])
        return float(np.mean((w.astype(np.float64) - target) ** 2))

    r = minimize(cost, x0, method="Nelder-Mead",
                 options={"xatol": 1e-7, "fatol": 1e-12, "maxiter": 4000, "maxfev": 8000})
    return r.x


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    out = os.path.join(HERE, "kidney")
    os.makedirs(out, exist_ok=True)
    for f in FILES:
        dst = os.path.join(out, f)
        if not os.path.exists(dst):
            shutil.copyfile(os.path.join(ref, "config", "Data", f), dst)
    us = np.fromfile(os.path.join(out, "US_KidneyPhantom.raw"), np.float32).reshape(310, 419)
    g_ct = O.geom(CT_DIMS, (CT_SP, CT_SP), (0, 0))
    g_us = O.geom(US_DIMS, (US_SP, US_SP), (0, 0))
    poses = {}
    for name, sp in (("moving.bin", 1.0), ("US_KidneyPhantom_virtual.bin", 0.9)):
        vd = O.virtual_dims(g_ct, (sp, sp), 2)
        g_v = O.geom(vd[:2], (sp, sp), (0, 0))
        tgt = np.fromfile(os.path.join(out, name), np.float32).reshape(vd[1], vd[0])
        best = None
        for x0 in ([-8.5, 9.7, -5.5], [-9.0, 9.7, -5.0], [0.0, 0.0, 0.0]):
            p = fit(us, g_ct, g_us, g_v, tgt.astype(np.float64), np.array(x0))
            m = rigid_matrix(p, g_ct, g_us, g_v)
            err = float(np.abs(O.warp_2d(us, g_us, g_v, m.ravel()[:6]) - tgt).max())
            if best is None or err < best[1]:
                best = (p, err)
        poses[name] = {"virtual_spacing": sp, "virtual_dims": list(vd[:2]),
                       "pose_tx_ty_rdeg": [float(v) for v in best[0]], "max_abs_err": best[1]}
        print(name, poses[name])
    manifest = {
        "source": "reference config/Data, geometry from config/config_backup/parameters2d.yaml",
        "fixed": {"file": "CT_KidneyPhantom.raw", "dims": CT_DIMS, "spacing": [CT_SP, CT_SP],
                  "origin": [0.0, 0.0]},
        "moving": {"file": "US_KidneyPhantom.raw", "dims": US_DIMS, "spacing": [US_SP, US_SP],
                   "origin": [0.0, 0.0]},
        "identity_dumps": {"fixed.bin": 1.0, "CT_KidneyPhantom_virtual.bin": 0.9},
        "center_overlaid": True,
        "rigid_dumps": poses,
    }
    with open(os.path.join(out, "kidney.json"), "w") as fh:
        json.dump(manifest, fh, indent=1)


if __name__ == "__main__":
    main()
