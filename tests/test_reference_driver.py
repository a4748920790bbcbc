"""A C++ driver written against the reference's public interface (configuration structs,
Interpolator*/Optimizers* constructors that take them, main.cu's call sequence) compiles
and links against include/lc2_registration.hpp and liblc2e.so (CPU), and on the GPU runs:
the interpolators' default patch sizes are the reference's (interpolator.cuh:55, :225), the
LC2 it reports matches the oracle, and the RegistrationInfo.log files have the reference's
layouts (optimizers.cu:610-623 and :709-720 for 3D; :1263-1272 and :1335-1340 for 2D->3D).
"""
import os
import subprocess

import numpy as np
import pytest

from lc2registration_amd import engine as E
from lc2registration_amd import synth
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "reference_driver.cpp")
OUT = os.path.join(ROOT, "tests", "cpp", "_build", "reference_driver")

HEADER_3D = ("Step,Transform[16],Translation_x,Translation_y,Translation_z,Rotation_x,Rotation_y,"
             "Rotation_z,Scale_x,Scale_y,Scale_z,Shear_x,Shear_y,Shear_z,Similarity,-log(sim),"
             "IteratorCounter")
HEADER_SLICE = ("Translation_x,Translation_y,Translation_z,Rotation_x,Rotation_y,Rotation_z,"
                "Scale_x,Scale_y,Scale_z,Shear_x,Shear_y,Shear_z,Similarity,IteratorCounter")


def build_driver():
    if not os.path.exists(E.library_path()):
        from lc2registration_amd import build
        build.build()
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    lib_dir = os.path.dirname(E.library_path())
    cmd = ["/opt/rocm/bin/hipcc", "-O1", "-std=c++17", "-x", "c++", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", "-I", os.path.join(ROOT, "include"), SRC, "-x", "none",
           "-L", lib_dir, "-llc2e", f"-Wl,-rpath,{lib_dir}", "-o", OUT + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    os.replace(OUT + ".tmp", OUT)
    return OUT


def test_reference_shaped_driver_compiles_and_links():
    assert os.access(build_driver(), os.X_OK)


@pytest.mark.gpu
def test_reference_shaped_driver_runs(tmp_path):
    exe = OUT if os.path.exists(OUT) else build_driver()
    w = synth.small3d()
    f, m = w.fixed, w.moving
    f.data.astype(np.float32).tofile(tmp_path / "fixed.raw")
    m.data.astype(np.float32).tofile(tmp_path / "moving.raw")
    rng = np.random.default_rng(3)
    pose = w.true_pose + np.r_[rng.uniform(-1, 1, 6), [0.0] * 6]
    vals = [*f.dims, *f.spacing, *f.origin, *m.dims, *m.spacing, *m.origin, *w.resampled_spacing,
            w.patch_radius, *pose]
    (tmp_path / "case.txt").write_text(" ".join(repr(float(v)) if isinstance(v, float) else str(v)
                                                for v in vals))
    print(f"reference driver on small3d in {tmp_path}", flush=True)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {ln.split()[0]: ln.split()[1:] for ln in r.stdout.splitlines() if ln.strip()}
    # LC2 at the starting pose through GetSimilarityMeasure() (default 3) and (patch) vs oracle
    fg, mg = O.geom(f.dims, f.spacing, f.origin), O.geom(m.dims, m.spacing, m.origin)
    vd = O.virtual_dims(fg, w.resampled_spacing, 3)
    vg = O.geom(vd, w.resampled_spacing, f.origin)
    p32 = pose.astype(np.float32)
    M = O.affine_3d(p32[:3], p32[3:6], p32[6:9], p32[9:12], O.center(fg, 3), np.zeros(3, np.float32))
    I = O.warp_3d(m.data.astype(np.float32), mg, vg, M[:3].ravel())
    U = O.warp_3d(f.data.astype(np.float32), fg, vg, [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0])
    G = O.sobel(I, "sep")
    assert abs(float(out["lc2_3d"][0]) - O.lc2_exact(U, I, G, w.patch_radius).lc2) < 1e-9
    assert abs(float(out["lc2_3d_default"][0]) - O.lc2_exact(U, I, G, 3).lc2) < 1e-9
    assert out["current_equals_m2f"] == ["1"]
    # Interpolator2DTo3D::GetSimilarityMeasure() defaults to patch 9
    assert out["lc2_slice_default"] == out["lc2_slice_9"]
    assert float(out["lc2_3d_final"][0]) >= float(out["lc2_3d"][0])
    assert 0 < int(out["nev_3d"][0]) < 2000 and int(out["nev_slice"][0]) > 0
    # log layouts
    log3 = (tmp_path / "3d" / "RegistrationInfo.log").read_text().splitlines()
    assert log3[0] == HEADER_3D
    assert len(log3[1].split(",")) == 1 + 16 + 12 + 2          # ShowCurrentResults: no -log
    assert all(len(ln.split(",")) == 1 + 16 + 12 + 3 for ln in log3[2:])
    assert len(log3) == int(out["nev_3d"][0]) + 2
    logs = (tmp_path / "slice" / "RegistrationInfo.log").read_text().splitlines()
    assert logs[0] == HEADER_SLICE and logs[1] == "Initial step"
    assert all(len(ln.split(",")) == 12 + 2 for ln in logs[2:])
    assert len(logs) == int(out["nev_slice"][0]) + 3
