"""Build recipe of the engine: hipcc for gfx950, in-tree output.

  python -m lc2registration_amd.build          # build liblc2e.so and the CLI
Outputs (git-ignored, they travel with the tree): lc2registration_amd/_lib/.
"""
from __future__ import annotations

import os
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB_DIR = os.path.join(PKG, "_lib")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LC2E_ARCH", "gfx950")

# -ffp-contract=off: FP32 stages use explicit fma only, reproducible against the oracle.
# -packed-fp32-ops off: no v_pk_*_f32 (known to misbehave on this fleet).
DEVICE_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                "-munsafe-fp-atomics", "-Wall", "-Wno-unused-result"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter"]

DEVICE_SOURCES = ["lc2e_kernels.hip"]
HOST_SOURCES = ["lc2e_engine.cpp", "lc2e_serve.cpp", "lc2e_sessions.cpp", "facade.cpp", "yaml_lite.cpp", "optim.cpp", "bobyqa.cpp", "igtl.cpp", "image_io.cpp",
                "lc2reg_capi.cpp"]
CLI_SOURCES = ["main_registration.cpp"]


def _run(cmd):
    t = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return time.time() - t


def _obj(src):
    return os.path.join(LIB_DIR, "obj", os.path.splitext(src)[0] + ".o")


def build(verbose: bool = False) -> str:
    os.makedirs(os.path.join(LIB_DIR, "obj"), exist_ok=True)
    inc = ["-I", INCLUDE, "-I", CSRC]
    jobs = []
    for s in DEVICE_SOURCES:
        jobs.append([HIPCC, *DEVICE_FLAGS, *inc, "-c", "-o", _obj(s), os.path.join(CSRC, s)])
    for s in HOST_SOURCES + CLI_SOURCES:
        if os.path.exists(os.path.join(CSRC, s)):
            # host-only C++ that calls the HIP runtime: hipcc without device code
            jobs.append([HIPCC, *HOST_FLAGS, "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", *inc, "-x", "c++", "-c", "-o",
                         _obj(s), os.path.join(CSRC, s)])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        times = list(ex.map(_run, jobs))
    if verbose:
        for j, t in zip(jobs, times):
            print(f"  {os.path.basename(j[-1])}: {t:.1f}s")
    objs = [_obj(s) for s in DEVICE_SOURCES + HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    lib = os.path.join(LIB_DIR, "liblc2e.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", lib + ".tmp", *objs])
    os.replace(lib + ".tmp", lib)
    cli_objs = [_obj(s) for s in CLI_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if cli_objs:
        exe = os.path.join(LIB_DIR, "lc2_registration")
        _run([HIPCC, f"--offload-arch={ARCH}", "-o", exe, *cli_objs, "-L", LIB_DIR, "-llc2e",
              f"-Wl,-rpath,{LIB_DIR}", "-Wl,-rpath,$ORIGIN"])
    return lib


if __name__ == "__main__":
    t0 = time.time()
    path = build(verbose=True)
    print(f"built {path} in {time.time() - t0:.1f}s")
