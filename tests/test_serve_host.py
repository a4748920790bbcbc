"""The serving loop's scheduling on the host (no GPU): lc2e_serve.cpp is compiled against a
fake HIP runtime and engine (tests/serve_host/fake_runtime.cpp) and driven by
serve_host/serve_test.cpp: merging, per-request results and status, errors, the budget
fallback, time-outs, admission by device share, destruction with work queued; and the
registrar (lc2e_sessions.cpp) by serve_host/registrar_test.cpp. The same scenarios run again
under ThreadSanitizer (host code only)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "lc2registration_amd", "csrc", "lc2e_serve.cpp"),
       os.path.join(ROOT, "tests", "serve_host", "fake_runtime.cpp"),
       os.path.join(ROOT, "tests", "serve_host", "serve_test.cpp")]
FLAGS = ["-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
         "-I" + os.path.join(ROOT, "include"), "-pthread"]


def _build_and_run(tmp_path, extra, cxx=None):
    cxx = cxx or shutil.which("g++")
    if cxx is None or not os.path.exists(cxx) or not os.path.isdir("/opt/rocm/include/hip"):
        pytest.skip("a host C++ compiler and the HIP headers are needed")
    exe = str(tmp_path / "serve_test")
    subprocess.run([cxx, *FLAGS, *extra, *SRC, "-o", exe], check=True, timeout=240)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all serving scenarios passed" in r.stdout
    return r


def test_serving_scenarios(tmp_path):
    _build_and_run(tmp_path, ["-O2", "-Wall", "-Wextra", "-Werror", "-Wno-deprecated-declarations"])


CSRC = os.path.join(ROOT, "lc2registration_amd", "csrc")
REG_SRC = [os.path.join(CSRC, f) for f in ("lc2e_serve.cpp", "lc2e_sessions.cpp", "bobyqa.cpp",
                                           "optim.cpp")] + \
          [os.path.join(ROOT, "tests", "serve_host", f) for f in ("fake_runtime.cpp",
                                                                  "registrar_test.cpp")]


def _build_and_run_registrar(tmp_path, extra, cxx=None):
    cxx = cxx or shutil.which("g++")
    if cxx is None or not os.path.exists(cxx) or not os.path.isdir("/opt/rocm/include/hip"):
        pytest.skip("a host C++ compiler and the HIP headers are needed")
    exe = str(tmp_path / "registrar_test")
    subprocess.run([cxx, *FLAGS, "-I" + CSRC, *extra, *REG_SRC, "-o", exe], check=True, timeout=240)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all registrar scenarios passed" in r.stdout
    return r


def test_registrar_scenarios(tmp_path):
    """Whole registrations as requests (lc2s_registrar) with the library's own optimizers:
    alone = in a mix bit for bit, admission by sessions and by HBM budget, a pair that never
    fits, destruction with registrations queued."""
    _build_and_run_registrar(tmp_path, ["-O2", "-Wall", "-Wextra", "-Werror",
                                        "-Wno-deprecated-declarations"])


def test_registrar_scenarios_thread_sanitizer(tmp_path):
    r = _build_and_run_registrar(tmp_path, ["-O1", "-g", "-fsanitize=thread"],
                                 "/opt/rocm/llvm/bin/clang++")
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr


def test_serving_scenarios_thread_sanitizer(tmp_path):
    # LLVM's runtime: it intercepts pthread_cond_clockwait, which libstdc++'s timed waits
    # use; GCC 11's does not, and reports every such wait as a race
    r = _build_and_run(tmp_path, ["-O1", "-g", "-fsanitize=thread"], "/opt/rocm/llvm/bin/clang++")
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr
