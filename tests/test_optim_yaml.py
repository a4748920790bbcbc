"""Registration layer on the CPU: the configuration reader/writer against PyYAML on the
reference's own parameters*.yaml files, and the batched optimizers (DIRECT-L against
scipy's DIRECT-L, the trust-region method on closed-form problems). No GPU calls."""
import ctypes as C
import glob
import os

import numpy as np
import pytest
import yaml
from scipy.optimize import direct

from lc2registration_amd import engine as E

HERE = os.path.dirname(os.path.abspath(__file__))
CONFIGS = sorted(glob.glob(os.path.join(HERE, "golden", "config", "*.yaml")) +
                 glob.glob(os.path.join(HERE, "golden", "config", "config_backup", "*.yaml")))

BATCH_FN = C.CFUNCTYPE(None, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                       C.c_void_p)


@pytest.fixture(scope="module")
def reg():
    if not os.path.exists(E.library_path()):
        from lc2registration_amd import build
        build.build()
    L = C.CDLL(E.library_path())
    dp = C.POINTER(C.c_double)
    L.lc2reg_last_error.restype = C.c_char_p
    L.lc2reg_yaml_roundtrip.argtypes = [C.c_char_p, C.c_char_p]
    L.lc2reg_minimize_direct.argtypes = [C.c_int, dp, dp, C.c_int, BATCH_FN, C.c_void_p, dp, dp,
                                         C.POINTER(C.c_int)]
    L.lc2reg_minimize_qmtr.argtypes = [C.c_int, dp, dp, dp, C.c_double, C.c_int, BATCH_FN,
                                       C.c_void_p, dp, dp, C.POINTER(C.c_int)]
    return L


@pytest.mark.parametrize("path", CONFIGS, ids=[os.path.basename(p) for p in CONFIGS])
def test_yaml_roundtrip_matches_pyyaml(reg, path, tmp_path):
    out = str(tmp_path / "out.yaml")
    assert reg.lc2reg_yaml_roundtrip(path.encode(), out.encode()) == 0, reg.lc2reg_last_error()
    with open(path) as fh:
        a = yaml.safe_load(fh)
    with open(out) as fh:
        b = yaml.safe_load(fh)
    assert a == b


def test_yaml_preserves_flow_maps_and_nested_sequences(reg, tmp_path):
    src = os.path.join(HERE, "golden", "config", "parameters3d_affine.yaml")
    out = str(tmp_path / "o.yaml")
    assert reg.lc2reg_yaml_roundtrip(src.encode(), out.encode()) == 0
    txt = open(out).read()
    assert "Dimension: {x: 394, y: 466, z: 378}" in txt
    assert "  TransformMatrix:\n    -\n      - 0.999353051" in txt


def _call(reg, fn, *args):
    rc = fn(*args)
    assert rc == 0, reg.lc2reg_last_error()


def _batch(f, log=None):
    def cb(n, d, X, F, _):
        x = np.ctypeslib.as_array(X, shape=(n, d))
        out = np.ctypeslib.as_array(F, shape=(n,))
        for i in range(n):
            out[i] = f(x[i])
            if log is not None:
                log.append(x[i].copy())
    return BATCH_FN(cb)


def _arr(v):
    a = np.ascontiguousarray(v, np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def branin(x):
    a, b, c, r, s, t = 1.0, 5.1 / (4 * np.pi ** 2), 5 / np.pi, 6.0, 10.0, 1 / (8 * np.pi)
    return a * (x[1] - b * x[0] ** 2 + c * x[0] - r) ** 2 + s * (1 - t) * np.cos(x[0]) + s


def test_direct_first_division_and_global_minimum(reg):
    lb, ub = np.array([-5.0, 0.0]), np.array([10.0, 15.0])
    log = []
    cb = _batch(branin, log)
    _, plb = _arr(lb)
    _, pub = _arr(ub)
    x, px = _arr(np.zeros(2))
    f = C.c_double()
    n = C.c_int()
    _call(reg, reg.lc2reg_minimize_direct, 2, plb, pub, 300, cb, None, px, C.byref(f), C.byref(n))
    # the first evaluations are the centre and centre +- (range/3) along each axis
    pts = np.array(log[:5])
    c = (lb + ub) / 2
    d = (ub - lb) / 3
    expect = {tuple(c)} | {tuple(c + s * d * e) for s in (-1, 1) for e in np.eye(2)}
    assert {tuple(np.round(p, 9)) for p in pts} == {tuple(np.round(p, 9)) for p in expect}
    ref = direct(branin, list(zip(lb, ub)), locally_biased=True, maxfun=300)
    assert n.value <= 300
    assert f.value < 0.397887 + 2e-3              # global minimum of Branin: 0.397887
    assert abs(f.value - ref.fun) < 2e-3


def test_direct_respects_maxeval(reg):
    cb = _batch(lambda x: float(np.sum((x - 0.3) ** 2)))
    _, plb = _arr(-np.ones(4))
    _, pub = _arr(np.ones(4))
    x, px = _arr(np.zeros(4))
    f, n = C.c_double(), C.c_int()
    _call(reg, reg.lc2reg_minimize_direct, 4, plb, pub, 57, cb, None, px, C.byref(f), C.byref(n))
    assert n.value <= 57
    assert f.value < 0.2


def test_qmtr_converges_on_quadratic_and_bound(reg):
    target = np.array([1.5, -2.0, 0.25, 3.0])
    scale = np.array([1.0, 10.0, 0.1, 2.0])
    cb = _batch(lambda x: float(np.sum(((x - target) / scale) ** 2) + 0.3 * (x[0] - target[0]) * (x[1] - target[1]) / 10))
    lb, ub = np.array([-10, -10, -1, -10.0]), np.array([10, 10, 1, 2.0])   # x3 bound active
    _, plb = _arr(lb)
    _, pub = _arr(ub)
    _, px0 = _arr(np.zeros(4))
    x, px = _arr(np.zeros(4))
    f, n = C.c_double(), C.c_int()
    _call(reg, reg.lc2reg_minimize_qmtr, 4, px0, plb, pub, 1e-6, 5000, cb, None, px, C.byref(f),
          C.byref(n))
    assert abs(x[3] - 2.0) < 1e-9                  # on the active upper bound
    assert np.allclose(x[:3], target[:3], atol=1e-3 * scale[:3] * 10)
    assert n.value < 5000


def test_qmtr_rosenbrock(reg):
    def rosen(x):
        return float(100 * (x[1] - x[0] ** 2) ** 2 + (1 - x[0]) ** 2)
    cb = _batch(rosen)
    _, plb = _arr(np.array([-2.0, -2.0]))
    _, pub = _arr(np.array([2.0, 2.0]))
    _, px0 = _arr(np.array([-1.2, 1.0]))
    x, px = _arr(np.zeros(2))
    f, n = C.c_double(), C.c_int()
    _call(reg, reg.lc2reg_minimize_qmtr, 2, px0, plb, pub, 1e-8, 20000, cb, None, px, C.byref(f),
          C.byref(n))
    assert f.value < 1e-4
    assert np.allclose(x, [1.0, 1.0], atol=2e-2)


def test_python_binding_and_config_writer(tmp_path):
    """lc2registration_amd.registration: the batch minimizers from Python and the
    reference-layout configuration writer (read back by PyYAML and by the C++ reader)."""
    from lc2registration_amd import registration as REG
    x, f, n = REG.minimize_qmtr(lambda X: ((X - [0.5, -1.0]) ** 2).sum(1), [0.0, 0.0],
                                [-2, -2], [2, 2], 1e-8, 2000)
    assert np.allclose(x, [0.5, -1.0], atol=1e-4) and f < 1e-8
    x, f, n = REG.minimize_direct(lambda X: np.array([branin(p) for p in X]), [-5, 0], [10, 15], 200)
    assert f < 0.397887 + 5e-3 and n <= 200
    cfg = str(tmp_path / "p3.yaml")
    REG.write_parameters_yaml(cfg, 3, ("/data/us.raw", "uint8", (192, 192, 160), (0.5,) * 3, (1, 2, 3)),
                              ("/data/mr.raw", "float32", (256, 256, 192), (1.0,) * 3, (0, 0, 0)),
                              "Moving", (0.5,) * 3, 3, str(tmp_path), False,
                              [0] * 6 + [1] * 3 + [0] * 3, [-5] * 6 + [0.9] * 3 + [-0.1] * 3,
                              [5] * 6 + [1.1] * 3 + [0.1] * 3, direct_maxeval=40,
                              rounds=[((True, True, False, False), 1e-4),
                                      ((True, True, True, True), 1e-5)])
    y = yaml.safe_load(open(cfg))
    ip = y["InputParameters"]
    assert ip["FixedVolume"]["Dimension"] == {"x": 192, "y": 192, "z": 160}
    assert ip["RegistrationParameters"]["HighResolutionModality"] == "Moving"
    assert ip["OptimizerSettings"]["DIRECT"] == {
        "Enabled": True, "MaxEvaluations": 40,
        "OptimizedEntries": {"Translation": True, "Rotation": False, "Scale": False, "Shear": False}}
    assert ip["OptimizerSettings"]["BOBYQA"][1]["AbsTol"] == 1e-5
    out = str(tmp_path / "rt.yaml")
    assert REG.lib().lc2reg_yaml_roundtrip(cfg.encode(), out.encode()) == 0
    assert yaml.safe_load(open(out)) == y
