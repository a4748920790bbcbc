"""Python binding of the registration layer (include/lc2reg.h) and a writer for the
reference's parameters*.yaml layout (config/parameters3d_affine.yaml,
config/parameters2d_affine.yaml, config/parameters2dTo3d.yaml).

    write_parameters_yaml(path, 3, fixed, moving, ...)   # reference-format configuration
    x, lc2, nevals, matrix = register_yaml("3D", path)   # DIRECT stage + BOBYQA rounds
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import engine as _E

DP = C.POINTER(C.c_double)
BATCH_FN = C.CFUNCTYPE(None, C.c_int, C.c_int, DP, DP, C.c_void_p)
KEYS_3D = ["TX", "TY", "TZ", "RX", "RY", "RZ", "ScaleX", "ScaleY", "ScaleZ", "ShearX", "ShearY",
           "ShearZ"]
KEYS_2D = ["TX", "TY", "R", "ScaleX", "ScaleY", "Shear"]
ENTRIES = ["Translation", "Rotation", "Scale", "Shear"]

_lib = None


def lib():
    global _lib
    if _lib is None:
        L = _E.lib()
        L.lc2reg_last_error.restype = C.c_char_p
        L.lc2reg_register_yaml.argtypes = [C.c_char_p, C.c_char_p, C.c_int, DP, DP,
                                           C.POINTER(C.c_int), C.POINTER(C.c_float)]
        L.lc2reg_evaluate_yaml.argtypes = [C.c_char_p, C.c_char_p, DP, C.c_int, DP]
        L.lc2reg_minimize_direct.argtypes = [C.c_int, DP, DP, C.c_int, BATCH_FN, C.c_void_p, DP,
                                             DP, C.POINTER(C.c_int)]
        L.lc2reg_minimize_qmtr.argtypes = [C.c_int, DP, DP, DP, C.c_double, C.c_int, BATCH_FN,
                                           C.c_void_p, DP, DP, C.POINTER(C.c_int)]
        L.lc2reg_minimize_nlopt.argtypes = [C.c_int, C.c_int, DP, DP, DP, DP, C.c_double, C.c_int,
                                            BATCH_FN, C.c_void_p, DP, DP, C.POINTER(C.c_int)]
        L.lc2reg_yaml_roundtrip.argtypes = [C.c_char_p, C.c_char_p]
        _lib = L
    return _lib


class RegistrationError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise RegistrationError(lib().lc2reg_last_error().decode())


def _p(a):
    return a.ctypes.data_as(DP)


def register_yaml(mode: str, path: str, write_images: bool = True):
    """Runs a whole registration from a parameters*.yaml (mode "2D", "3D" or "2DTo3D"):
    writes <cfg>_result.yaml, RegistrationInfo.log and, if asked, the registered images.
    Returns (parameters, last LC2, evaluations, moving-to-fixed matrix)."""
    n = 6 if mode == "2D" else 12
    x = np.zeros(n)
    lc2, nev = C.c_double(), C.c_int()
    mat = (C.c_float * 16)()
    _check(lib().lc2reg_register_yaml(mode.encode(), str(path).encode(), int(write_images), _p(x),
                                      C.byref(lc2), C.byref(nev), mat))
    md = 3 if mode == "2D" else 4
    return x, lc2.value, nev.value, np.array(mat[: md * md], np.float32).reshape(md, md)


def evaluate_yaml(mode: str, path: str, params) -> float:
    """LC2 of one parameter vector for the pair a parameters*.yaml describes."""
    q = np.ascontiguousarray(params, np.float64)
    out = C.c_double()
    _check(lib().lc2reg_evaluate_yaml(mode.encode(), str(path).encode(), _p(q), q.size,
                                      C.byref(out)))
    return out.value


def _objective(fn):
    def cb(n, d, X, F, _):
        xs = np.ctypeslib.as_array(X, shape=(n, d)).copy()
        np.ctypeslib.as_array(F, shape=(n,))[:] = fn(xs)
    return BATCH_FN(cb)


def minimize_direct(fn, lb, ub, maxeval: int):
    """DIRECT-L over the box [lb, ub]; fn maps an (n, d) array of points to n values."""
    lb, ub = (np.ascontiguousarray(v, np.float64) for v in (lb, ub))
    x, f, n = np.zeros(lb.size), C.c_double(), C.c_int()
    cb = _objective(fn)
    _check(lib().lc2reg_minimize_direct(lb.size, _p(lb), _p(ub), maxeval, cb, None, _p(x),
                                        C.byref(f), C.byref(n)))
    return x, f.value, n.value


def minimize_qmtr(fn, x0, lb, ub, xtol: float = 1e-4, maxeval: int = 100000):
    """Bound-constrained quadratic-model trust region from x0; fn as for minimize_direct."""
    x0, lb, ub = (np.ascontiguousarray(v, np.float64) for v in (x0, lb, ub))
    x, f, n = np.zeros(x0.size), C.c_double(), C.c_int()
    cb = _objective(fn)
    _check(lib().lc2reg_minimize_qmtr(x0.size, _p(x0), _p(lb), _p(ub), xtol, maxeval, cb, None,
                                      _p(x), C.byref(f), C.byref(n)))
    return x, f.value, n.value


GN_DIRECT_L, GN_ORIG_DIRECT_L, LN_BOBYQA = 0, 1, 2


def minimize_nlopt(algorithm: int, fn, lb, ub, x0=None, xtol_rel: float = 0.0, maxeval: int = 0,
                   step=None):
    """The nlopt::opt-shaped facade the Optimizers* classes use (GN_DIRECT_L,
    GN_ORIG_DIRECT_L, LN_BOBYQA); fn as for minimize_direct. Returns (x, f, evaluations,
    NLopt result code)."""
    lb, ub = (np.ascontiguousarray(v, np.float64) for v in (lb, ub))
    x0 = np.ascontiguousarray(lb if x0 is None else x0, np.float64)
    st = None if step is None else np.ascontiguousarray(step, np.float64)
    x, f, n = np.zeros(lb.size), C.c_double(), C.c_int()
    cb = _objective(fn)
    rc = lib().lc2reg_minimize_nlopt(algorithm, lb.size, _p(x0), _p(lb), _p(ub),
                                     None if st is None else _p(st), xtol_rel, maxeval, cb, None,
                                     _p(x), C.byref(f), C.byref(n))
    if rc < 0:
        raise RegistrationError(lib().lc2reg_last_error().decode())
    return x, f.value, n.value, rc


def _num(v):
    """A scalar as YAML text; exponent forms keep a '.' so YAML 1.1 readers see a float."""
    if isinstance(v, (bool, np.bool_)):
        return "true" if v else "false"
    if isinstance(v, (float, np.floating)):
        t = repr(float(v))
        if "e" in t and "." not in t.split("e")[0]:
            m, e = t.split("e")
            t = f"{m}.0e{e}"
        return t
    return str(v)


def _flow(keys, vals):
    return "{" + ", ".join(f"{k}: {_num(v)}" for k, v in zip(keys, vals)) + "}"


def _bools(entries):
    return _flow(ENTRIES, [bool(b) for b in entries])


def write_parameters_yaml(path, ndim, fixed, moving, high_res, resampled, patch, export_dir,
                          center_overlaid, init, lower, upper, direct_maxeval=0,
                          direct_entries=(True, False, False, False), rounds=()):
    """A parameters*.yaml in the reference's layout.

    fixed / moving: (file, data type, dims, spacing, origin); high_res: "Fixed" or "Moving";
    init / lower / upper: 12 (3D, 2DTo3D) or 6 (2D) values, translation and rotation bounds
    relative to init as the reference reads them; direct_maxeval 0 disables the DIRECT
    stage; rounds: [(entries (T, R, S, Sh), AbsTol), ...] BOBYQA rounds, all enabled."""
    ax = "xyz"[:ndim]
    kind = "Volume" if ndim == 3 else "Image"
    keys = KEYS_3D if len(init) == 12 else KEYS_2D
    lines = ["InputParameters:"]
    for tag, (fp, dt, dims, sp, org) in (("Fixed", fixed), ("Moving", moving)):
        lines += [f"  {tag}{kind}:", f"    Dimension: {_flow(ax, dims)}",
                  f"    Spacing: {_flow(ax, sp)}", f"    Origin: {_flow(ax, org)}",
                  f"    FilePath: {fp}", "    Format: raw", f"    DataType: {dt}"]
    lines += ["  RegistrationParameters:", f"    HighResolutionModality: {high_res}",
              "    Type: Affine", "    DisplayPattern: 0",
              f"    ResampledSpacing: {_flow(ax, resampled)}", f"    PatchSize: {patch}",
              f"    FinalRoundExportRootPath: {export_dir}",
              f"    CenterOverlaid: {'true' if center_overlaid else 'false'}",
              "  OptimizerSettings:", f"    InitialValues: {_flow(keys, init)}",
              f"    LowerBounds: {_flow(keys, lower)}", f"    UpperBounds: {_flow(keys, upper)}",
              "    DIRECT:", f"      Enabled: {'true' if direct_maxeval else 'false'}",
              f"      OptimizedEntries: {_bools(direct_entries)}",
              f"      MaxEvaluations: {direct_maxeval or 30}", "    BOBYQA:"]
    for i, (entries, tol) in enumerate(rounds):
        lines += [f"      {i}:", "        Enabled: true", f"        OptimizedEntries: {_bools(entries)}",
                  f"        AbsTol: {_num(float(tol))}"]
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return path
