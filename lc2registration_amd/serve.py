"""Python binding of the serving loop (include/lc2serve.h): queue pose batches for
resident pairs; a native worker thread merges the pending requests of a pair into one
fused launch."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import engine as _E

_lib = None
MAX_SLOTS = 8


class Counters(C.Structure):
    _fields_ = [("launches", C.c_longlong), ("requests", C.c_longlong), ("poses", C.c_longlong),
                ("queued", C.c_longlong), ("max_in_flight", C.c_int),
                ("graph_launches", C.c_longlong)]


class Registration(C.Structure):
    _fields_ = [("fixed", C.c_void_p), ("moving", C.c_void_p),
                ("resampled_spacing", C.c_float * 3), ("high_res_is_moving", C.c_int),
                ("center_overlaid", C.c_int), ("patch_radius", C.c_int), ("dim", C.c_int),
                ("n_frames", C.c_int), ("frames", C.POINTER(C.c_float)),
                ("frames_centre", C.c_float * 3), ("n_params", C.c_int),
                ("x0", C.c_double * 12), ("lower", C.c_double * 12), ("upper", C.c_double * 12),
                ("direct_maxeval", C.c_int), ("xtol_rel", C.c_double), ("maxeval", C.c_int)]


class RegistrationResult(C.Structure):
    _fields_ = [("x", C.c_double * 12), ("lc2", C.c_double), ("evaluations", C.c_int),
                ("status", C.c_int), ("admit_ms", C.c_double), ("run_ms", C.c_double)]


class RegistrarCounters(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("completed", C.c_longlong), ("failed", C.c_longlong),
                ("running", C.c_int), ("max_running", C.c_int), ("budget_waits", C.c_longlong)]


def lib():
    global _lib
    if _lib is None:
        L = _E.lib()
        vp, dp = C.c_void_p, C.POINTER(C.c_double)
        L.lc2s_create.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.POINTER(vp)]
        L.lc2s_destroy.argtypes = [vp]
        L.lc2s_destroy.restype = None
        L.lc2s_submit.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_longlong)]
        L.lc2s_wait.argtypes = [vp, C.c_longlong, C.c_double, dp, dp, dp]
        L.lc2s_stats.argtypes = [vp, C.POINTER(Counters)]
        L.lc2s_use_graphs.argtypes = [vp, C.c_int]
        L.lc2s_registrar_create.argtypes = [vp, vp, C.c_int, C.POINTER(vp)]
        L.lc2s_registrar_destroy.argtypes = [vp]
        L.lc2s_registrar_destroy.restype = None
        L.lc2s_register_submit.argtypes = [vp, C.POINTER(Registration), C.POINTER(C.c_longlong)]
        L.lc2s_register_wait.argtypes = [vp, C.c_longlong, C.c_double,
                                         C.POINTER(RegistrationResult)]
        L.lc2s_registrar_stats.argtypes = [vp, C.POINTER(RegistrarCounters)]
        _lib = L
    return _lib


class Server:
    """Requests are candidate pose sets for resident pairs; up to `slots` merged launches
    of different pairs run on the device at once."""

    def __init__(self, engine: "_E.Engine", max_poses: int = 64, window_us: float = 0.0,
                 slots: int = 4):
        self.h = C.c_void_p()
        _E._check(lib().lc2s_create(engine.h, max_poses, slots, window_us, C.byref(self.h)))
        self._n = {}

    def submit(self, pair: "_E.Pair", mats) -> int:
        m = np.ascontiguousarray(mats, np.float32)
        per = 12 * (pair.info.n_frames if pair.info.per_frame_matrices else 1)
        n = m.size // per
        t = C.c_longlong()
        _E._check(lib().lc2s_submit(self.h, pair.h, n, m.ctypes.data_as(C.POINTER(C.c_float)),
                                    C.byref(t)))
        self._n[t.value] = n
        return t.value

    def wait(self, ticket: int, timeout_ms: float = -1.0) -> "_E.BatchResult":
        n = self._n[ticket]
        l, sm, sw = (np.zeros(n) for _ in range(3))
        dp = C.POINTER(C.c_double)
        rc = lib().lc2s_wait(self.h, ticket, timeout_ms, l.ctypes.data_as(dp), sm.ctypes.data_as(dp),
                             sw.ctypes.data_as(dp))
        if rc < 0:
            _E._check(rc)
        del self._n[ticket]
        return _E.BatchResult(l, sm, sw, rc)

    def use_graphs(self, on: bool):
        """Replay captured launches or launch eagerly (the default)."""
        _E._check(lib().lc2s_use_graphs(self.h, int(on)))

    def stats(self):
        c = Counters()
        _E._check(lib().lc2s_stats(self.h, C.byref(c)))
        return {k: getattr(c, k) for k, _ in Counters._fields_}

    def close(self):
        if self.h:
            lib().lc2s_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Registrar:
    """Whole registrations as requests (lc2s_registrar): each builds its pair from two
    resident volumes when admitted (at most `max_sessions` at once, within the engine's HBM
    budget), runs DIRECT-L (optional) then BOBYQA on -log(LC2) with every candidate set
    evaluated through `server`, and frees the pair."""

    def __init__(self, server: Server, engine: "_E.Engine", max_sessions: int = 4):
        self.h = C.c_void_p()
        self._keep = {}
        _E._check(lib().lc2s_registrar_create(server.h, engine.h, max_sessions, C.byref(self.h)))

    def submit(self, fixed: "_E.Volume", moving: "_E.Volume", resampled_spacing,
               high_res_is_moving: bool, center_overlaid: bool, patch_radius: int, dim: int,
               x0, lower, upper, xtol_rel: float = 1e-4, maxeval: int = 0,
               direct_maxeval: int = 0, frames=None, frames_centre=None) -> int:
        r = Registration()
        r.fixed, r.moving = fixed.h, moving.h
        rs = list(resampled_spacing) + [0.0] * (3 - len(resampled_spacing))
        r.resampled_spacing = (C.c_float * 3)(*rs)
        r.high_res_is_moving, r.center_overlaid = int(high_res_is_moving), int(center_overlaid)
        r.patch_radius, r.dim = patch_radius, dim
        x0, lower, upper = (np.asarray(v, np.float64).ravel() for v in (x0, lower, upper))
        r.n_params = x0.size
        for name, v in (("x0", x0), ("lower", lower), ("upper", upper)):
            getattr(r, name)[:x0.size] = v.tolist()
        r.xtol_rel, r.maxeval, r.direct_maxeval = xtol_rel, maxeval, direct_maxeval
        fr = None
        if frames is not None:
            fr = np.ascontiguousarray(np.asarray(frames, np.float64)[:, :3, :4], np.float32)
            r.n_frames = fr.shape[0]
            r.frames = fr.ctypes.data_as(C.POINTER(C.c_float))
            r.frames_centre = (C.c_float * 3)(*[float(c) for c in frames_centre])
        t = C.c_longlong()
        _E._check(lib().lc2s_register_submit(self.h, C.byref(r), C.byref(t)))
        self._keep[t.value] = x0.size
        return t.value

    def wait(self, ticket: int, timeout_ms: float = -1.0):
        """(x, lc2, evaluations, admit_ms, run_ms); raises LC2Error on failure."""
        res = RegistrationResult()
        rc = lib().lc2s_register_wait(self.h, ticket, timeout_ms, C.byref(res))
        n = self._keep.pop(ticket, 12) if rc != -4 else None   # -4: time-out, still queued
        if rc < 0:
            _E._check(rc)
        return (np.array(res.x[:n]), res.lc2, res.evaluations, res.admit_ms, res.run_ms)

    def stats(self):
        c = RegistrarCounters()
        _E._check(lib().lc2s_registrar_stats(self.h, C.byref(c)))
        return {k: getattr(c, k) for k, _ in RegistrarCounters._fields_}

    def close(self):
        if self.h:
            lib().lc2s_registrar_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
