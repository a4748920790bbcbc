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
                ("queued", C.c_longlong), ("max_in_flight", C.c_int)]


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
