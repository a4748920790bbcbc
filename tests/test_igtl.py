"""OpenIGTLink TRANSFORM streaming (lc2igtl.h, the reference's pose server at
interpolator.cu:359-388 / :545-573) on the CPU: message bytes parsed back field by field,
the CRC-64 checked against an independent bitwise ECMA-182 implementation (pinned by its
published check value), and the server over a loopback socket."""
import ctypes as C
import os
import socket
import struct
import threading
import time

import numpy as np
import pytest

from lc2registration_amd import engine as E

HDR = struct.Struct(">H12s20sQQQ")  # version, type, device, time stamp, body size, CRC


def _lib():
    if not os.path.exists(E.library_path()):
        from lc2registration_amd import build
        build.build()
    L = C.CDLL(E.library_path())
    L.lc2igtl_last_error.restype = C.c_char_p
    L.lc2igtl_pack_transform.argtypes = [C.c_char_p, C.POINTER(C.c_float), C.c_double, C.c_int,
                                         C.c_uint, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
    L.lc2igtl_server_create.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_void_p)]
    L.lc2igtl_port.argtypes = [C.c_void_p]
    L.lc2igtl_wait_client.argtypes = [C.c_void_p, C.c_double]
    L.lc2igtl_send_transform.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_float)]
    L.lc2igtl_sent.argtypes = [C.c_void_p]
    L.lc2igtl_sent.restype = C.c_longlong
    L.lc2igtl_server_destroy.argtypes = [C.c_void_p]
    L.lc2igtl_server_destroy.restype = None
    return L


def crc64_ecma(data: bytes) -> int:
    """Bitwise CRC-64, polynomial 0x42F0E1EBA9EA3693, MSB first, init 0, no final xor."""
    crc = 0
    for b in data:
        crc ^= b << 56
        for _ in range(8):
            crc = ((crc << 1) ^ 0x42F0E1EBA9EA3693) if crc & (1 << 63) else (crc << 1)
            crc &= (1 << 64) - 1
    return crc


def test_crc64_check_value():
    assert crc64_ecma(b"123456789") == 0x6C40DF5F0B497347  # CRC-64/ECMA-182 check value


def parse(msg: bytes):
    v, typ, dev, ts, size, crc = HDR.unpack_from(msg, 0)
    body = msg[HDR.size:HDR.size + size]
    assert len(body) == size
    out = {"version": v, "type": typ.rstrip(b"\0"), "device": dev.rstrip(b"\0"), "ts": ts,
           "crc_ok": crc == crc64_ecma(body)}
    if v == 2:
        ext, meta_hdr, meta, mid = struct.unpack_from(">HHII", body, 0)
        content = body[ext:ext + 48]
        (count,) = struct.unpack_from(">H", body, ext + 48)
        out.update(ext=ext, meta_hdr=meta_hdr, meta=meta, id=mid, index_count=count)
    else:
        content = body
    f = struct.unpack(">12f", content)
    M = np.eye(4, dtype=np.float32)
    M[:3, 0], M[:3, 1], M[:3, 2], M[:3, 3] = f[0:3], f[3:6], f[6:9], f[9:12]
    out["matrix"] = M
    return out


def _matrix(seed):
    rng = np.random.default_rng(seed)
    M = np.eye(4, dtype=np.float32)
    M[:3, :4] = rng.normal(size=(3, 4)).astype(np.float32)
    return M


@pytest.mark.parametrize("version", [1, 2])
def test_pack_transform_fields(version):
    L = _lib()
    M = _matrix(version)
    n = C.c_size_t()
    mp = M.ravel().ctypes.data_as(C.POINTER(C.c_float))
    assert L.lc2igtl_pack_transform(b"RegistrationServer", mp, 0.0, version, 7, None, 0, C.byref(n)) == 0
    buf = (C.c_ubyte * n.value)()
    ts = 1700000000.25
    assert L.lc2igtl_pack_transform(b"RegistrationServer", mp, ts, version, 7, buf, n.value,
                                    C.byref(n)) == 0
    m = parse(bytes(buf))
    assert m["version"] == version and m["type"] == b"TRANSFORM"
    assert m["device"] == b"RegistrationServer" and m["crc_ok"]
    assert m["ts"] >> 32 == 1700000000 and (m["ts"] & 0xFFFFFFFF) == 1 << 30
    assert np.array_equal(m["matrix"], M)
    if version == 2:
        assert (m["ext"], m["meta_hdr"], m["meta"], m["id"], m["index_count"]) == (12, 2, 0, 7, 0)
        assert n.value == 58 + 12 + 48 + 2
    else:
        assert n.value == 58 + 48


def test_loopback_server_streams_transforms():
    L = _lib()
    h = C.c_void_p()
    assert L.lc2igtl_server_create(0, 1, C.byref(h)) == 0
    port = L.lc2igtl_port(h)
    assert port > 0
    assert L.lc2igtl_wait_client(h, 50.0) == 0  # nobody yet: time-out
    got = []

    def client():
        with socket.create_connection(("127.0.0.1", port), timeout=10) as c:
            data = b""
            while len(data) < 3 * 120:
                chunk = c.recv(4096)
                if not chunk:
                    break
                data += chunk
            got.append(data)

    t = threading.Thread(target=client)
    t.start()
    assert L.lc2igtl_wait_client(h, 5000.0) == 1
    mats = [_matrix(10 + k) for k in range(3)]
    t0 = time.time()
    for M in mats:
        assert L.lc2igtl_send_transform(h, b"RegistrationServer",
                                        M.ravel().ctypes.data_as(C.POINTER(C.c_float))) == 0
    t.join(10)
    assert L.lc2igtl_sent(h) == 3
    L.lc2igtl_server_destroy(h)
    data = got[0]
    assert len(data) == 3 * 120
    for k in range(3):
        m = parse(data[120 * k:120 * (k + 1)])
        assert m["crc_ok"] and m["id"] == k and m["version"] == 2
        assert abs((m["ts"] >> 32) - t0) < 5
        assert np.array_equal(m["matrix"], mats[k])

