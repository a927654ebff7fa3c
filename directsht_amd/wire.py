"""Binary request/response frames for the transform server.

The reference has no serving layer; a transform is a Python call
(sht/sht.py:112-150).  A remote caller of TransformServer sends one request
frame per transform and gets one response frame back.  All fields are
little-endian; arrays follow their header with no padding.

Request  (32-byte header + 24 N bytes):
    0  magic  b"DSHQ"
    4  u16    version (1)
    6  u16    flags (0)
    8  i32    Nl (2 .. 4096)
   12  i32    Nx (0: the server's default, nx_factor * Nl)
   16  i64    N, the number of points
   24  u64    request id (echoed in the response)
   32  f64[N] theta, then f64[N] phi, then f64[N] weights

Response (32-byte header + payload):
    0  magic  b"DSHR"
    4  u16    version (1)
    6  u16    status: 0 ok, 1 malformed request, 2 rejected (HBM budget),
              3 transform failed (e.g. theta out of range)
    8  i32    Nl
   12  i32    0
   16  i64    payload length: Nl(Nl+1)/2 complex128 alm (healpy order) if
              status is 0, else the bytes of a UTF-8 message
   24  u64    request id
   32  payload
"""
import struct

import numpy as np

REQ_MAGIC, RESP_MAGIC, VERSION = b"DSHQ", b"DSHR", 1
OK, MALFORMED, REJECTED, FAILED = 0, 1, 2, 3
_HDR = struct.Struct("<4sHHiiqQ")  # 32 bytes
HEADER_BYTES = _HDR.size
MAX_NL = 4096   # the library's transform band limit
MAX_NX = 1 << 20  # its grid limit


def encode_request(theta, phi, wt, Nl, Nx=0, rid=0):
    """Request frame for points (theta, phi, wt) at band limit Nl."""
    t, p, w = (np.ascontiguousarray(a, dtype="<f8").reshape(-1) for a in (theta, phi, wt))
    if not (t.size == p.size == w.size):
        raise ValueError("theta, phi and wt must have the same length")
    return b"".join([_HDR.pack(REQ_MAGIC, VERSION, 0, int(Nl), int(Nx), t.size, int(rid)),
                     t.tobytes(), p.tobytes(), w.tobytes()])


def decode_request(buf):
    """(rid, Nl, Nx, theta, phi, wt) of a request frame; the arrays are
    read-only views of buf.  Raises ValueError for a malformed frame."""
    mv = memoryview(buf)
    if len(mv) < HEADER_BYTES:
        raise ValueError("request shorter than its header")
    magic, ver, _flags, Nl, Nx, n, rid = _HDR.unpack_from(mv, 0)
    if magic != REQ_MAGIC:
        raise ValueError("not a request frame (bad magic)")
    if ver != VERSION:
        raise ValueError(f"unsupported request version {ver}")
    if not 2 <= Nl <= MAX_NL:
        raise ValueError(f"Nl={Nl} outside 2 .. {MAX_NL}")
    if Nx != 0 and not 2 <= Nx <= MAX_NX:
        raise ValueError(f"Nx={Nx} must be 0 (default) or in 2 .. {MAX_NX}")
    if n < 0 or len(mv) != HEADER_BYTES + 24 * n:
        raise ValueError(f"request length {len(mv)} does not hold {n} points")
    arr = np.frombuffer(mv, dtype="<f8", count=3 * n, offset=HEADER_BYTES)
    return rid, Nl, Nx, arr[:n], arr[n:2 * n], arr[2 * n:]


def encode_response(rid, Nl, alm):
    """Response frame carrying alm (Nl(Nl+1)/2 complex values)."""
    a = np.ascontiguousarray(alm, dtype="<c16").reshape(-1)
    if a.size != Nl * (Nl + 1) // 2:
        raise ValueError("alm size does not match Nl")
    return _HDR.pack(RESP_MAGIC, VERSION, OK, int(Nl), 0, a.size, int(rid)) + a.tobytes()


def encode_error(rid, status, message, Nl=0):
    """Response frame for a request that produced no alm."""
    msg = str(message).encode("utf-8")
    return _HDR.pack(RESP_MAGIC, VERSION, int(status), int(Nl), 0, len(msg), int(rid)) + msg


def decode_response(buf):
    """(rid, status, payload): alm (read-only complex128 view) when status is
    OK, else the error message.  Raises ValueError for a malformed frame."""
    mv = memoryview(buf)
    if len(mv) < HEADER_BYTES:
        raise ValueError("response shorter than its header")
    magic, ver, status, Nl, _z, n, rid = _HDR.unpack_from(mv, 0)
    if magic != RESP_MAGIC or ver != VERSION:
        raise ValueError("not a version-1 response frame")
    if status == OK:
        if n != Nl * (Nl + 1) // 2 or len(mv) != HEADER_BYTES + 16 * n:
            raise ValueError("response length does not match Nl")
        return rid, status, np.frombuffer(mv, dtype="<c16", count=n, offset=HEADER_BYTES)
    if len(mv) != HEADER_BYTES + n:
        raise ValueError("error response length does not match its message")
    return rid, status, bytes(mv[HEADER_BYTES:]).decode("utf-8", errors="replace")
