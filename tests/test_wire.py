"""Request/response frames of the transform server (directsht_amd.wire):
layout, round trips and malformed frames.  No GPU."""
import struct

import numpy as np
import pytest

from directsht_amd import wire


def test_request_round_trip_and_layout():
    rng = np.random.default_rng(0)
    t, p, w = rng.uniform(0, 3, 5), rng.uniform(0, 6, 5), rng.uniform(0, 1, 5)
    f = wire.encode_request(t, p, w, 64, 0, rid=77)
    assert len(f) == 32 + 24 * 5
    assert f[:4] == b"DSHQ" and struct.unpack_from("<HHiiqQ", f, 4) == (1, 0, 64, 0, 5, 77)
    assert np.frombuffer(f[32:72], "<f8").tolist() == t.tolist()
    rid, Nl, Nx, t2, p2, w2 = wire.decode_request(f)
    assert (rid, Nl, Nx) == (77, 64, 0)
    np.testing.assert_array_equal(t2, t)
    np.testing.assert_array_equal(p2, p)
    np.testing.assert_array_equal(w2, w)
    # big-endian input arrays are converted, empty requests are legal
    rid, Nl, Nx, t3, _, _ = wire.decode_request(wire.encode_request(t.astype(">f8"), p, w, 8, 16))
    np.testing.assert_array_equal(t3, t)
    assert Nx == 16
    assert wire.decode_request(wire.encode_request([], [], [], 4))[3].size == 0


def test_response_round_trip():
    Nl = 9
    a = np.arange(45) + 1j * np.arange(45)[::-1]
    rid, st, alm = wire.decode_response(wire.encode_response(5, Nl, a))
    assert (rid, st) == (5, wire.OK)
    np.testing.assert_array_equal(alm, a)
    rid, st, msg = wire.decode_response(wire.encode_error(6, wire.REJECTED, "too big", Nl))
    assert (rid, st, msg) == (6, wire.REJECTED, "too big")
    with pytest.raises(ValueError):
        wire.encode_response(1, Nl, a[:-1])


@pytest.mark.parametrize("mutate", ["magic", "version", "short", "long", "nl", "nx", "nx_huge",
                                    "header"])
def test_malformed_requests(mutate):
    f = bytearray(wire.encode_request(np.ones(3), np.ones(3), np.ones(3), 32))
    if mutate == "magic":
        f[:4] = b"XXXX"
    elif mutate == "version":
        struct.pack_into("<H", f, 4, 2)
    elif mutate == "short":
        f = f[:-8]
    elif mutate == "long":
        f += b"\0" * 8
    elif mutate == "nl":
        struct.pack_into("<i", f, 8, 4097)
    elif mutate == "nx":
        struct.pack_into("<i", f, 12, 1)
    elif mutate == "nx_huge":  # segment and pair counts of such a grid overflow int32
        struct.pack_into("<i", f, 12, 1 << 30)
    elif mutate == "header":
        f = f[:20]
    with pytest.raises(ValueError):
        wire.decode_request(bytes(f))
    with pytest.raises(ValueError):
        wire.encode_request(np.ones(3), np.ones(2), np.ones(3), 32)
