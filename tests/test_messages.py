"""View-request messages (CPU): the reference's libNetwork wire layout
(Message.hpp:18-60: a uint32 id, then the payload, at most 32768 bytes;
MessageCamera.cpp:24-45, MessageTransferFunction.cpp, MessageBoundingBox.cpp,
ids from MessageFactory.cpp) and their folding into a view request."""
import ctypes as C
import struct

import numpy as np
import pytest

import gcvol as G


def test_wire_layout(gcv):
    m = np.arange(12, dtype=np.float32) * 0.25 - 1.0
    cam = G.msg_camera(m)
    assert cam == struct.pack("<I12f", 3, *m.tolist())
    assert len(cam) == 4 + 12 * 4                              # MessageCamera::_size
    rng = np.random.default_rng(1)
    tf = rng.random((1024, 4)).astype(np.float32)
    msg = G.msg_transfer_function(tf)
    assert len(msg) == 4 + 256 * 4 * 16                         # MessageTransferFunction::_size
    assert msg[:4] == struct.pack("<I", 4) and msg[4:] == tf.astype("<f4").tobytes()
    box = G.msg_bounding_box((-1, -0.5, -0.25), (1, 0.5, 0.25))
    assert box == struct.pack("<I6f", 5, -1, -0.5, -0.25, 1, 0.5, 0.25)
    for msg_, want in ((cam, 3), (msg, 4), (box, 5)):
        i = G.u32(0)
        assert G.lib().gc_msg_id(msg_, len(msg_), C.byref(i)) == 0 and i.value == want
    # the largest transfer function a 32768-byte message holds
    assert len(G.msg_transfer_function(np.zeros((2047, 4), np.float32))) == 4 + 2047 * 16 <= 32768
    with pytest.raises(G.GcError):
        G.msg_transfer_function(np.zeros((2048, 4), np.float32))


def test_request_folds_messages_into_a_view(gcv):
    r = G.Request()
    with pytest.raises(G.GcError):
        r.view(64, 48)                                          # no camera yet
    cam = G.camera(20.0, 35.0, 2.5)
    tf = np.random.default_rng(2).random((256, 4)).astype(np.float32)
    for msg in (G.msg_camera(cam), G.msg_transfer_function(tf), G.msg_bounding_box((-1, -1, -0.5), (1, 0.75, 1)),
                struct.pack("<II", 1, 0), struct.pack("<I", 2) + bytes(100)):   # configuration, H.264: ignored
        r.apply(msg)
    v = r.view(64, 48, precision=64)
    direct = G.make_view(64, 48, cam, precision=64, bbox=((-1, -1, -0.5), (1, 0.75, 1)))
    assert bytes(v) == bytes(direct)
    assert np.array_equal(r.tf(), tf) and r.r.have == 7
    # a later camera replaces the earlier one
    cam2 = G.camera(-5.0, 100.0, 3.0)
    r.apply(G.msg_camera(cam2))
    assert bytes(r.view(64, 48, precision=64)) == bytes(G.make_view(64, 48, cam2, precision=64,
                                                                     bbox=((-1, -1, -0.5), (1, 0.75, 1))))


@pytest.mark.parametrize("bad", [b"", b"\x03\x00", struct.pack("<I11f", 3, *range(11)),
                                 struct.pack("<I", 4) + bytes(15), struct.pack("<I", 4),
                                 struct.pack("<I5f", 5, *range(5)), struct.pack("<I", 9) + bytes(8),
                                 struct.pack("<I", 3) + bytes(32768)])
def test_malformed_messages_leave_the_request_unchanged(gcv, bad):
    r = G.Request()
    r.apply(G.msg_camera(G.camera(1.0, 2.0, 3.0)))
    before = bytes(r.r)
    with pytest.raises(G.GcError):
        r.apply(bad)
    assert bytes(r.r) == before
