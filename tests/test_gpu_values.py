"""GPU per-value decoders (batched forms of decodeRenderTemporalControlValue,
decodeRenderTemporalLeafControlValue, decodeRenderTemporalLeafShellVoxelValue)
on the reference unit tests' synthetic streams, against the oracle bit for bit
and against the std::lower_bound reference within the reference's 2e-6."""
import ctypes

import numpy as np
import pytest

import oracle_py
import streams

hip = pytest.importorskip("vbt_amd.hip")

pytestmark = pytest.mark.gpu


def oracle_values(packed: bytes, s, frames, ts):
    O = oracle_py.oracle()
    buf = np.frombuffer(packed + b"\0\0\0\0", np.uint8)
    out = np.empty(len(ts), np.float32)
    val = ctypes.c_float()
    for i, t in enumerate(ts):
        assert O.vbto_decode_control_value(buf.ctypes.data, s.codec, s.count, s.bins, 1, len(s.keys), len(packed), 1,
                                           frames, 0, int(t), ctypes.byref(val), None) == 0
        out[i] = val.value
    return out


@pytest.mark.parametrize("codec", [0, 1])
@pytest.mark.parametrize("bins", [4, 8])
def test_control_values_all_frames_and_hints(codec, bins):
    for frames in range(1, 256):
        for dense, hint in [(True, -1), (False, -1), (False, 0), (False, 255)]:
            s = streams.make_stream(frames, bins, codec, dense)
            if hint >= 0:
                s.bin_index = [hint] * len(s.bin_index)
            packed = s.pack()
            ts = np.arange(-2, frames + 3, dtype=np.int32)
            got, st = hip.decode_control_values(packed, codec, 1, bins, len(s.keys), frames,
                                                np.zeros(len(ts), np.uint16), ts)
            assert (st == 0).all()
            exp = oracle_values(packed, s, frames, ts)
            np.testing.assert_array_equal(got.view(np.uint32), exp.view(np.uint32),
                                          err_msg=f"frames={frames} dense={dense} hint={hint}")
            if frames % 17 == 1 or frames > 250:
                ref = np.array([streams.lower_bound_reference(s.keys, frames, int(t)) for t in ts], np.float32)
                assert np.max(np.abs(got - ref)) <= 2e-6


def test_shell_voxels_match_shell_test():
    leaf, residual, voxels = streams.shell_test_leaf()
    vox = np.repeat(np.arange(512, dtype=np.uint16), 103)
    ts = np.tile(np.arange(-1, 102, dtype=np.int32), 512)
    got, st = hip.decode_leaf_values(leaf, 101, vox, ts, shell=True)
    assert (st == 0).all()
    for v in range(512):
        sl = slice(v * 103, (v + 1) * 103)
        if v in voxels:
            ref = np.array([streams.lower_bound_reference(residual.keys, 101, int(t)) for t in ts[sl]], np.float32)
            assert np.max(np.abs(got[sl] - ref)) <= 2e-6
        else:
            assert (got[sl] == 0).all()


def test_leaf_control_values_against_oracle():
    leaf, _, _ = streams.shell_test_leaf()
    buf = np.frombuffer(leaf, np.uint8)
    v = oracle_py.LeafView()
    O = oracle_py.oracle()
    assert O.vbto_parse_leaf(buf.ctypes.data, len(leaf), 1, ctypes.byref(v)) == 0
    ci = np.repeat(np.arange(64, dtype=np.uint16), 40)
    ts = np.tile(np.arange(-3, 37, dtype=np.int32) * 3, 64)
    got, st = hip.decode_leaf_values(leaf, 101, ci, ts)
    assert (st == 0).all()
    val = ctypes.c_float()
    for i in range(0, len(ci), 7):
        assert O.vbto_decode_leaf_control_value(buf.ctypes.data, ctypes.byref(v), 0, 1, 101, int(ci[i]), int(ts[i]),
                                                ctypes.byref(val), None) == 0
        assert np.float32(val.value).view(np.uint32) == got[i].view(np.uint32)


def test_error_statuses():
    s = streams.make_stream(20, 8, 0, False)
    packed = s.pack()
    # control index out of range -> per-request status, not a failed batch
    got, st = hip.decode_control_values(packed, 0, 1, 8, len(s.keys), 20, np.array([0, 1], np.uint16),
                                        np.array([3, 3], np.int32))
    assert st[0] == 0 and st[1] == 34 and np.isnan(got[1])
    assert hip.status_message(34) == "render temporal control index out of range"
    # stream too short for its declared layout
    got, st = hip.decode_control_values(packed, 0, 1, 8, len(s.keys), 20, np.array([0], np.uint16),
                                        np.array([3], np.int32), stream_limit=len(packed) - 4)
    assert st[0] == 10


def test_sweep_lerp_weights_for_every_gap():
    """The frame-cache sweep forms (frame - f0) / (f1 - f0) from a register
    reciprocal; every weight with gaps up to 65535 (fp32 key frames are 16-bit)
    must equal the correctly rounded float quotient of the reference."""
    assert hip.selftest_weights(65535) == 0
    assert hip.selftest_weights(1) == 0
