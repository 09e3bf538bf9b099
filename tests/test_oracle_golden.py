"""The oracle against the golden vectors: assets encoded by the reference
encoders and decoded by the reference decoder (tests/golden/make_golden.py)."""
import ctypes

import numpy as np
import pytest

import golden_io
import oracle_py

RENDER = ["smoke64", "liquid_mini"]
SCI = ["sci_mini", "sci_dense_mini", "sci_tile_mini", "sci_tile4_mini"]


@pytest.mark.parametrize("name", RENDER)
def test_render_controls_match_reference(name):
    f = oracle_py.OracleFile(golden_io.vbtp(name))
    exp = golden_io.expected(name)
    O = oracle_py.oracle()
    val = ctypes.c_float()
    F = f.header.frames
    for row in range(len(exp["leaf"])):
        leaf, t = int(exp["leaf"][row]), int(exp["frame"][row])
        b = np.ascontiguousarray(f.leaf_bytes(leaf))
        v = f.parse(leaf)
        cc = v.header.coarseResolution ** 3
        got = np.zeros(64, np.float32)
        for i in range(cc):
            assert O.vbto_decode_leaf_control_value(b.ctypes.data, ctypes.byref(v), 0, 1, F, i, t,
                                                    ctypes.byref(val), None) == 0
            got[i] = val.value
        np.testing.assert_array_equal(got.view(np.uint32), exp["coarse"][row].view(np.uint32))
        if v.header.mode == 3:
            gotf = np.zeros(216, np.float32)
            for i in range(v.header.fineResolution ** 3):
                assert O.vbto_decode_leaf_control_value(b.ctypes.data, ctypes.byref(v), 1, 1, F, i, t,
                                                        ctypes.byref(val), None) == 0
                gotf[i] = val.value
            np.testing.assert_array_equal(gotf.view(np.uint32), exp["fine"][row].view(np.uint32))
        if v.header.mode == 2:
            gots = np.zeros(512, np.float32)
            for vox in range(512):
                assert O.vbto_decode_shell_voxel_value(b.ctypes.data, ctypes.byref(v), 1, F, vox, t,
                                                       ctypes.byref(val), None) == 0
                gots[vox] = val.value
            np.testing.assert_array_equal(gots.view(np.uint32), exp["shell"][row].view(np.uint32))


@pytest.mark.parametrize("name", SCI)
def test_scientific_frames_and_queries_match_reference(name):
    f = oracle_py.OracleFile(golden_io.vbtp(name))
    exp = golden_io.expected(name)
    for i, t in enumerate(exp["frames"]):
        got = f.frame(int(t))
        np.testing.assert_array_equal(got.view(np.uint32), exp["dense"][i].view(np.uint32))
    got = f.query(exp["queries"])
    np.testing.assert_array_equal(got.view(np.uint32), exp["values"].view(np.uint32))


def test_golden_assets_cover_every_route():
    m = golden_io.manifest()
    render_modes = np.sum([m[n]["modes"] for n in RENDER], axis=0)
    sci_modes = np.sum([m[n]["modes"] for n in SCI], axis=0)
    assert all(render_modes > 0), render_modes  # EMPTY, GRID4, Mode2 shell, FINE6
    assert sci_modes[1] > 0 and sci_modes[2] > 0 and sci_modes[3] > 0  # sparse, dense grid3, dense grid4


def test_dense_subtypes_2_and_3_follow_the_reference_decoder():
    """The reference file decoder has no reader for the tile / patch temporal
    bases (dense subtypes 2 and 3): it reads them as per-frame BFP blocks.
    For the tile fixture that decodes (matched bit for bit above); the patch
    fixture's reads run past the payload pool, where the reference throws and
    the oracle reports an out-of-range read."""
    m = golden_io.manifest()
    for name, sub, mode in (("sci_tile_mini", 2, 2), ("sci_tile4_mini", 2, 3), ("sci_patch_mini", 3, 2),
                            ("sci_patch_b2_mini", 3, 2)):
        f = oracle_py.OracleFile(golden_io.vbtp(name))
        words = f.payload.view(np.uint32)
        hdr = words[f.offsets[:-1][f.offsets[1:] > f.offsets[:-1]]]
        assert ((hdr & 3) >= 2).all() and (((hdr >> 12) & 7) == sub).all(), name
        assert m[name]["modes"][mode] == 48
    for name in ("sci_patch_mini", "sci_patch_b2_mini"):
        exp = golden_io.expected(name)
        assert "range_check" in str(exp["error"])
        f = oracle_py.OracleFile(golden_io.vbtp(name))
        for i, t in enumerate(exp["frames"]):  # frames the reference decoded before throwing
            np.testing.assert_array_equal(f.frame(int(t)).view(np.uint32), exp["dense"][i].view(np.uint32))
        with pytest.raises(ValueError, match="out of range"):
            f.frame(int(exp["error_frame"]))

