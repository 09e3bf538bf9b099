"""The frame-cache build's traffic without its decode (vbt_asset_cache_ceiling,
the measurement aid behind bench.py's layout ceiling): it writes every slot of
every frame, rejects what it cannot run, and leaves the asset's own builds
unchanged."""
import numpy as np
import pytest

import golden_io
import oracle_py

hip = pytest.importorskip("vbt_amd.hip")

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tile", [0, 4, 5])
def test_ceiling_writes_every_slot_of_every_frame(tile):
    data = golden_io.vbtp("smoke64")
    a = hip.Asset(data)
    F, n = a.dims[3], a.info.controlsPerFrame
    out = hip.DeviceArray((F, n), np.float32, a.stream)
    for reads in (False, True):
        out.fill_bytes(0xFF)
        a.cache_ceiling_into(out, F, tile, reads=reads)
        a.stream.sync()
        got = out.host().view(np.uint32)
        if not reads:  # placeholder values: the lane's index in its workgroup + the frame
            assert (got < 0x10000).all()
        else:  # a staged word + the frame: all ones only by chance
            assert (got != 0xFFFFFFFF).mean() > 0.999
    # the asset's build is unaffected
    exp = oracle_py.OracleFile(data).cache_packed_frames(range(F), threads=4)
    a.cache_tile = tile
    np.testing.assert_array_equal(a.frame_cache(0, F).view(np.uint32), exp.view(np.uint32))
    out.free()
    a.close()


def test_ceiling_rejects_what_it_cannot_run():
    a = hip.Asset(golden_io.vbtp("smoke64"))
    out = hip.DeviceArray((1, a.info.controlsPerFrame), np.float32, a.stream)
    for nt, tile in ((0, 0), (a.dims[3] + 1, 0), (1, 99), (1, -1)):
        with pytest.raises(hip.VbtError):
            a.cache_ceiling_into(out, nt, tile)
    s = hip.Asset(golden_io.vbtp("sci_mini"))
    with pytest.raises(hip.VbtError):
        s.cache_ceiling_into(out, 1, 0)
    out.free()
    a.close()
    s.close()
