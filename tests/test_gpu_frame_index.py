"""One- and two-frame packed-cache builds through the frame index
(vbt_asset_build_frame_index) equal the builds without it and the oracle,
bit for bit, at every frame: on assets with Grid4, Fine6 and Mode2 leaves,
fp32 keyframes, permuted keyframes (leaves that keep the bracket search) and
a poisoned leaf (the error report, the other slots unchanged)."""
import numpy as np
import pytest

import golden_io
import oracle_py
from test_gpu_assets import shuffle_keyframes
from test_gpu_server import poisoned_liquid
from vbt_amd import synth

hip = pytest.importorskip("vbt_amd.hip")

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def levelset():
    return synth.Workload("ls", synth.LEVELSET, (40, 40, 32, 24), 8, (3.0, 9.0), 15).encode(threads=8)


def temp_fp32():
    return synth.Workload("tf", synth.TEMPERATURE, (48, 40, 24, 20), 8, (2.0, 6.0), 4, burst=0.3,
                          codec=1).encode(threads=8)


def shuffled():
    data = synth.Workload("sd", synth.DENSITY, (48, 48, 32, 30), 10, (2.0, 8.0), 9).encode(threads=8)
    return shuffle_keyframes(data, oracle_py.OracleFile(data), 5)


ASSETS = {
    "smoke64": lambda: golden_io.vbtp("smoke64"),
    "liquid_mini": lambda: golden_io.vbtp("liquid_mini"),
    "levelset40": levelset,
    "temp_fp32": temp_fp32,
    "shuffled": shuffled,
}


@pytest.mark.parametrize("name", sorted(ASSETS))
def test_indexed_builds_equal_the_bracket_search(name):
    data = ASSETS[name]()
    a = hip.Asset(data)
    o = oracle_py.OracleFile(data)
    F = a.dims[3]
    plain = [a.frame_cache(t, 1)[0] for t in range(F)]
    pair = a.frame_cache(F // 2, 2)
    nbytes = a.build_frame_index()
    assert nbytes > 0
    for t in range(F):
        got = a.frame_cache(t, 1)[0]
        assert np.array_equal(bits(got), bits(plain[t])), (name, t)
        exp, _ = o.cache_packed(t)
        assert np.array_equal(bits(got), bits(exp)), (name, t)
    assert np.array_equal(bits(a.frame_cache(F // 2, 2)), bits(pair))
    assert np.array_equal(bits(a.frame_cache(F - 2, 2)), bits(np.stack(plain[F - 2:])))
    # rebuilding and dropping leave the results unchanged
    assert a.build_frame_index() == nbytes
    a.drop_frame_index()
    assert np.array_equal(bits(a.frame_cache(1, 1)[0]), bits(plain[1]))
    a.close()


def test_poisoned_leaf_with_index():
    data, victim, h = poisoned_liquid()
    a = hip.Asset(data)
    F = a.dims[3]
    ref = []
    for t in (0, F // 2, F - 1):
        with pytest.raises(hip.VbtError):
            a.frame_cache(t, 1)
        ref.append(a_cache_no_check(a, t))
    a.build_frame_index()
    for k, t in enumerate((0, F // 2, F - 1)):
        with pytest.raises(hip.VbtError):
            a.frame_cache(t, 1)
        got = a_cache_no_check(a, t)
        assert np.array_equal(bits(got), bits(ref[k]))
    a.close()


def a_cache_no_check(a, t):
    """A one-frame packed cache whose error return is ignored (every slot of
    the leaves that decode holds its value)."""
    out = hip.DeviceArray((1, a.info.controlsPerFrame), np.float32, a.stream)
    try:
        a.frame_cache_into(out, t, 1, hip.CACHE_PACKED)
    except hip.VbtError:
        pass
    return out.host()[0]


@pytest.mark.parametrize("name", ["smoke256", "liquid320"])
def test_full_size_every_frame_against_the_oracle(name):
    """Every frame of the full-size asset built one at a time from the index
    (and in pairs), bit for bit against the oracle and the builds without it."""
    from workloads import THREADS, encoded

    data = encoded(name)
    a = hip.Asset(data)
    o = oracle_py.OracleFile(data)
    F = a.dims[3]
    frames = list(range(F)) if name == "smoke256" else list(range(0, F, 7)) + [F - 1]
    exp = o.cache_packed_frames(frames, threads=THREADS)
    plain = {t: a.frame_cache(t, 1)[0] for t in frames[::5]}
    assert a.build_frame_index() > 0
    for i, t in enumerate(frames):
        got = a.frame_cache(t, 1)[0]
        assert np.array_equal(bits(got), bits(exp[i])), (name, t)
        if t in plain:
            assert np.array_equal(bits(got), bits(plain[t])), (name, t)
    two = a.frame_cache(F - 2, 2)
    assert np.array_equal(bits(two[1]), bits(exp[-1])), name
    a.close()
