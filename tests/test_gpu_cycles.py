"""The Cycles resident blob packed on the device (vbt_cycles_pack_blob) and
its two density samplers (vbt_cycles_sample_density), byte for byte and bit
for bit against the oracle's restatement of cycles_native_vbt/ and
blender_patch/native_cycles_vbt.patch:688-929 (itself pinned against the
reference's CPU probe by tests/test_cycles_oracle.py)."""
import ctypes

import numpy as np
import pytest

import golden_io
import oracle_py
from vbt_amd import synth

hip = pytest.importorskip("vbt_amd.hip")

pytestmark = pytest.mark.gpu

FORMS = {"standalone": 0, "standalone-plain": hip.CYCLES_PLAIN, "kernel": hip.CYCLES_KERNEL,
         "kernel-plain": hip.CYCLES_KERNEL | hip.CYCLES_PLAIN}


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def points(W, H, D, n, seed):
    rng = np.random.default_rng(seed)
    p = rng.random((n, 3)) * [W + 2, H + 2, D + 2] - 1
    q = n // 5
    p[:q] = np.floor(p[:q])
    p[q:2 * q] = np.round(p[q:2 * q] / 8) * 8 + rng.choice([-1e-3, 0.0, 1e-3], size=(q, 3))
    return p.astype(np.float32)


ASSETS = {
    "smoke64": lambda: golden_io.vbtp("smoke64"),
    "liquid_mini": lambda: golden_io.vbtp("liquid_mini"),
    "levelset40": lambda: synth.Workload("ls", synth.LEVELSET, (40, 40, 32, 24), 8, (3.0, 9.0), 15).encode(threads=8),
    "temp_fp32": lambda: synth.Workload("tf", synth.TEMPERATURE, (48, 40, 24, 20), 8, (2.0, 6.0), 4, burst=0.3,
                                        codec=1).encode(threads=8),
}


@pytest.fixture(scope="module", params=sorted(ASSETS))
def asset(request):
    data = ASSETS[request.param]()
    a = hip.Asset(data)
    yield request.param, data, a
    a.close()


def test_blob_bytes_equal_the_oracle(asset):
    name, data, a = asset
    W, H, D, F = a.dims
    for frame in (-4, 0, F // 2, F + 3):
        dev = a.cycles_blob(frame).host()
        exp = oracle_py.cycles_blob(data, frame)
        assert dev.shape == exp.shape and np.array_equal(dev, exp), (name, frame)


@pytest.mark.parametrize("form", sorted(FORMS))
def test_samples_equal_the_oracle(asset, form):
    name, data, a = asset
    W, H, D, F = a.dims
    flags = FORMS[form]
    for frame in (0, F // 3, F - 1):
        blob = a.cycles_blob(frame)
        hb = oracle_py.cycles_blob(data, frame)
        p = points(W, H, D, 5000, frame + 11)
        got = hip.cycles_sample(blob, p, flags)
        exp = oracle_py.cycles_sample(hb, p, flags, threads=8)
        bad = np.flatnonzero(bits(got) != bits(exp))
        assert len(bad) == 0, (name, form, frame, len(bad), p[bad[:3]], got[bad[:3]], exp[bad[:3]])
        assert (exp != 0).mean() > 0.02


def test_world_flag_and_bad_headers():
    data = golden_io.vbtp("smoke64")
    a = hip.Asset(data)
    blob = a.cycles_blob(9)
    hb = blob.host()
    hb[64:76] = np.frombuffer(np.array([-12.5, 3.25, 100.0], np.float32).tobytes(), np.uint8)
    dblob = hip.DeviceArray.from_host(hb, a.stream)
    p = points(64, 64, 64, 3000, 4) + np.array([-12.5, 3.25, 100.0], np.float32)
    for flags in (hip.CYCLES_WORLD, hip.CYCLES_WORLD | hip.CYCLES_KERNEL, hip.CYCLES_WORLD | hip.CYCLES_PLAIN):
        assert np.array_equal(bits(hip.cycles_sample(dblob, p, flags)),
                              bits(oracle_py.cycles_sample(hb, p, flags, threads=8)))
    # a header that fails the view's checks samples to 0 in every form
    for pos, val in ((0, 0), (4, 2), (56, 255)):
        bad = hb.copy()
        bad[pos] = val
        dbad = hip.DeviceArray.from_host(bad, a.stream)
        for flags in FORMS.values():
            assert not hip.cycles_sample(dbad, p[:500], flags).any()
    # a blob size that cuts the payload short is refused the same way
    L = hip.lib()
    dp = hip.DeviceArray.from_host(p[:500], a.stream)
    out = hip.DeviceArray(500, np.float32, a.stream)
    out.fill_bytes(0x7F)
    hip.check(L.vbt_cycles_sample_density(blob.ptr, blob.nbytes - 4, 0, dp.ptr, 500, out.ptr, a.stream.h))
    assert not out.host().any()
    with pytest.raises(hip.VbtError):
        hip.check(L.vbt_cycles_pack_blob(a.h, 0, blob.ptr, 100, a.stream.h))
    with pytest.raises(hip.VbtError):
        hip.check(L.vbt_cycles_sample_density(blob.ptr, blob.nbytes, 64, None, 1, None, a.stream.h))
    assert L.vbt_cycles_sample_density(blob.ptr, blob.nbytes, 0, None, 0, None, a.stream.h) == 0
    n = ctypes.c_uint64()
    hip.check(L.vbt_cycles_blob_bytes(a.h, ctypes.byref(n)))
    assert n.value == len(hb)
    a.close()


@pytest.mark.parametrize("name", ["smoke256", "liquid320"])
def test_full_size_assets(name):
    """The benchmark assets at full size: 20,000 points per form, bit for bit."""
    w = synth.WORKLOADS[name]
    data = w.encode(threads=16)
    a = hip.Asset(data)
    W, H, D, F = a.dims
    frame = F // 2
    blob = a.cycles_blob(frame)
    hb = blob.host()
    assert np.array_equal(hb, oracle_py.cycles_blob(data, frame))
    p = points(W, H, D, 20000, 5)
    for flags in FORMS.values():
        got = hip.cycles_sample(blob, p, flags)
        exp = oracle_py.cycles_sample(hb, p, flags, threads=16)
        assert np.array_equal(bits(got), bits(exp)), (name, flags)
    a.close()
