"""The benchmark workloads at full size, against the oracle: the Mode2 liquid
level set (320^3 x 150 frames, every leaf active) with its phi sign
checked on every |phi| < 2 voxel, and the fire triple (384x384x256 x 96
frames each) built as one batch.  Sampled frames are compared bit for bit;
the rest is covered by the smaller assets of the other GPU tests."""
import os

import numpy as np
import pytest

import oracle_py
from vbt_amd import synth

hip = pytest.importorskip("vbt_amd.hip")

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def bitexact(a, b, what):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
    assert len(bad) == 0, f"{what}: {len(bad)} mismatches, first at {bad[0]}"


def test_liquid320_levelset_sign_and_parity():
    w = synth.WORKLOADS["liquid320"]
    data = w.encode(threads=THREADS)
    a = hip.Asset(data)
    o = oracle_py.OracleFile(data)
    W, H, D, F = a.dims
    assert (W, H, D, F) == (320, 320, 320, 150)
    modes = np.array(a.info.mode_counts[:])
    assert modes[0] == 0 and modes[2] > 0  # no empty leaves; Mode2 shells present
    field = w.field()
    for t in (0, 75, 149):
        g = a.dense(t)
        e = o.frame(t, threads=THREADS)
        bitexact(g, e, f"liquid320 dense t={t}")
        band = np.abs(e) < 2.0
        assert band.sum() > 100000
        assert (np.signbit(g[band]) == np.signbit(e[band])).all()
        # the encoded asset against the field it was made from: signs agree
        # wherever |phi| exceeds the encoder's keyframe error
        src = field.frame(t, threads=THREADS)
        sb = np.abs(src) < 2.0
        flips = np.signbit(g[sb]) != np.signbit(src[sb])
        assert not (flips & (np.abs(src[sb]) > 0.25)).any(), f"t={t}"
    # packed cache: a run of frames in one launch
    cpf = a.info.controlsPerFrame
    out = hip.DeviceArray((8, cpf), np.float32)
    a.frame_cache_into(out, 71, 8)
    for i in (0, 4, 7):
        bitexact(out.host_slice(i * cpf, cpf), o.cache_packed(71 + i)[0], f"liquid320 cache t={71 + i}")
    a.close()


def test_fire_triple_batch_parity():
    names = ["fire_density", "fire_flames", "fire_temperature"]
    datas = {n: synth.WORKLOADS[n].encode(threads=THREADS) for n in names}
    assets = {n: hip.Asset(d) for n, d in datas.items()}
    jobs, outs = [], {}
    for n in names:
        a = assets[n]
        F = a.dims[3]
        assert a.dims == (384, 384, 256, 96)
        outs[n] = hip.DeviceArray((F, a.info.controlsPerFrame), np.float32)
        jobs.append((a, 0, F, hip.CACHE_PACKED, outs[n]))
    b = hip.CacheBatch(jobs)
    b.autotune(reps=1)
    b.launch()
    b.stream.sync()
    for n in names:
        o = oracle_py.OracleFile(datas[n])
        cpf = assets[n].info.controlsPerFrame
        for t in (0, 47, 95):
            bitexact(outs[n].host_slice(t * cpf, cpf), o.cache_packed(t)[0], f"{n} batch frame {t}")
    b.close()
    for a in assets.values():
        a.close()
