"""Batched frame-cache builds (vbt_cache_batch_*): one graph replay decodes
several assets' frame ranges -- render jobs through the persistent work-queue
kernel, scientific jobs on a concurrent branch -- and every job's output must
equal the oracle bit for bit, under every queue shape, on every replay of the
graph (with other work between replays), including a poisoned leaf."""
import struct

import numpy as np
import pytest

import golden_io
import oracle_py
from vbt_amd import synth

hip = pytest.importorskip("vbt_amd.hip")

pytestmark = pytest.mark.gpu


def bitexact(a, b, what):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    bad = np.argwhere(a.view(np.uint32) != b.view(np.uint32))
    assert len(bad) == 0, f"{what}: {len(bad)} mismatches, first {tuple(bad[0])}: {a[tuple(bad[0])]} vs {b[tuple(bad[0])]}"


def scramble(data: bytes, o) -> bytes:
    """Swap the first two keyframes of the first multi-key coarse control of
    every 3rd active leaf (those leaves take the exact bracket search)."""
    buf = bytearray(data)
    base = 52 + 4 * o.header.coarseAcScaleCount + 4 * (o.header.leafCount + 1)
    compact, _ = o.page_table()
    for leaf in np.nonzero(compact != 0xFFFFFFFF)[0][::3]:
        v = o.parse(int(leaf))
        lb = base + 4 * int(o.offsets[leaf])
        kb = 8 if v.header.coarseCodec == 1 else 3
        key0 = lb + v.layout.coarse.keyframeOffset
        for c in range(v.header.coarseResolution ** 3):
            ks, kc = struct.unpack_from("<HB", buf, lb + 8 + 4 * c)
            if kc >= 2:
                a0, a1 = key0 + kb * ks, key0 + kb * (ks + 1)
                r0, r1 = bytes(buf[a0:a0 + kb]), bytes(buf[a1:a1 + kb])
                buf[a0:a0 + kb], buf[a1:a1 + kb] = r1, r0
                break
    return bytes(buf)


@pytest.fixture(scope="module")
def world():
    """Assets of every route and codec, and the jobs over them:
    (name, t0, nt, layout)."""
    datas = {n: golden_io.vbtp(n) for n in ["smoke64", "liquid_mini", "sci_mini", "sci_dense_mini"]}
    datas["fp32_bursts"] = synth.Workload("d32", synth.DENSITY, (40, 32, 24, 300), 5, (2.0, 7.0), 12, support=2.0,
                                          burst=0.3, codec=1).encode(threads=8)
    flames = synth.Workload("fl", synth.FLAMES, (48, 40, 32, 30), 8, (3.0, 9.0), 13, support=2.0,
                            burst=0.25).encode(threads=8)
    datas["flames"] = flames
    datas["flames_scrambled"] = scramble(flames, oracle_py.OracleFile(flames))
    assets = {n: (hip.Asset(d), oracle_py.OracleFile(d)) for n, d in datas.items()}
    F = {n: a.dims[3] for n, (a, _) in assets.items()}
    jobs = [
        ("smoke64", 0, F["smoke64"], hip.CACHE_PACKED),
        ("smoke64", 3, 5, hip.CACHE_STRIDE576),
        ("liquid_mini", 0, F["liquid_mini"], hip.CACHE_PACKED),
        ("liquid_mini", F["liquid_mini"] // 2, F["liquid_mini"] - F["liquid_mini"] // 2, hip.CACHE_STRIDE576),
        ("sci_mini", 0, F["sci_mini"], hip.CACHE_PACKED),
        ("sci_dense_mini", 1, F["sci_dense_mini"] - 1, hip.CACHE_PACKED),
        ("fp32_bursts", 0, 300, hip.CACHE_PACKED),
        ("flames", 2, 27, hip.CACHE_PACKED),
        ("flames_scrambled", 0, F["flames_scrambled"], hip.CACHE_PACKED),
    ]
    expected = []
    for name, t0, nt, layout in jobs:
        a, o = assets[name]
        if layout == hip.CACHE_PACKED:
            expected.append(np.stack([o.cache_packed(t)[0] for t in range(t0, t0 + nt)]))
        else:
            expected.append(np.stack([o.cache_576(t).ravel() for t in range(t0, t0 + nt)]))
    yield assets, jobs, expected
    for a, _ in assets.values():
        a.close()


def make_batch(world, shape=None):
    assets, jobs, _ = world
    outs, spec = [], []
    for name, t0, nt, layout in jobs:
        a = assets[name][0]
        n = a.info.controlsPerFrame if layout == hip.CACHE_PACKED else a.info.activeLeaves * 576
        out = hip.DeviceArray((nt, n), np.float32)
        outs.append(out)
        spec.append((a, t0, nt, layout, out))
    return hip.CacheBatch(spec, shape), outs


def check_outputs(world, outs, what):
    _, jobs, expected = world
    for (name, t0, nt, layout), out, exp in zip(jobs, outs, expected):
        bitexact(out.host(), exp, f"{what}: {name} [{t0},{t0 + nt}) layout {layout}")


def shape(**kw):
    s = hip.default_batch_shape()
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def test_fixture_covers_every_route(world):
    assets, _, _ = world
    modes = np.zeros(4, np.int64)
    for name in ["smoke64", "liquid_mini", "flames"]:
        modes += np.array(assets[name][0].info.mode_counts[:], np.int64)
    assert modes[1] > 0 and modes[2] > 0 and modes[3] > 0  # Grid4, Mode2, Fine6
    a = assets["flames_scrambled"][0]
    assert 0 < a.info.fast_leaves < a.info.activeLeaves  # sweep and exact bracket leaves
    assert assets["fp32_bursts"][0].dims[3] == 300  # fp32 key gaps above 255 frames


@pytest.mark.parametrize("tile", range(10))
def test_every_tile_shape(world, tile):
    b, outs = make_batch(world, shape(tile=tile))
    for o in outs:
        o.fill_bytes(0xFF)
    b.launch()
    b.stream.sync()
    check_outputs(world, outs, f"tile {tile}")
    b.close()


@pytest.mark.parametrize("tile", range(10))
def test_one_shot_kernel_every_tile_shape(world, tile):
    """The one-shot frame-cache kernel (vbt_decode_frame_cache, the kernel the
    benchmark times; its staging depth differs from the batch kernels') at
    every tile shape the autotuner may pick, on every render job of the
    world: fp32 bursts, permuted keys, Mode2 shells, both layouts."""
    assets, jobs, expected = world
    for (name, t0, nt, layout), exp in zip(jobs, expected):
        a = assets[name][0]
        if a.header.profileType == 2:
            continue  # scientific caches have one kernel shape
        assert nt > 2  # more than two frames: the staged sweep kernel, not the exact one
        n = a.info.controlsPerFrame if layout == hip.CACHE_PACKED else a.info.activeLeaves * 576
        out = hip.DeviceArray((nt, n), np.float32, a.stream)
        out.fill_bytes(0xFF)
        keep = a.cache_tile
        a.cache_tile = tile
        try:
            a.frame_cache_into(out, t0, nt, layout)
        finally:
            a.cache_tile = keep
        bitexact(out.host(), exp, f"one-shot tile {tile}: {name} [{t0},{t0 + nt}) layout {layout}")
        out.free()


@pytest.mark.parametrize("kw", [
    dict(ranges=1, chunk=1, order=0, queues=1, persistent=1),
    dict(ranges=3, chunk=4, order=1, queues=8),
    dict(ranges=8, chunk=2, order=0, queues=8),
    dict(ranges=64, chunk=1, order=1, queues=8),   # one- and two-frame items: exact path
    dict(ranges=5, chunk=64, order=1, queues=1, wgPerCU=1),
    dict(persistent=0, ranges=1, chunk=1, order=1),
    dict(persistent=0, ranges=7, chunk=3, order=0),
    dict(persistent=0, tile=6, heavyTile=1, order=1),             # 64 slots: 4 KiB light, 16 KiB heavy window
    dict(persistent=0, tile=7, heavyTile=3, ranges=2, chunk=2),   # 128 slots: 8 KiB light, 32 KiB heavy
    dict(persistent=2),                                           # each job's one-shot build, two branches
], ids=lambda kw: "-".join(f"{k}{v}" for k, v in kw.items()))
def test_queue_shapes_and_replays(world, kw):
    """Every replay of the captured graph is checked, with eager decode work
    and host synchronisation between replays; outputs are poisoned first."""
    b, outs = make_batch(world, shape(**{"tile": 2, **kw}))
    assets = world[0]
    side = hip.DeviceArray((2, assets["smoke64"][0].info.controlsPerFrame), np.float32)
    items, wgs = b.plan_info()
    assert items > 0 and (0 < wgs <= items or (kw.get("persistent") == 2 and wgs == 0))
    for rep in range(3):
        for o in outs:
            o.fill_bytes(0xFF)
        b.launch()
        b.stream.sync()
        check_outputs(world, outs, f"{kw} replay {rep}")
        assets["smoke64"][0].frame_cache_into(side, rep, 2)  # eager work between replays
        assets["smoke64"][0].stream.sync()
    # the same work as plain launches
    for o in outs:
        o.fill_bytes(0xFF)
    b.launch(graph=False)
    b.stream.sync()
    check_outputs(world, outs, f"{kw} direct")
    b.close()


def test_autotune_keeps_outputs_exact(world):
    b, outs = make_batch(world)
    best, ms, trials = b.autotune(reps=1)
    assert ms > 0 and len(trials) >= 10
    assert best in [t for t, _ in trials]
    assert b.shape.as_dict() == best
    for o in outs:
        o.fill_bytes(0xFF)
    b.launch()
    b.stream.sync()
    check_outputs(world, outs, f"autotuned {best}")
    b.close()


def test_poisoned_leaf_fails_only_its_slots(world):
    w = synth.Workload("ls", synth.LEVELSET, (40, 40, 32, 24), 8, (3.0, 9.0), 15)
    data = bytearray(w.encode(threads=8))
    o = oracle_py.OracleFile(bytes(data))
    base = 52 + 4 * (o.header.leafCount + 1)
    shell = [l for l in range(o.header.leafCount) if o.parse(l).header.mode == 2]
    victim = shell[len(shell) // 2]
    v = o.parse(victim)
    off = base + 4 * int(o.offsets[victim]) + v.layout.fine.descriptorOffset + 64 + 2
    data[off:off + 2] = struct.pack("<H", 999)  # inconsistent group prefix
    bad = hip.Asset(bytes(data))
    assert bad.info.badLeaves == 1
    good_a, good_o = world[0]["smoke64"]
    F = bad.dims[3]
    out_bad = hip.DeviceArray((F, bad.info.controlsPerFrame), np.float32)
    out_good = hip.DeviceArray((4, good_a.info.controlsPerFrame), np.float32)
    b = hip.CacheBatch([(bad, 0, F, hip.CACHE_PACKED, out_bad), (good_a, 1, 4, hip.CACHE_PACKED, out_good)],
                       shape(tile=0, ranges=2))
    with pytest.raises(hip.VbtError) as e:
        b.launch()
    assert e.value.code == 2 and "shell group prefix" in e.value.msg
    b.stream.sync()
    ref = hip.DeviceArray((F, bad.info.controlsPerFrame), np.float32)
    with pytest.raises(hip.VbtError):
        bad.frame_cache_into(ref, 0, F)  # the one-shot kernel: NaN slots for the poisoned leaf
    bad.stream.sync()
    got, exp = out_bad.host(), ref.host()
    bitexact(got, exp, "batch vs one-shot around a poisoned leaf")
    offs = bad.cache_offsets()
    c = int(bad.page_table()[victim])
    assert np.isnan(got[:, offs[c]:offs[c + 1]]).all()
    assert not np.isnan(np.delete(got, np.s_[offs[c]:offs[c + 1]], axis=1)).any()
    bitexact(out_good.host(), np.stack([good_o.cache_packed(t)[0] for t in range(1, 5)]), "good job beside it")
    b.close()
    bad.close()


def test_invalid_jobs_and_shapes(world):
    a = world[0]["smoke64"][0]
    s = world[0]["sci_mini"][0]
    F = a.dims[3]
    out = hip.DeviceArray((F, a.info.activeLeaves * 576), np.float32)
    with pytest.raises(hip.VbtError, match="frame range outside the asset"):
        hip.CacheBatch([(a, 1, F, hip.CACHE_PACKED, out)])
    with pytest.raises(hip.VbtError, match="packed cache only"):
        hip.CacheBatch([(s, 0, 1, hip.CACHE_STRIDE576, out)])
    for kw in [dict(tile=10), dict(ranges=0), dict(chunk=65), dict(order=2), dict(queues=9), dict(persistent=3),
               dict(persistent=2, tile=0, heavyTile=1),    # no heavy kernel in the per-job form
               dict(heavyTile=1),                          # heavy tiles need the static form
               dict(persistent=0, tile=0, heavyTile=2),    # different slots per workgroup
               dict(persistent=0, tile=1, heavyTile=0)]:   # smaller window
        with pytest.raises(hip.VbtError, match="batch shape"):
            hip.CacheBatch([(a, 0, 1, hip.CACHE_PACKED, out)], shape(**kw))
