"""vbt_cache_batch_run_noalloc, the batched frame-cache build that allocates
no device memory (its work queue lives in a file-scope device array), called
through the C-ABI: two runs in a row -- the second over a smaller set of jobs
and frame ranges -- each equal bit for bit to the same jobs through
vbt_cache_batch_create + vbt_cache_batch_launch (which the oracle pins), with
outputs poisoned before every run; and its size limits."""
import numpy as np
import pytest

import golden_io
import oracle_py
from vbt_amd import synth

hip = pytest.importorskip("vbt_amd.hip")

pytestmark = pytest.mark.gpu


def bitexact(a, b, what):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
    assert len(bad) == 0, f"{what}: {len(bad)} mismatches, first at {bad[0]}"


@pytest.fixture(scope="module")
def assets():
    datas = {n: golden_io.vbtp(n) for n in ["smoke64", "liquid_mini", "sci_mini", "sci_dense_mini"]}
    datas["fp32_bursts"] = synth.Workload("d32", synth.DENSITY, (40, 32, 24, 300), 5, (2.0, 7.0), 12, support=2.0,
                                          burst=0.3, codec=1).encode(threads=8)
    datas["flames"] = synth.Workload("fl", synth.FLAMES, (48, 40, 32, 30), 8, (3.0, 9.0), 13, support=2.0,
                                     burst=0.25).encode(threads=8)
    out = {n: (hip.Asset(d), d) for n, d in datas.items()}
    yield out
    for a, _ in out.values():
        a.close()


def outputs(assets, spec, stream):
    jobs = []
    for name, t0, nt, layout in spec:
        a = assets[name][0]
        n = a.info.controlsPerFrame if layout == hip.CACHE_PACKED else a.info.activeLeaves * 576
        out = hip.DeviceArray((nt, n), np.float32, stream)
        out.fill_bytes(0xFF)
        jobs.append((a, t0, nt, layout, out))
    return jobs


def shape(**kw):
    s = hip.noalloc_batch_shape()
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def run_both(assets, spec, sh):
    """The jobs through the no-alloc entry point and through a created batch
    (same shape); returns both outputs per job."""
    s = hip.Stream()
    mine = outputs(assets, spec, s)
    s.sync()
    hip.cache_batch_run_noalloc(mine, sh, s)
    s.sync()
    ref_jobs = outputs(assets, spec, s)
    s.sync()
    b = hip.CacheBatch(ref_jobs, sh, s)
    b.launch()
    s.sync()
    got = [(j[4].host(), r[4].host()) for j, r in zip(mine, ref_jobs)]
    b.close()
    return got


def test_two_runs_match_the_batch(assets):
    F = {n: a.dims[3] for n, (a, _) in assets.items()}
    big = [
        ("smoke64", 0, F["smoke64"], hip.CACHE_PACKED),
        ("smoke64", 3, 5, hip.CACHE_STRIDE576),
        ("liquid_mini", 0, F["liquid_mini"], hip.CACHE_PACKED),
        ("liquid_mini", 4, F["liquid_mini"] - 4, hip.CACHE_STRIDE576),
        ("sci_mini", 0, F["sci_mini"], hip.CACHE_PACKED),
        ("sci_dense_mini", 1, F["sci_dense_mini"] - 1, hip.CACHE_PACKED),
        ("fp32_bursts", 0, 300, hip.CACHE_PACKED),
        ("flames", 2, 27, hip.CACHE_PACKED),
    ]
    first = run_both(assets, big, shape(tile=2, ranges=3, chunk=2, queues=8, persistent=1))
    for (name, t0, nt, layout), (got, exp) in zip(big, first):
        bitexact(got, exp, f"first run: {name} [{t0},{t0 + nt}) layout {layout}")
    # and against the oracle for the packed jobs of one asset
    o = oracle_py.OracleFile(assets["flames"][1])
    bitexact(first[7][0], o.cache_packed_frames(range(2, 29)), "first run: flames vs oracle")
    # a smaller second run right after: fewer jobs, shorter frame ranges, the
    # static form with another tile (the array still holds the first run's queue)
    small = [
        ("flames", 10, 4, hip.CACHE_PACKED),
        ("smoke64", 30, 2, hip.CACHE_PACKED),
        ("liquid_mini", 0, 3, hip.CACHE_STRIDE576),
    ]
    second = run_both(assets, small, shape(tile=0, ranges=1, chunk=1, queues=1, persistent=0))
    for (name, t0, nt, layout), (got, exp) in zip(small, second):
        bitexact(got, exp, f"second run: {name} [{t0},{t0 + nt}) layout {layout}")
    # the per-job form (each job's one-shot build) uses no queue at all
    third = run_both(assets, small, shape(tile=0, persistent=2))
    for (name, t0, nt, layout), (got, exp) in zip(small, third):
        bitexact(got, exp, f"per-job form: {name} [{t0},{t0 + nt}) layout {layout}")


def test_default_shape_and_scientific_only(assets):
    F = assets["sci_mini"][0].dims[3]
    got = run_both(assets, [("sci_mini", 0, F, hip.CACHE_PACKED)], None)
    bitexact(got[0][0], got[0][1], "scientific job alone, default shape")
    assert hip.noalloc_batch_shape().order == 0 and hip.noalloc_batch_shape().heavyTile == -1


def test_limits_are_rejected(assets):
    a = assets["smoke64"][0]
    s = hip.Stream()
    out = hip.DeviceArray((1, a.info.controlsPerFrame), np.float32, s)
    with pytest.raises(hip.VbtError, match="more than 32 jobs"):
        hip.cache_batch_run_noalloc([(a, 0, 1, hip.CACHE_PACKED, out)] * 33, None, s)
    with pytest.raises(hip.VbtError, match="order 0"):
        hip.cache_batch_run_noalloc([(a, 0, 1, hip.CACHE_PACKED, out)], shape(order=1), s)
    with pytest.raises(hip.VbtError, match="order 0"):
        hip.cache_batch_run_noalloc([(a, 0, 1, hip.CACHE_PACKED, out)], shape(persistent=0, tile=6, heavyTile=1), s)
    with pytest.raises(hip.VbtError, match="frame range outside the asset"):
        hip.cache_batch_run_noalloc([(a, 1, a.dims[3], hip.CACHE_PACKED, out)], None, s)
    s.sync()


def test_tile_limit_is_rejected():
    """More leaf tiles than the scratch array holds: one-frame jobs over
    fire_temperature at 64 slots per tile (7 x 79,571 tiles > 2^19).  The
    outputs are full size, so even a run that went ahead would stay in
    bounds."""
    from workloads import encoded

    a = hip.Asset(encoded("fire_temperature"))
    tiles = (a.info.controlsPerFrame + 63) // 64
    n = (1 << 19) // tiles + 1
    assert n <= 32
    s = hip.Stream()
    outs = [hip.DeviceArray((1, a.info.controlsPerFrame), np.float32, s) for _ in range(n)]
    with pytest.raises(hip.VbtError, match="exceed the scratch maximum"):
        hip.cache_batch_run_noalloc([(a, 0, 1, hip.CACHE_PACKED, o) for o in outs], shape(tile=0), s)
    # one fewer job fits, and its output matches the one-shot kernel
    fit = n - 1
    if fit >= 1:
        sel = outs[:2]
        hip.cache_batch_run_noalloc([(a, 7 * i, 1, hip.CACHE_PACKED, o) for i, o in enumerate(sel)], shape(tile=0), s)
        s.sync()
        for i, o in enumerate(sel):
            bitexact(o.host()[0], a.frame_cache(7 * i, 1)[0], f"fire_temperature frame {7 * i}")
    for o in outs:
        o.free()
    a.close()
