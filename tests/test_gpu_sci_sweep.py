"""The scientific frame sweep (packed caches of a frame range: one workgroup
per leaf and run of frames, the leaf staged in LDS when its reads fit, else
read from the payload; sparse/empty and dense leaves in separate launches),
bit for bit against the oracle on every scientific fixture, over frame ranges
that start and end inside a frame group and a run; leaves too large to stage
(64 DCT terms); single-term leaves; poisoned leaves (NaN, error raised, the
other leaves still exact); and the benchmarked 200-frame sci192 build."""
import numpy as np
import pytest

import golden_io
import oracle_py
from vbt_amd import synth
from workloads import THREADS, encoded

hip = pytest.importorskip("vbt_amd.hip")

pytestmark = pytest.mark.gpu


def bitexact(a, b, what):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape, what
    bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
    assert len(bad) == 0, f"{what}: {len(bad)} mismatches, first {bad[0]}: {a.flat[bad[0]]} vs {b.flat[bad[0]]}"


def synth_sci(keep, spikes=0.0005, dims=(40, 24, 32, 70), seed=0x5C15EE):
    w = synth.Workload("sweep", synth.SCIENTIFIC, dims, 10, (0.0, 0.0), seed, spike_fraction=spikes)
    o = synth.default_encode_opts(8)
    o.sciKeep = keep
    return w.field().encode(o)


ASSETS = {
    "sci_mini": lambda: golden_io.vbtp("sci_mini"),
    "sci_dense_mini": lambda: golden_io.vbtp("sci_dense_mini"),
    "sci_tile_mini": lambda: golden_io.vbtp("sci_tile_mini"),
    "keep64": lambda: synth_sci(64),
    "keep1": lambda: synth_sci(1),
    "dense_keep64": lambda: synth_sci(64, spikes=0.0, seed=0x5C15EF),
}


def ranges(F):
    r = {(0, F), (0, 1), (1, 7), (3, 8), (5, 9), (2, 17), (F - 9, 9), (F - 1, 1), (F // 2, F - F // 2)}
    return sorted((t0, nt) for t0, nt in r if 0 <= t0 and nt >= 1 and t0 + nt <= F)


@pytest.mark.parametrize("name", sorted(ASSETS))
def test_sweep_equals_the_oracle_over_frame_ranges(name):
    data = ASSETS[name]()
    a = hip.Asset(data)
    o = oracle_py.OracleFile(data)
    assert a.header.profileType == 2
    F = a.dims[3]
    exp = o.cache_packed_frames(range(F), threads=THREADS)
    for t0, nt in ranges(F):
        bitexact(a.frame_cache(t0, nt), exp[t0:t0 + nt], f"{name} frames {t0}..{t0 + nt - 1}")
    a.close()


def test_fixture_modes_cover_both_launches():
    """The fixtures above hold staged sparse leaves, dense leaves (the second
    launch) and leaves whose reads do not fit the stage (64 DCT terms: 4164
    control bytes)."""
    modes = {}
    for name in ["sci_dense_mini", "keep64", "dense_keep64"]:
        o = oracle_py.OracleFile(ASSETS[name]())
        W, H, D, F = o.dims
        L = ((W + 7) // 8) * ((H + 7) // 8) * ((D + 7) // 8)
        words = [int.from_bytes(bytes(o.leaf_bytes(l)[:4]), "little") for l in range(L)]
        modes[name] = {w & 3 for w in words}
        keeps = {((w >> 2) & 0x3F) + 1 for w in words}
        if name != "sci_dense_mini":
            assert keeps == {64}, name
    assert {2, 3} & modes["sci_dense_mini"] and {0, 1} & modes["sci_dense_mini"]
    assert 1 in modes["keep64"]
    assert {2, 3} & modes["dense_keep64"]


def test_poisoned_leaves_are_nan_and_the_rest_exact():
    data = golden_io.vbtp("sci_patch_mini")
    a = hip.Asset(data)
    o = oracle_py.OracleFile(data)
    assert a.info.badLeaves > 0
    F = a.dims[3]
    out = hip.DeviceArray((F, a.info.controlsPerFrame), np.float32, a.stream)
    with pytest.raises(hip.VbtError):
        a.frame_cache_into(out, 0, F)
    got = out.host()
    out.free()
    offs = a.cache_offsets()
    st = a.leaf_status()
    compact = a.page_table()
    bad = np.zeros(a.info.controlsPerFrame, bool)
    for leaf in np.flatnonzero((st & 0xFF) != 0):
        c = int(compact[leaf])
        bad[int(offs[c]):int(offs[c + 1])] = True
    assert bad.any() and (~bad).any()
    assert np.isnan(got[:, bad]).all()
    checked = 0
    for t in range(F):
        try:
            exp, _ = o.cache_packed(t)
        except ValueError:
            continue  # the reference throws at the frames that read past the pool
        bitexact(got[t, ~bad], exp[~bad], f"sci_patch_mini readable leaves, frame {t}")
        checked += 1
    assert checked > 0
    a.close()


def test_sci192_benchmark_build_every_frame():
    """The benchmarked launch: all 200 frames of sci192 (13,824 sparse-event
    leaves, 12 DCT terms) in one call, every frame against the oracle."""
    data = encoded("sci192")
    a = hip.Asset(data)
    o = oracle_py.OracleFile(data)
    W, H, D, F = a.dims
    assert (W, H, D, F) == (192, 192, 192, 200)
    cpf = a.info.controlsPerFrame
    out = hip.DeviceArray((F, cpf), np.float32, a.stream)
    a.frame_cache_into(out, 0, F)
    for t0 in range(0, F, 25):
        exp = o.cache_packed_frames(range(t0, t0 + 25), threads=THREADS)
        bitexact(out.host_slice(t0 * cpf, 25 * cpf).reshape(25, cpf), exp, f"sci192 frames {t0}..{t0 + 24}")
    out.free()
    a.close()


def unsort_events(data: bytes, leaves: int = 6) -> tuple[bytes, int]:
    """Break the key order of one event bin in each of the first `leaves`
    leaves that have a bin of two or more events: swap its first two keys,
    or (every other leaf) repeat the first key."""
    o = oracle_py.OracleFile(data)
    out = bytearray(data)
    base = 52 + 4 * o.header.coarseAcScaleCount + 4 * (o.header.leafCount + 1)  # payload start
    done = 0
    for leaf in range(o.header.leafCount):
        lb = bytes(o.leaf_bytes(leaf))
        w = int.from_bytes(lb[:4], "little")
        if w & 3 != 1:
            continue
        keep, count = ((w >> 2) & 0x3F) + 1, (w >> 16) & 0xFFF
        bins = 4 + 64 * (keep + 1) + 68
        coords = bins + 32
        at = base + 4 * int(o.offsets[leaf])
        for b in range(16):
            start = int.from_bytes(lb[bins + 2 * b:bins + 2 * b + 2], "little")
            end = int.from_bytes(lb[bins + 2 * b + 2:bins + 2 * b + 4], "little") if b < 15 else count
            if end - start >= 2:
                k0 = lb[coords + 2 * start:coords + 2 * start + 2]
                k1 = lb[coords + 2 * start + 2:coords + 2 * start + 4]
                first = at + coords + 2 * start
                out[first:first + 4] = (k1 + k0) if done % 2 == 0 else (k0 + k0)
                done += 1
                break
        if done == leaves:
            break
    return bytes(out), done


def test_unsorted_event_bins_take_the_search_path():
    """Leaves whose event keys are out of order or repeated are decoded by
    each voxel's binary search, as the reference does (it can miss or pick
    among such events), the sorted ones by the scatter: both bit-exact."""
    data, broken = unsort_events(golden_io.vbtp("sci_mini"))
    assert broken == 6
    a = hip.Asset(data)
    o = oracle_py.OracleFile(data)
    assert a.info.fast_leaves == a.info.activeLeaves - broken
    F = a.dims[3]
    exp = o.cache_packed_frames(range(F), threads=THREADS)
    for t0, nt in ranges(F):
        bitexact(a.frame_cache(t0, nt), exp[t0:t0 + nt], f"unsorted events, frames {t0}..{t0 + nt - 1}")
    a.close()
