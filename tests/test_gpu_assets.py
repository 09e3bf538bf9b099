"""GPU parity on assets made by this repository's encoder (every field kind,
both keyframe codecs, keyframe bursts, Mode2 shells, every scientific mode),
on leaves whose keyframes are out of order (the exact per-frame bracket path
instead of the sweep), and on corrupted leaves (errors stay per leaf)."""
import struct

import numpy as np
import pytest

import oracle_py
from vbt_amd import synth

hip = pytest.importorskip("vbt_amd.hip")

pytestmark = pytest.mark.gpu

SMALL = {
    "density_fp16": synth.Workload("d16", synth.DENSITY, (40, 32, 24, 40), 5, (2.0, 7.0), 11, support=2.0),
    "density_fp32_bursts": synth.Workload("d32", synth.DENSITY, (40, 32, 24, 300), 5, (2.0, 7.0), 12, support=2.0,
                                          burst=0.3, codec=1),
    "flames": synth.Workload("fl", synth.FLAMES, (48, 40, 32, 30), 8, (3.0, 9.0), 13, support=2.0, burst=0.25),
    "temperature": synth.Workload("te", synth.TEMPERATURE, (32, 32, 24, 30), 6, (3.0, 9.0), 14, support=2.0),
    "levelset": synth.Workload("ls", synth.LEVELSET, (40, 40, 32, 24), 8, (3.0, 9.0), 15),
    "scientific": synth.Workload("sc", synth.SCIENTIFIC, (48, 40, 16, 30), 12, (0.0, 0.0), 0xABC123,
                                 spike_fraction=0.00005),
}


def bitexact(a, b, what):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    bad = np.argwhere(a.view(np.uint32) != b.view(np.uint32))
    assert len(bad) == 0, f"{what}: {len(bad)} mismatches, first {tuple(bad[0])}: {a[tuple(bad[0])]} vs {b[tuple(bad[0])]}"


@pytest.fixture(scope="module", params=sorted(SMALL))
def pair(request):
    data = SMALL[request.param].encode(threads=8)
    a = hip.Asset(data)
    yield request.param, a, oracle_py.OracleFile(data), data
    a.close()


def test_encoder_asset_parity(pair):
    name, a, o, _ = pair
    W, H, D, F = a.dims
    compact, n = o.page_table()
    np.testing.assert_array_equal(a.page_table(), compact)
    got = a.frame_cache(0, F)
    for t in sorted({0, 1, F // 2, F - 1}):
        exp, _ = o.cache_packed(t)
        bitexact(got[t], exp, f"{name} cache t={t}")
    for t in sorted({0, F // 3, F - 1}):
        bitexact(a.dense(t), o.frame(t), f"{name} dense t={t}")
    rng = np.random.default_rng(5)
    q = np.stack([rng.integers(0, W, 4000), rng.integers(0, H, 4000), rng.integers(0, D, 4000),
                  rng.integers(0, F, 4000)], axis=1).astype(np.uint32)
    bitexact(a.query(q), o.query(q), f"{name} queries")
    bitexact(a.query(q, plain=True), o.query(q), f"{name} plain queries")
    if o.header.profileType == 1:
        assert a.info.fast_leaves == a.info.activeLeaves  # the encoder writes increasing keyframes


def scramble_keyframes(data: bytes, o) -> bytes:
    """Swap the first two keyframe records of the first multi-key control of
    every 3rd active render leaf: frames stop increasing there."""
    buf = bytearray(data)
    base = 52 + 4 * o.header.coarseAcScaleCount + 4 * (o.header.leafCount + 1)
    compact, _ = o.page_table()
    for leaf in np.nonzero(compact != 0xFFFFFFFF)[0][::3]:
        v = o.parse(int(leaf))
        lb = base + 4 * int(o.offsets[leaf])
        cc = v.header.coarseResolution ** 3
        kb = 8 if v.header.coarseCodec == 1 else 3
        key0 = lb + v.layout.coarse.keyframeOffset
        for c in range(cc):
            ks, kc = struct.unpack_from("<HB", buf, lb + 8 + 4 * c)
            if kc >= 2:
                a0, a1 = key0 + kb * ks, key0 + kb * (ks + 1)
                r0, r1 = bytes(buf[a0:a0 + kb]), bytes(buf[a1:a1 + kb])
                buf[a0:a0 + kb], buf[a1:a1 + kb] = r1, r0
                break
    return bytes(buf)


def test_unsorted_keyframes_take_exact_path():
    w = SMALL["density_fp16"]
    data = w.encode(threads=8)
    o0 = oracle_py.OracleFile(data)
    bad = scramble_keyframes(data, o0)
    a = hip.Asset(bad)
    o = oracle_py.OracleFile(bad)
    assert 0 < a.info.fast_leaves < a.info.activeLeaves
    F = a.dims[3]
    got = a.frame_cache(0, F)
    for t in range(F):
        exp, _ = o.cache_packed(t)
        bitexact(got[t], exp, f"scrambled cache t={t}")
    bitexact(a.dense(7), o.frame(7), "scrambled dense")
    a.close()


def shuffle_keyframes(data: bytes, o, seed: int) -> bytes:
    """Permute the keyframe records of every multi-key coarse control of every
    2nd active render leaf (seeded): bin hints then point past or short of
    the frame and the binary search runs over unsorted frames."""
    rng = np.random.default_rng(seed)
    buf = bytearray(data)
    base = 52 + 4 * o.header.coarseAcScaleCount + 4 * (o.header.leafCount + 1)
    compact, _ = o.page_table()
    for leaf in np.nonzero(compact != 0xFFFFFFFF)[0][::2]:
        v = o.parse(int(leaf))
        lb = base + 4 * int(o.offsets[leaf])
        kb = 8 if v.header.coarseCodec == 1 else 3
        key0 = lb + v.layout.coarse.keyframeOffset
        for c in range(v.header.coarseResolution ** 3):
            ks, kc = struct.unpack_from("<HB", buf, lb + 8 + 4 * c)
            if kc >= 2:
                a0 = key0 + kb * ks
                recs = [bytes(buf[a0 + kb * i:a0 + kb * (i + 1)]) for i in range(kc)]
                buf[a0:a0 + kb * kc] = b"".join(recs[i] for i in rng.permutation(kc))
    return bytes(buf)


@pytest.mark.parametrize("name", ["density_fp16", "density_fp32_bursts", "flames"])
def test_unsorted_keyframes_queries_and_short_caches(name):
    """Batched and plain point queries and short frame caches equal the
    oracle (the reference's bracket search) on streams whose frames are
    permuted."""
    data = SMALL[name].encode(threads=8)
    bad = shuffle_keyframes(data, oracle_py.OracleFile(data), 7)
    a = hip.Asset(bad)
    o = oracle_py.OracleFile(bad)
    assert a.info.fast_leaves < a.info.activeLeaves
    W, H, D, F = a.dims
    compact, _ = o.page_table()
    act = np.nonzero(compact != 0xFFFFFFFF)[0]
    rng = np.random.default_rng(11)
    n = 30000
    leaves = act[rng.integers(0, len(act), n)]
    lcx, lcy = (W + 7) // 8, (H + 7) // 8
    q = np.stack([np.minimum((leaves % lcx) * 8 + rng.integers(0, 8, n), W - 1),
                  np.minimum(((leaves // lcx) % lcy) * 8 + rng.integers(0, 8, n), H - 1),
                  np.minimum((leaves // (lcx * lcy)) * 8 + rng.integers(0, 8, n), D - 1),
                  rng.integers(0, F + 3, n)], axis=1).astype(np.uint32)
    exp = o.query(q)
    bitexact(a.query(q), exp, f"{name} shuffled queries")
    bitexact(a.query(q, plain=True), exp, f"{name} shuffled plain queries")
    bitexact(a.dense(F // 2), o.frame(F // 2), f"{name} shuffled dense")
    # one- and two-frame caches take the per-slot exact kernel, longer runs the
    # sweep with its exact fallback for these leaves
    for t0, nt in ((0, 1), (F // 3, 1), (F - 2, 2), (1, 5)):
        got = a.frame_cache(t0, nt)
        for i in range(nt):
            bitexact(got[i], o.cache_packed(t0 + i)[0], f"{name} shuffled cache t={t0 + i} (nt={nt})")
    a.close()


def test_corrupt_leaf_fails_alone():
    data = bytearray(SMALL["levelset"].encode(threads=8))
    o = oracle_py.OracleFile(bytes(data))
    base = 52 + 4 * (o.header.leafCount + 1)
    shell = [l for l in range(o.header.leafCount) if o.parse(l).header.mode == 2]
    victim = shell[len(shell) // 2]
    v = o.parse(victim)
    off = base + 4 * int(o.offsets[victim]) + v.layout.fine.descriptorOffset + 64 + 2
    data[off:off + 2] = struct.pack("<H", 999)  # inconsistent group prefix
    a = hip.Asset(bytes(data))
    st = a.leaf_status()
    assert a.info.badLeaves == 1 and (st[victim] & 0xFF) == 5
    assert hip.status_message(5) == "render temporal shell group prefix is inconsistent"
    W, H, D, F = a.dims
    with pytest.raises(hip.VbtError):
        a.frame_cache(0, F)
    # point queries: only the ones inside the poisoned leaf fail
    lcx, lcy = (W + 7) // 8, (H + 7) // 8
    bx, by, bz = victim % lcx, (victim // lcx) % lcy, victim // (lcx * lcy)
    q = np.array([[bx * 8, by * 8, bz * 8, 3], [0, 0, 0, 3], [W - 1, H - 1, D - 1, 5]], np.uint32)
    if victim == 0:
        q[1] = [W - 1, 0, 0, 3]
    vals, status = a.query(q, with_status=True)
    assert status[0] == 5 and np.isnan(vals[0])
    assert (status[1:] == 0).all()
    good = oracle_py.OracleFile(SMALL["levelset"].encode(threads=8))
    bitexact(vals[1:], good.query(q[1:]), "queries beside a poisoned leaf")
    a.close()


def all_halves_asset():
    """5 Grid4 leaves whose 64 controls hold a keyframe at every frame 0..254
    with raw half bit patterns covering all 65536 values (zeros, subnormals,
    normals, infinities, NaN payloads): at frame f the cache holds key f."""
    F = 255
    leaves = []
    for leaf in range(5):
        desc = b"".join(struct.pack("<HBB", c * F, F, 32) for c in range(64))
        bins = b"".join(bytes([min(255, (b * F // 8))]) for _ in range(64) for b in range(8))
        keys = bytearray()
        for c in range(64):
            for j in range(F):
                h = (leaf * 64 * F + c * F + j) % 65536
                keys += struct.pack("<BH", j, h)
        keys += b"\0" * ((-len(keys)) % 4)
        hdr = (1 | (0 << 2) | (4 << 6))
        leaves.append(struct.pack("<IHH", hdr, 64 * F, 0) + desc + bins + bytes(keys))
    offs = [0]
    for lb in leaves:
        offs.append(offs[-1] + len(lb) // 4)
    return (b"VBTPACK4" + struct.pack("<11I", 4, 40, 8, 8, F, 8, 4, 0, 5, 1, 0) +
            struct.pack(f"<{len(offs)}I", *offs) + b"".join(leaves))


def test_every_half_pattern_through_the_sweep():
    data = all_halves_asset()
    o = oracle_py.OracleFile(data)
    a = hip.Asset(data)
    assert a.info.fast_leaves == 5
    got = a.frame_cache(0, 255)
    for t in range(0, 255, 1):
        exp, _ = o.cache_packed(t)
        np.testing.assert_array_equal(got[t].view(np.uint32), exp.view(np.uint32), err_msg=f"t={t}")
    a.close()
