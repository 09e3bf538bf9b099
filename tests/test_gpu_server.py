"""GPU tests of the continuous-batching server (vbt_server_* through the C-ABI).

Every served result is compared bit-for-bit with the one-shot entry points
(vbt_query_points / vbt_decode_frame_cache), which are themselves pinned to
the oracle and the reference's vectors, and with the oracle directly for the
probe requests.  The central property: a request's result is the same when it
runs alone and when it runs inside a mix of other requests, assets and kinds.
"""
import struct
import threading
import time

import numpy as np
import pytest

import golden_io
import oracle_py

hip = pytest.importorskip("vbt_amd.hip")

pytestmark = pytest.mark.gpu

NAMES = ["smoke64", "liquid_mini", "sci_mini", "sci_dense_mini"]


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def same(got, exp, what):
    got, exp = np.asarray(got, np.float32), np.asarray(exp, np.float32)
    assert got.shape == exp.shape, (what, got.shape, exp.shape)
    bad = np.nonzero(bits(got) != bits(exp))[0]
    assert len(bad) == 0, f"{what}: {len(bad)} mismatches, first {bad[0]}: {got[bad[0]]!r} vs {exp[bad[0]]!r}"


def rand_queries(rng, dims, n):
    W, H, D, F = dims
    return np.stack([rng.integers(0, W, n), rng.integers(0, H, n), rng.integers(0, D, n),
                     rng.integers(0, F, n)], axis=1).astype(np.uint32)


@pytest.fixture(scope="module")
def assets():
    out = {}
    for n in NAMES:
        data = golden_io.vbtp(n)
        out[n] = (hip.Asset(data), oracle_py.OracleFile(data))
    yield out
    for a, _ in out.values():
        a.close()


def expected(assets, req):
    kind, name, arg = req
    a, _ = assets[name]
    if kind == "q":
        return a.query(arg, with_status=True)
    return a.frame_cache(arg, 1)[0], None


def submit(srv, ids, req):
    kind, name, arg = req
    return srv.submit_query(ids[name], arg) if kind == "q" else srv.submit_cache(ids[name], arg)


@pytest.mark.parametrize("in_loop,frame_index", [(0, 1), (0, 0), (1, 0)],
                         ids=["frames-beside-loop-indexed", "frames-beside-loop", "frames-in-loop"])
def test_alone_equals_mix(assets, in_loop, frame_index):
    rng = np.random.default_rng(2024)
    probes = []
    for n in NAMES:
        dims = assets[n][0].dims
        probes.append(("q", n, rand_queries(rng, dims, 3000)))
        probes.append(("c", n, int(rng.integers(0, dims[3]))))
    mix = []
    for _ in range(240):
        n = NAMES[int(rng.integers(0, len(NAMES)))]
        dims = assets[n][0].dims
        if rng.random() < 0.2:
            mix.append(("c", n, int(rng.integers(0, dims[3]))))
        else:
            size = int(np.exp(rng.uniform(0, np.log(20000))))
            mix.append(("q", n, rand_queries(rng, dims, max(1, size))))
    ref = {i: expected(assets, r) for i, r in enumerate(probes + mix)}

    srv = hip.Server(frameBuildsInLoop=in_loop, frameIndex=frame_index)
    ids = {n: srv.register(assets[n][0]) for n in NAMES}
    # probes alone, one at a time
    alone = []
    for i, r in enumerate(probes):
        rid = submit(srv, ids, r)
        vals, st, res = srv.fetch(rid, with_status=True)
        assert res.status == 0
        alone.append(vals)
        same(vals, ref[i][0], f"alone {r[0]} {r[1]}")
        if r[0] == "q":
            same(vals, assets[r[1]][1].query(r[2]), f"alone {r[1]} queries vs oracle")
            assert (st == 0).all()
        else:
            exp, _ = assets[r[1]][1].cache_packed(r[2])
            same(vals, exp, f"alone {r[1]} cache vs oracle")
    # the same probes inside a mix submitted at once
    order = list(range(len(mix)))
    slots = sorted(rng.choice(len(mix), size=len(probes), replace=False))
    rids = []
    pi = 0
    for j in order:
        while pi < len(probes) and slots[pi] == j:
            rids.append(("p", pi, submit(srv, ids, probes[pi])))
            pi += 1
        rids.append(("m", j, submit(srv, ids, mix[j])))
    for kind, i, rid in rids:
        vals = srv.fetch(rid)
        if kind == "p":
            same(vals, alone[i], f"probe {i} in the mix vs alone")
        else:
            same(vals, ref[len(probes) + i][0], f"mix request {i} ({mix[i][0]} {mix[i][1]})")
    st = srv.stats()
    assert st["completed"] == st["submitted"] == len(probes) * 2 + len(mix)
    assert st["rejected"] == 0 and st["launches"] >= 1
    assert st["maxInflight"] > 1  # requests really were on the device together
    srv.close()


def test_sign_rules_through_the_server(assets):
    a, o = assets["liquid_mini"]
    rng = np.random.default_rng(5)
    q = rand_queries(rng, a.dims, 5000)
    srv = hip.Server()
    i = srv.register(a)
    for rule, bg in ((hip.SIGN_NONE, 0.0), (hip.SIGN_CLAMP_ZERO, 0.0), (hip.SIGN_BACKGROUND, 3.5)):
        comp = hip.compose(rule, bg)
        got = srv.fetch(srv.submit_query(i, q, comp))
        same(got, a.query(q, comp), f"sign rule {rule}")
        same(got, o.query(q, oracle_py.compose(rule, bg)), f"sign rule {rule} vs oracle")
    srv.close()


def test_admission_waits_for_arena_space(assets):
    a, _ = assets["smoke64"]
    rng = np.random.default_rng(11)
    qs = [rand_queries(rng, a.dims, 2048) for _ in range(12)]
    exp = [a.query(q) for q in qs]
    # each request holds 2 x 8 KiB of outputs: four fit in the arena at once
    srv = hip.Server(arenaBytes=64 * 1024)
    i = srv.register(a)
    rids = [srv.submit_query(i, q) for q in qs]
    for rid, e in zip(rids, exp):
        same(srv.fetch(rid), e, "arena-limited request")
    st = srv.stats()
    assert st["completed"] == 12 and 1 <= st["maxInflight"] <= 4
    assert st["arenaInUse"] == 0
    srv.close()


def test_hbm_budget_rejections(assets):
    a, _ = assets["smoke64"]
    b, _ = assets["liquid_mini"]
    need = a.info.deviceBytes
    srv = hip.Server(hbmBudgetBytes=need + 64 * 1024, arenaBytes=64 * 1024, frameIndex=0)
    i = srv.register(a)
    with pytest.raises(hip.VbtError) as e:
        srv.register(b)
    assert e.value.code == 5
    # 8192 queries need 2 x 32 KiB of outputs: fits; 16384 do not
    q = rand_queries(np.random.default_rng(1), a.dims, 16384)
    same(srv.fetch(srv.submit_query(i, q[:8192])), a.query(q[:8192]), "largest admissible request")
    with pytest.raises(hip.VbtError) as e:
        srv.submit_query(i, q)
    assert e.value.code == 5
    assert srv.stats()["rejected"] == 1
    with pytest.raises(hip.VbtError):
        srv.submit_query(7, q[:10])
    with pytest.raises(hip.VbtError):
        srv.submit_cache(i, a.dims[3])
    srv.close()


def poisoned_liquid():
    data = bytearray(golden_io.vbtp("liquid_mini"))
    o = oracle_py.OracleFile(bytes(data))
    base = 52 + 4 * (o.header.leafCount + 1)
    shell = [l for l in range(o.header.leafCount) if o.parse(l).header.mode == 2]
    victim = shell[len(shell) // 2]
    v = o.parse(victim)
    off = base + 4 * int(o.offsets[victim]) + v.layout.fine.descriptorOffset + 64 + 2
    data[off:off + 2] = struct.pack("<H", 999)  # inconsistent shell group prefix
    return bytes(data), victim, o.header


def test_poisoned_leaf_fails_only_its_requests(assets):
    data, victim, h = poisoned_liquid()
    bad = hip.Asset(data)
    good, _ = assets["smoke64"]
    lcx, lcy = (h.width + 7) // 8, (h.height + 7) // 8
    bx, by, bz = victim % lcx, (victim // lcx) % lcy, victim // (lcx * lcy)
    rng = np.random.default_rng(3)
    q = rand_queries(rng, bad.dims, 400)
    q[17] = [bx * 8 + 1, by * 8 + 2, bz * 8 + 3, 4]
    exp_vals, exp_st = bad.query(q, with_status=True)
    assert exp_st[17] == 5
    qg = rand_queries(rng, good.dims, 1000)
    srv = hip.Server()
    ib, ig = srv.register(bad), srv.register(good)
    r1 = srv.submit_query(ib, q)
    r2 = srv.submit_cache(ib, 3)
    r3 = srv.submit_query(ig, qg)
    r4 = srv.submit_cache(ig, 5)
    vals, st, res = srv.fetch(r1, with_status=True)
    assert res.status == 2
    np.testing.assert_array_equal(st, exp_st)
    same(vals, exp_vals, "queries on a poisoned asset")
    # a whole-frame build on an asset with a poisoned leaf fails as a whole, as
    # the one-shot vbt_decode_frame_cache does (the leaf owns no cache slots)
    vals, _, res = srv.fetch(r2, with_status=True)
    assert res.status == 2 and len(vals) == bad.info.controlsPerFrame
    with pytest.raises(hip.VbtError):
        bad.frame_cache(3, 1)
    _, _, res = srv.fetch(r3, with_status=True)
    assert res.status == 0
    vals, _, res = srv.fetch(r4, with_status=True)
    assert res.status == 0
    same(vals, good.frame_cache(5, 1)[0], "cache beside a poisoned asset")
    srv.close()
    bad.close()


def test_relaunch_after_idle_and_lease(assets):
    a, _ = assets["smoke64"]
    rng = np.random.default_rng(9)
    srv = hip.Server(idleUs=200, leaseMs=2)
    i = srv.register(a)
    for k in range(3):
        q = rand_queries(rng, a.dims, 100)
        same(srv.fetch(srv.submit_query(i, q)), a.query(q), f"after idle {k}")
        time.sleep(0.02)  # the device loop leaves after 200 us without work
    assert srv.stats()["launches"] >= 3
    qs = [rand_queries(rng, a.dims, int(rng.integers(1, 300))) for _ in range(1500)]
    rids = []
    for q in qs:
        rids.append(srv.submit_query(i, q))
    for rid, q in zip(rids, qs):
        same(srv.fetch(rid), a.query(q), "burst across leases")
    srv.close()


def test_concurrent_submitters(assets):
    a, _ = assets["smoke64"]
    c, _ = assets["sci_mini"]
    srv = hip.Server()
    ia, ic = srv.register(a), srv.register(c)
    errors = []

    def client(seed):
        try:
            rng = np.random.default_rng(seed)
            for _ in range(60):
                asset, idx = (a, ia) if rng.random() < 0.5 else (c, ic)
                q = rand_queries(rng, asset.dims, int(rng.integers(1, 2000)))
                got = srv.fetch(srv.submit_query(idx, q))
                if not np.array_equal(bits(got), bits(asset.query(q))):
                    errors.append(seed)
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(repr(e))

    th = [threading.Thread(target=client, args=(s,)) for s in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    assert srv.stats()["completed"] == 240
    srv.close()


def test_register_while_serving(assets):
    a, _ = assets["smoke64"]
    b, _ = assets["sci_dense_mini"]
    rng = np.random.default_rng(21)
    srv = hip.Server()
    ia = srv.register(a)
    qs = [rand_queries(rng, a.dims, 5000) for _ in range(20)]
    rids = [srv.submit_query(ia, q) for q in qs]
    ib = srv.register(b)
    qb = rand_queries(rng, b.dims, 3000)
    rb = srv.submit_query(ib, qb)
    rc = srv.submit_cache(ib, 7)
    for rid, q in zip(rids, qs):
        same(srv.fetch(rid), a.query(q), "request queued before a registration")
    same(srv.fetch(rb), b.query(qb), "request on the new asset")
    same(srv.fetch(rc), b.frame_cache(7, 1)[0], "cache on the new asset")
    srv.close()


def test_largest_query_batch_and_item_shapes(assets):
    a, o = assets["smoke64"]
    q = rand_queries(np.random.default_rng(4), a.dims, 65536)
    exp = a.query(q)
    same(exp[:5000], o.query(q[:5000]), "one-shot vs oracle")
    for qi, ci, bpc in ((32, 256, 1), (128, 1024, 2), (512, 4096, 4)):
        srv = hip.Server(queryItem=qi, cacheItem=ci, blocksPerCU=bpc, frameBuildsInLoop=1)
        i = srv.register(a)
        same(srv.fetch(srv.submit_query(i, q)), exp, f"65536 queries, item {qi}")
        same(srv.fetch(srv.submit_cache(i, 9)), a.frame_cache(9, 1)[0], f"cache, item {ci}")
        srv.close()


def test_server_frame_index_counts_as_resident(assets):
    """With the server's frame index (the default), each registered render
    asset costs its device bytes plus the index's; scientific assets have
    none; a budget without room for the index refuses the asset."""
    a, _ = assets["smoke64"]
    sci, _ = assets["sci_mini"]
    probe = hip.Asset(golden_io.vbtp("smoke64"))
    index_bytes = probe.build_frame_index()  # the same build the server runs
    probe.close()
    assert index_bytes > 0
    srv = hip.Server()
    srv.register(a)
    srv.register(sci)
    assert srv.stats()["hbmInUse"] == a.info.deviceBytes + index_bytes + sci.info.deviceBytes
    srv.close()
    tight = hip.Server(hbmBudgetBytes=a.info.deviceBytes + index_bytes // 2, arenaBytes=4096)
    with pytest.raises(hip.VbtError) as e:
        tight.register(a)
    assert e.value.code == 5
    tight.close()
    plain = hip.Server(hbmBudgetBytes=a.info.deviceBytes + index_bytes // 2, arenaBytes=4096, frameIndex=0)
    plain.register(a)
    plain.close()
