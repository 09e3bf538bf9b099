"""The serving mix at full size: the six benchmark assets (smoke256, the
three fire fields, liquid320, sci192) resident in one server, a seeded mixed
trace of point-query batches (log-uniform 1..65536 queries over all x, y, z,
t) and whole-frame packed-cache builds at random t.  Every result is
compared bit for bit with the oracle, and each probe request's result is the
same alone and inside the mix -- with the default 80 GiB budget and with a
budget that leaves room for only a few request outputs beside the resident
assets, so admission waits on memory."""
import numpy as np
import pytest

import oracle_py
from workloads import THREADS, encoded

hip = pytest.importorskip("vbt_amd.hip")

pytestmark = pytest.mark.gpu

NAMES = ["smoke256", "fire_density", "fire_flames", "fire_temperature", "liquid320", "sci192"]


def same(got, exp, what):
    got = np.ascontiguousarray(got, np.float32)
    exp = np.ascontiguousarray(exp, np.float32)
    assert got.shape == exp.shape, (what, got.shape, exp.shape)
    bad = np.flatnonzero(got.view(np.uint32) != exp.view(np.uint32))
    assert len(bad) == 0, f"{what}: {len(bad)} mismatches, first {bad[0]}: {got[bad[0]]!r} vs {exp[bad[0]]!r}"


def rand_queries(rng, dims, n):
    W, H, D, F = dims
    return np.stack([rng.integers(0, W, n), rng.integers(0, H, n), rng.integers(0, D, n),
                     rng.integers(0, F, n)], axis=1).astype(np.uint32)


@pytest.fixture(scope="module")
def resident():
    out = {}
    for n in NAMES:
        data = encoded(n)
        out[n] = (hip.Asset(data), oracle_py.OracleFile(data))
    yield out
    for a, _ in out.values():
        a.close()


@pytest.fixture(scope="module")
def trace(resident):
    """Probes (one query batch and one frame build per asset) and a mixed
    trace of 150 requests, with the oracle's result of each."""
    rng = np.random.default_rng(0x5E5F)
    probes, mix = [], []
    for n in NAMES:
        dims = resident[n][0].dims
        probes.append(("q", n, rand_queries(rng, dims, 4096)))
        probes.append(("c", n, int(rng.integers(0, dims[3]))))
    for _ in range(150):
        n = NAMES[int(rng.integers(0, len(NAMES)))]
        dims = resident[n][0].dims
        if rng.random() < 0.12:
            mix.append(("c", n, int(rng.integers(0, dims[3]))))
        else:
            size = int(np.exp(rng.uniform(0.0, np.log(65536.0))))
            mix.append(("q", n, rand_queries(rng, dims, max(1, size))))
    ref = []
    for kind, n, arg in probes + mix:
        o = resident[n][1]
        ref.append(o.query(arg, threads=THREADS) if kind == "q" else o.cache_packed(arg)[0])
    return probes, mix, ref


def submit(srv, ids, req):
    kind, name, arg = req
    return srv.submit_query(ids[name], arg) if kind == "q" else srv.submit_cache(ids[name], arg)


@pytest.mark.parametrize("budget", ["default", "tight"])
def test_full_resident_set_alone_equals_mix(resident, trace, budget):
    probes, mix, ref = trace
    # resident bytes as the server counts them (assets and their frame indexes)
    probe = hip.Server()
    for n in NAMES:
        probe.register(resident[n][0])
    resident_bytes = probe.stats()["hbmInUse"]
    probe.close()
    assert resident_bytes > sum(a.info.deviceBytes for a, _ in resident.values())
    cfg = {}
    if budget == "tight":
        # room for about three of the largest frame outputs beside the assets
        largest = max(4 * a.info.controlsPerFrame for a, _ in resident.values())
        cfg = dict(hbmBudgetBytes=resident_bytes + 3 * largest)
    srv = hip.Server(**cfg)
    ids = {n: srv.register(resident[n][0]) for n in NAMES}
    alone = []
    for i, r in enumerate(probes):
        vals, st, res = srv.fetch(submit(srv, ids, r), with_status=True)
        assert res.status == 0, (r[0], r[1], res.status)
        same(vals, ref[i], f"alone {r[0]} {r[1]} vs oracle")
        if r[0] == "q":
            assert (st == 0).all()
        alone.append(vals)
    # the probes again, interleaved with the mix, everything submitted at once
    rng = np.random.default_rng(7)
    slots = set(int(x) for x in rng.choice(len(mix), size=len(probes), replace=False))
    rids, pi = [], 0
    for j, r in enumerate(mix):
        if j in slots and pi < len(probes):
            rids.append(("p", pi, submit(srv, ids, probes[pi])))
            pi += 1
        rids.append(("m", j, submit(srv, ids, r)))
    while pi < len(probes):
        rids.append(("p", pi, submit(srv, ids, probes[pi])))
        pi += 1
    for kind, i, rid in rids:
        vals, _, res = srv.fetch(rid, with_status=True, timeout_ms=120000)
        assert res.status == 0, (kind, i, res.status)
        if kind == "p":
            same(vals, alone[i], f"probe {i} ({probes[i][0]} {probes[i][1]}) in the mix vs alone")
        else:
            same(vals, ref[len(probes) + i], f"mix request {i} ({mix[i][0]} {mix[i][1]}) vs oracle")
    st = srv.stats()
    assert st["completed"] == st["submitted"] == 2 * len(probes) + len(mix)
    assert st["rejected"] == 0
    if budget == "tight":
        assert st["hbmBudget"] == resident_bytes + 3 * max(4 * a.info.controlsPerFrame for a, _ in resident.values())
    srv.close()
