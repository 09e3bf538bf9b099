"""Scientific dense subtypes 2 (tile) and 3 (patch) decoded as the reference
encoder reconstructs them (assets opened with VBT_OPEN_SCI_DENSE_BASIS), on
the GPU: whole frames, packed caches over frame ranges (the frame sweep),
point queries, the server and a cache batch, bit for bit against the oracle
(pinned against the encoder's own reconstruction in tests/test_sci_basis.py)
and against the committed vectors; a malformed block poisons its leaf only;
without the option the same files decode as the reference file decoder
reads them."""
import numpy as np
import pytest

import golden_io
import oracle_py
from test_sci_basis import BASIS, MALFORMED, mutated
from workloads import THREADS

hip = pytest.importorskip("vbt_amd.hip")

pytestmark = pytest.mark.gpu


def bitexact(a, b, what):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape, what
    bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
    assert len(bad) == 0, f"{what}: {len(bad)} mismatches, first {bad[0]}: {a.flat[bad[0]]} vs {b.flat[bad[0]]}"


def rand_queries(rng, dims, n):
    W, H, D, F = dims
    return np.stack([rng.integers(0, W, n), rng.integers(0, H, n), rng.integers(0, D, n), rng.integers(0, F, n)],
                    axis=1).astype(np.uint32)


@pytest.mark.parametrize("name", BASIS)
def test_frames_and_queries_match_the_vectors(name):
    a = hip.Asset(golden_io.vbtp(name), dense_basis=True)
    assert a.info.badLeaves == 0
    exp = golden_io.basis(name)
    for i, t in enumerate(exp["frames"]):
        bitexact(a.dense(int(t)), exp["dense"][i], f"{name} frame {t}")
    bitexact(a.query(exp["queries"]), exp["values"], f"{name} queries")
    a.close()


@pytest.mark.parametrize("name", BASIS)
def test_sweep_and_queries_equal_the_oracle(name):
    data = golden_io.vbtp(name)
    a = hip.Asset(data, dense_basis=True)
    o = oracle_py.OracleFile(data, dense_basis=True)
    F = a.dims[3]
    exp = o.cache_packed_frames(range(F), threads=THREADS)
    for t0, nt in [(0, F), (0, 1), (3, 8), (F - 9, 9), (F - 1, 1), (5, 17)]:
        bitexact(a.frame_cache(t0, nt), exp[t0:t0 + nt], f"{name} frames {t0}..{t0 + nt - 1}")
    q = rand_queries(np.random.default_rng(11), a.dims, 20000)
    bitexact(a.query(q), o.query(q), f"{name} random queries")
    a.close()


def test_without_the_option_the_file_decoder_reading_holds():
    data = golden_io.vbtp("sci_tile_mini")
    plain = hip.Asset(data)
    basis = hip.Asset(data, dense_basis=True)
    exp = golden_io.expected("sci_tile_mini")
    vec = golden_io.basis("sci_tile_mini")
    for i, t in enumerate(exp["frames"]):
        bitexact(plain.dense(int(t)), exp["dense"][i], f"file-decoder reading, frame {t}")
        bitexact(basis.dense(int(t)), vec["dense"][i], f"encoder reading, frame {t}")
    assert not np.array_equal(exp["dense"][1], vec["dense"][1])
    plain.close()
    basis.close()


@pytest.mark.parametrize("case", sorted(MALFORMED))
def test_malformed_block_poisons_its_leaf_only(case):
    name, fn, _ = MALFORMED[case]
    good = oracle_py.OracleFile(golden_io.vbtp(name), dense_basis=True)
    leaf = good.header.leafCount - 1
    data = mutated(name, leaf, fn)
    print(f"[sci basis] malformed {case}: leaf {leaf} of {name}", flush=True)
    a = hip.Asset(data, dense_basis=True)
    assert a.info.badLeaves == 1
    st = a.leaf_status() & 0xFF
    assert st[leaf] == (11 if case == "range" else 12) and (np.delete(st, leaf) == 0).all()
    with pytest.raises(hip.VbtError):
        a.dense(5)
    F = a.dims[3]
    out = hip.DeviceArray((F, a.info.controlsPerFrame), np.float32, a.stream)
    with pytest.raises(hip.VbtError):
        a.frame_cache_into(out, 0, F)
    got = out.host()
    out.free()
    offs = a.cache_offsets()
    c = int(a.page_table()[leaf])
    bad = np.zeros(a.info.controlsPerFrame, bool)
    bad[int(offs[c]):int(offs[c + 1])] = True
    assert np.isnan(got[:, bad]).all()
    for t in (0, 17, F - 1):
        exp, _ = good.cache_packed(t)
        bitexact(got[t, ~bad], exp[~bad], f"{case}: other leaves, frame {t}")
    a.close()


def test_server_and_batch_read_the_option():
    names = ["sci_tile4_mini", "sci_patch_b2_mini"]
    assets = {n: hip.Asset(golden_io.vbtp(n), dense_basis=True) for n in names}
    oracles = {n: oracle_py.OracleFile(golden_io.vbtp(n), dense_basis=True) for n in names}
    srv = hip.Server()
    ids = {n: srv.register(assets[n]) for n in names}
    rng = np.random.default_rng(12)
    for n in names:
        q = rand_queries(rng, assets[n].dims, 7000)
        bitexact(srv.fetch(srv.submit_query(ids[n], q)), oracles[n].query(q), f"server queries {n}")
        exp, _ = oracles[n].cache_packed(21)
        bitexact(srv.fetch(srv.submit_cache(ids[n], 21)), exp, f"server cache {n}")
    srv.close()
    outs = {n: hip.DeviceArray((7, assets[n].info.controlsPerFrame), np.float32) for n in names}
    batch = hip.CacheBatch([(assets[n], 30, 7, hip.CACHE_PACKED, outs[n]) for n in names])
    batch.launch()
    batch.stream.sync()
    for n in names:
        bitexact(outs[n].host(), oracles[n].cache_packed_frames(range(30, 37), threads=THREADS), f"batch {n}")
        outs[n].free()
    batch.close()
    for a in assets.values():
        a.close()
